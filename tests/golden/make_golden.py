"""Generate tests/golden/lagrangian_chains.json: joint accelerations of small
revolute/prismatic chains and trees derived from the Lagrangian, independently
of any spatial-algebra code.

Only forward kinematics (body poses from joint values, Rodrigues rotations)
and energies are written here:
    T(q, qd) = sum_i 1/2 m_i |v_com,i|^2 + 1/2 w_i^T (R_i I_com,i R_i^T) w_i
    V(q)     = sum_i m_i g z_com,i            (gravity -g z)
with velocities from central differences of the poses along q + s qd, all in
50-digit mpmath. Then M_ij = d2T/dqd_i dqd_j (T is quadratic in qd:
polarisation), h = Mdot qd - dT/dq + dV/dq and
    qdd = M^-1 (tau + tau_spring - h).
Truncation of the nested central differences is below 1e-18 relative.

Conventions of the mechanism (the reference's, mechanism.rs:41-58): body i is
the child of joint i; the joint's successor-to-predecessor transform is
x_parent = R_t (R_J(q) x_child + a q [prismatic]) + p_t; the moment J is
about the body origin and c = m * com.

    python tests/golden/make_golden.py      (about a minute)
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50
G = mp.mpf("9.81")
H_IN = mp.mpf("1e-20")
H_OUT = mp.mpf("1e-12")


def rodrigues(a, th):
    a = mp.matrix(a)
    K = mp.matrix([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return mp.eye(3) + mp.sin(th) * K + (1 - mp.cos(th)) * (K * K)


def quat_to_R(qx, qy, qz, qw):
    return mp.matrix([[1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
                      [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
                      [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)]])


def poses(model, q):
    out = []
    for i, b in enumerate(model["bodies"]):
        Rt = quat_to_R(*[mp.mpf(x) for x in b["x_tree_quat"]])
        pt = mp.matrix([mp.mpf(x) for x in b["x_tree_trans"]])
        a = [mp.mpf(x) for x in b["axis"]]
        if b["jtype"] == 1:
            RJ, tJ = rodrigues(a, q[i]), mp.matrix([0, 0, 0])
        else:
            RJ, tJ = mp.eye(3), mp.matrix(a) * q[i]
        if b["parent"] < 0:
            Rp, tp = mp.eye(3), mp.matrix([0, 0, 0])
        else:
            Rp, tp = out[b["parent"]]
        R = Rp * Rt * RJ
        t = tp + Rp * (Rt * tJ + pt)
        out.append((R, t))
    return out


def com_and_inertia(b):
    m = mp.mpf(b["mass"])
    c = mp.matrix([mp.mpf(x) for x in b["mc"]]) / m
    J = b["moment"]
    J = mp.matrix([[J[0], J[3], J[4]], [J[3], J[1], J[5]], [J[4], J[5], J[2]]])
    cc = (c.T * c)[0]
    Icom = J - m * (cc * mp.eye(3) - c * c.T)
    return m, c, Icom


def kinetic(model, q, qd):
    n = len(q)
    qp = [q[k] + H_IN * qd[k] for k in range(n)]
    qm = [q[k] - H_IN * qd[k] for k in range(n)]
    P0, Pp, Pm = poses(model, q), poses(model, qp), poses(model, qm)
    T = mp.mpf(0)
    for i, b in enumerate(model["bodies"]):
        m, c, Icom = com_and_inertia(b)
        R, t = P0[i]
        Rd = (Pp[i][0] - Pm[i][0]) / (2 * H_IN)
        W = Rd * R.T
        w = mp.matrix([W[2, 1], W[0, 2], W[1, 0]])
        cp = Pp[i][1] + Pp[i][0] * c
        cm = Pm[i][1] + Pm[i][0] * c
        vc = (cp - cm) / (2 * H_IN)
        T += m * (vc.T * vc)[0] / 2 + (w.T * (R * Icom * R.T) * w)[0] / 2
    return T


def potential(model, q):
    P = poses(model, q)
    V = mp.mpf(0)
    for i, b in enumerate(model["bodies"]):
        m, c, _ = com_and_inertia(b)
        R, t = P[i]
        V += m * G * (t + R * c)[2]
    return V


def mass_matrix(model, q):
    n = len(q)
    e = lambda k: [mp.mpf(1) if j == k else mp.mpf(0) for j in range(n)]
    Tii = [kinetic(model, q, e(k)) for k in range(n)]
    M = mp.matrix(n, n)
    for i in range(n):
        M[i, i] = 2 * Tii[i]
        for j in range(i + 1, n):
            Tij = kinetic(model, q, [a + b for a, b in zip(e(i), e(j))])
            M[i, j] = M[j, i] = Tij - Tii[i] - Tii[j]
    return M


def qdd_lagrange(model, q, qd, tau):
    n = len(q)
    q = [mp.mpf(x) for x in q]
    qd = [mp.mpf(x) for x in qd]
    M = mass_matrix(model, q)
    Mp = mass_matrix(model, [q[k] + H_OUT * qd[k] for k in range(n)])
    Mm = mass_matrix(model, [q[k] - H_OUT * qd[k] for k in range(n)])
    Mdot = (Mp - Mm) / (2 * H_OUT)
    dT, dV = [], []
    for k in range(n):
        qp = list(q); qp[k] += H_OUT
        qm = list(q); qm[k] -= H_OUT
        dT.append((kinetic(model, qp, qd) - kinetic(model, qm, qd)) / (2 * H_OUT))
        dV.append((potential(model, qp) - potential(model, qm)) / (2 * H_OUT))
    rhs = mp.matrix(n, 1)
    qdv = mp.matrix(qd)
    Mq = Mdot * qdv
    for k in range(n):
        b = model["bodies"][k]
        spring = -mp.mpf(b["spring_k"]) * (q[k] - mp.mpf(b["spring_l"])) if b["jtype"] == 2 else 0
        rhs[k] = mp.mpf(tau[k]) + spring - (Mq[k] - dT[k] + dV[k])
    return mp.lu_solve(M, rhs)


def random_model(rng, n, tree=False):
    bodies = []
    for i in range(n):
        parent = -1 if i == 0 else (int(rng.integers(0, i)) if tree else i - 1)
        jtype = 1 if rng.random() < 0.7 else 2
        axis = rng.normal(size=3)
        if rng.random() < 0.5:  # coordinate axes, as in the workloads
            axis = np.eye(3)[rng.integers(0, 3)] * (1 if rng.random() < 0.5 else -1)
        axis = axis / np.linalg.norm(axis)
        qt = rng.normal(size=4)
        qt /= np.linalg.norm(qt)
        m = float(rng.uniform(0.5, 3.0))
        com = rng.uniform(-0.2, 0.2, 3)
        A = rng.normal(size=(3, 3))
        Icom = A @ A.T * 0.05 + np.eye(3) * 0.01
        J = Icom + m * (com @ com * np.eye(3) - np.outer(com, com))
        bodies.append(dict(parent=parent, jtype=jtype, axis=axis.tolist(), x_tree_quat=qt.tolist(),
                           x_tree_trans=rng.uniform(-0.5, 0.5, 3).tolist(),
                           moment=[J[0, 0], J[1, 1], J[2, 2], J[0, 1], J[0, 2], J[1, 2]], mc=(m * com).tolist(),
                           mass=m, spring_k=float(rng.uniform(0, 20)) if jtype == 2 else 0.0,
                           spring_l=float(rng.uniform(-0.2, 0.2))))
    return dict(bodies=bodies, gravity=9.81)


def main():
    rng = np.random.default_rng(20261202)
    cases = []
    for n, tree, count in ((1, False, 6), (2, False, 8), (3, False, 8), (3, True, 6)):
        for _ in range(count):
            model = random_model(rng, n, tree)
            q = rng.uniform(-np.pi, np.pi, n).tolist()
            qd = rng.normal(0, 1.0, n).tolist()
            tau = rng.normal(0, 1.0, n).tolist()
            a = qdd_lagrange(model, q, qd, tau)
            cases.append(dict(model=model, q=q, v=qd, tau=tau, qdd=[mp.nstr(a[k], 25) for k in range(n)]))
            print(n, tree, [mp.nstr(a[k], 12) for k in range(n)], flush=True)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lagrangian_chains.json")
    json.dump(dict(generator="tests/golden/make_golden.py (mpmath %s, 50 digits)" % mp.__version__, cases=cases),
              open(out, "w"), indent=1)
    print("wrote", out, len(cases), "cases")


if __name__ == "__main__":
    main()
