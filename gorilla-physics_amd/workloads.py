"""Synthetic workloads (seeded, counter-based) for the four batch shapes of the
batched step, plus the acrobot CPU case.

Every random field is drawn from its own Philox stream keyed by
(seed, workload id, field id), one row per instance, so instance i gets the same
values whatever the batch size (a prefix property the tests rely on).

  acrobot   1024 planar 2-link revolute chains (build_double_pendulum pattern,
            helpers.rs:49; uniform rod as in simulate.rs:129-136), zero torque
  arms      65,536 serial 7-DOF revolute arms, random links/axes/torques
  mixed     32,768 chains of 1-12 revolute/prismatic DOF (75%) and
            floating-base quadruped trees, 6 + 4 x 3 DOF (25%)
  cubes     16,384 floating unit cubes with 8 corner contacts over z = 0
            (build_cube, helpers.rs:151; HalfSpace::new_with_params)
"""
import math

import numpy as np

from . import abi
from .mechanism import (GRAVITY, Isometry3, ModelDesc, SpatialInertia, UnitQuaternion, body_desc, contact_desc,
                        halfspace_desc, HalfSpace)

WORKLOAD_IDS = {"acrobot": 1, "arms": 2, "mixed": 3, "cubes": 4, "thrown_cubes": 5, "so101": 6}


def stream(seed, workload, field):
    """Independent Philox stream for (seed, workload, field)."""
    return np.random.Generator(np.random.Philox(key=[int(seed), WORKLOAD_IDS[workload] * 1000 + int(field)]))


def _axis(k):
    a = np.zeros(3)
    a[int(k)] = 1.0
    return a


# ----------------------------------------------------------------- acrobot
def acrobot_desc(m=1.0, l=1.0):
    inertia = SpatialInertia("rod", np.diag([0.0, m * l * l / 3.0, m * l * l / 3.0]), [m * l / 2.0, 0.0, 0.0], m)
    y = [0.0, 1.0, 0.0]
    bodies = [body_desc(-1, abi.GP_JOINT_REVOLUTE, y, Isometry3.identity(), inertia),
              body_desc(0, abi.GP_JOINT_REVOLUTE, y, Isometry3.translation_xyz(l, 0.0, 0.0), inertia)]
    return ModelDesc(bodies, [], [])


def acrobot(n=1024, seed=0):
    d = acrobot_desc()
    q = stream(seed, "acrobot", 1).uniform(-math.pi, math.pi, (n, 2))
    v = stream(seed, "acrobot", 2).normal(0.0, 0.5, (n, 2))
    return dict(models=[d], counts=[n], q=q.reshape(-1), v=v.reshape(-1), tau=np.zeros(2 * n),
                dt=1e-3, steps=2000, name="acrobot")


# ----------------------------------------------------------------- serial chains
def chain_desc(n, kinds, lengths, axes, masses, widths, depths, comx, comy):
    """Serial chain: X_tree_0 = identity, X_tree_i = Trans(0, 0, l_{i-1}); body i
    is a box (w, d, h = l_i) with com (cx, cy, l_i / 2) (inertia.rs:56-74)."""
    bodies = []
    for i in range(n):
        iso = Isometry3.identity() if i == 0 else Isometry3.translation_xyz(0.0, 0.0, lengths[i - 1])
        inertia = SpatialInertia.cuboid_at([comx[i], comy[i], lengths[i] / 2.0], masses[i], widths[i], depths[i],
                                           lengths[i], f"b{i}")
        jt = abi.GP_JOINT_REVOLUTE if kinds[i] == 0 else abi.GP_JOINT_PRISMATIC
        bodies.append(body_desc(i - 1, jt, _axis(axes[i]), iso, inertia))
    return ModelDesc(bodies, [], [])


def _chain_fields(seed, workload, n, nb):
    return dict(
        l=stream(seed, workload, 10).uniform(0.1, 0.5, (n, nb)),
        ax=stream(seed, workload, 11).integers(0, 3, (n, nb)),
        m=stream(seed, workload, 12).uniform(0.5, 3.0, (n, nb)),
        w=stream(seed, workload, 13).uniform(0.03, 0.08, (n, nb)),
        d=stream(seed, workload, 14).uniform(0.03, 0.08, (n, nb)),
        cx=stream(seed, workload, 15).uniform(-0.02, 0.02, (n, nb)),
        cy=stream(seed, workload, 16).uniform(-0.02, 0.02, (n, nb)),
        tau=stream(seed, workload, 17).normal(0.0, 1.0, (n, nb)),
        q=stream(seed, workload, 18).uniform(-math.pi, math.pi, (n, nb)),
        v=stream(seed, workload, 19).normal(0.0, 0.5, (n, nb)),
    )


def arms(n=65536, seed=0, dof=7):
    f = _chain_fields(seed, "arms", n, dof)
    models = [chain_desc(dof, [0] * dof, f["l"][i], f["ax"][i], f["m"][i], f["w"][i], f["d"][i], f["cx"][i],
                         f["cy"][i]) for i in range(n)]
    return dict(models=models, counts=[1] * n, q=f["q"].reshape(-1), v=f["v"].reshape(-1),
                tau=f["tau"].reshape(-1), dt=1e-3, steps=1000, name="arms")


# ----------------------------------------------------------------- quadruped trees
def quadruped_desc(p):
    """Cuboid base on a floating joint + 4 legs of [hip-x, hip-y, knee-y]
    revolute joints (13 bodies, nq = 19, nv = 18), cf. build_quadruped
    (helpers.rs:423)."""
    bw, bd, bh, bm = p["bw"], p["bd"], p["bh"], p["bm"]
    bodies = [body_desc(-1, abi.GP_JOINT_FLOATING, [0, 0, 0], Isometry3.identity(),
                        SpatialInertia.cuboid_at([0, 0, 0], bm, bw, bd, bh, "base"))]
    for leg, (sx, sy) in enumerate(((1, -1), (1, 1), (-1, -1), (-1, 1))):
        lt, ls = p["lt"][leg], p["ls"][leg]
        hx = len(bodies)
        bodies.append(body_desc(0, abi.GP_JOINT_REVOLUTE, [1, 0, 0],
                                Isometry3.translation_xyz(sx * bw / 2, sy * bd / 2, 0.0),
                                SpatialInertia.cuboid_at([0, 0, 0], 0.2, 0.05, 0.05, 0.05, "hx")))
        hy = len(bodies)
        bodies.append(body_desc(hx, abi.GP_JOINT_REVOLUTE, [0, 1, 0], Isometry3.identity(),
                                SpatialInertia.cuboid_at([0, 0, -lt / 2], p["mt"][leg], 0.05, 0.05, lt, "th")))
        bodies.append(body_desc(hy, abi.GP_JOINT_REVOLUTE, [0, 1, 0], Isometry3.translation_xyz(0.0, 0.0, -lt),
                                SpatialInertia.cuboid_at([0, 0, -ls / 2], p["ms"][leg], 0.04, 0.04, ls, "sh")))
    return ModelDesc(bodies, [], [])


def mixed(n=32768, seed=0, tree_fraction=0.25):
    """75% chains of 1-12 revolute (p = 0.7) / prismatic DOF, 25% quadrupeds."""
    is_tree = stream(seed, "mixed", 1).random(n) < tree_fraction
    nb = stream(seed, "mixed", 2).integers(1, 13, n)
    kind = (stream(seed, "mixed", 3).random((n, 12)) >= 0.7).astype(int)  # 0 revolute, 1 prismatic
    f = _chain_fields(seed, "mixed", n, 12)
    tb = dict(bw=stream(seed, "mixed", 30).uniform(0.8, 1.5, n), bd=stream(seed, "mixed", 31).uniform(0.3, 0.6, n),
              bh=stream(seed, "mixed", 32).uniform(0.2, 0.4, n), bm=stream(seed, "mixed", 33).uniform(3.0, 6.0, n),
              lt=stream(seed, "mixed", 34).uniform(0.2, 0.5, (n, 4)), ls=stream(seed, "mixed", 35).uniform(0.2, 0.5, (n, 4)),
              mt=stream(seed, "mixed", 36).uniform(0.3, 1.0, (n, 4)), ms=stream(seed, "mixed", 37).uniform(0.3, 1.0, (n, 4)))
    leg_q = stream(seed, "mixed", 40).uniform(-0.5, 0.5, (n, 12))
    leg_v = stream(seed, "mixed", 41).normal(0.0, 0.5, (n, 12))
    base_v = stream(seed, "mixed", 42).normal(0.0, 0.5, (n, 6))
    models, qs, vs, taus = [], [], [], []
    for i in range(n):
        if is_tree[i]:
            p = {k: tb[k][i] for k in tb}
            models.append(quadruped_desc(p))
            qs.append(np.concatenate([[0, 0, 0, 1, 0, 0, 1.0], leg_q[i]]))
            vs.append(np.concatenate([base_v[i], leg_v[i]]))
            taus.append(np.concatenate([np.zeros(6), f["tau"][i]]))
        else:
            k = int(nb[i])
            models.append(chain_desc(k, kind[i][:k], f["l"][i], f["ax"][i], f["m"][i], f["w"][i], f["d"][i],
                                     f["cx"][i], f["cy"][i]))
            q = np.where(kind[i][:k] == 0, f["q"][i][:k], 0.1 * f["q"][i][:k] / math.pi)
            qs.append(q)
            vs.append(f["v"][i][:k])
            taus.append(f["tau"][i][:k])
    return dict(models=models, counts=[1] * n, q=np.concatenate(qs), v=np.concatenate(vs),
                tau=np.concatenate(taus), dt=1e-3, steps=1000, name="mixed")


# ----------------------------------------------------------------- cubes on a plane
def cube_desc(k_point=12500.0, alpha=1.0 / 150.0, mu=0.5, m=1.0, l=1.0):
    """build_cube(1, 1) (helpers.rs:151) with corner contacts of stiffness
    k_point over HalfSpace::new_with_params(z, 0, alpha, mu). With the
    half-space's own 50e3 in series, k_point = 12,500 gives k = 1e4 and
    alpha = 1/150 gives lambda = 1.5 alpha k = 1e2 (contact.rs:274-276)."""
    mx = m * l * l / 6.0
    inertia = SpatialInertia("cube", np.diag([mx, mx, mx]), np.zeros(3), m)
    bodies = [body_desc(-1, abi.GP_JOINT_FLOATING, [0, 0, 0], Isometry3.identity(), inertia)]
    h = l / 2.0
    corners = ([h, h, -h], [h, -h, -h], [-h, h, -h], [-h, -h, -h], [h, h, h], [h, -h, h], [-h, h, h], [-h, -h, h])
    contacts = [contact_desc(0, p, k_point) for p in corners]
    hs = [halfspace_desc(HalfSpace.new_with_params([0.0, 0.0, 1.0], 0.0, alpha, mu))]
    return ModelDesc(bodies, contacts, hs)


def so101(n=65536, seed=0):
    """SO-101 arms (builders/mod.rs:252: welded base, six revolutes, the
    URDF's inertias): joint angles U(-pi, pi), velocities N(0, 0.5), held
    torques N(0, 0.02) N m, RK4 at dt 1e-3 for 1000 steps. The arm has no
    compiled topology (its root joint is fixed), so it runs on the
    runtime-topology FP64 kernel."""
    from . import builders
    d = builders.build_so101().model_desc()
    q = stream(seed, "so101", 1).uniform(-np.pi, np.pi, 6 * n)
    v = stream(seed, "so101", 2).normal(0.0, 0.5, 6 * n)
    tau = stream(seed, "so101", 3).normal(0.0, 0.02, 6 * n)
    return dict(models=[d], counts=[n], q=q, v=v, tau=tau, dt=1e-3, steps=1000, name="so101", integrator=2)


def thrown_cubes(n=16384, seed=0):
    """Cubes of side 0.1 thrown at a plane and stepped by velocity stepping with
    continuous collision detection, as the reference's random-cube test
    (collision/ccd/mod.rs:284): positions U(-100, 100)^2 x U(l, 2l), uniform
    random rotation, linear velocity U(-1, 1)^3, angular U(-5, 5)^3 (world
    frame), dt = 1/120, 120 steps (one second)."""
    from . import builders
    from .mechanism import HalfSpace
    l = 0.1
    st = builders.build_cube(1.0, l)
    st.add_halfspace(HalfSpace.new([0, 0, 1.0], 0.0))
    d = st.model_desc()
    qr = stream(seed, "thrown_cubes", 1).normal(size=(n, 4))
    qr /= np.linalg.norm(qr, axis=1, keepdims=True)
    pos = stream(seed, "thrown_cubes", 2).uniform(-100.0, 100.0, (n, 3))
    pos[:, 2] = stream(seed, "thrown_cubes", 3).uniform(l, 2 * l, n)
    lin = stream(seed, "thrown_cubes", 4).uniform(-1.0, 1.0, (n, 3))
    ang = stream(seed, "thrown_cubes", 5).uniform(-5.0, 5.0, (n, 3))
    q = np.zeros((n, 7))
    v = np.zeros((n, 6))
    for i in range(n):
        R = UnitQuaternion(*qr[i]).to_matrix()
        q[i, :4] = qr[i]
        q[i, 4:] = pos[i]
        v[i, :3] = R.T @ ang[i]
        v[i, 3:] = R.T @ lin[i]
    return dict(models=[d], counts=[n], q=q.reshape(-1), v=v.reshape(-1), tau=np.zeros(6 * n),
                dt=1.0 / 120.0, steps=120, name="thrown_cubes", integrator=4)


def cubes(n=16384, seed=0):
    d = cube_desc()
    qr = stream(seed, "cubes", 1).normal(size=(n, 4))
    qr /= np.linalg.norm(qr, axis=1, keepdims=True)
    h = stream(seed, "cubes", 2).uniform(0.5, 2.0, n)
    w = stream(seed, "cubes", 3).normal(0.0, 1.0, (n, 3))
    corners = np.array([[sx * 0.5, sy * 0.5, sz * 0.5] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])
    q = np.zeros((n, 7))
    v = np.zeros((n, 6))
    for i in range(n):
        R = UnitQuaternion(*qr[i]).to_matrix()
        low = float(np.min(corners @ R.T[:, 2]))
        q[i, :4] = qr[i]
        q[i, 6] = h[i] - low
        v[i, :3] = w[i]
    return dict(models=[d], counts=[n], q=q.reshape(-1), v=v.reshape(-1), tau=np.zeros(6 * n),
                dt=1e-4, steps=5000, name="cubes")

