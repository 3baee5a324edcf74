"""Shared helpers of the GPU parity tests: the per-step error metric and
teacher-forced comparison of one batched step against the oracle."""
import numpy as np


def rel_err(x, ref):
    """e(x) = max_i |x_i - ref_i| / max(max_i |ref_i|, 1)."""
    x = np.asarray(x, dtype=float)
    ref = np.asarray(ref, dtype=float)
    if ref.size == 0:
        return 0.0
    return float(np.max(np.abs(x - ref)) / max(float(np.max(np.abs(ref))), 1.0))


def instance_slices(models_per_inst):
    qo, vo = [0], [0]
    for d in models_per_inst:
        qo.append(qo[-1] + d.nq)
        vo.append(vo[-1] + d.nv)
    return qo, vo


def quat_fix(q_gpu, q_ref, desc):
    """Align the sign of each floating quaternion of q_gpu with q_ref."""
    q = np.array(q_gpu, dtype=float, copy=True)
    k = 0
    for i in range(desc.n_bodies):
        jt = desc._bodies[i].jtype
        if jt == 3:
            if np.dot(q[k:k + 4], q_ref[k:k + 4]) < 0:
                q[k:k + 4] = -q[k:k + 4]
            k += 7
        elif jt != 0:
            k += 1
    return q


def step_errors(oracle, descs, q0, v0, tau, q1, v1, a1, dt, integrator=0):
    """Per-instance teacher-forced errors of one step: for each instance run
    the oracle's step from (q0, v0) and compare (q1, v1, a1)."""
    qo, vo = instance_slices(descs)
    eq, ev, ea = [], [], []
    for i, d in enumerate(descs):
        qs, qe = qo[i], qo[i + 1]
        vs, ve = vo[i], vo[i + 1]
        rq, rv, ra = oracle.step(d, q0[qs:qe], v0[vs:ve], tau[vs:ve], dt, integrator)
        eq.append(rel_err(quat_fix(q1[qs:qe], rq, d), rq))
        ev.append(rel_err(v1[vs:ve], rv))
        ea.append(rel_err(a1[vs:ve], ra))
    return np.array(eq), np.array(ev), np.array(ea)
