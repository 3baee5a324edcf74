"""Welded-base arms on the compiled kernels (TopoWelded: a body fixed to the
world and a chain of 1-7 one-DOF joints behind it, the shape of the SO-101,
builders/mod.rs:252): every step against the oracle from the same state,
FP64 within 1e-10 (SIE, RK2, RK4), mixed precision within the mixed bounds;
batches are built with GP_KERNEL_FAST, which fails unless every bucket runs
on a compiled kernel. Random welded arms have a tilted, offset base (gravity
seen at an angle in the base frame), revolute and sprung prismatic joints."""
import numpy as np
import pytest

from gpu_helpers import step_errors

pytestmark = pytest.mark.gpu

F64_TOL = 1e-10
MIXED_Q_TOL, MIXED_VA_TOL = 1e-5, 2e-3


def welded_arm_desc(rng, k):
    from gorilla_physics_amd import abi
    from gorilla_physics_amd.mechanism import (Isometry3, JointSpring, ModelDesc, SpatialInertia, UnitQuaternion,
                                               body_desc)
    axis = rng.normal(size=3)
    base_iso = Isometry3(UnitQuaternion.from_axis_angle(axis / np.linalg.norm(axis), rng.uniform(0.2, 1.0)),
                         rng.uniform(-0.3, 0.3, 3))
    bodies = [body_desc(-1, abi.GP_JOINT_FIXED, [0, 0, 1], base_iso,
                        SpatialInertia.cuboid_at([0, 0, 0.05], 2.0, 0.2, 0.2, 0.1, "base"))]
    for i in range(1, k + 1):
        l = rng.uniform(0.1, 0.4)
        prism = rng.random() < 0.25
        jt = abi.GP_JOINT_PRISMATIC if prism else abi.GP_JOINT_REVOLUTE
        ax = np.zeros(3)
        ax[rng.integers(0, 3)] = 1.0
        spring = JointSpring(rng.uniform(50, 300), 0.05) if prism else None
        bodies.append(body_desc(i - 1, jt, ax, Isometry3.translation_xyz(0.0, 0.0, 0.1 if i == 1 else l),
                                SpatialInertia.cuboid_at(rng.uniform(-0.02, 0.02, 3) + [0, 0, l / 2],
                                                         rng.uniform(0.3, 2.0), 0.05, 0.05, l, f"l{i}"), spring))
    return ModelDesc(bodies, [], [])


def _workload(gp, n_each=24, seed=4):
    from gorilla_physics_amd import builders
    rng = np.random.default_rng(seed)
    descs = [builders.build_so101().model_desc()] * n_each
    for k in range(1, 8):
        descs += [welded_arm_desc(rng, k) for _ in range(n_each)]
    nq = sum(d.nq for d in descs)
    return dict(models=descs, counts=[1] * len(descs), q=rng.uniform(-1.5, 1.5, nq), v=rng.normal(0, 0.5, nq),
                tau=rng.normal(0, 0.05, nq), dt=1e-3)


@pytest.mark.parametrize("precision,integrator", [(0, 0), (0, 1), (0, 2), (1, 0), (1, 2)])
def test_welded_arms_step_parity(gp, oracle, precision, integrator):
    from gorilla_physics_amd import abi, batch
    w = _workload(gp)
    print(f"case welded arms precision {precision} integrator {integrator}: {len(w['models'])}", flush=True)
    b, _ = batch.batch_from_workload(w, precision=precision, kernel=abi.GP_KERNEL_FAST, steps_per_launch=1)
    q, v = b.get_state()
    worst = np.zeros(3)
    for _ in range(40):
        b.step(w["dt"], integrator, 1)
        q1, v1 = b.get_state()
        a1 = b.get_vdot()
        eq, ev, ea = step_errors(oracle, w["models"], q, v, w["tau"], q1, v1, a1, w["dt"], integrator)
        worst = np.maximum(worst, [eq.max(), ev.max(), ea.max()])
        q, v = q1, v1
    assert not b.faults().any()
    print(f"worst per-step e(q') {worst[0]:.3g} e(v') {worst[1]:.3g} e(qdd) {worst[2]:.3g}", flush=True)
    if precision == abi.GP_PREC_F64:
        assert worst.max() <= F64_TOL
    else:
        assert worst[0] <= MIXED_Q_TOL and worst[1] <= MIXED_VA_TOL and worst[2] <= MIXED_VA_TOL


def test_so101_workload_at_full_size_on_the_compiled_kernel(gp, oracle):
    """65,536 SO-101 arms (the benchmark's workload), RK4, 100 steps in one
    launch on the compiled kernel; a sample of 256 arms against the oracle's
    100-step rollouts from the same start (FP64: within 1e-9)."""
    from gorilla_physics_amd import abi, batch, workloads
    w = workloads.so101(n=65536)
    print("case SO-101 x 65536 RK4 compiled", flush=True)
    b, _ = batch.batch_from_workload(w, precision=0, kernel=abi.GP_KERNEL_FAST, steps_per_launch=100)
    b.step(w["dt"], 2, 100)
    q, v = b.get_state()
    assert not b.faults().any()
    idx = np.random.default_rng(0).choice(65536, 256, replace=False)
    d = w["models"][0]
    qs = np.concatenate([w["q"][6 * i:6 * i + 6] for i in idx])
    vs = np.concatenate([w["v"][6 * i:6 * i + 6] for i in idx])
    ts = np.concatenate([w["tau"][6 * i:6 * i + 6] for i in idx])
    rq, rv, fails = oracle.rollout([d] * 256, qs, vs, ts, w["dt"], 100, integrator=2, threads=8)
    assert fails == 0
    gq = np.concatenate([q[6 * i:6 * i + 6] for i in idx])
    gv = np.concatenate([v[6 * i:6 * i + 6] for i in idx])
    e = max(np.abs(gq - rq).max() / max(1, np.abs(rq).max()), np.abs(gv - rv).max() / max(1, np.abs(rv).max()))
    print(f"100-step rollout difference {e:.3g}", flush=True)
    assert e <= 1e-9
