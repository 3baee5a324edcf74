"""GPU parity: one batched step on the MI355X against the CPU oracle (the
reference's CRBA + RNEA + LU restated in FP64), teacher-forced from the same
state. Metric e(x) = max|x - x_ref| / max(max|x_ref|, 1) per instance for
x in {q', v', qdd}; quaternions compared up to sign.

Bounds (written here, stated per workload):
  FP64 kernels (plain and fast):            e <= 1e-10 (measured <= 2.3e-12)
  mixed kernels (FP32 sweeps, FP64 D + q/v): e(q') <= 1e-5, e(v'), e(qdd) <= 2e-3,
      p99 e(qdd) <= 1e-4 (measured on 512 arms: max 2e-4, p99 1.5e-5; the
      FP32 cancellation in I^A - U U^T / D grows with cond(M), up to 2.5e5 here)
"""
import numpy as np
import pytest

from gpu_helpers import rel_err, step_errors

pytestmark = pytest.mark.gpu

F64_TOL = 1e-10
MIXED_Q_TOL = 1e-5
MIXED_VA_TOL = 2e-3
MIXED_P99_TOL = 1e-4


def _run_step(gp, w, n_pick=None, **opts):
    from gorilla_physics_amd import batch
    b, models = batch.batch_from_workload(w, **opts)
    q0, v0 = b.get_state()
    b.step(w["dt"], 0, 1)
    q1, v1 = b.get_state()
    a1 = b.get_vdot()
    assert not b.faults().any()
    return q0, v0, q1, v1, a1


def _descs(w):
    out = []
    for d, c in zip(w["models"], w["counts"]):
        out.extend([d] * int(c))
    return out


@pytest.fixture(scope="module")
def wl(gp):
    from gorilla_physics_amd import workloads
    return dict(
        acrobot=workloads.acrobot(n=256),
        arms=workloads.arms(n=512),
        mixed=workloads.mixed(n=512),
        cubes=workloads.cubes(n=256),
    )


@pytest.mark.parametrize("name", ["acrobot", "arms", "mixed", "cubes"])
@pytest.mark.parametrize("kernel", ["plain", "fast"])
def test_f64_step_parity(gp, oracle, wl, name, kernel):
    from gorilla_physics_amd import abi
    w = wl[name]
    k = abi.GP_KERNEL_PLAIN if kernel == "plain" else abi.GP_KERNEL_FAST
    print(f"case f64 {name} {kernel}", flush=True)
    q0, v0, q1, v1, a1 = _run_step(gp, w, precision=abi.GP_PREC_F64, kernel=k)
    eq, ev, ea = step_errors(oracle, _descs(w), q0, v0, w["tau"], q1, v1, a1, w["dt"])
    print(f"{name} {kernel}: max e(q') {eq.max():.3g} e(v') {ev.max():.3g} e(qdd) {ea.max():.3g}")
    assert eq.max() <= F64_TOL and ev.max() <= F64_TOL and ea.max() <= F64_TOL


@pytest.mark.parametrize("name", ["acrobot", "arms", "mixed", "cubes"])
def test_mixed_step_bound(gp, oracle, wl, name):
    from gorilla_physics_amd import abi
    w = wl[name]
    print(f"case mixed {name}", flush=True)
    q0, v0, q1, v1, a1 = _run_step(gp, w, precision=abi.GP_PREC_MIXED, kernel=abi.GP_KERNEL_FAST)
    eq, ev, ea = step_errors(oracle, _descs(w), q0, v0, w["tau"], q1, v1, a1, w["dt"])
    print(f"{name} mixed: max e(q') {eq.max():.3g} e(v') {ev.max():.3g} e(qdd) {ea.max():.3g} "
          f"p99 e(qdd) {np.percentile(ea, 99):.3g}")
    assert eq.max() <= MIXED_Q_TOL and ev.max() <= MIXED_VA_TOL and ea.max() <= MIXED_VA_TOL
    assert np.percentile(ea, 99) <= MIXED_P99_TOL


def _per_instance(x, offs):
    return [x[offs[i]:offs[i + 1]] for i in range(len(offs) - 1)]


@pytest.mark.parametrize("name,n", [("arms", 65536), ("cubes", 16384), ("mixed", 32768)])
def test_full_size_step_parity(gp, oracle, name, n):
    """The full batch sizes of the benchmark workloads: one teacher-forced
    step of every instance, FP64 fast kernel (<= 1e-10) and mixed kernel
    (bounds above), against the oracle run on all host threads."""
    import os
    from gorilla_physics_amd import abi, workloads
    w = getattr(workloads, name)(n=n)
    descs = _descs(w)
    print(f"case full {name} n={n}", flush=True)
    rq, rv, ra, fails = oracle.rollout(descs, w["q"], w["v"], w["tau"], w["dt"], 1,
                                       threads=min(16, os.cpu_count() or 1), want_vdot=True)
    assert fails == 0
    qo = np.concatenate([[0], np.cumsum([d.nq for d in descs])])
    vo = np.concatenate([[0], np.cumsum([d.nv for d in descs])])
    for prec, tq, tva in ((abi.GP_PREC_F64, F64_TOL, F64_TOL), (abi.GP_PREC_MIXED, MIXED_Q_TOL, MIXED_VA_TOL)):
        q0, v0, q1, v1, a1 = _run_step(gp, w, precision=prec, kernel=abi.GP_KERNEL_AUTO)
        eq = np.array([rel_err(a, b) for a, b in zip(_per_instance(q1, qo), _per_instance(rq, qo))])
        ev = np.array([rel_err(a, b) for a, b in zip(_per_instance(v1, vo), _per_instance(rv, vo))])
        ea = np.array([rel_err(a, b) for a, b in zip(_per_instance(a1, vo), _per_instance(ra, vo))])
        print(f"{name} n={n} prec={prec}: max e(q') {eq.max():.3g} e(v') {ev.max():.3g} e(qdd) {ea.max():.3g} "
              f"p99 e(qdd) {np.percentile(ea, 99):.3g}", flush=True)
        assert eq.max() <= tq and ev.max() <= tva and ea.max() <= tva


def test_reference_known_answers_on_gpu(gp):
    """The reference's closed-form dynamics tests, run through the HIP path
    via the MechanismState mirror (dynamics.rs:883-1088, floating.rs:82-129)."""
    from gorilla_physics_amd import builders
    g = gp.GRAVITY
    m, l = 5.0, 7.0
    rod = builders.build_pendulum(m, np.diag([0, m * l * l / 3, m * l * l / 3]), [m * l / 2, 0, 0],
                                  gp.Isometry3.identity(), [0, 1, 0])
    assert abs(gp.dynamics_continuous(rod, [0.0])[0] - 3 * g / (2 * l)) < 1e-12
    rot = builders.build_pendulum(m, np.diag([0, m * l * l / 3, m * l * l / 3]), [m * l / 2, 0, 0],
                                  gp.Isometry3.rotation_scaled_axis([np.pi / 2, 0, 0]), [0, 0, 1])
    assert abs(gp.dynamics_continuous(rot, [0.0])[0] + 3 * g / (2 * l)) < 1e-12
    hold = builders.build_pendulum(m, np.diag([0, m * l * l / 3, m * l * l / 3]), [m * l / 2, 0, 0],
                                   gp.Isometry3.identity(), [0, 1, 0])
    assert abs(gp.dynamics_continuous(hold, [-m * g * l / 2])[0]) < 1e-6
    ball = gp.RigidBody(gp.SpatialInertia("ball", np.diag([2.0, 2.0, 2.0]), [0, 0, 0], 5.0))
    st = gp.MechanismState([gp.FloatingJoint(gp.Transform3D.identity("ball", gp.WORLD_FRAME))], [ball])
    st.update([gp.Pose(gp.UnitQuaternion.from_euler_angles(0.1, 0.2, 0.3), [1, 2, 3])],
              [gp.SpatialVector([1, 2, 3], [4, 5, 6])])
    acc = gp.dynamics_continuous(st, [gp.SpatialVector.zero()])[0]
    got = np.concatenate([acc.angular, acc.linear])
    assert np.allclose(got, [0, 0, 0, 4.948946, -6.959844, -6.566419], atol=1e-5)  # RigidBodyDynamics.jl


def test_lane_independence(gp):
    """An instance's result does not depend on its lane or its batch mates:
    the same arm run alone and at lane 77 of a mixed batch is bit-identical."""
    from gorilla_physics_amd import abi, batch, workloads
    w = workloads.arms(n=128)
    for prec in (abi.GP_PREC_F64, abi.GP_PREC_MIXED):
        b, _ = batch.batch_from_workload(w, precision=prec, kernel=abi.GP_KERNEL_FAST)
        b.step(w["dt"], 0, 50)
        q_all, v_all = b.get_state()
        i = 77
        one = dict(models=[w["models"][i]], counts=[1], q=w["q"][7 * i:7 * i + 7], v=w["v"][7 * i:7 * i + 7],
                   tau=w["tau"][7 * i:7 * i + 7], dt=w["dt"])
        b1, _ = batch.batch_from_workload(one, precision=prec, kernel=abi.GP_KERNEL_FAST)
        b1.step(w["dt"], 0, 50)
        q1, v1 = b1.get_state()
        assert np.array_equal(q1, q_all[7 * i:7 * i + 7]) and np.array_equal(v1, v_all[7 * i:7 * i + 7])


def _quadrupeds(n, prismatic_knee):
    """n floating quadrupeds of the mixed workload; with prismatic_knee, the
    knee of leg 2 of every other instance slides along its axis against a
    spring (k = 40 N/m, rest 0.05 m), so the tree kernel's per-lane
    revolute/prismatic selection runs."""
    from gorilla_physics_amd import abi, workloads
    w = workloads.mixed(n=n, seed=11, tree_fraction=1.0)
    if prismatic_knee:
        q = w["q"].copy()
        for i, d in enumerate(w["models"]):
            if i % 2 == 0:
                d._bodies[9].jtype = abi.GP_JOINT_PRISMATIC
                d._bodies[9].spring_k, d._bodies[9].spring_l = 40.0, 0.05
                q[19 * i + 7 + 8] = 0.02 * (i % 5)
        w["q"] = q
    return w


@pytest.mark.parametrize("prismatic_knee", [False, True])
@pytest.mark.parametrize("lanes_per_tree", [1, 4])
def test_quadruped_tree_parity(gp, oracle, lanes_per_tree, prismatic_knee):
    """Floating quadrupeds through the compiled tree kernels: one lane per
    instance (lanes_per_tree = 1) and one lane per leg with the legs' base
    contributions summed across lanes (4). FP64 within 1e-10 of the oracle,
    mixed within the mixed bounds; 40 steps of the FP64 rollout against the
    oracle's."""
    from gorilla_physics_amd import abi
    w = _quadrupeds(384, prismatic_knee)
    descs = _descs(w)
    print(f"case quadruped tree lanes={lanes_per_tree} prismatic={prismatic_knee}", flush=True)
    for prec in (abi.GP_PREC_F64, abi.GP_PREC_MIXED):
        q0, v0, q1, v1, a1 = _run_step(gp, w, precision=prec, kernel=abi.GP_KERNEL_FAST,
                                       lanes_per_tree=lanes_per_tree)
        eq, ev, ea = step_errors(oracle, descs, q0, v0, w["tau"], q1, v1, a1, w["dt"])
        print(f"prec={prec}: max e(q') {eq.max():.3g} e(v') {ev.max():.3g} e(qdd) {ea.max():.3g}", flush=True)
        if prec == abi.GP_PREC_F64:
            assert eq.max() <= F64_TOL and ev.max() <= F64_TOL and ea.max() <= F64_TOL
        else:
            assert eq.max() <= MIXED_Q_TOL and ev.max() <= MIXED_VA_TOL and ea.max() <= MIXED_VA_TOL
    from gorilla_physics_amd import batch
    b, _ = batch.batch_from_workload(w, precision=abi.GP_PREC_F64, kernel=abi.GP_KERNEL_FAST,
                                     lanes_per_tree=lanes_per_tree, steps_per_launch=16)
    b.step(w["dt"], 0, 40)
    qg, vg = b.get_state()
    rq, rv, fails = oracle.rollout(descs, w["q"], w["v"], w["tau"], w["dt"], 40)
    assert fails == 0 and not b.faults().any()
    assert rel_err(qg, rq) < 1e-9 and rel_err(vg, rv) < 1e-8


def test_quadruped_tree_and_serial_agree_in_a_mix(gp):
    """The tree kernel inside a grouped launch (quadrupeds beside chains) and
    in a launch of its own give the same bits; the serial kernel agrees to
    rounding."""
    from gorilla_physics_amd import abi, batch, workloads
    w = workloads.mixed(n=1024, seed=4)
    print("case quadruped tree grouped vs separate", flush=True)
    out = {}
    for name, kw in (("tree_grouped", {}), ("tree_separate", dict(separate_launches=1)),
                     ("serial", dict(lanes_per_tree=1))):
        b, _ = batch.batch_from_workload(w, precision=abi.GP_PREC_F64, kernel=abi.GP_KERNEL_FAST, **kw)
        b.step(w["dt"], 0, 30)
        out[name] = b.get_state()
    assert np.array_equal(out["tree_grouped"][0], out["tree_separate"][0], equal_nan=True)
    assert np.array_equal(out["tree_grouped"][1], out["tree_separate"][1], equal_nan=True)
    qt, qs = out["tree_grouped"][0], out["serial"][0]
    fin = np.isfinite(qt)
    assert np.array_equal(fin, np.isfinite(qs)) and fin.mean() > 0.99
    assert rel_err(qt[fin], qs[fin]) < 1e-10
