"""gp_step_noalloc: the stepping entry point that allocates nothing (its
kernels' scratch is a static device array of the library), through the C
ABI, against gp_step on identical batches, bit for bit, on two calls in a
row, the second on a smaller batch; and its refusal of a batch whose scratch
would not fit. The batches run the kernels that use scratch: the
runtime-topology kernel (RK4 over the reference's builder models touching a
plane, the SO-101's welded base) and velocity stepping with continuous
collision detection (thrown cubes), beside compiled-topology buckets."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib(gp):
    return gp.abi.load_library()  # signatures from the binding's table, gp_step_noalloc included


def _pair(gp, w, **opts):
    from gorilla_physics_amd import batch
    b1, _ = batch.batch_from_workload(w, **opts)
    b2, _ = batch.batch_from_workload(w, **opts)
    return b1, b2


def _builders_workload(gp, copies):
    from gorilla_physics_amd import builders
    from test_gpu_builders import _on_ground
    from test_oracle_golden import builder_models
    rng = np.random.default_rng(7)
    descs, qs, vs = [], [], []
    for _ in range(copies):
        for name in builder_models(gp):
            d, q, v = _on_ground(gp, name, rng)
            descs.append(d)
            qs.append(q)
            vs.append(v)
        descs.append(builders.build_so101().model_desc())
        qs.append(rng.uniform(-2, 2, 6))
        vs.append(rng.normal(size=6))
    v = np.concatenate(vs)
    return dict(models=descs, counts=[1] * len(descs), q=np.concatenate(qs), v=v, tau=rng.normal(0.0, 0.3, v.size),
                dt=1e-3)


def _check(gp, lib, w, integrator, steps, label, **opts):
    b_ref, b_na = _pair(gp, w, **opts)
    print(f"case no-alloc {label}: {b_ref.n_instances} instances, integrator {integrator}", flush=True)
    b_ref.step(w["dt"], integrator, steps)
    st = lib.gp_step_noalloc(b_na.handle, float(w["dt"]), int(integrator), int(steps), None)
    assert st == 0, lib.gp_last_error()
    b_na.sync()
    q1, v1 = b_ref.get_state()
    q2, v2 = b_na.get_state()
    assert np.array_equal(q1, q2) and np.array_equal(v1, v2)
    assert np.array_equal(b_ref.get_vdot(), b_na.get_vdot())
    assert np.array_equal(b_ref.faults(), b_na.faults())


def test_noalloc_matches_gp_step_twice_second_smaller(gp):
    from gorilla_physics_amd import abi, workloads
    lib = _lib(gp)
    # first call: builder models (plain kernel, RK4) beside a compiled-topology bucket of arms
    w = _builders_workload(gp, 40)
    arms = workloads.arms(n=64)
    w = dict(models=w["models"] + arms["models"], counts=w["counts"] + arms["counts"],
             q=np.concatenate([w["q"], arms["q"]]), v=np.concatenate([w["v"], arms["v"]]),
             tau=np.concatenate([w["tau"], arms["tau"]]), dt=1e-3)
    _check(gp, lib, w, abi.GP_RUNGE_KUTTA_4, 30, "builders + arms, RK4", precision=0, steps_per_launch=10)
    # second call, smaller: thrown cubes by CCD velocity stepping (the scratch's last contents differ)
    t = workloads.thrown_cubes(n=96)
    _check(gp, lib, t, abi.GP_CCD_VELOCITY_STEPPING, 60, "thrown cubes, CCD", precision=0, steps_per_launch=60)
    # third call, smaller again: the runtime-topology kernel over the scratch the cubes left
    _check(gp, lib, _builders_workload(gp, 5), abi.GP_RUNGE_KUTTA_4, 30, "builders, RK4", precision=0,
           steps_per_launch=10)


def test_noalloc_refuses_a_batch_beyond_the_static_scratch(gp):
    from gorilla_physics_amd import abi, batch, workloads
    lib = _lib(gp)
    w = workloads.thrown_cubes(n=8192)  # about 137 MB of velocity-stepping scratch
    b, _ = batch.batch_from_workload(w, precision=0)
    q0, v0 = b.get_state()
    st = lib.gp_step_noalloc(b.handle, float(w["dt"]), abi.GP_CCD_VELOCITY_STEPPING, 1, None)
    assert st == abi.GP_ERR_UNSUPPORTED
    assert b"static array" in lib.gp_last_error()
    q1, v1 = b.get_state()
    assert np.array_equal(q0, q1) and np.array_equal(v0, v1)  # nothing ran
