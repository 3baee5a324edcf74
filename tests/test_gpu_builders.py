"""The reference's model builders (helpers.rs, builders/) through the HIP
path: every model in one batch, each instance started in contact with a
ground plane with random joint velocities, and every step of every instance
compared with one oracle step from the same state (FP64: <= 1e-10, the
bound of the other FP64 parity tests); the biped standing as the browser
client creates it (interface/biped.rs:16-56), stepped by velocity stepping
with continuous collision detection, against the oracle at its 1e-5."""
import numpy as np
import pytest

from gpu_helpers import step_errors
from test_gpu_velocity_stepping import CCD, VS, Stepper
from test_oracle_golden import builder_models

pytestmark = pytest.mark.gpu

F64_TOL = 1e-10


def _on_ground(gp, name, rng):
    """The builder's model (the browser client's standing pose where it has
    one) on a plane tilted 0.1 rad about y whose offset puts the lowest
    contact point 1 mm inside it, with random velocities."""
    from gorilla_physics_amd import interface
    create = {"leg": interface.create_leg, "leg_from_foot": interface.create_leg_from_foot,
              "biped": interface.create_biped}
    st = create[name]() if name in create else builder_models(gp)[name][0]
    n = np.array([np.sin(0.1), 0.0, np.cos(0.1)])
    lowest = (interface.InterfaceSimulator(st).contact_positions().reshape(-1, 3) @ n).min()
    st.add_halfspace(gp.HalfSpace.new_with_params(n, lowest + 1e-3, 0.9, 0.5))
    return st.model_desc(), st.flat_q(), st.flat_v() + rng.normal(size=st.flat_v().size) * 0.3


@pytest.mark.parametrize("integrator", [0, 2])
def test_builder_models_step_parity(gp, oracle, integrator):
    from gorilla_physics_amd import batch
    rng = np.random.default_rng(11)
    descs, q, v = [], [], []
    for name in builder_models(gp):
        for _ in range(3):
            d, qi, vi = _on_ground(gp, name, rng)
            descs.append(d)
            q.append(qi)
            v.append(vi)
    print(f"case builder models, integrator {integrator}, {len(descs)} instances", flush=True)
    b = batch.Batch([batch.Model(d) for d in descs], [1] * len(descs), precision=0)
    q, v = np.concatenate(q), np.concatenate(v)
    tau = np.zeros(v.size)
    b.set_state(q, v)
    worst = 0.0
    for k in range(200):
        b.step(1e-3, integrator, 1)
        q1, v1 = b.get_state()
        a1 = b.get_vdot()
        if k % 5 == 0:
            eq, ev, ea = step_errors(oracle, descs, q, v, tau, q1, v1, a1, 1e-3, integrator)
            worst = max(worst, eq.max(), ev.max(), ea.max())
        q, v = q1, v1
    print(f"worst per-step difference {worst:.3g}", flush=True)
    assert not b.faults().any()
    assert worst <= F64_TOL


def test_interface_biped_velocity_stepping(gp, oracle):
    """The standing biped of createBiped on a plane through its feet's
    bottom faces, pushed sideways, stepped by velocity stepping at 1/60 s
    against the oracle (and the same number of collision-loop rounds).

    With continuous collision detection the reference's loop diverges on
    this model (feet of 0.1 kg links, each cut adding gravity for the rest of
    the step again, integrators.rs:56-60): the oracle's velocities pass 1e269
    within the second step at 1/60 s, and at 1/240 to 1/1000 s it runs away or
    exceeds 64 rounds within 300 steps. The GPU must flag that instance at
    the step where the oracle's state turns non-finite, and it is stepped
    without CCD here."""
    from gorilla_physics_amd import interface
    def standing():
        st = interface.create_biped()
        st.add_halfspace(gp.HalfSpace.new([0, 0, 1.0], -0.025))
        v = st.flat_v()
        v[3] = 0.3
        return st, v
    st, v = standing()
    print("case interface biped, velocity stepping", flush=True)
    s = Stepper(oracle, [st.model_desc()], st.flat_q(), v, integrator=VS)
    for _ in range(60):
        s.step(1.0 / 60.0)
    print(f"worst per-step difference {s.worst:.3g}", flush=True)
    assert not s.b.faults().any()
    st, v = standing()
    print("case interface biped, CCD velocity stepping: the fault at step 1", flush=True)
    s = Stepper(oracle, [st.model_desc()], st.flat_q(), v, integrator=CCD)
    s.step(1.0 / 60.0)
    assert not s.b.faults().any()
    q1, v1 = s.q, s.v
    rq, rv, _ = oracle.vs_step(st.model_desc(), q1, v1, None, 1.0 / 60.0, True)
    assert not (np.isfinite(rq).all() and np.isfinite(rv).all())
    s.step(1.0 / 60.0, check=False)
    assert s.b.faults().all()


@pytest.mark.parametrize("integrator", [0, 2, 3])
def test_fixed_joint_arm_step_parity(gp, oracle, integrator):
    """An arm with a fixed joint between its two revolutes (the oracle's
    fixed-joint handling is pinned by the merged-body identity), 32 random
    instances, some touching the plane: every step against the oracle
    (velocity stepping at its 1e-5, the others at 1e-10)."""
    from gorilla_physics_amd import batch
    from test_oracle_golden import welded_arm
    print(f"case welded arm, integrator {integrator}", flush=True)
    d = welded_arm(gp).model_desc()
    rng = np.random.default_rng(8)
    n = 32
    q, v = rng.uniform(-2, 2, 2 * n), rng.normal(size=2 * n)
    if integrator == 3:
        s = Stepper(oracle, [d] * n, q, v, integrator=VS)
        for _ in range(50):
            s.step(1e-2)
        print(f"worst per-step difference {s.worst:.3g}", flush=True)
        assert not s.b.faults().any()
        return
    b = batch.Batch([batch.Model(d)], [n], precision=0)
    b.set_state(q, v)
    tau = np.zeros(2 * n)
    worst = 0.0
    for k in range(100):
        b.step(1e-3, integrator, 1)
        q1, v1 = b.get_state()
        a1 = b.get_vdot()
        eq, ev, ea = step_errors(oracle, [d] * n, q, v, tau, q1, v1, a1, 1e-3, integrator)
        worst = max(worst, eq.max(), ev.max(), ea.max())
        q, v = q1, v1
    print(f"worst per-step difference {worst:.3g}", flush=True)
    assert not b.faults().any() and worst <= F64_TOL


def test_fixed_joint_no_move_through_the_mirror(gp):  # joint/fixed.rs:45,65
    """A cube welded to the world has no coordinates; a second of velocity
    stepping (and of semi-implicit Euler) through the HIP path leaves it
    where the joint puts it and returns empty states."""
    from test_oracle_golden import fixed_cube
    print("case fixed cube", flush=True)
    for integ in (gp.Integrator.VelocityStepping, gp.Integrator.SemiImplicitEuler):
        st, iso = fixed_cube(gp)
        for _ in range(100):
            q, v = gp.step(st, 1e-2, [], integ)
        assert st.flat_q().size == 0 and st.flat_v().size == 0
        pose = st.poses()[0]
        assert np.allclose(pose.translation, iso.translation, atol=1e-12)
        assert np.allclose(pose.rotation.coords(), iso.rotation.coords(), atol=1e-12)


def test_ball_fall(gp, oracle):  # joint/floating.rs:133
    from gorilla_physics_amd import batch
    from test_oracle_golden import ball_fall, check_ball_fall
    print("case reference ball_fall", flush=True)

    def gpu_simulate(d, final_time, dt, q0, v0, tau, integrator):
        b = batch.Batch([batch.Model(d)], [1], precision=0)
        b.set_state(np.asarray(q0, float), np.asarray(v0, float))
        return b.simulate(final_time, dt, None, integrator)

    q0, v0, qs = ball_fall(gp, gpu_simulate)
    check_ball_fall(qs[-1])
    _, _, rq = ball_fall(gp, oracle.simulate)
    assert qs.shape == rq.shape and np.abs(qs - rq).max() < 1e-12


@pytest.mark.parametrize("integrator", [0, 2, 3])
def test_so101_arm_step_parity(gp, oracle, integrator):  # builders/mod.rs:252
    """64 SO-101 arms with random configurations, velocities and held
    torques, every step against the oracle (velocity stepping at its 1e-5,
    the others at 1e-10)."""
    from gorilla_physics_amd import batch, builders
    print(f"case SO-101 arms, integrator {integrator}", flush=True)
    d = builders.build_so101().model_desc()
    rng = np.random.default_rng(12)
    n = 64
    q, v, tau = rng.uniform(-2, 2, 6 * n), rng.normal(size=6 * n), rng.normal(size=6 * n) * 0.02
    if integrator == 3:
        s = Stepper(oracle, [d] * n, q, v, tau, integrator=VS)
        for _ in range(50):
            s.step(1e-3)
        print(f"worst per-step difference {s.worst:.3g}", flush=True)
        assert not s.b.faults().any()
        return
    b = batch.Batch([batch.Model(d)], [n], precision=0)
    b.set_state(q, v)
    b.set_tau(tau)
    worst = 0.0
    for k in range(100):
        b.step(1e-3, integrator, 1)
        q1, v1 = b.get_state()
        a1 = b.get_vdot()
        eq, ev, ea = step_errors(oracle, [d] * n, q, v, tau, q1, v1, a1, 1e-3, integrator)
        worst = max(worst, eq.max(), ev.max(), ea.max())
        q, v = q1, v1
    print(f"worst per-step difference {worst:.3g}", flush=True)
    assert not b.faults().any() and worst <= F64_TOL


def test_plain_kernel_steps_in_one_launch_are_bit_identical(gp):
    """The runtime-topology kernel with 7 steps per launch (the last launch
    partial) against one launch per step, 30 steps, bit for bit: SLIP
    hoppers (per-instance spring-contact state, SIE), SO-101 arms (RK4),
    welded arms (RK2)."""
    from gorilla_physics_amd import batch, builders
    from test_oracle_golden import welded_arm
    rng = np.random.default_rng(21)
    slip = builders.build_SLIP(0.54, 0.1, 0.2, 0.3, 2e3)
    slip.add_halfspace(gp.HalfSpace.new([0, 0, 1.0], 0.0))
    cases = [(slip.model_desc(), 0, np.tile([0, 0, 0, 1.0, 0, 0, 0.25], 16),
              np.concatenate([[0, 0, 0, u, 0, -1.0] for u in rng.uniform(0.5, 2.0, 16)])),
             (builders.build_so101().model_desc(), 2, rng.uniform(-2, 2, 6 * 16), rng.normal(size=6 * 16)),
             (welded_arm(gp).model_desc(), 1, rng.uniform(-2, 2, 2 * 16), rng.normal(size=2 * 16))]
    for d, integrator, q0, v0 in cases:
        print(f"case plain kernel steps per launch, integrator {integrator}, nq {d.nq}", flush=True)
        out = []
        for spl in (1, 7):
            b = batch.Batch([batch.Model(d)], [16], precision=0, kernel=1, steps_per_launch=spl)
            b.set_state(q0, v0)
            b.step(1e-3, integrator, 30)
            _, launches = b.last_timing()
            q, v = b.get_state()
            out.append((np.concatenate([q, v, b.get_vdot()]), b.faults(), launches))
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
        assert (out[0][2], out[1][2]) == (30, 5)
