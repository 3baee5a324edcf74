"""The reference's own contact and spring tests (contact.rs, dynamics.rs),
run through the HIP path: each model is a batch of one instance stepped by
gp_simulate with RK4 (the reference's float-accumulated loop), FP64. The
reference's assertions are applied to the GPU trajectory, and the trajectory
is compared with the oracle's: every tenth step of it against one oracle step
from the same state within 1e-10; within 1e-9 over the first 10 steps; over
the whole rollout within 1e-3 on positions (a tenth of the reference's own
1e-2 position tolerance) and 5e-2 on velocities, the final velocities within
1e-3. The regularised Coulomb friction (slip velocity 1e-3 m/s,
contact.rs:287) and the stiff contact springs amplify last-bit differences
between the two algorithms over thousands of steps, most during stick-slip
transitions; for the rimless wheel (12,000 steps of spoke impacts) and the
standing compass gait (sliding on the verge of sticking) the two rollouts
part by more than that, so those tests keep the per-step check and the
reference's own assertions on both trajectories."""
import math

import numpy as np
import pytest

from gpu_helpers import quat_fix, rel_err
from test_oracle_golden import (G, ball_with_point, check_compass_gait, check_cube_at_rest, check_rimless_wheel,
                                check_spring_forward_drop, compass_gait, rimless_wheel, rod_model, spring_forward_drop,
                                two_cubes)

pytestmark = pytest.mark.gpu

RK4 = 2


def _simulate(gp, st, final_time, dt, q0, v0, tau=None, integrator=RK4):
    from gorilla_physics_amd import batch
    m = batch.Model(st.model_desc())
    b = batch.Batch([m], [1], precision=0)
    b.set_state(np.asarray(q0, float), np.asarray(v0, float))
    control = None if tau is None else (lambda q, v, t: tau)
    return b.simulate(final_time, dt, control, integrator)


def _same_as_oracle(oracle, st, final_time, dt, q0, v0, qs, vs, tau=None, integrator=RK4, rollout=True):
    """Returns the oracle's rollout."""
    d = st.model_desc()
    rq, rv = oracle.simulate(d, final_time, dt, q0, v0, tau, integrator)
    assert qs.shape == rq.shape
    worst = 0.0
    for k in range(0, len(qs) - 1, 10):
        sq, sv, _ = oracle.step(d, qs[k], vs[k], tau, dt, integrator)
        worst = max(worst, rel_err(quat_fix(qs[k + 1], sq, d), sq), rel_err(vs[k + 1], sv))
    eq, ev = np.abs(qs - rq).max(axis=1), np.abs(vs - rv).max(axis=1)
    print(f"vs oracle: per step {worst:.3g}, first 10 steps {max(eq[:11].max(), ev[:11].max()):.3g}, "
          f"whole rollout q {eq.max():.3g} v {ev.max():.3g}", flush=True)
    assert worst <= 1e-10
    assert eq[:11].max() < 1e-9 and ev[:11].max() < 1e-9
    if rollout:
        assert eq.max() < 1e-3 and ev.max() < 5e-2 and ev[-1] < 1e-3
    return rq, rv


def test_cube_fall_ground(gp, oracle):  # contact.rs:408
    from gorilla_physics_amd import builders
    print("case reference cube_fall_ground", flush=True)
    st = builders.build_cube(3.0, 1.0)
    st.add_halfspace(gp.HalfSpace.new([0, 0, 1], -10.0))
    q0, v0 = [0, 0, 0, 1, 0, 0, 0.0], [1, 1, 1, 1, 1, 1.0]
    qs, vs = _simulate(gp, st, 5.0, 1e-3, q0, v0)
    assert abs(qs[-1, 6] - (-10.0 + 0.5)) < 1e-2
    _same_as_oracle(oracle, st, 5.0, 1e-3, q0, v0, qs, vs)


def test_cube_slide_ground(gp, oracle):  # contact.rs:444
    from gorilla_physics_amd import builders
    print("case reference cube_slide_ground", flush=True)
    st = builders.build_cube(3.0, 1.0)
    st.add_halfspace(gp.HalfSpace.new_with_params([0, 0, 1], -0.5, 1.0, 0.5))
    q0, v0 = [0, 0, 0, 1, 0, 0, 0.0], [0, 0, 0, 1, 0, 0.0]
    qs, vs = _simulate(gp, st, 2.0, 1e-3, q0, v0)
    assert abs(qs[-1, 6]) < 1e-2
    assert np.linalg.norm(vs[-1, 3:]) < 5e-3 and np.linalg.norm(vs[-1, :3]) < 1e-2
    t = 1.0 / (G * 0.5)
    assert abs(qs[-1, 4] - (t - 0.5 * G * 0.5 * t * t)) < 1e-2
    _same_as_oracle(oracle, st, 2.0, 1e-3, q0, v0, qs, vs)


def test_cube_hit_ground(gp, oracle):  # contact.rs:493
    from gorilla_physics_amd import builders
    print("case reference cube_hit_ground", flush=True)
    st = builders.build_cube(3.0, 1.0)
    st.add_halfspace(gp.HalfSpace.new([0, 0, 1], -10.0))
    q0, v0 = [0, 0, 0, 1, 0, 0, 0.0], [0, 5.0, 0, 1.0, 0, 0]
    qs, vs = _simulate(gp, st, 5.0, 1e-3, q0, v0)
    check_cube_at_rest(qs[-1], vs[-1])
    _same_as_oracle(oracle, st, 5.0, 1e-3, q0, v0, qs, vs)


def test_two_cubes_hit_ground(gp, oracle):  # contact.rs:533
    print("case reference two_cubes_hit_ground", flush=True)
    st = two_cubes(gp)
    q0 = [0, 0, 0, 1, 2.0, 0, 0, 0, 0, 0, 1, -2.0, 0, 0]
    v0 = [0, 5.0, 0, 1.0, 0, 0, 0, -5.0, 0, -1.0, 0, 0]
    qs, vs = _simulate(gp, st, 5.0, 1e-3, q0, v0)
    check_cube_at_rest(qs[-1, :7], vs[-1, :6])
    check_cube_at_rest(qs[-1, 7:], vs[-1, 6:])
    _same_as_oracle(oracle, st, 5.0, 1e-3, q0, v0, qs, vs)


def test_rimless_wheel(gp, oracle):  # contact.rs:679
    print("case reference rimless_wheel", flush=True)
    st = rimless_wheel(gp)
    q0, v0 = [0, 0, 0, 1, 0, 0, 0.0], [0, 0, 0, 1.0, 0, 0]
    qs, vs = _simulate(gp, st, 20.0, 1.0 / 600.0, q0, v0)
    check_rimless_wheel(vs)
    rq, rv = _same_as_oracle(oracle, st, 20.0, 1.0 / 600.0, q0, v0, qs, vs, rollout=False)
    check_rimless_wheel(rv)


def test_compass_gait_standing(gp, oracle):  # contact.rs:735
    print("case reference compass_gait_standing", flush=True)
    st, q0 = compass_gait(gp)
    qs, vs = _simulate(gp, st, 2.0, 1e-3, q0, np.zeros(8))
    print("final hip velocity", vs[-1, :6].tolist(), flush=True)
    check_compass_gait(qs[-1], vs[-1])
    rq, rv = _same_as_oracle(oracle, st, 2.0, 1e-3, q0, np.zeros(8), qs, vs, rollout=False)
    check_compass_gait(rq[-1], rv[-1])


def test_spring_forward_drop(gp, oracle):  # contact.rs:991
    print("case reference spring_forward_drop", flush=True)
    st, q0, v0 = spring_forward_drop(gp)
    qs, vs = _simulate(gp, st, 1.5, 1e-3, q0, v0)
    k = check_spring_forward_drop(qs, vs)
    rq, rv = _same_as_oracle(oracle, st, 1.5, 1e-3, q0, v0, qs, vs)
    assert check_spring_forward_drop(rq, rv) == k


def test_mass_hit_ground(gp, oracle):  # contact.rs:607
    print("case reference mass_hit_ground", flush=True)
    st = ball_with_point(gp, m=1.0, k=50e3, alpha=1.0, mu=0.5)
    st.halfspaces[0] = gp.HalfSpace.new_with_params([0, 0, 1], -0.3, 1.0, 0.5)
    q0, v0 = [0, 0, 0, 1, 0, 0, 0.0], [0, 0, 0, 2.0, 0, 0]
    qs, vs = _simulate(gp, st, 2.0, 1e-3, q0, v0)
    assert abs(qs[-1, 6] + 0.3) < 1e-2
    assert np.linalg.norm(vs[-1, 3:]) < 1e-2 and np.linalg.norm(vs[-1, :3]) < 1e-3
    th = math.sqrt(0.3 * 2 / G)
    vx = 2.0 - G * th * 0.5
    ts = vx / (G * 0.5)
    assert abs(qs[-1, 4] - (th * 2.0 + vx * ts - 0.5 * G * 0.5 * ts * ts)) < 1e-2
    _same_as_oracle(oracle, st, 2.0, 1e-3, q0, v0, qs, vs)


def test_pendulum_hit_ground(gp, oracle):  # contact.rs:371
    print("case reference pendulum_hit_ground", flush=True)
    m, l = 1.5, 10.0
    st = rod_model(gp, m, l, gp.Isometry3.identity(), [0, 1, 0], cross=[m * l, 0, 0],
                   moment=np.diag([0, m * l * l, m * l * l]))
    st.add_contact_point(gp.ContactPoint.new("rod", [l, 0, 0]))
    st.add_halfspace(gp.HalfSpace.new([0, 0, 1], -5.0))
    qs, vs = _simulate(gp, st, 5.0, 1e-2, [0.0], [0.0], tau=[0.0])
    assert abs(qs[-1, 0] - math.radians(30)) < 1e-3
    _same_as_oracle(oracle, st, 5.0, 1e-2, [0.0], [0.0], qs, vs, tau=[0.0])


def test_spring_on_frictionless_ground(gp, oracle):  # dynamics.rs:1133
    print("case reference spring on frictionless ground", flush=True)
    A = gp.RigidBody.new_sphere(1.0, 0.1, "A")
    B = gp.RigidBody.new_sphere(1.0, 0.1, "B")
    st = gp.MechanismState([gp.FloatingJoint(gp.Transform3D.identity("A", gp.WORLD_FRAME)),
                            gp.PrismaticJoint.new_with_spring(gp.Transform3D.identity("B", "A"), [1, 0, 0],
                                                              gp.JointSpring(50.0, 2.0 / 3.0))], [A, B])
    st.add_halfspace(gp.HalfSpace.new_with_params([0, 0, 1], 0.0, 1.0, 0.0))
    q0, v0 = [0, 0, 0, 1, 0, 0, 0, 2.0], np.zeros(7)
    qs, vs = _simulate(gp, st, 1.1, 1e-3, q0, v0)
    xA = qs[-1, 4]
    xB = xA + qs[-1, 7]
    assert abs((xA + xB) / 2 - 1.0) < 1e-5 and xA > 0 and xB < 2.0
    _same_as_oracle(oracle, st, 1.1, 1e-3, q0, v0, qs, vs)


def test_spring_drop(gp, oracle):  # contact.rs:908
    """A floating sphere with a sprung prismatic leg (k = 100, rest 0) dropped
    onto the ground (alpha 1, mu 0): the foot settles at the ground within twice
    the static penetration and the body stays upright."""
    print("case reference spring_drop", flush=True)
    m_body = m_foot = 1.0
    l_leg, h_ground = 1.0, -2.0
    body = gp.RigidBody.new_sphere(m_body, 1.0, "body")
    foot = gp.RigidBody.new_sphere(m_foot, 1.0, "foot")
    st = gp.MechanismState([gp.FloatingJoint(gp.Transform3D.identity("body", gp.WORLD_FRAME)),
                            gp.PrismaticJoint.new_with_spring(
                                gp.Transform3D("foot", "body", gp.Isometry3.translation_xyz(0.0, 0.0, -l_leg)),
                                [0, 0, -1], gp.JointSpring(100.0, 0.0))], [body, foot])
    st.add_contact_point(gp.ContactPoint.new("foot", [0, 0, 0]))
    st.add_halfspace(gp.HalfSpace.new_with_params([0, 0, 1], h_ground, 1.0, 0.0))
    q0, v0 = [0, 0, 0, 1, 0, 0, 0, 0.0], np.zeros(7)
    qs, vs = _simulate(gp, st, 4.0, 1e-3, q0, v0)
    # foot position: body translation + R_body (0, 0, -l_leg - q_spring)
    zb, qspring = qs[-1, 6], qs[-1, 7]
    z_foot = zb - l_leg - qspring
    z_tol = ((m_body + m_foot) * G / 50e3) ** (2.0 / 3.0)
    assert abs(z_foot - h_ground) < 2 * z_tol
    assert np.allclose(qs[-1, :3], 0.0, atol=1e-12) and abs(qs[-1, 3] - 1.0) < 1e-12  # upright
    _same_as_oracle(oracle, st, 4.0, 1e-3, q0, v0, qs, vs)


def test_SLIP_hopping(gp, oracle):  # contact.rs:836
    """The reference's SLIP test through gp_step (semi-implicit Euler, one step
    per call, FP64), four instances at different forward speeds in one batch.
    At every apex the leg is swung back to 45 degrees forward with
    gp_batch_set_spring_contact, as the reference's test does. Every step is
    checked against the oracle stepped from the same state (q, v and the spring
    contact's registration), within 1e-10; the reference's assertions are
    applied to the GPU trajectory of the 5 m/s instance. At 3 m/s the hopper
    falls back onto a flat leg and the spring would pull the body into the
    ground, where the reference panics (contact.rs:161-163): that instance's
    fault flag is raised at the very step where the oracle reports it."""
    from gorilla_physics_amd import batch, builders
    print("case reference SLIP_hopping", flush=True)
    m, r, l_rest, angle, k = 0.54, 0.1, 0.2, math.radians(45.0), 500.0
    direction = np.array([math.sin(angle), 0.0, -math.cos(angle)])
    st = builders.build_SLIP(m, r, l_rest, angle, k)
    st.add_halfspace(gp.HalfSpace.new([0, 0, 1], -0.3))
    d = st.model_desc()
    speeds = [5.0, 4.5, 5.5, 3.0]
    n, dt, steps = len(speeds), 1.0 / 2000.0, int(3.0 / (1.0 / 2000.0))
    b = batch.Batch([batch.Model(d)], [n], precision=0)
    q = np.tile([0, 0, 0, 1.0, 0, 0, 0], n)
    v = np.concatenate([[0, 0, 0, s, 0, 0] for s in speeds])
    b.set_state(q, v)
    e0 = sum(oracle.energy(d, q[:7], v[:6]))
    ostate = [oracle.spring_state(d) for _ in range(n)]
    live = [True] * n
    vz_prev = np.zeros(n)
    energies, hs, worst, fault_step = [], [], 0.0, None
    for step in range(steps):
        b.step(dt, 0, 1)
        qn, vn = b.get_state()
        faults = b.faults()
        for i in range(n):
            if not live[i]:
                continue
            try:
                rq, rv, _ = oracle.step_spring(d, q[7 * i:7 * i + 7], v[6 * i:6 * i + 6], ostate[i], None, dt)
            except RuntimeError as e:
                assert "pulls into the half-space" in str(e) and faults[i] == 1, i
                live[i], fault_step = False, step
                continue
            assert faults[i] == 0, (i, step)
            worst = max(worst, np.abs(qn[7 * i:7 * i + 7] - rq).max(), np.abs(vn[6 * i:6 * i + 6] - rv).max())
            reg, h, p, dr, lr = b.get_spring_contact(i, 0, with_halfspace=True)
            assert reg == bool(ostate[i][0]) and h == (0 if reg else -1)
            if reg:
                assert np.abs(p - ostate[i][1:4]).max() < 1e-10
                ostate[i][1:4] = p  # the oracle continues from the GPU's state
            else:
                assert np.abs(dr - ostate[i][5:8]).max() < 1e-10 and lr == ostate[i][8]
                ostate[i][5:8] = dr
            vz = vn[6 * i + 5]
            if vz_prev[i] > 0.0 and vz <= 0.0:
                if i == 0:
                    energies.append(sum(oracle.energy(d, qn[:7], vn[:6])))
                    hs.append(qn[6])
                b.set_spring_contact(i, 0, direction, l_rest)
                ostate[i][5:8] = direction
                ostate[i][8] = l_rest
            vz_prev[i] = vz
        q, v = qn, vn
    print(f"worst per-step difference {worst:.3g}, {len(hs)} apexes, fault at step {fault_step}", flush=True)
    assert worst < 1e-10
    assert live == [True, True, True, False]
    assert len(energies) > 2
    for e in energies:
        assert abs(e - e0) < 1e-2
    assert abs(hs[-3] - hs[-2]) < 1e-3 and abs(hs[-2] - hs[-1]) < 1e-3


def test_spring_contacts_reject_runge_kutta(gp):
    """Spring contacts change their state inside the dynamics evaluation (the
    reference notes it does not work with Runge-Kutta, contact.rs:131-132): a
    Runge-Kutta step of a batch that holds them is refused."""
    from gorilla_physics_amd import batch, builders
    st = builders.build_SLIP(0.54, 0.1, 0.2, math.radians(45.0), 500.0)
    st.add_halfspace(gp.HalfSpace.new([0, 0, 1], -0.3))
    b = batch.Batch([batch.Model(st.model_desc())], [1], precision=0)
    b.set_state(np.array([0, 0, 0, 1.0, 0, 0, 0]), np.zeros(6))
    for integrator in (1, RK4):
        with pytest.raises(gp.abi.GPError) as e:
            b.step(1e-3, integrator, 1)
        assert e.value.code == gp.abi.GP_ERR_UNSUPPORTED
    b.step(1e-3, 0, 1)


def test_SLIP_hopping_through_the_mechanism_mirror(gp, oracle):  # contact.rs:836
    """The same test written as the reference writes it, against the Python
    mirror of MechanismState / step: the leg is reset by assigning
    bodies[0].spring_contacts[0].direction and .l_rest at every apex, and the
    registration is visible on the SpringContact between steps."""
    from gorilla_physics_amd import builders
    print("case reference SLIP_hopping (mechanism mirror)", flush=True)
    m, r, l_rest, angle, k = 0.54, 0.1, 0.2, math.radians(45.0), 500.0
    direction = np.array([math.sin(angle), 0.0, -math.cos(angle)])
    state = builders.build_SLIP(m, r, l_rest, angle, k)
    state.add_halfspace(gp.HalfSpace.new([0, 0, 1], -0.3))
    state.update([gp.Pose(gp.UnitQuaternion.identity(), [0.0, 0.0, 0.0])],
                 [gp.SpatialVector([0.0, 0.0, 0.0], [5.0, 0.0, 0.0])])
    def energy():  # as the reference's test computes it (contact.rs:882)
        return state.kinetic_energy() + state.gravitational_energy()

    e0 = energy()
    dt = 1.0 / 2000.0
    v_z_prev, energies, hs, stance = 0.0, [], [], 0
    for _ in range(int(3.0 / dt)):
        q, v = gp.step(state, dt, [], gp.Integrator.SemiImplicitEuler)
        sc = state.bodies[0].spring_contacts[0]
        if sc.registered_contact is not None:
            stance += 1
            assert sc.registered_halfspace is state.halfspaces[0]
        v_z = v[0].linear[2]
        if v_z_prev > 0.0 and v_z <= 0.0:
            energies.append(energy())
            hs.append(q[0].translation[2])
            state.bodies[0].spring_contacts[0].direction = direction
            state.bodies[0].spring_contacts[0].l_rest = l_rest
        v_z_prev = v_z
    assert len(energies) > 2 and stance > 0
    for e in energies:
        assert abs(e - e0) < 1e-2
    assert abs(hs[-3] - hs[-2]) < 1e-3 and abs(hs[-2] - hs[-1]) < 1e-3
