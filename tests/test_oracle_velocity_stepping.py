"""The oracle's velocity-stepping path (integrators.rs:42-175,
dynamics.rs:366-742) pinned on the CPU: the contact impulse program by its
optimality conditions and by projected gradient on small cases, the step by
closed forms, and the reference's own velocity-stepping and continuous
collision tests (sphere.rs, integrators.rs, collision/ccd/mod.rs)."""
import math

import numpy as np
import pytest

MU = 0.95
G_ACC = 9.81


def _skew(p):
    return np.array([[0, -p[2], p[1]], [p[2], 0, -p[0]], [-p[1], p[0], 0]])


def _frame(n):
    t = np.cross(n, [1.0, 0, 0])
    if np.linalg.norm(t) == 0:
        t = np.cross(n, [0, 1.0, 0])
    t /= np.linalg.norm(t)
    b = np.cross(n, t)
    return np.array([n, t, b / np.linalg.norm(b)])


def _body_problem(rng, n_contacts):
    """G, g of one free body (world twist coordinates) touching a tilted plane
    at random points: the shape of every single-body contact step."""
    m = rng.uniform(0.5, 3.0)
    c = rng.normal(size=3) * 0.1
    A = rng.normal(size=(3, 3))
    Jc = A @ A.T * 0.1 + 0.05 * np.eye(3)
    M = np.zeros((6, 6))
    M[:3, :3] = Jc + m * (c @ c * np.eye(3) - np.outer(c, c))
    M[:3, 3:] = m * _skew(c)
    M[3:, :3] = -m * _skew(c)
    M[3:, 3:] = m * np.eye(3)
    ang = rng.uniform(-0.3, 0.3)
    n = np.array([math.sin(ang), 0.0, math.cos(ang)])
    J = np.vstack([_frame(n) @ np.hstack([-_skew(c + rng.normal(size=3) * 0.5), np.eye(3)])
                   for _ in range(n_contacts)])
    vfree = rng.normal(size=6)
    Minv = np.linalg.inv(M)
    return J @ Minv @ J.T, J @ vfree, M, J, vfree


def _cone_residuals(G, g, x):
    """(primal infeasibility, dual infeasibility, complementarity) of x for
    min 1/2 x'Gx + g'x, (mu x_n, x_t) in the cone: s = D x, z = D^-1 (G x + g)."""
    d = np.tile([MU, 1.0, 1.0], len(g) // 3)
    s, z = d * x, (G @ x + g) / d
    sc, zc = s.reshape(-1, 3), z.reshape(-1, 3)
    pf = max(0.0, (np.linalg.norm(sc[:, 1:], axis=1) - sc[:, 0]).max())
    df = max(0.0, (np.linalg.norm(zc[:, 1:], axis=1) - zc[:, 0]).max())
    return pf, df, abs(s @ z)


def test_cone_program_optimality_conditions(oracle):
    """Every returned impulse is optimal by the conic KKT conditions: D x in
    the cone, D^-1 (G x + g) in the (self-dual) cone, orthogonal."""
    rng = np.random.default_rng(0)
    worst = [0.0, 0.0, 0.0]
    for _ in range(200):
        G, g, *_ = _body_problem(rng, int(rng.integers(1, 9)))
        x, it = oracle.cone_program(G, g)
        r = _cone_residuals(G, g, x)
        scale = (1 + np.abs(g).max()) * (1 + np.abs(x).max())
        worst = [max(worst[0], r[0]), max(worst[1], r[1] / (1 + np.abs(g).max())), max(worst[2], r[2] / scale)]
        assert it < 50
    print("worst primal / dual / gap", worst)
    assert worst[0] < 1e-12 and worst[1] < 1e-9 and worst[2] < 1e-9


def _project_cone(x):
    """Euclidean projection onto {(mu x_n, x_t) in SOC} per contact, via the
    scaled variable y = D x (closed form of the SOC projection)."""
    out = np.empty_like(x)
    for c in range(len(x) // 3):
        y = np.array([MU * x[3 * c], x[3 * c + 1], x[3 * c + 2]])
        t, u = y[0], y[1:]
        nu = np.linalg.norm(u)
        if nu <= t:
            p = y
        elif nu <= -t:
            p = np.zeros(3)
        else:
            a = (t + nu) / 2
            p = np.array([a, *(a * u / nu)])
        out[3 * c:3 * c + 3] = [p[0] / MU, p[1], p[2]]
    return out


def test_cone_program_matches_projected_gradient(oracle):
    """One and two contacts with a positive definite G: the minimiser is
    unique. It is a fixed point of the projected-gradient map (optimality by
    Euclidean projection onto the cone, independent of the KKT form above),
    and projected gradient from zero (a different algorithm) approaches it."""
    rng = np.random.default_rng(1)
    for trial in range(40):
        G, g, *_ = _body_problem(rng, 1 + trial % 2)
        if np.linalg.eigvalsh(G).min() < 1e-3:
            continue
        x, _ = oracle.cone_program(G, g)
        D = np.diag(np.tile([MU, 1.0, 1.0], len(g) // 3))
        Di = np.linalg.inv(D)
        H, h = Di @ G @ Di, Di @ g  # objective in y = D x
        step = 1.0 / np.linalg.eigvalsh(H).max()

        def pg(y):
            return D @ _project_cone(Di @ (y - step * (H @ y + h)))
        y = D @ x
        assert np.abs(pg(y) - y).max() < 1e-6 * (1 + np.abs(y).max()), trial
        y = np.zeros(len(g))
        for _ in range(5000):
            y = pg(y)
        assert np.abs(Di @ y - x).max() < 1e-6 * (1 + np.abs(x).max()), trial


def _particle(gp, mass=1.0):
    """A floating body whose single contact point is its centre of mass: the
    contact impulse does not turn it, so the step has a closed form."""
    body = gp.RigidBody(gp.SpatialInertia("p", np.eye(3) * 0.1, np.zeros(3), mass))
    st = gp.MechanismState([gp.FloatingJoint(gp.Transform3D.identity("p", gp.WORLD_FRAME))], [body])
    st.add_contact_point(gp.ContactPoint.new("p", [0, 0, 0]))
    st.add_halfspace(gp.HalfSpace.new([0, 0, 1.0], 0.0))
    return st


@pytest.mark.parametrize("speed", [0.0, 0.004, 1.0, 3.0])
def test_particle_on_ground_closed_form(gp, oracle, speed):
    """Resting on the plane with horizontal speed u: while u > mu g dt it
    slides, and the relaxed cone of dynamics.rs:668-704 gives slip
    s = (u - mu g dt) / (1 + mu^2) with separation velocity mu s; below that it
    sticks (velocity zero)."""
    st = _particle(gp)
    d = st.model_desc()
    dt = 1e-3
    q = np.array([0, 0, 0, 1.0, 0, 0, 0])
    v = np.array([0, 0, 0, speed * 0.6, speed * 0.8, 0])
    q1, v1, _ = oracle.vs_step(d, q, v, None, dt)
    if speed > MU * G_ACC * dt:
        s = (speed - MU * G_ACC * dt) / (1 + MU * MU)
        assert abs(np.linalg.norm(v1[3:5]) - s) < 1e-10
        assert abs(v1[5] - MU * s) < 1e-10
        assert abs(v1[3] * 0.8 - v1[4] * 0.6) < 1e-10  # slips along its velocity
    else:
        assert np.abs(v1).max() < 1e-10
    assert np.abs(v1[:3]).max() == 0.0
    assert np.allclose(q1[4:], q[4:] + v1[3:] * dt, atol=1e-15)


def test_free_flight_is_semi_implicit_euler(gp, oracle):
    """Without a touching contact the velocity-stepping step is the
    semi-implicit Euler step of the contact-free dynamics."""
    from gorilla_physics_amd import builders
    st = builders.build_cube(1.0, 1.0)
    st.add_halfspace(gp.HalfSpace.new([0, 0, 1.0], -10.0))
    d = st.model_desc()
    rng = np.random.default_rng(3)
    for _ in range(5):
        q = np.array([*(lambda u: u / np.linalg.norm(u))(rng.normal(size=4)), *rng.normal(size=3)])
        v = rng.normal(size=6)
        q1, v1, _ = oracle.vs_step(d, q, v, None, 1e-3)
        q2, v2, _ = oracle.step(d, q, v, None, 1e-3, 0)
        assert np.abs(q1 - q2).max() < 1e-13 and np.abs(v1 - v2).max() < 1e-12


# ------------------------------------------------------------------ the reference's tests
def _run(oracle, d, q, v, steps, dt, tau=None, ccd=False, each=None):
    for _ in range(steps):
        q, v, _ = oracle.vs_step(d, q, v, tau, dt, ccd)
        if each is not None:
            each(q, v)
    return q, v


def test_sphere_hit_ground(gp, oracle):  # collision/sphere.rs:58
    from gorilla_physics_amd import builders
    st = builders.build_sphere(1.0, 1.0)
    st.add_halfspace(gp.HalfSpace.new([0, 0, 1.0], 0.0))
    q, _ = _run(oracle, st.model_desc(), np.array([0, 0, 0, 1.0, 0, 0, 1.5]), np.zeros(6), 2000, 1e-3)
    assert np.abs(q[4:] - [0, 0, 1]).max() < 1e-3


def _euler(qxyzw):
    R = __import__("gorilla_physics_amd").UnitQuaternion(*qxyzw).to_matrix()
    return math.atan2(R[2, 1], R[2, 2]), math.asin(-R[2, 0]), math.atan2(R[1, 0], R[0, 0])


def test_sphere_roll_forward(gp, oracle):  # collision/sphere.rs:96
    from gorilla_physics_amd import builders
    r = 0.5
    st = builders.build_sphere(1.0, r)
    st.add_halfspace(gp.HalfSpace.new([0, 0, 1.0], 0.0))
    q, _ = _run(oracle, st.model_desc(), np.array([0, 0, 0, 1.0, 0, 0, r]), np.zeros(6), 2000, 1e-3,
                tau=[0, 0.1, 0, 0, 0, 0])
    roll, pitch, yaw = _euler(q[:4])
    assert roll == 0.0 and yaw == 0.0 and q[5] == 0.0
    assert abs(q[6] - r) < 1e-5
    assert abs(r * pitch - q[4]) < 1e-4


def test_velocity_stepping_shoot_box(gp, oracle):  # integrators.rs:340
    from gorilla_physics_amd import builders
    l = 1.0
    st = builders.build_cube(1.0, l)
    rot = gp.UnitQuaternion.from_axis_angle([0, 1.0, 0], math.pi / 2)
    inv = gp.UnitQuaternion(-rot.x, -rot.y, -rot.z, rot.w)
    angle = math.radians(-10.0)
    normal = np.array([math.sin(angle), 0.0, math.cos(angle)])
    st.add_halfspace(gp.HalfSpace.new(normal, 0.0))
    q = np.array([*rot.coords(), 0, 0, 1.0])
    v = np.array([0, 0, 0, *inv * np.array([1.0, 0, 0])])
    q, v = _run(oracle, st.model_desc(), q, v, int(2.0 / (1.0 / 600.0)), 1.0 / 600.0)
    assert abs(q[4:] @ normal - l / 2) < 1e-2 and abs(q[5]) < 1e-5
    assert np.abs(v).max() < 1e-5


def test_velocity_stepping_drop_quadruped(gp, oracle):  # integrators.rs:379
    from gorilla_physics_amd import builders
    st = builders.build_quadruped()
    st.add_halfspace(gp.HalfSpace.new([0, 0, 1.0], 0.0))
    d = st.model_desc()
    q = np.zeros(d.nq)
    q[3], q[6] = 1.0, 5.0
    q, v = _run(oracle, d, q, np.zeros(d.nv), int(3.0 / 1e-2), 1e-2)
    assert q[6] > -1e-3
    assert np.abs(v[:6]).max() < 1e-3


def _sphere_penetration(q, r, normal):
    return max(0.0, -(q[4:7] @ normal - r))


@pytest.mark.parametrize("tilt", [0.0, 10.0])
def test_ccd_sphere_ground(gp, oracle, tilt):  # collision/ccd/mod.rs:38,90
    from gorilla_physics_amd import builders
    r = 0.1
    st = builders.build_sphere(1.0, r)
    a = math.radians(tilt)
    normal = np.array([math.sin(a), 0.0, math.cos(a)])
    st.add_halfspace(gp.HalfSpace.new(normal, 0.0))
    worst = [0.0]
    q, v = _run(oracle, st.model_desc(), np.array([0, 0, 0, 1.0, 0, 0, 10 * r]), np.array([0, 0, 0, 0, 0, -5.0]),
                int((1.0 if tilt == 0 else 0.5) / (1.0 / 60.0)), 1.0 / 60.0, ccd=True,
                each=lambda q, v: worst.__setitem__(0, max(worst[0], _sphere_penetration(q, r, normal))))
    assert worst[0] < 1e-8
    if tilt == 0:
        assert abs(q[6] - r) < 1e-8
    else:
        R = gp.UnitQuaternion(*q[:4]).to_matrix()
        vel = R @ v[3:]  # centre velocity (body origin at the centre)
        assert vel[0] > 1.0 and abs(vel[1]) < 1e-5


def test_ccd_three_sphere_ground(gp, oracle):  # collision/ccd/mod.rs:152
    from gorilla_physics_amd import builders
    r = 0.1
    st = builders.build_n_spheres(1.0, r, 3)
    st.add_halfspace(gp.HalfSpace.new([0, 0, 1.0], 0.0))
    ys = [0.0, 1.0, 2.0]
    q = np.concatenate([[0, 0, 0, 1.0, 0, y, 10 * r + 1e-8 * k] for k, y in enumerate(ys)])
    v = np.tile([0, 0, 0, 0, 0, -5.0], 3)
    worst = [0.0]

    def pen(q, v):
        for k in range(3):
            worst[0] = max(worst[0], _sphere_penetration(q[7 * k:7 * k + 7], r, np.array([0, 0, 1.0])))

    q, v = _run(oracle, st.model_desc(), q, v, int(1.0 / (1.0 / 60.0)), 1.0 / 60.0, ccd=True, each=pen)
    for k, y in enumerate(ys):
        assert abs(q[7 * k + 6] - r) < 1e-8
        assert q[7 * k + 4] == 0.0 and q[7 * k + 5] == y
    assert worst[0] < 1e-8


def test_ccd_cube_ground_stationary(gp, oracle):  # collision/ccd/mod.rs:229
    """Ten of the reference's hundred random resting cubes (its choices of
    position and of roll / pitch / yaw among multiples of pi/2)."""
    from gorilla_physics_amd import builders
    l = 0.1
    st = builders.build_cube(1.0, l)
    st.add_halfspace(gp.HalfSpace.new([0, 0, 1.0], 0.0))
    d = st.model_desc()
    rng = np.random.default_rng(7)
    for _ in range(10):
        x, y = rng.uniform(-100, 100, 2)
        roll, pitch, yaw = (rng.choice([-1.0, -0.5, 0.0, 0.5, 1.0]) * math.pi for _ in range(3))
        rot = gp.UnitQuaternion.from_euler_angles(roll, pitch, yaw)
        q, _ = _run(oracle, d, np.array([*rot.coords(), x, y, l / 2]), np.zeros(6), int(1.0 / (1 / 60.0)),
                    1 / 60.0, ccd=True)
        assert abs(q[6] - l / 2) < 1e-6 and abs(q[4] - x) < 1e-6 and abs(q[5] - y) < 1e-6


def test_ccd_cube_ground_dynamic(gp, oracle):  # collision/ccd/mod.rs:284
    """Ten of the reference's random thrown cubes: each ends resting on a face."""
    from gorilla_physics_amd import builders
    l = 0.1
    st = builders.build_cube(1.0, l)
    st.add_halfspace(gp.HalfSpace.new([0, 0, 1.0], 0.0))
    d = st.model_desc()
    rng = np.random.default_rng(8)
    for _ in range(10):
        x, y = rng.uniform(-100, 100, 2)
        z = rng.uniform(l, 2 * l)
        axis = rng.normal(size=3)
        rot = gp.UnitQuaternion.from_axis_angle(axis / np.linalg.norm(axis), rng.uniform(0, 2 * math.pi))
        lin = rng.uniform(-1, 1, 3)
        ang = rng.uniform(-5, 5, 3)
        R = rot.to_matrix()
        v = np.array([*(R.T @ ang), *(R.T @ lin)])  # world vectors into the body frame
        q, _ = _run(oracle, d, np.array([*rot.coords(), x, y, z]), v, int(1.0 / (1 / 60.0 / 2.0)), 1 / 60.0 / 2.0,
                    ccd=True)
        assert abs(q[6] - l / 2) < 1e-6


def test_rollout_runs_velocity_stepping(gp, oracle):
    """oracle_rollout with the velocity-stepping integrators is the per-step
    loop, instance by instance (the CPU baseline of the thrown-cube workload)."""
    from gorilla_physics_amd import workloads
    w = workloads.thrown_cubes(n=6)
    d = w["models"][0]
    q, v, fails = oracle.rollout([d] * 6, w["q"], w["v"], w["tau"], w["dt"], 30, integrator=4, threads=3)
    assert fails == 0
    for i in range(6):
        qi, vi = w["q"][7 * i:7 * i + 7], w["v"][6 * i:6 * i + 6]
        for _ in range(30):
            qi, vi, _ = oracle.vs_step(d, qi, vi, None, w["dt"], True)
        assert np.array_equal(qi, q[7 * i:7 * i + 7]) and np.array_equal(vi, v[6 * i:6 * i + 6])
