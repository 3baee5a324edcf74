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

WORKLOAD_IDS = {"acrobot": 1, "arms": 2, "mixed": 3, "cubes": 4, "thrown_cubes": 5, "so101": 6, "trees": 7}


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
    torques N(0, 0.02) N m, RK4 at dt 1e-3 for 1000 steps. Its welded base
    and six-joint chain run on the welded-base compiled kernels."""
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



# ----------------------------------------------------------------- trees of 16+ bodies
def lowest_contact_z(d, q):
    """World height of the lowest contact point of a model description at
    configuration q (forward kinematics of the joint chain, mechanism.rs:153-170)."""
    world, qo = [], 0
    for i in range(d.n_bodies):
        b = d._bodies[i]
        iso = Isometry3(UnitQuaternion(*b.x_tree_quat), np.array(b.x_tree_trans))
        axis = np.array(b.axis)
        if b.jtype == abi.GP_JOINT_REVOLUTE:
            iso = iso * Isometry3(UnitQuaternion.from_axis_angle(axis, q[qo]))
            qo += 1
        elif b.jtype == abi.GP_JOINT_PRISMATIC:
            iso = iso * Isometry3(None, axis * q[qo])
            qo += 1
        elif b.jtype == abi.GP_JOINT_FLOATING:
            iso = iso * Isometry3(UnitQuaternion(*q[qo:qo + 4]), np.array(q[qo + 4:qo + 7]))
            qo += 7
        world.append(iso if b.parent < 0 else world[b.parent] * iso)
    zs = [(world[c.body].rotation.to_matrix() @ np.array(c.p) + world[c.body].translation)[2]
          for c in (d._contacts[k] for k in range(d.desc.n_contacts))]
    return min(zs) if zs else 0.0


def hexapod_desc(p, ground=True):
    """Cuboid base on a floating joint and six legs of [hip-z, hip-y, knee-y]
    revolute joints (19 bodies, nq = 25, nv = 24), one contact point at each
    shank's tip, over the half-space z = 0."""
    bw, bd, bh, bm = p["bw"], p["bd"], p["bh"], p["bm"]
    bodies = [body_desc(-1, abi.GP_JOINT_FLOATING, [0, 0, 0], Isometry3.identity(),
                        SpatialInertia.cuboid_at([0, 0, 0], bm, bw, bd, bh, "base"))]
    contacts = []
    for leg in range(6):
        sx, sy = (-1, 0, 1)[leg % 3], (-1, 1)[leg // 3]
        lt, ls = p["lt"][leg], p["ls"][leg]
        hz = len(bodies)
        bodies.append(body_desc(0, abi.GP_JOINT_REVOLUTE, [0, 0, 1],
                                Isometry3.translation_xyz(sx * bw / 2, sy * bd / 2, 0.0),
                                SpatialInertia.cuboid_at([0, 0, 0], 0.1, 0.04, 0.04, 0.04, "hz")))
        hy = len(bodies)
        bodies.append(body_desc(hz, abi.GP_JOINT_REVOLUTE, [0, 1, 0], Isometry3.identity(),
                                SpatialInertia.cuboid_at([0, sy * lt / 2, 0], p["mt"][leg], 0.03, lt, 0.03, "th")))
        kn = len(bodies)
        bodies.append(body_desc(hy, abi.GP_JOINT_REVOLUTE, [0, 1, 0], Isometry3.translation_xyz(0.0, sy * lt, 0.0),
                                SpatialInertia.cuboid_at([0, 0, -ls / 2], p["ms"][leg], 0.03, 0.03, ls, "sh")))
        contacts.append(contact_desc(kn, [0.0, 0.0, -ls], 20e3))
    hs = [halfspace_desc(HalfSpace.new_with_params([0.0, 0.0, 1.0], 0.0, 0.9, 0.5))] if ground else []
    return ModelDesc(bodies, contacts if ground else [], hs)


def humanoid_desc(p):
    """The reference's biped (builders/biped_builder.rs:12: floating base, two
    legs of pelvis, hip, thigh, calf, ankle and foot, the feet's eight
    corners as contact points) with two arms of [shoulder-x, shoulder-y,
    elbow-y] revolute joints on the base: 19 bodies, over z = 0."""
    from . import builders
    st = builders.build_biped()
    st.add_halfspace(HalfSpace.new_with_params([0.0, 0.0, 1.0], 0.0, 0.9, 0.5))
    d = st.model_desc()
    bodies = [d._bodies[i] for i in range(d.n_bodies)]
    for side in (-1.0, 1.0):
        sx = len(bodies)
        bodies.append(body_desc(0, abi.GP_JOINT_REVOLUTE, [1, 0, 0],
                                Isometry3.translation_xyz(side * 0.06, 0.0, 0.08),
                                SpatialInertia.cuboid_at([0, 0, 0], 0.05, 0.03, 0.03, 0.03, "sx")))
        sy = len(bodies)
        bodies.append(body_desc(sx, abi.GP_JOINT_REVOLUTE, [0, 1, 0], Isometry3.identity(),
                                SpatialInertia.cuboid_at([0, 0, -p["la"] / 2], p["ma"], 0.03, 0.03, p["la"], "ua")))
        bodies.append(body_desc(sy, abi.GP_JOINT_REVOLUTE, [0, 1, 0], Isometry3.translation_xyz(0.0, 0.0, -p["la"]),
                                SpatialInertia.cuboid_at([0, 0, -p["lf"] / 2], p["mf"], 0.03, 0.03, p["lf"], "fa")))
    contacts = [d._contacts[i] for i in range(d.desc.n_contacts)]
    hs = [d._halfspaces[i] for i in range(d.desc.n_halfspaces)]
    return ModelDesc(bodies, contacts, hs)


def random_tree_shape(rng, nb):
    """Topology of a random tree of nb bodies: each body after the first hung
    from one of the bodies before it; joints revolute (70 %), prismatic
    (15 %) or fixed (15 %); a contact point on every fourth body."""
    parents = [-1] + [int(rng.integers(0, i)) for i in range(1, nb)]
    u = rng.random(nb)
    jtypes = [abi.GP_JOINT_FLOATING] + [abi.GP_JOINT_REVOLUTE if u[i] < 0.7 else
                                        (abi.GP_JOINT_PRISMATIC if u[i] < 0.85 else abi.GP_JOINT_FIXED)
                                        for i in range(1, nb)]
    return parents, jtypes


def random_tree_desc(rng, nb, shape=None):
    """A random tree of nb bodies (random_tree_shape, or the given (parents,
    joint types)) on a floating root: random coordinate axes, random boxes
    with offset centres of mass, prismatic joints with springs, a contact
    point at the far end of every fourth body, over z = 0."""
    from .mechanism import JointSpring
    parents, jtypes = shape if shape is not None else random_tree_shape(rng, nb)
    bodies = [body_desc(-1, abi.GP_JOINT_FLOATING, [0, 0, 0], Isometry3.identity(),
                        SpatialInertia.cuboid_at([0, 0, 0], rng.uniform(2, 5), 0.4, 0.3, 0.2, "root"))]
    contacts = []
    for i in range(1, nb):
        jt = jtypes[i]
        l = rng.uniform(0.05, 0.15)
        iso = Isometry3.translation_xyz(*rng.uniform(-0.15, 0.15, 2), -l)
        inertia = SpatialInertia.cuboid_at([rng.uniform(-0.02, 0.02), rng.uniform(-0.02, 0.02), -l / 2],
                                           rng.uniform(0.1, 1.0), 0.03, 0.03, l, f"b{i}")
        spring = JointSpring(rng.uniform(50, 500), rng.uniform(-0.05, 0.05)) if jt == abi.GP_JOINT_PRISMATIC else None
        bodies.append(body_desc(parents[i], jt, _axis(rng.integers(0, 3)), iso, inertia, spring))
        if i % 4 == 0:
            contacts.append(contact_desc(i, [0.0, 0.0, -l], 20e3))
    hs = [halfspace_desc(HalfSpace.new_with_params([0.0, 0.0, 1.0], 0.0, 0.9, 0.5))]
    return ModelDesc(bodies, contacts, hs)


def trees(n=16384, seed=0, random_fraction=0.25):
    """Trees of 16 or more bodies on a ground plane: hexapods (19 bodies, six
    feet), humanoids (the biped plus two arms, 19 bodies, 16 foot corners),
    half each of the non-random share, and random trees of 16-40 bodies
    (eight random topologies, each instance with its own masses, lengths and
    axes); joint angles N(0, 0.3) (prismatic N(0, 0.02)), the base raised
    so that the lowest contact point starts between 3 mm inside the ground and
    1 cm above it, velocities N(0, 0.3), held torques N(0, 0.02); dt 1e-4
    (at 1e-3 the reference's explicit contact law, stiff normal force and
    regularised Coulomb friction on 0.1 kg feet, diverges in the oracle itself
    on most trees within a thousand steps)."""
    kind = stream(seed, "trees", 1).random(n)
    lt = stream(seed, "trees", 2).uniform(0.1, 0.25, (n, 6))
    ls = stream(seed, "trees", 3).uniform(0.15, 0.3, (n, 6))
    mt = stream(seed, "trees", 4).uniform(0.1, 0.4, (n, 6))
    ms = stream(seed, "trees", 5).uniform(0.1, 0.4, (n, 6))
    base = stream(seed, "trees", 6).uniform(0.3, 0.6, n)
    arms = stream(seed, "trees", 7).uniform(0.1, 0.3, (n, 4))
    rng = np.random.Generator(np.random.Philox(key=[int(seed), WORKLOAD_IDS["trees"] * 1000 + 50]))
    srng = np.random.Generator(np.random.Philox(key=[int(seed), WORKLOAD_IDS["trees"] * 1000 + 51]))
    shapes = [random_tree_shape(srng, int(srng.integers(16, 41))) for _ in range(8)]
    pick = stream(seed, "trees", 10).integers(0, len(shapes), n)
    rq = stream(seed, "trees", 8)
    rv = stream(seed, "trees", 9)
    models, qs, vs, taus = [], [], [], []
    for i in range(n):
        if kind[i] < random_fraction:
            d = random_tree_desc(rng, len(shapes[pick[i]][0]), shapes[pick[i]])
        elif kind[i] < random_fraction + (1 - random_fraction) / 2:
            d = hexapod_desc(dict(bw=0.5, bd=0.3, bh=0.1, bm=3.0, lt=lt[i], ls=ls[i], mt=mt[i], ms=ms[i]))
        else:
            d = humanoid_desc(dict(la=arms[i, 0], lf=arms[i, 1], ma=arms[i, 2], mf=arms[i, 3]))
        q = np.zeros(d.nq)
        q[3] = 1.0
        q[7:] = rq.normal(0.0, 0.3, d.nq - 7)
        q[7:] *= [0.02 / 0.3 if d._bodies[b].jtype == abi.GP_JOINT_PRISMATIC else 1.0
                  for b in range(1, d.n_bodies) if d._bodies[b].jtype != abi.GP_JOINT_FIXED]
        q[6] = -lowest_contact_z(d, q) + (base[i] - 0.3) / 0.3 * 0.013 - 0.003
        models.append(d)
        qs.append(q)
        vs.append(rv.normal(0.0, 0.3, d.nv))
        taus.append(np.concatenate([np.zeros(6), rv.normal(0.0, 0.02, d.nv - 6)]))
    return dict(models=models, counts=[1] * n, q=np.concatenate(qs), v=np.concatenate(vs),
                tau=np.concatenate(taus), dt=1e-4, steps=1000, name="trees")
