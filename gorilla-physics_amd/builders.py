"""Model builders mirroring the reference's helpers.rs and builders/ (same
names and arguments), used by the tests and the workload generators. Masses,
sizes, frames and joint axes are the reference's: parity is defined on these
models."""
import numpy as np

from .mechanism import (WORLD_FRAME, ContactPoint, FloatingJoint, Isometry3, JointSpring, MechanismState,
                        PrismaticJoint, RevoluteJoint, RigidBody, SpatialInertia, SpringContact, Transform3D)


def build_pendulum(mass, moment, cross_part, rod_to_world, axis):  # helpers.rs:24
    rod = RigidBody(SpatialInertia("rod", moment, cross_part, mass))
    j = RevoluteJoint(Transform3D.new("rod", WORLD_FRAME, rod_to_world), axis)
    return MechanismState([j], [rod])


def build_double_pendulum(mass, moment, cross_part, rod1_to_world, rod2_to_rod1, axis):  # helpers.rs:49
    joints = [RevoluteJoint(Transform3D.new("rod1", WORLD_FRAME, rod1_to_world), axis),
              RevoluteJoint(Transform3D.new("rod2", "rod1", rod2_to_rod1), axis)]
    bodies = [RigidBody(SpatialInertia("rod1", moment, cross_part, mass)),
              RigidBody(SpatialInertia("rod2", moment, cross_part, mass))]
    return MechanismState(joints, bodies)


def build_cart(mass, moment, cross_part, axis):  # helpers.rs:86
    j = PrismaticJoint(Transform3D.identity("cart", WORLD_FRAME), axis)
    return MechanismState([j], [RigidBody(SpatialInertia("cart", moment, cross_part, mass))])


def build_cart_pole(mass_cart, mass_pole, moment_cart, moment_pole, cross_part_cart, cross_part_pole,
                    axis_pole):  # helpers.rs:111
    joints = [PrismaticJoint(Transform3D.identity("cart", WORLD_FRAME), [1.0, 0.0, 0.0]),
              RevoluteJoint(Transform3D.identity("pole", "cart"), axis_pole)]
    bodies = [RigidBody(SpatialInertia("cart", moment_cart, cross_part_cart, mass_cart)),
              RigidBody(SpatialInertia("pole", moment_pole, cross_part_pole, mass_pole))]
    return MechanismState(joints, bodies)


def add_cube_contacts(state, frame, l):  # helpers.rs:388
    h = l / 2.0
    for p in ([h, h, -h], [h, -h, -h], [-h, h, -h], [-h, -h, -h], [h, h, h], [h, -h, h], [-h, h, h], [-h, -h, h]):
        state.add_contact_point(ContactPoint.new(frame, p))


def build_cube(mass, length):  # helpers.rs:151
    cube = RigidBody.new_cube(mass, length, "cube")
    state = MechanismState([FloatingJoint(Transform3D.identity("cube", WORLD_FRAME))], [cube])
    add_cube_contacts(state, "cube", length)
    return state


def build_sphere(mass, radius):  # helpers.rs:812
    """A floating solid sphere with its sphere collider."""
    body = RigidBody.new_sphere(mass, radius, "sphere")
    return MechanismState([FloatingJoint(Transform3D.identity("sphere", WORLD_FRAME))], [body])


def build_n_spheres(mass, radius, n):  # helpers.rs:823
    bodies, joints = [], []
    for i in range(1, n + 1):
        frame = f"sphere {i}"
        bodies.append(RigidBody.new_sphere(mass, radius, frame))
        joints.append(FloatingJoint(Transform3D.identity(frame, WORLD_FRAME)))
    return MechanismState(joints, bodies)


def build_SLIP(m, r, l_rest, angle, k_spring):  # helpers.rs:308
    """Spring-loaded inverted pendulum: a floating solid sphere (moment
    2/5 m r^2) with one spring leg at `angle` from straight down, forward."""
    mx = 2.0 / 5.0 * m * r * r
    body = RigidBody(SpatialInertia("body", np.diag([mx, mx, mx]), np.zeros(3), m))
    state = MechanismState([FloatingJoint(Transform3D.identity("body", WORLD_FRAME))], [body])
    state.add_spring_contact(SpringContact("body", l_rest, [np.sin(angle), 0.0, -np.cos(angle)], k_spring))
    return state


def build_quadruped():  # helpers.rs:423 (moments as written there)
    m_body, w_body, d_body, h_body = 5.0, 1.5, 0.5, 0.5
    moment_body = np.diag([(d_body * d_body + h_body + h_body) * m_body / 12.0,
                           (w_body * w_body + h_body + h_body) * m_body / 12.0,
                           (w_body * w_body + d_body * d_body) * m_body / 12.0])
    l_leg = 1.0
    body = RigidBody(SpatialInertia("body", moment_body, np.zeros(3), m_body))
    m_hip, l_hip, m_knee, l_knee = 0.5, 0.2, 0.5, 0.2
    hip_axis = [0.0, -1.0, 0.0]
    knee_axis = [0.0, -1.0, 0.0]
    bodies = [body]
    joints = [FloatingJoint(Transform3D.identity("body", WORLD_FRAME))]
    for name, sx, sy in (("fr", 1, -1), ("fl", 1, 1), ("br", -1, -1), ("bl", -1, 1)):
        hip, knee = f"{name}_hip", f"{name}_knee"
        bodies.append(RigidBody.new_cube(m_hip, l_hip, hip))
        joints.append(RevoluteJoint(Transform3D.new(hip, "body", Isometry3.translation_xyz(
            sx * w_body / 2.0, sy * d_body / 2.0, 0.0)), hip_axis))
        bodies.append(RigidBody.new_cube(m_knee, l_knee, knee))
        joints.append(RevoluteJoint(Transform3D.move_z(knee, hip, -l_leg / 2.0), knee_axis))
    state = MechanismState(joints, bodies)
    for name in ("fr", "fl", "br", "bl"):
        state.add_contact_point(ContactPoint.new(f"{name}_hip", [0.0, 0.0, 0.0]))
    for name in ("fr", "fl", "br", "bl"):
        state.add_contact_point(ContactPoint.new(f"{name}_knee", [0.0, 0.0, 0.0]))
    for name in ("fr", "fl", "br", "bl"):
        state.add_contact_point(ContactPoint.new_with_k(f"{name}_knee", [0.0, 0.0, -l_leg / 2.0], 10e3))
    return state


def build_rimless_wheel(m_body, r_body, l, n_foot):  # helpers.rs:168
    """A solid-sphere body (moment 2/5 m r^2, no collider) on a floating
    joint with n_foot spoke tips at distance l, spaced 2 pi / n_foot about y,
    the first straight down."""
    mx = 2.0 / 5.0 * m_body * r_body * r_body
    body = RigidBody(SpatialInertia("body", np.diag([mx, mx, mx]), np.zeros(3), m_body))
    state = MechanismState([FloatingJoint(Transform3D.identity("body", WORLD_FRAME))], [body])
    for i in range(n_foot):
        a = i * 2.0 * np.pi / n_foot
        state.add_contact_point(ContactPoint.new("body", [-l * np.sin(a), 0.0, -l * np.cos(a)]))
    return state


def build_2d_hopper(m_body, w_body, h_body, m_hip, r_hip, body_hip_length, m_piston, r_piston,
                    hip_piston_length, m_leg, l_leg, piston_leg_length):  # helpers.rs:203
    """Floating body, revolute hip about y, prismatic piston and prismatic leg
    along -z, one contact point at the leg's frame origin. The body's z
    moment is (w^2 + w^2) m / 12 as the reference writes it."""
    moment_body = np.diag([(w_body * w_body + h_body * h_body) * m_body / 12.0,
                           (w_body * w_body + h_body * h_body) * m_body / 12.0,
                           (w_body * w_body + w_body * w_body) * m_body / 12.0])

    def ball(frame, m, r):
        mx = 2.0 / 5.0 * m * r * r
        return RigidBody(SpatialInertia(frame, np.diag([mx, mx, mx]), np.zeros(3), m))

    ml = 1.0 / 3.0 * m_leg * l_leg * l_leg
    bodies = [RigidBody(SpatialInertia("body", moment_body, np.zeros(3), m_body)), ball("hip", m_hip, r_hip),
              ball("piston", m_piston, r_piston),
              RigidBody(SpatialInertia("leg", np.diag([ml, ml, 0.0]), np.array([0.0, 0.0, m_leg * l_leg / 2.0]),
                                       m_leg))]
    joints = [FloatingJoint(Transform3D.identity("body", WORLD_FRAME)),
              RevoluteJoint(Transform3D.move_z("hip", "body", -body_hip_length), [0.0, 1.0, 0.0]),
              PrismaticJoint(Transform3D.move_z("piston", "hip", -hip_piston_length), [0.0, 0.0, -1.0]),
              PrismaticJoint(Transform3D.move_z("leg", "piston", -piston_leg_length), [0.0, 0.0, -1.0])]
    state = MechanismState(joints, bodies)
    state.add_contact_point(ContactPoint.new("leg", [0.0, 0.0, 0.0]))
    return state


def build_hopper(m_foot, r_foot, m_hip, r_hip, m_body, r_body, l_foot_to_hip):  # helpers.rs:345
    """Three spheres (with their colliders): a floating foot, a hip on a
    sprung prismatic leg along z (k 1e3, rest length 0), a body on a revolute
    joint about y; the foot's contact point has k = 75e3."""
    joints = [FloatingJoint(Transform3D.identity("foot", WORLD_FRAME)),
              PrismaticJoint.new_with_spring(Transform3D.move_z("hip", "foot", l_foot_to_hip), [0.0, 0.0, 1.0],
                                             JointSpring(1e3, 0.0)),
              RevoluteJoint(Transform3D.identity("body", "hip"), [0.0, 1.0, 0.0])]
    bodies = [RigidBody.new_sphere(m_foot, r_foot, "foot"), RigidBody.new_sphere(m_hip, r_hip, "hip"),
              RigidBody.new_sphere(m_body, r_body, "body")]
    state = MechanismState(joints, bodies)
    state.add_contact_point(ContactPoint.new_with_k("foot", [0.0, 0.0, 0.0], 75e3))
    return state


_L1, _L2, _M_LINK = 0.05, 0.2, 0.1  # link thickness, link length, link mass of the biped family


def _biped_leg(side, sign, with_ankle, limb_contacts, foot_axis):
    """One leg of build_biped / build_leg (biped_builder.rs:12-188,
    leg_builder.rs:8-109): pelvis (revolute z) at x = sign (L1 + L1) / 2
    from the base, hip (revolute -y) below the pelvis, thigh (revolute x) at
    the hip's outer end, calf (revolute x), an optional ankle (revolute x),
    foot with its eight corners as contact points."""
    L1, L2, m = _L1, _L2, _M_LINK
    down = [0.0, 0.0, -L2 / 2.0]
    names = [f"{n}_{side}" for n in ("pelvis", "hip", "thigh", "calf", "ankle", "foot")]
    pelvis, hip, thigh, calf, ankle, foot = names
    bodies = [RigidBody.new_cuboid_at(down, m, L1, L1, L2, pelvis),
              RigidBody.new_cuboid_at([sign * L2 / 2.0, 0.0, 0.0], m, L2, L1, L1, hip),
              RigidBody.new_cuboid_at(down, m, L1, L1, L2, thigh),
              RigidBody.new_cuboid_at(down, m, L1, L1, L2, calf)]
    joints = [RevoluteJoint(Transform3D.move_x(pelvis, "base", sign * (L1 + L1) / 2.0), [0.0, 0.0, 1.0]),
              RevoluteJoint(Transform3D.move_xyz(hip, pelvis, 0.0, 0.0, -L2), [0.0, -1.0, 0.0]),
              RevoluteJoint(Transform3D.move_x(thigh, hip, sign * L2), [1.0, 0.0, 0.0]),
              RevoluteJoint(Transform3D.move_z(calf, thigh, -L2), [1.0, 0.0, 0.0])]
    if limb_contacts:
        for b in bodies[2:4]:
            b.add_cuboid_contacts_with(down, L1, L1, L2)
    above = calf
    if with_ankle:
        bodies.append(RigidBody.new_cuboid(m, L1, L2, L1, ankle))
        joints.append(RevoluteJoint(Transform3D.move_z(ankle, calf, -L2), [1.0, 0.0, 0.0]))
        above = ankle
    f = RigidBody.new_cuboid(m, 0.2, L2, L1, foot)
    f.add_cuboid_contacts(0.2, L2, L1)
    bodies.append(f)
    offset = -(L1 + L1) / 2.0 if with_ankle else -L2
    joints.append(RevoluteJoint(Transform3D.move_z(foot, above, offset), foot_axis))
    return joints, bodies


def _biped_base():
    return (FloatingJoint(Transform3D.identity("base", WORLD_FRAME)),
            RigidBody.new_cuboid(_M_LINK, _L1, _L1, _L2, "base"))


def build_biped():  # builders/biped_builder.rs:12
    """Floating base and two six-joint legs (pelvis, hip, thigh, calf,
    ankle, foot); 13 bodies, 16 contact points (the feet's corners)."""
    jb, bb = _biped_base()
    jl, bl = _biped_leg("left", 1.0, True, False, [0.0, -1.0, 0.0])
    jr, br = _biped_leg("right", -1.0, True, False, [0.0, -1.0, 0.0])
    return MechanismState([jb] + jl + jr, [bb] + bl + br)


def build_leg():  # builders/leg_builder.rs:8
    """Floating base and one five-joint leg (pelvis, hip, thigh, calf, foot
    about x), with contact points on the thigh, calf and foot corners."""
    jb, bb = _biped_base()
    jl, bl = _biped_leg("left", 1.0, False, True, [1.0, 0.0, 0.0])
    return MechanismState([jb] + jl, [bb] + bl)


def build_leg_from_foot():  # builders/leg_builder.rs:111
    """The leg rooted at a floating foot: calf, thigh (revolute x), hip
    (revolute x), pelvis (revolute -y), base (revolute z), each link's centre
    of mass on the side away from the foot; contact points on the foot,
    calf and thigh corners."""
    L1, L2, m = _L1, _L2, _M_LINK
    up = [0.0, 0.0, L2 / 2.0]
    foot = RigidBody.new_cuboid(m, 0.2, L2, L1, "foot_left")
    foot.add_cuboid_contacts(0.2, L2, L1)
    calf = RigidBody.new_cuboid_at(up, m, L1, L1, L2, "calf_left")
    calf.add_cuboid_contacts_with(up, L1, L1, L2)
    thigh = RigidBody.new_cuboid_at(up, m, L1, L1, L2, "thigh_left")
    thigh.add_cuboid_contacts_with(up, L1, L1, L2)
    hip = RigidBody.new_cuboid_at([-L2 / 2.0, 0.0, 0.0], m, L2, L1, L1, "hip_left")
    pelvis = RigidBody.new_cuboid_at(up, m, L1, L1, L2, "pelvis_left")
    base = RigidBody.new_cuboid_at([-(L1 + L1) / 2.0, 0.0, 0.0], m, L1, L1, L2, "base")
    joints = [FloatingJoint(Transform3D.identity("foot_left", WORLD_FRAME)),
              RevoluteJoint(Transform3D.identity("calf_left", "foot_left"), [1.0, 0.0, 0.0]),
              RevoluteJoint(Transform3D.move_z("thigh_left", "calf_left", L2), [1.0, 0.0, 0.0]),
              RevoluteJoint(Transform3D.move_z("hip_left", "thigh_left", L2), [1.0, 0.0, 0.0]),
              RevoluteJoint(Transform3D.move_x("pelvis_left", "hip_left", -L2), [0.0, -1.0, 0.0]),
              RevoluteJoint(Transform3D.move_z("base", "pelvis_left", L2), [0.0, 0.0, 1.0])]
    return MechanismState(joints, [foot, calf, thigh, hip, pelvis, base])


# SO-101 links (builders/mod.rs:19-249, from the arm's URDF): mass, centre of
# mass, and the inertia about it as (ixx, ixy, ixz, iyy, iyz, izz)
_SO101_LINKS = (
    ("base", 0.147, (0.0137179, -5.19711e-05, 0.0334843),
     (0.000114686, -4.59787e-07, 4.97151e-06, 0.000136117, 9.75275e-08, 0.000130364)),
    ("shoulder", 0.100006, (-0.0307604, -1.66727e-05, -0.0252713),
     (8.3759e-05, 7.55525e-08, -1.16342e-06, 8.10403e-05, 1.54663e-07, 2.39783e-05)),
    ("upper_arm", 0.103, (-0.0898471, -0.00838224, 0.0184089),
     (4.08002e-05, -1.97819e-05, -4.03016e-08, 0.000147318, 8.97326e-09, 0.000142487)),
    ("lower_arm", 0.104, (-0.0980701, 0.00324376, 0.0182831),
     (2.87438e-05, 7.41152e-06, 1.26409e-06, 0.000159844, -4.90188e-08, 0.00014529)),
    ("wrist", 0.079, (-0.000103312, -0.0386143, 0.0281156),
     (3.68263e-05, 1.7893e-08, -5.28128e-08, 2.5391e-05, 3.6412e-06, 2.1e-05)),
    ("gripper", 0.087, (0.000213627, 0.000245138, -0.025187),
     (2.75087e-05, -3.35241e-07, -5.7352e-06, 4.33657e-05, -5.17847e-08, 3.45059e-05)),
    ("jaw", 0.012, (-0.00157495, -0.0300244, 0.0192755),
     (6.61427e-06, -3.19807e-07, -5.90717e-09, 1.89032e-06, -1.09945e-07, 5.28738e-06)),
)
# each revolute joint (about the child's z) as xyz and roll-pitch-yaw of the
# child frame in its parent's (builders/mod.rs:266-320)
_SO101_JOINTS = (
    ((0.0388353, -8.97657e-09, 0.0624), (3.14159, 4.18253e-17, -3.14159)),
    ((-0.0303992, -0.0182778, -0.054), (-1.5708, -1.5708, 0.0)),
    ((-0.11257, -0.028, 1.73763e-16), (-3.63608e-16, 8.74301e-16, 1.5708)),
    ((-0.1349, 0.0052, 3.62355e-17), (4.02456e-15, 8.67362e-16, -1.5708)),
    ((5.55112e-17, -0.0611, 0.0181), (1.5708, 0.0486795, 3.14159)),
    ((0.0202, 0.0188, -0.0234), (1.5708, -5.24284e-08, -1.41553e-15)),
)


def build_so101(*meshes):  # builders/mod.rs:252
    """The SO-101 arm: a base welded to the world and six revolute joints
    about z (shoulder pan, shoulder lift, elbow, wrist flex, wrist roll,
    jaw). The reference passes each link's mesh, which only draws it (its
    collider is disabled), so any mesh arguments are accepted and unused."""
    from .mechanism import FixedJoint, UnitQuaternion
    bodies = []
    for frame, m, com, (ixx, ixy, ixz, iyy, iyz, izz) in _SO101_LINKS:
        c = np.asarray(com, dtype=float)
        j_com = np.array([[ixx, ixy, ixz], [ixy, iyy, iyz], [ixz, iyz, izz]])
        bodies.append(RigidBody(SpatialInertia(frame, j_com + m * (c.dot(c) * np.eye(3) - np.outer(c, c)), m * c, m)))
    names = [link[0] for link in _SO101_LINKS]
    joints = [FixedJoint(Transform3D.identity("base", WORLD_FRAME))]
    for k, (xyz, rpy) in enumerate(_SO101_JOINTS):
        iso = Isometry3(UnitQuaternion.from_euler_angles(*rpy), xyz)
        joints.append(RevoluteJoint(Transform3D.new(names[k + 1], names[k], iso), [0.0, 0.0, 1.0]))
    return MechanismState(joints, bodies)
