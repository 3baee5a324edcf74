"""Model builders mirroring the reference's helpers.rs (same names and
arguments), used by the tests and the workload generators."""
import numpy as np

from .mechanism import (WORLD_FRAME, ContactPoint, FloatingJoint, Isometry3, MechanismState, PrismaticJoint,
                        RevoluteJoint, RigidBody, SpatialInertia, SpringContact, Transform3D)


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
