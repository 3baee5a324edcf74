"""Host-side mirror of the reference's MechanismState / step / simulate API.

Names, argument meaning and error behaviour follow the reference crate so a
caller (or a test) reads the same:
  MechanismState::new            mechanism.rs:62
  MechanismState::update         mechanism.rs:209
  add_halfspace / add_contact_point  mechanism.rs:379,384
  step                           simulate.rs:20
  simulate                       simulate.rs:87
  dynamics_continuous            dynamics.rs:322
The reference panics where this module raises ValueError / AssertionError.
Every numerical call goes through the HIP library (a batch of one instance);
nothing here computes dynamics on the CPU.
"""
import ctypes as C
import math

import numpy as np

from . import abi

WORLD_FRAME = "world"  # lib.rs:44
GRAVITY = 9.81  # lib.rs:39


# --------------------------------------------------------------------- geometry
class UnitQuaternion:
    """Rotation as a unit quaternion with nalgebra's coords order (x, y, z, w)."""

    __slots__ = ("x", "y", "z", "w")

    def __init__(self, x, y, z, w):
        self.x, self.y, self.z, self.w = float(x), float(y), float(z), float(w)

    @staticmethod
    def identity():
        return UnitQuaternion(0.0, 0.0, 0.0, 1.0)

    @staticmethod
    def from_axis_angle(axis, angle):
        a = np.asarray(axis, dtype=float)
        s = math.sin(angle / 2.0)
        return UnitQuaternion(a[0] * s, a[1] * s, a[2] * s, math.cos(angle / 2.0))

    @staticmethod
    def from_scaled_axis(v):
        v = np.asarray(v, dtype=float)
        n = float(np.linalg.norm(v))
        if n == 0.0:
            return UnitQuaternion.identity()
        return UnitQuaternion.from_axis_angle(v / n, n)

    @staticmethod
    def from_euler_angles(roll, pitch, yaw):
        """R = Rz(yaw) Ry(pitch) Rx(roll) (nalgebra's from_euler_angles)."""
        cr, sr = math.cos(roll * 0.5), math.sin(roll * 0.5)
        cp, sp = math.cos(pitch * 0.5), math.sin(pitch * 0.5)
        cy, sy = math.cos(yaw * 0.5), math.sin(yaw * 0.5)
        return UnitQuaternion(sr * cp * cy - cr * sp * sy, cr * sp * cy + sr * cp * sy,
                              cr * cp * sy - sr * sp * cy, cr * cp * cy + sr * sp * sy)

    @staticmethod
    def from_quaternion(x, y, z, w):
        n = math.sqrt(x * x + y * y + z * z + w * w)
        return UnitQuaternion(x / n, y / n, z / n, w / n)

    def coords(self):
        return [self.x, self.y, self.z, self.w]

    def __mul__(self, o):
        if isinstance(o, UnitQuaternion):
            a, b = self, o
            return UnitQuaternion(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
                                  a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
                                  a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
                                  a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z)
        v = np.asarray(o, dtype=float)
        u = np.array([self.x, self.y, self.z])
        t = 2.0 * np.cross(u, v)
        return v + self.w * t + np.cross(u, t)

    def to_matrix(self):
        x, y, z, w = self.x, self.y, self.z, self.w
        return np.array([
            [w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (w * y + x * z)],
            [2 * (w * z + x * y), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)],
            [2 * (x * z - w * y), 2 * (w * x + y * z), w * w - x * x - y * y + z * z]])


class Isometry3:
    """x_to = R x_from + t (nalgebra Isometry3)."""

    def __init__(self, rotation=None, translation=(0.0, 0.0, 0.0)):
        self.rotation = rotation or UnitQuaternion.identity()
        self.translation = np.asarray(translation, dtype=float).copy()

    @staticmethod
    def identity():
        return Isometry3()

    @staticmethod
    def translation_xyz(x, y, z):
        return Isometry3(None, (x, y, z))

    @staticmethod
    def rotation_scaled_axis(v):
        return Isometry3(UnitQuaternion.from_scaled_axis(v))

    @staticmethod
    def new(translation, scaled_axis):
        return Isometry3(UnitQuaternion.from_scaled_axis(scaled_axis), translation)

    def __mul__(self, o):
        return Isometry3(self.rotation * o.rotation, self.translation + self.rotation * o.translation)


class Transform3D:
    """Frame-tagged isometry successor -> predecessor (transform.rs:54)."""

    def __init__(self, from_, to, iso=None):
        self.from_ = from_
        self.to = to
        self.iso = iso or Isometry3.identity()

    @staticmethod
    def identity(from_, to):
        return Transform3D(from_, to)

    @staticmethod
    def new(from_, to, iso):
        return Transform3D(from_, to, iso)

    @staticmethod
    def move_x(from_, to, amount):
        return Transform3D(from_, to, Isometry3.translation_xyz(amount, 0.0, 0.0))

    @staticmethod
    def move_z(from_, to, amount):
        return Transform3D(from_, to, Isometry3.translation_xyz(0.0, 0.0, amount))

    @staticmethod
    def move_xyz(from_, to, x, y, z):
        return Transform3D(from_, to, Isometry3.translation_xyz(x, y, z))


class Pose:
    """Floating-joint position (pose.rs:6): rotation + translation."""

    def __init__(self, rotation=None, translation=(0.0, 0.0, 0.0)):
        self.rotation = rotation or UnitQuaternion.identity()
        self.translation = np.asarray(translation, dtype=float).copy()

    @staticmethod
    def identity():
        return Pose()

    def to_array(self):  # pose.rs:43-56
        return self.rotation.coords() + list(self.translation)


class SpatialVector:
    def __init__(self, angular=(0.0, 0.0, 0.0), linear=(0.0, 0.0, 0.0)):
        self.angular = np.asarray(angular, dtype=float).copy()
        self.linear = np.asarray(linear, dtype=float).copy()

    @staticmethod
    def zero():
        return SpatialVector()


# --------------------------------------------------------------------- bodies
class SpatialInertia:
    """moment about the frame origin, cross_part = m * com (inertia.rs:32)."""

    def __init__(self, frame, moment, cross_part, mass):
        self.frame = frame
        self.moment = np.asarray(moment, dtype=float).reshape(3, 3).copy()
        self.cross_part = np.asarray(cross_part, dtype=float).copy()
        self.mass = float(mass)

    @staticmethod
    def cuboid_at(com, m, w, d, h, frame):  # inertia.rs:56-74
        com = np.asarray(com, dtype=float)
        mc = np.diag([m * (d * d + h * h) / 12.0, m * (w * w + h * h) / 12.0, m * (w * w + d * d) / 12.0])
        moment = mc + m * (com.dot(com) * np.eye(3) - np.outer(com, com))
        return SpatialInertia(frame, moment, m * com, m)


class ContactPoint:
    def __init__(self, frame, location, k=50e3):  # contact.rs:24-38
        self.frame = frame
        self.location = np.asarray(location, dtype=float).copy()
        self.k = float(k)

    @staticmethod
    def new(frame, location):
        return ContactPoint(frame, location)

    @staticmethod
    def new_with_k(frame, location, k):
        return ContactPoint(frame, location, k)


class SpringContact:
    """A stateful spring leg from a body's origin (contact.rs:74-94): direction
    in the body frame (unit), rest length, stiffness, and the registration the
    dynamics keeps (the world contact point and the half-space, None while the
    leg is in the air). step / simulate / dynamics_continuous send direction
    and rest length to the device when the caller changed them and read the
    registration, direction and rest length back afterwards."""

    def __init__(self, frame, l_rest, direction, k):
        self.frame = frame
        self.l_rest = float(l_rest)
        d = np.asarray(direction, dtype=float)
        self.direction = d / np.linalg.norm(d)
        self.k = float(k)
        self.registered_contact = None
        self.registered_halfspace = None

    @staticmethod
    def new(frame, l_rest, direction, k):
        return SpringContact(frame, l_rest, direction, k)


class HalfSpace:
    def __init__(self, point, normal, alpha=0.9, mu=0.5):  # halfspace.rs:6-36
        self.point = np.asarray(point, dtype=float).copy()
        self.normal = np.asarray(normal, dtype=float).copy()
        self.alpha = float(alpha)
        self.mu = float(mu)

    @staticmethod
    def new(normal, distance):
        n = np.asarray(normal, dtype=float)
        return HalfSpace(n * distance, n)

    @staticmethod
    def new_with_params(normal, distance, alpha, mu):
        n = np.asarray(normal, dtype=float)
        return HalfSpace(n * distance, n, alpha, mu)

    def has_inside(self, point):  # halfspace.rs:39-44
        return float(np.dot(np.asarray(point) - self.point, self.normal)) <= 1e-8


class Sphere:
    """Sphere collider (collision/sphere.rs:6-20): centre at base_translation
    in the body frame."""

    def __init__(self, radius):
        self.base_translation = np.zeros(3)
        self.radius = float(radius)

    @staticmethod
    def new(radius):
        return Sphere(radius)


class Collider:
    """A body's collision geometry (rigid_body.rs:11-37); only spheres reach
    the device (the velocity-stepping integrators use them)."""

    def __init__(self, geometry):
        self.geometry = geometry
        self.enabled = True

    @staticmethod
    def new_sphere(sphere):
        return Collider(sphere)


class RigidBody:
    def __init__(self, inertia):
        self.inertia = inertia
        self.contact_points = []
        self.spring_contacts = []
        self.collider = None

    @staticmethod
    def new(inertia):
        return RigidBody(inertia)

    @staticmethod
    def new_sphere(m, r, frame):  # rigid_body.rs:114
        mx = 2.0 / 5.0 * m * r * r
        body = RigidBody(SpatialInertia(frame, np.diag([mx, mx, mx]), np.zeros(3), m))
        body.collider = Collider.new_sphere(Sphere(r))
        return body

    @staticmethod
    def new_cube(m, l, frame):  # rigid_body.rs:147
        mx = m * l * l / 6.0
        return RigidBody(SpatialInertia(frame, np.diag([mx, mx, mx]), np.zeros(3), m))

    @staticmethod
    def new_cuboid(m, w, d, h, frame):  # rigid_body.rs:160
        return RigidBody(SpatialInertia(frame, np.diag([m * (d * d + h * h) / 12.0, m * (w * w + h * h) / 12.0,
                                                        m * (w * w + d * d) / 12.0]), np.zeros(3), m))

    @staticmethod
    def new_cuboid_at(com, m, w, d, h, frame):  # rigid_body.rs:175
        return RigidBody(SpatialInertia.cuboid_at(com, m, w, d, h, frame))

    def add_contact_point(self, cp):  # rigid_body.rs:205
        if self.inertia.frame != cp.frame:
            raise ValueError(f"Contact point frame {cp.frame} does not match body frame {self.inertia.frame}")
        self.contact_points.append(cp)

    def add_spring_contact(self, sc):  # rigid_body.rs:266
        if self.inertia.frame != sc.frame:
            raise ValueError(f"Spring contact point frame {sc.frame} does not match body frame {self.inertia.frame}")
        self.spring_contacts.append(sc)

    def add_sphere_collider(self, sphere):  # rigid_body.rs:201
        self.collider = Collider.new_sphere(sphere)

    def add_cuboid_contacts(self, w, d, h):  # rigid_body.rs:216
        """The eight corners of a w x d x h box centred on the frame origin:
        the top face (z = +h/2) first, each face in the order (-x, +y),
        (+x, +y), (-x, -y), (+x, -y)."""
        for z in (h / 2.0, -h / 2.0):
            for y in (d / 2.0, -d / 2.0):
                for x in (-w / 2.0, w / 2.0):
                    self.add_contact_point(ContactPoint.new(self.inertia.frame, [x, y, z]))

    def add_cuboid_contacts_with(self, com, w, d, h):  # rigid_body.rs:251
        """The eight corners of a w x d x h box centred on com, x slowest and
        z fastest."""
        c = np.asarray(com, dtype=float)
        for i in (-1.0, 1.0):
            for j in (-1.0, 1.0):
                for k in (-1.0, 1.0):
                    self.add_contact_point(ContactPoint.new(self.inertia.frame,
                                                            c + [i * w / 2.0, j * d / 2.0, k * h / 2.0]))


# --------------------------------------------------------------------- joints
class JointSpring:
    def __init__(self, k, l):
        self.k = float(k)
        self.l = float(l)


class RevoluteJoint:
    jtype = abi.GP_JOINT_REVOLUTE

    def __init__(self, transform, axis):
        self.transform = transform
        self.axis = np.asarray(axis, dtype=float).copy()


class PrismaticJoint:
    jtype = abi.GP_JOINT_PRISMATIC

    def __init__(self, transform, axis, spring=None):
        self.transform = transform
        self.axis = np.asarray(axis, dtype=float).copy()
        self.spring = spring

    @staticmethod
    def new_with_spring(transform, axis, spring):
        return PrismaticJoint(transform, axis, spring)


class FloatingJoint:
    jtype = abi.GP_JOINT_FLOATING

    def __init__(self, transform):
        self.transform = transform
        self.axis = np.zeros(3)


class FixedJoint:
    jtype = abi.GP_JOINT_FIXED

    def __init__(self, transform):
        self.transform = transform
        self.axis = np.zeros(3)


class Integrator:  # integrators.rs:17
    SemiImplicitEuler = abi.GP_SEMI_IMPLICIT_EULER
    RungeKutta2 = abi.GP_RUNGE_KUTTA_2
    RungeKutta4 = abi.GP_RUNGE_KUTTA_4
    VelocityStepping = abi.GP_VELOCITY_STEPPING
    CCDVelocityStepping = abi.GP_CCD_VELOCITY_STEPPING


# --------------------------------------------------------------------- model description
class ModelDesc:
    """Owner of the POD arrays behind one gp_model_desc."""

    def __init__(self, bodies, contacts, halfspaces, gravity=GRAVITY, spring_contacts=(), spheres=()):
        self.n_bodies = len(bodies)
        self._bodies = (abi.gp_body_desc * max(1, len(bodies)))(*bodies)
        self._contacts = (abi.gp_contact_desc * max(1, len(contacts)))(*contacts)
        self._halfspaces = (abi.gp_halfspace_desc * max(1, len(halfspaces)))(*halfspaces)
        self._springs = (abi.gp_spring_contact_desc * max(1, len(spring_contacts)))(*spring_contacts)
        self._spheres = (abi.gp_sphere_desc * max(1, len(spheres)))(*spheres)
        self.desc = abi.gp_model_desc(len(bodies), len(contacts), len(halfspaces), 0,
                                      self._bodies, self._contacts, self._halfspaces, gravity,
                                      len(spring_contacts), 0, self._springs, len(spheres), 0, self._spheres)
        self.nq = sum(7 if b.jtype == abi.GP_JOINT_FLOATING else (0 if b.jtype == abi.GP_JOINT_FIXED else 1)
                      for b in bodies)
        self.nv = sum(6 if b.jtype == abi.GP_JOINT_FLOATING else (0 if b.jtype == abi.GP_JOINT_FIXED else 1)
                      for b in bodies)

    def ptr(self):
        return C.pointer(self.desc)


def body_desc(parent, jtype, axis, iso, inertia, spring=None):
    b = abi.gp_body_desc()
    b.parent = parent
    b.jtype = jtype
    b.axis[:] = [float(x) for x in axis]
    b.x_tree_quat[:] = iso.rotation.coords()
    b.x_tree_trans[:] = [float(x) for x in iso.translation]
    J = inertia.moment
    b.moment[:] = [J[0, 0], J[1, 1], J[2, 2], J[0, 1], J[0, 2], J[1, 2]]
    b.mc[:] = [float(x) for x in inertia.cross_part]
    b.mass = inertia.mass
    if spring is not None:
        b.spring_k, b.spring_l = spring.k, spring.l
    return b


def contact_desc(body, p, k):
    c = abi.gp_contact_desc()
    c.body = body
    c.p[:] = [float(x) for x in p]
    c.k = float(k)
    return c


def spring_contact_desc(body, sc):
    d = abi.gp_spring_contact_desc()
    d.body = body
    d.direction[:] = [float(x) for x in sc.direction]
    d.l_rest, d.k = sc.l_rest, sc.k
    return d


def sphere_desc(body, collider):
    d = abi.gp_sphere_desc()
    d.body = body
    d.enabled = 1 if collider.enabled else 0
    d.center[:] = [float(x) for x in collider.geometry.base_translation]
    d.radius = collider.geometry.radius
    return d


def halfspace_desc(h):
    d = abi.gp_halfspace_desc()
    d.point[:] = [float(x) for x in h.point]
    d.normal[:] = [float(x) for x in h.normal]
    d.alpha, d.mu = h.alpha, h.mu
    return d


def flatten_q(q):
    out = []
    for qi in q:
        if qi is None:
            continue
        if isinstance(qi, Pose):
            out.extend(qi.to_array())
        else:
            out.append(float(qi))
    return out


def flatten_v(v):
    out = []
    for vi in v:
        if vi is None:
            continue
        if isinstance(vi, SpatialVector):
            out.extend(list(vi.angular) + list(vi.linear))
        else:
            out.append(float(vi))
    return out


def unflatten_q(flat, joints):
    out, k = [], 0
    for j in joints:
        if j.jtype == abi.GP_JOINT_FLOATING:
            out.append(Pose(UnitQuaternion(*flat[k:k + 4]), flat[k + 4:k + 7]))
            k += 7
        elif j.jtype == abi.GP_JOINT_FIXED:
            out.append(None)
        else:
            out.append(float(flat[k]))
            k += 1
    return out


def unflatten_v(flat, joints):
    out, k = [], 0
    for j in joints:
        if j.jtype == abi.GP_JOINT_FLOATING:
            out.append(SpatialVector(flat[k:k + 3], flat[k + 3:k + 6]))
            k += 6
        elif j.jtype == abi.GP_JOINT_FIXED:
            out.append(None)
        else:
            out.append(float(flat[k]))
            k += 1
    return out


# --------------------------------------------------------------------- MechanismState
class MechanismState:
    """A single mechanism (mechanism.rs:41). Joint i's child is body i."""

    def __init__(self, treejoints, bodies, gravity=GRAVITY):
        self.treejoints = list(treejoints)
        self.bodies = list(bodies)
        self.parents = []  # 0 = world, else 1-based body number (mechanism.rs:45)
        self.q, self.v = [], []
        for index, j in enumerate(self.treejoints):
            if j.jtype == abi.GP_JOINT_FLOATING:
                self.q.append(Pose.identity())
                self.v.append(SpatialVector.zero())
            elif j.jtype == abi.GP_JOINT_FIXED:
                self.q.append(None)
                self.v.append(None)
            else:
                self.q.append(0.0)
                self.v.append(0.0)
            if j.transform.from_ != self.bodies[index].inertia.frame:  # mechanism.rs:91-96
                raise ValueError(f"joint {index + 1}'s child body is not body {index + 1}")
            if j.transform.to == WORLD_FRAME:
                self.parents.append(0)
            else:
                for bi, b in enumerate(self.bodies):
                    if j.transform.to == b.inertia.frame:
                        self.parents.append(bi + 1)
                        break
                else:  # mechanism.rs:110-115
                    raise ValueError(f"joint {index + 1} has no parent body of frame {j.transform.to}")
        self.halfspaces = []
        self.gravity = gravity
        self._desc = None
        self._model = None
        self._batch = None
        self._spring_seen = None

    # -- model edits (invalidate the device copy)
    def add_halfspace(self, halfspace):
        self.halfspaces.append(halfspace)
        self._invalidate()

    def add_contact_point(self, point):
        for b in self.bodies:
            if b.inertia.frame == point.frame:
                b.add_contact_point(point)
                break
        self._invalidate()

    def add_spring_contact(self, sc):  # mechanism.rs:394
        for b in self.bodies:
            if b.inertia.frame == sc.frame:
                b.add_spring_contact(sc)
        self._invalidate()

    def _invalidate(self):
        self._batch = None
        self._model = None
        self._desc = None
        self._spring_seen = None

    def spring_contacts(self):
        """Spring contacts in description order (body by body)."""
        return [sc for b in self.bodies for sc in b.spring_contacts]

    # -- state
    def update(self, q, v):
        self.q = list(q)
        self.v = list(v)

    def set_joint_q(self, jointid, q):
        if self.treejoints[jointid - 1].jtype == abi.GP_JOINT_FIXED:
            raise ValueError("Should not try to set a fixed joint's position")
        self.q[jointid - 1] = q

    def set_joint_v(self, jointid, v):
        self.v[jointid - 1] = v

    def kinetic_energy(self):  # mechanism.rs:334
        b = self._device_batch()
        b.set_state(self.flat_q(), self.flat_v())
        return float(b.energy()[0][0])

    def gravitational_energy(self):  # mechanism.rs:352 (height of each body frame's origin)
        b = self._device_batch()
        b.set_state(self.flat_q(), self.flat_v())
        return float(b.energy()[1][0])

    def poses(self):
        """World pose of every body at the current q (mechanism.rs:403-418)."""
        world, out = [], []
        for i, j in enumerate(self.treejoints):
            iso, q = j.transform.iso, self.q[i]
            if j.jtype == abi.GP_JOINT_REVOLUTE:
                iso = iso * Isometry3(UnitQuaternion.from_axis_angle(j.axis, q))
            elif j.jtype == abi.GP_JOINT_PRISMATIC:
                iso = iso * Isometry3(None, j.axis * q)
            elif j.jtype == abi.GP_JOINT_FLOATING:
                iso = iso * Isometry3(q.rotation, q.translation)
            b2r = iso if self.parents[i] == 0 else world[self.parents[i] - 1] * iso
            world.append(b2r)
            out.append(Pose(b2r.rotation, b2r.translation))
        return out

    def flat_q(self):
        return np.array(flatten_q(self.q), dtype=float)

    def flat_v(self):
        return np.array(flatten_v(self.v), dtype=float)

    # -- model description in the C ABI's terms
    def model_desc(self):
        if self._desc is None:
            bodies, contacts, springs, spheres = [], [], [], []
            for i, (j, b) in enumerate(zip(self.treejoints, self.bodies)):
                spring = getattr(j, "spring", None)
                bodies.append(body_desc(self.parents[i] - 1, j.jtype, j.axis, j.transform.iso, b.inertia, spring))
                for cp in b.contact_points:
                    contacts.append(contact_desc(i, cp.location, cp.k))
                for sc in b.spring_contacts:
                    springs.append(spring_contact_desc(i, sc))
                if b.collider is not None and isinstance(b.collider.geometry, Sphere):
                    spheres.append(sphere_desc(i, b.collider))
            hs = [halfspace_desc(h) for h in self.halfspaces]
            self._desc = ModelDesc(bodies, contacts, hs, self.gravity, springs, spheres)
        return self._desc

    def _device_batch(self, opts=None):
        from .batch import Batch, Model
        if self._batch is None:
            self._model = Model(self.model_desc())
            self._batch = Batch([self._model], [1], opts)
            self._spring_seen = [(tuple(sc.direction), sc.l_rest) for sc in self.spring_contacts()]
        return self._batch

    def _push_springs(self, b):
        """Send leg directions / rest lengths the caller changed since the
        device last reported them (contact.rs:887-888 writes them directly)."""
        for j, sc in enumerate(self.spring_contacts()):
            key = (tuple(np.asarray(sc.direction, dtype=float)), float(sc.l_rest))
            if key != self._spring_seen[j]:
                b.set_spring_contact(0, j, key[0], key[1])
                self._spring_seen[j] = key

    def _pull_springs(self, b):
        for j, sc in enumerate(self.spring_contacts()):
            reg, h, p, d, lr = b.get_spring_contact(0, j, with_halfspace=True)
            sc.registered_contact = p if reg else None
            sc.registered_halfspace = self.halfspaces[h] if reg else None
            sc.direction, sc.l_rest = d, lr
            self._spring_seen[j] = (tuple(d), lr)


def _tau_flat(state, tau):
    """Empty tau = zeros (simulate.rs:27-37); else its length must match v."""
    if tau is None or len(tau) == 0:
        return np.zeros(len(flatten_v(state.v)))
    if len(tau) != len(state.v):  # simulate.rs:39-45
        raise AssertionError(
            f"joint torques vector τ length {len(tau)}  and Joint velocity vector v length {len(state.v)} differ!")
    return np.array(flatten_v(tau), dtype=float)


def dynamics_continuous(state, tau):
    """Joint accelerations at the state's (q, v) (dynamics.rs:322)."""
    b = state._device_batch()
    b.set_state(state.flat_q(), state.flat_v())
    state._push_springs(b)
    vdot = b.dynamics(_tau_flat(state, tau))
    state._pull_springs(b)
    return unflatten_v(vdot, state.treejoints)


def step(state, dt, tau, integrator):
    """Advance the state by dt (simulate.rs:20); returns the new (q, v)."""
    b = state._device_batch()
    b.set_state(state.flat_q(), state.flat_v())
    b.set_tau(_tau_flat(state, tau))
    state._push_springs(b)
    b.step(dt, integrator, 1)
    state._pull_springs(b)
    qf, vf = b.get_state()
    q, v = unflatten_q(qf, state.treejoints), unflatten_v(vf, state.treejoints)
    state.update(q, v)
    return q, v


def simulate(state, final_time, dt, control_fn, integrator):
    """Float-accumulated loop `while t < final_time` recording every state
    (simulate.rs:87-112). control_fn(state) returns the torques of each step."""
    b = state._device_batch()
    b.set_state(state.flat_q(), state.flat_v())
    joints = state.treejoints

    def cb(q, v, t):
        state.update(unflatten_q(q, joints), unflatten_v(v, joints))
        state._pull_springs(b)  # a controller may read and move the legs (SLIP_control.rs:120-121)
        tau = _tau_flat(state, control_fn(state))
        state._push_springs(b)
        return tau

    state._push_springs(b)
    qs, vs = b.simulate(final_time, dt, cb, integrator)
    state._pull_springs(b)
    state.update(unflatten_q(qs[-1], joints), unflatten_v(vs[-1], joints))
    return ([unflatten_q(r, joints) for r in qs], [unflatten_v(r, joints) for r in vs])
