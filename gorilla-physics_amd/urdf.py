"""URDF ingest: a robot description as a MechanismState of the reference's
joint types, ready for gp_model_create (MechanismState.model_desc()).

The reference builds its URDF robots by hand from the description's numbers
(build_so101, builders/mod.rs:250-330) and, in its hybrid engine, from an
urdf-rs Robot (build_rigid / build_joint / build_fixed_joint,
hybrid/rigid/helper.rs:13-87). This reader follows the same conversions:

- a link's <inertial> (mass, centre of mass c = origin xyz, the inertia
  tensor about c) becomes SpatialInertia(moment about the link frame's
  origin = I_c + m (|c|^2 1 - c c^T), cross part m c, m) (helper.rs:17-34,
  inertia.rs:17-37); an inertial origin rpy rotates the tensor first
  (R I_c R^T; the reference reads only xyz, which is the same for rpy = 0);
- a joint's <origin xyz rpy> becomes its init_iso, Transform3D::new_xyz_rpy
  (transform.rs:77-94: rotation UnitQuaternion::from_euler_angles(r, p, y),
  the URDF's fixed-axis roll-pitch-yaw); <axis> (default x) is the joint axis
  in the child link's frame;
- revolute and continuous joints become RevoluteJoint, prismatic
  PrismaticJoint, fixed FixedJoint, floating FloatingJoint (planar joints are
  not in the reference and are refused);
- the root link (the one that is no joint's child) is welded to the world
  (FixedJoint, as build_so101 does) or, with floating_base, a FloatingJoint;
- a <collision> sphere becomes the link's sphere collider (rigid_body.rs:201);
  meshes only draw a link in the reference and are ignored;
- snap_angles rounds every joint-origin angle within 1e-3 of 0, +-pi/2 or
  +-pi to that value, the correction helper.rs:60-75 applies to revolute
  joints (off by default: build_so101 keeps the description's digits).

Bodies are ordered so that every parent precedes its children (depth first
in the document's joint order), as MechanismState requires (mechanism.rs:98).
"""
import math
import xml.etree.ElementTree as ET

import numpy as np

from .mechanism import (WORLD_FRAME, FixedJoint, FloatingJoint, Isometry3, PrismaticJoint, RevoluteJoint, RigidBody,
                        Sphere, SpatialInertia, Transform3D, UnitQuaternion, MechanismState)


class UrdfError(ValueError):
    pass


def _floats(text, n, default):
    if text is None:
        return list(default)
    vals = [float(x) for x in text.split()]
    if len(vals) != n:
        raise UrdfError(f"expected {n} numbers, got {text!r}")
    return vals


def _origin(el, snap=False):
    o = el.find("origin") if el is not None else None
    xyz = _floats(o.get("xyz") if o is not None else None, 3, (0.0, 0.0, 0.0))
    rpy = _floats(o.get("rpy") if o is not None else None, 3, (0.0, 0.0, 0.0))
    if snap:
        for k, x in enumerate(rpy):
            for v in (math.pi, -math.pi, math.pi / 2, -math.pi / 2, 0.0):
                if abs(x - v) < 1e-3:
                    rpy[k] = v
                    break
    return np.array(xyz), rpy


def _link_body(link):
    name = link.get("name")
    inertial = link.find("inertial")
    if inertial is None:
        m, c, I_c = 0.0, np.zeros(3), np.zeros((3, 3))
    else:
        mass = inertial.find("mass")
        m = float(mass.get("value")) if mass is not None else 0.0
        c, rpy = _origin(inertial)
        it = inertial.find("inertia")
        g = (lambda k: float(it.get(k, "0"))) if it is not None else (lambda k: 0.0)
        I_c = np.array([[g("ixx"), g("ixy"), g("ixz")], [g("ixy"), g("iyy"), g("iyz")], [g("ixz"), g("iyz"), g("izz")]])
        R = UnitQuaternion.from_euler_angles(*rpy).to_matrix()
        I_c = R @ I_c @ R.T
    moment = I_c + m * (float(c @ c) * np.eye(3) - np.outer(c, c))
    body = RigidBody(SpatialInertia(name, moment, m * c, m))
    for col in link.findall("collision"):
        sph = col.find("geometry/sphere")
        if sph is not None:
            center, _ = _origin(col)
            sphere = Sphere(float(sph.get("radius")))
            sphere.base_translation = center
            body.add_sphere_collider(sphere)
    return body


def parse_urdf(source, floating_base=False, snap_angles=False):
    """A MechanismState from URDF text (or a path to a .urdf file). Returns
    (state, joint_names): joint_names[i] names the URDF joint of body i (the
    root's is "world")."""
    text = source
    if not source.lstrip().startswith("<"):
        with open(source) as f:
            text = f.read()
    root = ET.fromstring(text)
    if root.tag != "robot":
        raise UrdfError("not a <robot> description")
    links = {l.get("name"): l for l in root.findall("link")}
    joints = root.findall("joint")
    children = {}
    child_of = {}
    for j in joints:
        p, c = j.find("parent").get("link"), j.find("child").get("link")
        if p not in links or c not in links:
            raise UrdfError(f"joint {j.get('name')} names an unknown link")
        if c in child_of:
            raise UrdfError(f"link {c} has two parent joints (closed loops are not a tree)")
        child_of[c] = j
        children.setdefault(p, []).append(j)
    roots = [n for n in links if n not in child_of]
    if len(roots) != 1:
        raise UrdfError(f"expected one root link, found {roots}")
    order = []  # (link name, joint element or None), parents first

    def visit(name, joint):
        order.append((name, joint))
        for j in children.get(name, []):
            visit(j.find("child").get("link"), j)

    visit(roots[0], None)
    treejoints, bodies, names = [], [], []
    for name, j in order:
        bodies.append(_link_body(links[name]))
        if j is None:
            tr = Transform3D.identity(name, WORLD_FRAME)
            treejoints.append(FloatingJoint(tr) if floating_base else FixedJoint(tr))
            names.append("world")
            continue
        xyz, rpy = _origin(j, snap_angles)
        iso = Isometry3(UnitQuaternion.from_euler_angles(*rpy), xyz)
        tr = Transform3D.new(name, j.find("parent").get("link"), iso)
        kind = j.get("type")
        ax = j.find("axis")
        axis = np.array(_floats(ax.get("xyz") if ax is not None else None, 3, (1.0, 0.0, 0.0)))
        if kind in ("revolute", "continuous", "prismatic"):
            nrm = float(np.linalg.norm(axis))
            if nrm == 0.0:
                raise UrdfError(f"joint {j.get('name')} has a zero axis")
            axis = axis / nrm
        if kind in ("revolute", "continuous"):
            treejoints.append(RevoluteJoint(tr, axis))
        elif kind == "prismatic":
            treejoints.append(PrismaticJoint(tr, axis))
        elif kind == "fixed":
            treejoints.append(FixedJoint(tr))
        elif kind == "floating":
            treejoints.append(FloatingJoint(tr))
        else:
            raise UrdfError(f"joint type {kind!r} of {j.get('name')} is not supported")
        names.append(j.get("name"))
    return MechanismState(treejoints, bodies), names
