"""URDF ingest (gorilla-physics_amd/urdf.py): the SO-101's description read
from URDF text gives the same model as the reference's hand-written
build_so101 (builders/mod.rs:250-330, the mirror's builders.build_so101), and
the reader's conventions are pinned by closed forms: the rpy rotation (a
pendulum whose joint origin is pitched a quarter turn), the inertia about the
frame origin (parallel axes), axis normalisation and joint kinds, a
floating base, sphere collisions, parents before children, and refusals."""
import math

import numpy as np
import pytest

G = 9.81

# The SO-101's links and joints as its URDF lists them (the numbers
# builders/mod.rs:250-330 transcribes); visuals and limits left out.
SO101_URDF = """<?xml version="1.0"?>
<robot name="so101">
  <link name="base"><inertial><origin xyz="0.0137179 -5.19711e-05 0.0334843" rpy="0 0 0"/><mass value="0.147"/>
    <inertia ixx="0.000114686" ixy="-4.59787e-07" ixz="4.97151e-06" iyy="0.000136117" iyz="9.75275e-08"
     izz="0.000130364"/></inertial></link>
  <link name="shoulder"><inertial><origin xyz="-0.0307604 -1.66727e-05 -0.0252713"/><mass value="0.100006"/>
    <inertia ixx="8.3759e-05" ixy="7.55525e-08" ixz="-1.16342e-06" iyy="8.10403e-05" iyz="1.54663e-07"
     izz="2.39783e-05"/></inertial></link>
  <link name="upper_arm"><inertial><origin xyz="-0.0898471 -0.00838224 0.0184089"/><mass value="0.103"/>
    <inertia ixx="4.08002e-05" ixy="-1.97819e-05" ixz="-4.03016e-08" iyy="0.000147318" iyz="8.97326e-09"
     izz="0.000142487"/></inertial></link>
  <link name="lower_arm"><inertial><origin xyz="-0.0980701 0.00324376 0.0182831"/><mass value="0.104"/>
    <inertia ixx="2.87438e-05" ixy="7.41152e-06" ixz="1.26409e-06" iyy="0.000159844" iyz="-4.90188e-08"
     izz="0.00014529"/></inertial></link>
  <link name="wrist"><inertial><origin xyz="-0.000103312 -0.0386143 0.0281156"/><mass value="0.079"/>
    <inertia ixx="3.68263e-05" ixy="1.7893e-08" ixz="-5.28128e-08" iyy="2.5391e-05" iyz="3.6412e-06"
     izz="2.1e-05"/></inertial></link>
  <link name="gripper"><inertial><origin xyz="0.000213627 0.000245138 -0.025187"/><mass value="0.087"/>
    <inertia ixx="2.75087e-05" ixy="-3.35241e-07" ixz="-5.7352e-06" iyy="4.33657e-05" iyz="-5.17847e-08"
     izz="3.45059e-05"/></inertial></link>
  <link name="jaw"><inertial><origin xyz="-0.00157495 -0.0300244 0.0192755"/><mass value="0.012"/>
    <inertia ixx="6.61427e-06" ixy="-3.19807e-07" ixz="-5.90717e-09" iyy="1.89032e-06" iyz="-1.09945e-07"
     izz="5.28738e-06"/></inertial></link>
  <joint name="shoulder_pan" type="revolute"><parent link="base"/><child link="shoulder"/>
    <origin xyz="0.0388353 -8.97657e-09 0.0624" rpy="3.14159 4.18253e-17 -3.14159"/><axis xyz="0 0 1"/></joint>
  <joint name="shoulder_lift" type="revolute"><parent link="shoulder"/><child link="upper_arm"/>
    <origin xyz="-0.0303992 -0.0182778 -0.054" rpy="-1.5708 -1.5708 0"/><axis xyz="0 0 1"/></joint>
  <joint name="elbow_flex" type="revolute"><parent link="upper_arm"/><child link="lower_arm"/>
    <origin xyz="-0.11257 -0.028 1.73763e-16" rpy="-3.63608e-16 8.74301e-16 1.5708"/><axis xyz="0 0 1"/></joint>
  <joint name="wrist_flex" type="revolute"><parent link="lower_arm"/><child link="wrist"/>
    <origin xyz="-0.1349 0.0052 3.62355e-17" rpy="4.02456e-15 8.67362e-16 -1.5708"/><axis xyz="0 0 1"/></joint>
  <joint name="wrist_roll" type="revolute"><parent link="wrist"/><child link="gripper"/>
    <origin xyz="5.55112e-17 -0.0611 0.0181" rpy="1.5708 0.0486795 3.14159"/><axis xyz="0 0 1"/></joint>
  <joint name="gripper" type="revolute"><parent link="gripper"/><child link="jaw"/>
    <origin xyz="0.0202 0.0188 -0.0234" rpy="1.5708 -5.24284e-08 -1.41553e-15"/><axis xyz="0 0 1"/></joint>
</robot>
"""


def _urdf(gp):
    import importlib
    return importlib.import_module("gorilla_physics_amd.urdf")


def _bodies(d):
    return [d._bodies[i] for i in range(d.n_bodies)]


def test_so101_from_urdf_equals_the_builder(gp, oracle):
    from gorilla_physics_amd import builders
    st, names = _urdf(gp).parse_urdf(SO101_URDF)
    assert names == ["world", "shoulder_pan", "shoulder_lift", "elbow_flex", "wrist_flex", "wrist_roll", "gripper"]
    a, b = st.model_desc(), builders.build_so101().model_desc()
    for x, y in zip(_bodies(a), _bodies(b)):
        assert (x.parent, x.jtype) == (y.parent, y.jtype)
        for f in ("axis", "x_tree_quat", "x_tree_trans", "moment", "mc"):
            assert np.allclose(list(getattr(x, f)), list(getattr(y, f)), rtol=0, atol=1e-15), f
        assert x.mass == y.mass
    rng = np.random.default_rng(1)
    for _ in range(5):
        q, v, t = rng.uniform(-2, 2, 6), rng.normal(size=6), rng.normal(size=6) * 0.02
        assert np.allclose(oracle.dynamics(a, q, v, t), oracle.dynamics(b, q, v, t), rtol=1e-13, atol=1e-13)


PENDULUM = """<robot name="p">
  <link name="top"/>
  <link name="bob"><inertial><origin xyz="0 0 -{l}" rpy="0.3 -0.2 0.7"/><mass value="{m}"/>
    <inertia ixx="{i}" ixy="0" ixz="0" iyy="{i}" iyz="0" izz="{i}"/></inertial></link>
  <joint name="hinge" type="{kind}"><parent link="top"/><child link="bob"/>
    <origin xyz="0.1 0.2 0.3" rpy="0 {pitch} 0"/><axis xyz="0 {ay} 0"/></joint>
</robot>"""


@pytest.mark.parametrize("kind", ["revolute", "continuous"])
def test_pendulum_closed_form(gp, oracle, kind):
    """A point-like bob (inertia i about its centre, any inertial rpy: the
    tensor is isotropic) a distance l below a hinge whose origin is pitched
    by a quarter turn: the hinge axis y stays horizontal, and at angle q the
    bob's lever is l sin(q + pi/2)-like: qdd = -m g l cos(q) / (i + m l^2) for
    a hinge rotated +pi/2 about y (the rod then points along world +x at q = 0,
    rpy = from_euler_angles). An axis given as (0, 2, 0) is normalised."""
    m, l, i = 1.3, 0.7, 0.01
    st, _ = _urdf(gp).parse_urdf(PENDULUM.format(l=l, m=m, i=i, kind=kind, pitch=math.pi / 2, ay=2.0))
    d = st.model_desc()
    assert d._bodies[1].jtype == 1 and list(d._bodies[1].axis) == [0.0, 1.0, 0.0]
    for q in (-1.0, 0.0, 0.4, 2.0):
        qdd = oracle.dynamics(d, [q], [0.0], [0.0])[0]
        # rod direction in world: Ry(pi/2) Ry(q) (0, 0, -1) = (-sin(pi/2 + q), 0, -cos(pi/2 + q))
        # torque of gravity about y: tau_y = (r x F)_y = r_z F_x - r_x F_z = -r_x (-m g) ... with r = l dir
        rx = -l * math.sin(math.pi / 2 + q)
        expect = (rx * m * G) / (i + m * l * l)
        assert abs(qdd - expect) < 1e-12 * max(1.0, abs(expect)), (q, qdd, expect)


def test_inertia_about_the_frame_origin(gp):
    """Parallel axes: the moment about the link origin of a body whose
    inertial origin is c with an inertia tensor rotated by the inertial rpy is
    R I_c R^T + m (|c|^2 1 - c c^T), and the cross part is m c."""
    U = _urdf(gp)
    text = """<robot name="b"><link name="a"><inertial><origin xyz="0.1 -0.2 0.3" rpy="0.5 0 0"/>
      <mass value="2"/><inertia ixx="1" ixy="0" ixz="0" iyy="2" iyz="0" izz="3"/></inertial></link></robot>"""
    st, names = U.parse_urdf(text, floating_base=True)
    b = st.model_desc()._bodies[0]
    assert names == ["world"] and b.jtype == 3
    c = np.array([0.1, -0.2, 0.3])
    ca, sa = math.cos(0.5), math.sin(0.5)
    R = np.array([[1, 0, 0], [0, ca, -sa], [0, sa, ca]])
    J = R @ np.diag([1.0, 2.0, 3.0]) @ R.T + 2.0 * (c @ c * np.eye(3) - np.outer(c, c))
    assert np.allclose(list(b.moment), [J[0, 0], J[1, 1], J[2, 2], J[0, 1], J[0, 2], J[1, 2]], atol=1e-15)
    assert np.allclose(list(b.mc), 2.0 * c, atol=1e-16) and b.mass == 2.0


def test_tree_order_kinds_and_spheres(gp):
    """Bodies come parents first whatever the document's order; prismatic,
    fixed and floating joints map to the reference's kinds; a collision
    sphere becomes the body's sphere collider."""
    U = _urdf(gp)
    text = """<robot name="t">
      <joint name="j3" type="fixed"><parent link="b"/><child link="d"/></joint>
      <link name="d"/><link name="c"/>
      <joint name="j2" type="prismatic"><parent link="a"/><child link="c"/><axis xyz="0 0 3"/></joint>
      <link name="a"><inertial><mass value="1"/><inertia ixx="1" iyy="1" izz="1"/></inertial></link>
      <joint name="j1" type="floating"><parent link="a"/><child link="b"/></joint>
      <link name="b"><inertial><mass value="1"/><inertia ixx="1" iyy="1" izz="1"/></inertial>
        <collision><origin xyz="0 0 0.25"/><geometry><sphere radius="0.1"/></geometry></collision></link>
    </robot>"""
    st, names = U.parse_urdf(text)
    d = st.model_desc()
    order = [(d._bodies[i].parent, d._bodies[i].jtype) for i in range(d.n_bodies)]
    assert names == ["world", "j2", "j1", "j3"]  # depth first, joints in document order
    assert order == [(-1, 0), (0, 2), (0, 3), (2, 0)]
    assert list(d._bodies[1].axis) == [0.0, 0.0, 1.0]
    assert d.desc.n_spheres == 1 and d._spheres[0].body == 2 and d._spheres[0].radius == 0.1
    assert list(d._spheres[0].center) == [0.0, 0.0, 0.25]


def test_refusals(gp):
    U = _urdf(gp)
    with pytest.raises(U.UrdfError):
        U.parse_urdf('<robot name="x"><link name="a"/><link name="b"/></robot>')  # two roots
    with pytest.raises(U.UrdfError):
        U.parse_urdf('<robot name="x"><link name="a"/><link name="b"/><joint name="j" type="planar">'
                     '<parent link="a"/><child link="b"/></joint></robot>')
    with pytest.raises(U.UrdfError):
        U.parse_urdf('<robot name="x"><link name="a"/><joint name="j" type="fixed"><parent link="a"/>'
                     '<child link="zz"/></joint></robot>')
