"""The reference's browser-client interface (interface/mod.rs:76-230), as a
client of this library: the same calls and array formats, stepping through
gp_step with continuous-collision velocity stepping as the reference's
InterfaceSimulator does (interface/mod.rs:92-118).

A controller is anything with control(state, input) -> list of joint
torques (the Controller trait, control/mod.rs:32-38); NullController returns
no torques (zero torques).
"""
import numpy as np

from .mechanism import HalfSpace, Integrator, Pose, flatten_q, step


class ControlInput:  # control/mod.rs:22-30
    def __init__(self, floats):
        self.floats = [float(x) for x in floats]


class NullController:  # control/mod.rs:40-48
    def control(self, state, input=None):
        return []


def _standing(state, q_joints, translation):
    """Set the floating root to the identity rotation at `translation` and
    the remaining joints to q_joints."""
    state.set_joint_q(1, Pose(None, translation))
    for k, a in enumerate(q_joints):
        state.set_joint_q(k + 2, float(a))
    return state


def create_biped():  # interface/biped.rs:16-56
    """build_biped with both legs bent (thigh -pi/4, calf pi/2, ankle -pi/4),
    lowered so that the last body's (the right foot's) frame origin is at
    z = 0."""
    from .builders import build_biped
    leg = [0.0, 0.0, -np.pi / 4.0, np.pi / 2.0, -np.pi / 4.0, 0.0]
    state = _standing(build_biped(), leg + leg, [0.0, 0.0, 0.0])
    foot = state.poses()[-1].translation
    return _standing(state, leg + leg, [0.0, 0.0, -float(foot[2])])


def create_leg():  # interface/biped.rs:59-95
    """build_leg bent the same way (thigh -pi/4, calf pi/2, foot -pi/4),
    moved so that the foot's frame origin is at the world origin."""
    from .builders import build_leg
    leg = [0.0, 0.0, -np.pi / 4.0, np.pi / 2.0, -np.pi / 4.0]
    state = _standing(build_leg(), leg, [0.0, 0.0, 0.0])
    foot = np.asarray(state.poses()[-1].translation, dtype=float)
    return _standing(state, leg, list(-foot))


def create_leg_from_foot():  # interface/biped.rs:98-124
    """build_leg_from_foot with the foot at the origin and calf pi/4 - 0.4,
    thigh -pi/2, hip pi/4, pelvis 0.1, base pi/2."""
    from .builders import build_leg_from_foot
    return _standing(build_leg_from_foot(), [np.pi / 4.0 - 0.4, -np.pi / 2.0, np.pi / 4.0, 0.1, np.pi / 2.0],
                     [0.0, 0.0, 0.0])


class InterfaceSimulator:
    """InterfaceSimulator (interface/mod.rs:76-160): one mechanism, stepped
    by CCD velocity stepping; poses and contact positions as flat arrays."""

    def __init__(self, state, controller=None):
        self.state = state
        self.controller = controller or NullController()

    def step(self, dt, control_input=()):
        """One step of dt with the controller's torques for this input;
        returns the new q flattened (interface/mod.rs:92-118)."""
        tau = self.controller.control(self.state, ControlInput(control_input))
        q, _ = step(self.state, dt, tau, Integrator.CCDVelocityStepping)
        return np.array(flatten_q(q), dtype=float)

    def poses(self):
        """[qx, qy, qz, qw, tx, ty, tz] of every body's world pose
        (InterfaceMechanismState::poses, interface/mod.rs:172-194)."""
        out = []
        for p in self.state.poses():
            out.extend(p.rotation.coords())
            out.extend(float(x) for x in p.translation)
        return np.array(out, dtype=float)

    def contact_positions(self):
        """World position of every contact point, body by body
        (interface/mod.rs:197-220)."""
        out = []
        for pose, body in zip(self.state.poses(), self.state.bodies):
            R = pose.rotation.to_matrix()
            t = np.asarray(pose.translation, dtype=float)
            for cp in body.contact_points:
                out.extend(R @ np.asarray(cp.location, dtype=float) + t)
        return np.array(out, dtype=float)

    def addHalfSpace(self, normal, distance):  # interface/mod.rs:222-230
        n = np.asarray(normal, dtype=float)
        self.state.add_halfspace(HalfSpace.new(n / np.linalg.norm(n), float(distance)))
