"""gorilla-physics_amd — MI355X-native batched rigid-body step.

A drop-in for the forward-dynamics path of the reference crate
(MechanismState / step / simulate / dynamics_continuous): hand-written HIP
kernels for gfx950 behind the C ABI of include/gorilla_amd.h, with this
package as the host-side mirror of the reference interface.

The directory name is not a Python identifier; load it with `load()` below
(or importlib) under the module name `gorilla_physics_amd`.
"""
from . import abi  # noqa: F401
from .mechanism import (GRAVITY, WORLD_FRAME, Collider, ContactPoint, FixedJoint, FloatingJoint,  # noqa: F401
                        HalfSpace, Integrator, Isometry3, JointSpring, MechanismState, Pose, PrismaticJoint,
                        RevoluteJoint, RigidBody, SpatialInertia, SpatialVector, Sphere, SpringContact, Transform3D,
                        UnitQuaternion, dynamics_continuous, simulate, step)


def load():
    import sys
    return sys.modules[__name__]
