"""gp_batch_energy (MechanismState::kinetic_energy / gravitational_energy,
mechanism.rs:334-362) on the GPU against the oracle: kinetic energy of the
world-frame twists and inertias (inertia.rs:182-202), gravitational energy of
each body frame's origin, on random states of every workload shape."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _instances(w):
    descs = []
    for d, c in zip(w["models"], w["counts"]):
        descs.extend([d] * int(c))
    return descs


@pytest.mark.parametrize("name", ["arms", "mixed", "cubes", "acrobot"])
def test_energy_matches_the_oracle(gp, oracle, name):
    from gorilla_physics_amd import batch, workloads
    w = getattr(workloads, name)(n=48)
    print(f"case energy {name}", flush=True)
    models = [batch.Model(d) for d in w["models"]]
    b = batch.Batch(models, w["counts"], precision=0)
    b.set_state(w["q"], w["v"])
    ke, pe = b.energy()
    descs = _instances(w)
    qo = vo = 0
    worst = 0.0
    for i, d in enumerate(descs):
        q, v = w["q"][qo:qo + d.nq], w["v"][vo:vo + d.nv]
        rk, _ = oracle.energy(d, q, v)
        rp = oracle.gravitational_energy(d, q)
        worst = max(worst, abs(ke[i] - rk) / max(1.0, abs(rk)), abs(pe[i] - rp) / max(1.0, abs(rp)))
        qo += d.nq
        vo += d.nv
    print(f"worst relative difference {worst:.3g}", flush=True)
    assert worst < 1e-12


def test_mirror_energy_methods(gp, oracle):
    """MechanismState.kinetic_energy / gravitational_energy of the Python
    mirror: a cube spinning and moving at height h."""
    from gorilla_physics_amd import builders
    st = builders.build_cube(2.0, 1.0)
    st.update([gp.Pose(gp.UnitQuaternion.from_axis_angle([0, 0, 1.0], 0.3), [1.0, 2.0, 3.0])],
              [gp.SpatialVector([0.1, 0.2, 0.3], [1.0, -1.0, 0.5])])
    ke, pe = st.kinetic_energy(), st.gravitational_energy()
    # cube: J = m l^2 / 6 I about its centre, the body origin
    assert abs(ke - 0.5 * (2.0 * (1.0 + 1.0 + 0.25) + 2.0 / 6.0 * (0.01 + 0.04 + 0.09))) < 1e-12
    assert abs(pe - 2.0 * 9.81 * 3.0) < 1e-12
