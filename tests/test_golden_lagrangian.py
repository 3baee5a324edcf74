"""The oracle against brute-force Lagrangian dynamics (tests/golden/
lagrangian_chains.json, 50-digit mpmath from forward kinematics and energies
only): 28 random revolute/prismatic chains and trees of 1-3 bodies with
general tree rotations, axes, inertias and prismatic springs."""
import json
import os

import numpy as np
import pytest

from oracle.oracle import ABA, CRBA

HERE = os.path.dirname(os.path.abspath(__file__))


def load_cases():
    return json.load(open(os.path.join(HERE, "golden", "lagrangian_chains.json")))["cases"]


def desc_from_fixture(m):
    from conftest import load_package
    load_package()
    from gorilla_physics_amd import abi
    from gorilla_physics_amd.mechanism import ModelDesc
    bodies = []
    for b in m["bodies"]:
        d = abi.gp_body_desc()
        d.parent, d.jtype = b["parent"], b["jtype"]
        d.axis[:] = b["axis"]
        d.x_tree_quat[:] = b["x_tree_quat"]
        d.x_tree_trans[:] = b["x_tree_trans"]
        d.moment[:] = b["moment"]
        d.mc[:] = b["mc"]
        d.mass, d.spring_k, d.spring_l = b["mass"], b["spring_k"], b["spring_l"]
        bodies.append(d)
    return ModelDesc(bodies, [], [], m["gravity"])


@pytest.mark.parametrize("method", [CRBA, ABA])
def test_oracle_matches_lagrangian(gp, oracle, method):
    cases = load_cases()
    assert len(cases) == 28
    worst = 0.0
    for c in cases:
        d = desc_from_fixture(c["model"])
        a = oracle.dynamics(d, c["q"], c["v"], c["tau"], method)
        ref = np.array([float(x) for x in c["qdd"]])
        worst = max(worst, np.abs(a - ref).max() / max(np.abs(ref).max(), 1.0))
    assert worst < 1e-12, worst


def test_lagrangian_fixture_catches_a_dropped_term(gp, oracle):
    """A plausible mistake (gravity with the wrong sign) is far outside the
    tolerance on every case: the fixture pins what it is meant to pin."""
    for c in load_cases()[:6]:
        m = dict(c["model"], gravity=-c["model"]["gravity"])
        a = oracle.dynamics(desc_from_fixture(m), c["q"], c["v"], c["tau"], CRBA)
        ref = np.array([float(x) for x in c["qdd"]])
        assert np.abs(a - ref).max() > 1e-3
