"""GPU tests of gp_batch_autotune: it leaves the state as it found it, picks
valid options, and the tuned batch still meets the parity bounds."""
import numpy as np
import pytest

from gpu_helpers import rel_err, step_errors

pytestmark = pytest.mark.gpu


def _descs(w):
    out = []
    for d, c in zip(w["models"], w["counts"]):
        out.extend([d] * int(c))
    return out


@pytest.mark.parametrize("name", ["cubes", "mixed"])
def test_autotune_restores_state_and_keeps_parity(gp, oracle, name):
    from gorilla_physics_amd import abi, batch, workloads
    w = getattr(workloads, name)(n=2048)
    print(f"case autotune {name}", flush=True)
    b, _ = batch.batch_from_workload(w, precision=abi.GP_PREC_F64, use_graph=1)
    q0, v0 = b.get_state()
    chosen, ms = b.autotune(w["dt"], 0, trial_steps=50)
    print(f"chosen lanes_per_mech={chosen.lanes_per_mech} lanes_per_tree={chosen.lanes_per_tree} "
          f"block={chosen.block_size} steps_per_launch={chosen.steps_per_launch} trial {ms:.3f} ms", flush=True)
    assert ms > 0 and chosen.use_graph == 1 and chosen.precision == abi.GP_PREC_F64
    assert chosen.lanes_per_mech in (0, 1, 2, 4, 8) and chosen.lanes_per_tree in (0, 1, 4)
    assert chosen.block_size in (0, 64, 128, 256) and chosen.steps_per_launch in (0, 25, 50, 100, 200)
    o = b.get_opts()
    assert (o.lanes_per_mech, o.lanes_per_tree, o.block_size, o.steps_per_launch) == \
        (chosen.lanes_per_mech, chosen.lanes_per_tree, chosen.block_size, chosen.steps_per_launch)
    q1, v1 = b.get_state()
    assert np.array_equal(q0, q1) and np.array_equal(v0, v1) and not b.faults().any()
    b.step(w["dt"], 0, 1)
    q2, v2 = b.get_state()
    a2 = b.get_vdot()
    eq, ev, ea = step_errors(oracle, _descs(w), q0, v0, w["tau"], q2, v2, a2, w["dt"])
    assert eq.max() <= 1e-10 and ev.max() <= 1e-10 and ea.max() <= 1e-10


def test_autotuned_options_do_not_change_bits_of_block_and_chunk(gp):
    """Workgroup size and steps per launch never change a result bit."""
    from gorilla_physics_amd import abi, batch, workloads
    w = workloads.arms(n=1000)
    print("case autotune bits", flush=True)
    out = []
    for block, spl in ((64, 25), (256, 200), (128, 100)):
        b, _ = batch.batch_from_workload(w, precision=abi.GP_PREC_MIXED, block_size=block, steps_per_launch=spl)
        b.step(w["dt"], 0, 200)
        out.append(b.get_state())
    for q, v in out[1:]:
        assert np.array_equal(q, out[0][0]) and np.array_equal(v, out[0][1])
