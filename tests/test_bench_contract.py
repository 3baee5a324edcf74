"""bench.py's output line: the fields and units a caller reads, and that
its throughput agrees with its own per-step time. The GPU runs use small
instance counts so that they take seconds."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args, timeout=240, rc=0):
    p = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                       capture_output=True, text=True, timeout=timeout)
    lines = [l for l in p.stdout.splitlines() if l.startswith("{")]
    assert p.returncode == rc and lines, (p.returncode, p.stdout[-2000:], p.stderr[-2000:])
    return json.loads(lines[-1])


def test_more_than_one_gpu_is_refused_with_a_message():
    out = _bench("--gpus", "2", timeout=120, rc=2)
    assert "one GPU" in out["error"]


def _check_line(out, instances, dtype):
    for key, typ in (("metric", str), ("value", float), ("unit", str), ("n_gpus", int), ("steps", int),
                     ("warmup", int), ("ms_per_step", float), ("higher_is_better", bool), ("scaling", str),
                     ("dtype", str), ("data", str), ("config", dict)):
        assert isinstance(out[key], typ), key
    assert out["unit"] == "env-steps/s" and out["n_gpus"] == 1 and out["higher_is_better"] is True
    assert out["dtype"] == dtype and "workload" in out["config"] and out["config"]["instances"] == instances
    assert out["vs_baseline"] is None
    per_step = instances / (out["ms_per_step"] / 1e3)
    assert abs(per_step - out["value"]) <= 1e-6 * out["value"]


@pytest.mark.gpu
def test_arms_line_with_roofline(gp):
    print("case bench arms 4096", flush=True)
    out = _bench("--instances", "4096", "--steps", "200", "--warmup", "200", "--no-serve", "--no-cpu-baseline")
    _check_line(out, 4096, "f64")
    for r in (out["roofline"], out["mixed_precision"]["roofline"]):
        assert set(r) >= {"bound", "achieved", "peak", "unit", "frac", "traffic"}
        assert r["unit"] == "TFLOP/s" and 0 < r["frac"] < 1 and abs(r["achieved"] / r["peak"] - r["frac"]) < 1e-12
    # FP64 kernel: rated at the FP64 peak; the mix between the two peaks
    assert out["roofline"]["peak"] == 78.6
    assert 78.6 < out["mixed_precision"]["roofline"]["peak"] < 157.3
    assert out["valid_value"] <= out["value"]


@pytest.mark.gpu
def test_timed_call_replays_prepared_graphs(gp):
    """The driver's shape (20 timed steps after 5 warmup steps): the graphs
    of both are built before either runs, so the HIP-event time of the timed
    launch is the kernel's, not a capture's."""
    print("case bench arms 4096 steps 20 warmup 5", flush=True)
    out = _bench("--instances", "4096", "--steps", "20", "--warmup", "5", "--no-serve", "--no-cpu-baseline",
                 "--no-compare")
    r = out["roofline"]
    assert r["launches"] == 1 and r["steps_per_launch"] == 20
    # wall-clock per step includes one launch and a synchronisation, not a capture or an instantiation
    assert r["launch_ms"] <= out["ms_per_step"] * 20


@pytest.mark.gpu
def test_velocity_stepping_line(gp):
    print("case bench thrown cubes 256", flush=True)
    out = _bench("--workload", "thrown_cubes", "--instances", "256", "--steps", "10", "--warmup", "2",
                 "--no-serve", "--no-cpu-baseline")
    _check_line(out, 256, "f64")
    assert out["faults"] == 0
