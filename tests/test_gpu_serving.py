"""GPU tests of the serving loop: continuous batching, memory-based admission,
results independent of the mix, parity with the oracle."""
import os

import numpy as np
import pytest

from gpu_helpers import rel_err

pytestmark = pytest.mark.gpu


def _requests(gp, n, seed=0):
    """A mix of chain / quadruped mechanisms and free cubes with horizons of
    different lengths (the serving workload shape, scaled down)."""
    from gorilla_physics_amd import batch, workloads
    mix = workloads.mixed(n=n, seed=seed)
    cub = workloads.cubes(n=max(1, n // 4), seed=seed)
    cube_model = batch.Model(cub["models"][0])
    rng = np.random.default_rng(seed)
    reqs = []
    qo = vo = 0
    for d in mix["models"]:
        m = batch.Model(d)
        reqs.append(dict(model=m, desc=d, q=mix["q"][qo:qo + d.nq], v=mix["v"][vo:vo + d.nv],
                         tau=mix["tau"][vo:vo + d.nv], dt=1e-3, horizon=int(rng.integers(20, 160))))
        qo += d.nq
        vo += d.nv
    for i in range(len(cub["q"]) // 7):
        reqs.append(dict(model=cube_model, desc=cub["models"][0], q=cub["q"][7 * i:7 * i + 7],
                         v=cub["v"][6 * i:6 * i + 6], tau=np.zeros(6), dt=1e-4, horizon=int(rng.integers(20, 160))))
    order = rng.permutation(len(reqs))
    return [reqs[i] for i in order]


def _serve(gp, reqs, precision, stagger=0.0, **kw):
    from gorilla_physics_amd import serve
    s = serve.Server(precision=precision, chunk_steps=kw.pop("chunk_steps", 16), slots_per_pool=kw.pop("slots", 64),
                     **kw)
    t0 = s.clock()
    ids = [s.submit(r["model"], r["q"], r["v"], r["tau"], r["horizon"], r["dt"], t0 + stagger * k)
           for k, r in enumerate(reqs)]
    s.run()
    assert s.pending() == (0, 0)
    done = set(s.completed().tolist())
    assert done == set(ids)
    out = [s.result(i) for i in ids]
    return out, s


@pytest.mark.parametrize("precision", [0, 1])
def test_result_is_the_same_alone_and_in_the_mix(gp, precision):
    reqs = _requests(gp, 40)
    print(f"case serving alone-vs-mix precision={precision} requests={len(reqs)}", flush=True)
    mixed_out, s = _serve(gp, reqs, precision, stagger=2e-4)
    st = s.stats()
    assert st["completed"] == len(reqs) and st["rounds"] > 1
    for k in range(len(reqs)):
        alone, _ = _serve(gp, [reqs[k]], precision)
        qa, va, _, fa = alone[0]
        qm, vm, _, fm = mixed_out[k]
        assert not fa and not fm
        assert np.array_equal(qa, qm) and np.array_equal(va, vm), k


def test_served_fp64_results_match_the_oracle(gp, oracle):
    reqs = _requests(gp, 24, seed=3)
    print(f"case serving vs oracle requests={len(reqs)}", flush=True)
    out, _ = _serve(gp, reqs, 0, stagger=1e-4)
    for r, (q, v, lat, fault) in zip(reqs, out):
        rq, rv, fails = oracle.rollout([r["desc"]], r["q"], r["v"], r["tau"], r["dt"], r["horizon"])
        assert fails == 0 and not fault
        assert rel_err(q, rq) < 1e-9 and rel_err(v, rv) < 1e-8
        assert lat > 0


def test_more_pools_than_one_launch_table(gp):
    """One slot per pool, every request arriving at once: more pools than one
    grouped-launch table holds (32), so each round admits, steps and retires
    through several tables. The results equal those of the same requests
    served in a few large pools, bit for bit."""
    reqs = _requests(gp, 48, seed=5)
    print(f"case serving one slot per pool requests={len(reqs)}", flush=True)
    out_small, s = _serve(gp, reqs, 0, slots=1)
    assert s.stats()["pools"] > 32
    out_big, _ = _serve(gp, reqs, 0, slots=64)
    for (qa, va, _, fa), (qb, vb, _, fb) in zip(out_small, out_big):
        assert not fa and not fb
        assert np.array_equal(qa, qb) and np.array_equal(va, vb)


def test_memory_budget_admits_and_evicts(gp):
    """With room for a single pool, requests of two topologies still all
    complete: the second topology waits until the first pool is idle and
    released."""
    from gorilla_physics_amd import batch, serve, workloads
    arms = workloads.arms(n=6)
    cube = workloads.cubes(n=3)
    s = serve.Server(precision=1, chunk_steps=10, slots_per_pool=64, mem_budget_bytes=1)
    # a budget of 1 byte cannot hold any pool
    m = batch.Model(arms["models"][0])
    s.submit(m, arms["q"][:7], arms["v"][:7], arms["tau"][:7], 30, 1e-3)
    with pytest.raises(gp.abi.GPError) as e:
        s.run()
    assert e.value.code == gp.abi.GP_ERR_BUSY
    # room for the 7-DOF arm pool (~150 kB at 64 slots) but not for a cube pool beside it
    s = serve.Server(precision=1, chunk_steps=10, slots_per_pool=64, mem_budget_bytes=160_000)
    ids = []
    for i in range(6):
        ids.append(s.submit(batch.Model(arms["models"][i]), arms["q"][7 * i:7 * i + 7], arms["v"][7 * i:7 * i + 7],
                            arms["tau"][7 * i:7 * i + 7], 30, 1e-3))
    cm = batch.Model(cube["models"][0])
    for i in range(3):
        ids.append(s.submit(cm, cube["q"][7 * i:7 * i + 7], cube["v"][6 * i:6 * i + 6], None, 25, 1e-4))
    s.run()
    st = s.stats()
    assert st["completed"] == 9 and st["pools"] == 1 and st["pool_bytes"] <= 160_000
    for i in ids:
        q, v, lat, fault = s.result(i)
        assert np.isfinite(q).all() and not fault


def test_poisson_arrivals_latency(gp):
    from gorilla_physics_amd import serve
    reqs = _requests(gp, 200, seed=5)
    rng = np.random.default_rng(1)
    gaps = rng.exponential(1.0 / 20000.0, len(reqs))
    s = serve.Server(precision=1, chunk_steps=25, slots_per_pool=256)
    t0 = s.clock() + 0.01
    arrivals = t0 + np.cumsum(gaps)
    ids = [s.submit(r["model"], r["q"], r["v"], r["tau"], r["horizon"], r["dt"], float(a))
           for r, a in zip(reqs, arrivals)]
    s.run()
    lat = np.array([s.result(i)[2] for i in ids])
    print(f"p50 {np.percentile(lat, 50) * 1e3:.3f} ms p99 {np.percentile(lat, 99) * 1e3:.3f} ms", flush=True)
    assert (lat > 0).all() and np.percentile(lat, 99) < 5.0
