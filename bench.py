"""Benchmark of the batched rigid-body step on one MI355X.

Default workload: 65,536 serial 7-DOF revolute arms (random links, axes,
masses, inertias, constant torques), dt = 1e-3, semi-implicit Euler, mixed
precision (FP32 sweeps, FP64 D solve and q/v accumulation). One "step" is one
dt of every instance. Prints one JSON line with
env-steps/s, the dominant kernel's VALU roofline (algorithmic flops from the
ABA op count, per-launch duration from HIP events on the batch's stream),
the plain-FP64 and fast-FP64 kernels measured on the same inputs, and the
CPU baseline (the oracle's restatement of the reference algorithm, timed on
the host's cores over a bounded sample).

    python bench.py [--steps K] [--warmup W] [--workload arms|mixed|cubes|acrobot]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))

PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}  # MI355X published vector peaks
WORK_TABLE = os.path.join(ROOT, "gorilla-physics_amd", "algorithmic_work.json")


def load_package():
    name = "gorilla_physics_amd"
    if name in sys.modules:
        return sys.modules[name]
    pkg = os.path.join(ROOT, "gorilla-physics_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name, n):
    from gorilla_physics_amd import workloads
    if name == "arms":
        return workloads.arms(n=n or 65536)
    if name == "mixed":
        return workloads.mixed(n=n or 32768)
    if name == "cubes":
        return workloads.cubes(n=n or 16384)
    if name == "acrobot":
        return workloads.acrobot(n=n or 1024)
    if name == "thrown_cubes":
        return workloads.thrown_cubes(n=n or 16384)
    if name == "so101":
        return workloads.so101(n=n or 65536)
    raise SystemExit(f"unknown workload {name}")


def run_variant(w, models, steps, warmup, autotune=False, **opts):
    """Returns env-steps/s, ms per step, lib-timed ms, launches, faults and the
    options the batch ran with (after gp_batch_autotune when asked; the tuning
    runs before the warmup, outside the timed region, and restores the state)."""
    from gorilla_physics_amd import batch
    b = batch.Batch(models, w["counts"], **opts)
    b.set_state(w["q"], w["v"])
    b.set_tau(w["tau"])
    tuned_ms = None
    integrator = w.get("integrator", 0)
    if autotune:
        # trials long enough that launch-to-launch noise does not pick the options
        _, tuned_ms = b.autotune(w["dt"], integrator, trial_steps=min(steps, 1000))
    o = b.get_opts()
    chosen = dict(block_size=o.block_size, lanes_per_mech=o.lanes_per_mech, lanes_per_tree=o.lanes_per_tree,
                  steps_per_launch=o.steps_per_launch, use_graph=o.use_graph)
    if warmup:
        b.step(w["dt"], integrator, warmup)
    b.sync()
    b.set_state(w["q"], w["v"])
    b.clear_faults()
    b.sync()
    t0 = time.perf_counter()
    b.step(w["dt"], integrator, steps)
    b.sync()
    t1 = time.perf_counter()
    lib_ms, launches = b.last_timing()
    faults = int(b.faults().sum())
    n = b.n_instances
    sec = t1 - t0
    del b
    return dict(value=n * steps / sec, ms_per_step=1e3 * sec / steps, lib_ms=lib_ms, launches=launches,
                faults=faults, instances=n, opts=chosen, autotune_trial_ms=tuned_ms)


def serving_requests(n, seed=0):
    """Requests mixing the mixed-workload mechanisms (chains and quadrupeds,
    dt 1e-3) and free cubes on a plane (dt 1e-4), horizons U{100..2000}."""
    from gorilla_physics_amd import batch, workloads
    n_cubes = n // 4
    mix = workloads.mixed(n=n - n_cubes, seed=seed)
    cub = workloads.cubes(n=max(1, n_cubes), seed=seed)
    cube_model = batch.Model(cub["models"][0])
    rng = np.random.default_rng(seed + 17)
    reqs, qo, vo = [], 0, 0
    for d in mix["models"]:
        reqs.append(dict(model=batch.Model(d), q=mix["q"][qo:qo + d.nq], v=mix["v"][vo:vo + d.nv],
                         tau=mix["tau"][vo:vo + d.nv], dt=1e-3))
        qo += d.nq
        vo += d.nv
    for i in range(n_cubes):
        reqs.append(dict(model=cube_model, q=cub["q"][7 * i:7 * i + 7], v=cub["v"][6 * i:6 * i + 6],
                         tau=np.zeros(6), dt=1e-4))
    order = rng.permutation(len(reqs))
    reqs = [reqs[i] for i in order]
    for r in reqs:
        r["horizon"] = int(rng.integers(100, 2001))
    return reqs


def serve_run(reqs, rate, chunk=50, slots=4096, seed=1):
    """Poisson arrivals at `rate` requests/s into one running server; returns
    throughput and latency percentiles. The server is warm, as a long-running
    one is: one short request per distinct mechanism shape runs first
    (untimed), so its pools exist before the measured arrivals."""
    from gorilla_physics_amd import abi, serve
    s = serve.Server(precision=abi.GP_PREC_MIXED, chunk_steps=chunk, slots_per_pool=slots)
    seen = set()
    for r in reqs:
        d = r["model"].desc
        all_rev = all(d._bodies[i].jtype != abi.GP_JOINT_PRISMATIC for i in range(d.n_bodies))
        key = (r["model"].signature(), r["dt"], all_rev)
        if key not in seen:
            seen.add(key)
            s.submit(r["model"], r["q"], r["v"], r["tau"], 1, r["dt"])
    s.run()
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, len(reqs))
    st0 = s.stats()
    t0 = s.clock() + 0.02
    arrivals = t0 + np.cumsum(gaps)
    ids = [s.submit(r["model"], r["q"], r["v"], r["tau"], r["horizon"], r["dt"], float(a))
           for r, a in zip(reqs, arrivals)]
    s.run()
    t_end = s.clock()
    lat = np.array([s.result(i)[2] for i in ids])
    st = s.stats()
    steps = float(sum(r["horizon"] for r in reqs))
    span = t_end - arrivals[0]
    return dict(rate=rate, requests=len(reqs), env_steps_per_s=steps / span, requests_per_s=len(reqs) / span,
                p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                rounds=int(st["rounds"] - st0["rounds"]), pools=int(st["pools"]), gpu_busy_s=st["busy_s"] - st0["busy_s"],
                makespan_s=span)


def serve_baseline_plain(reqs, rate, seed=1):
    """Plain FP64 per-request launches: each request alone in its own batch,
    one plain-kernel launch per step, served first-in first-out. Service times
    are measured; the FIFO completion times replay the same Poisson arrivals."""
    from gorilla_physics_amd import abi, batch
    service = []
    for r in reqs:
        b = batch.Batch([r["model"]], [1], precision=abi.GP_PREC_F64, kernel=abi.GP_KERNEL_PLAIN)
        b.set_state(r["q"], r["v"])
        b.set_tau(r["tau"])
        b.sync()
        t0 = time.perf_counter()
        b.step(r["dt"], 0, r["horizon"])
        b.get_state()
        service.append(time.perf_counter() - t0)
        del b
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, len(reqs))
    arrivals = np.cumsum(gaps)
    done, lat = 0.0, []
    for a, sv in zip(arrivals, service):
        done = max(a, done) + sv
        lat.append(done - a)
    steps = float(sum(r["horizon"] for r in reqs))
    return dict(rate=rate, requests=len(reqs), env_steps_per_s=steps / (done - arrivals[0]),
                p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                mean_service_ms=float(np.mean(service) * 1e3))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(w, budget_s=12.0):
    """The oracle (CRBA + RNEA + LU, the reference's algorithm) over a bounded
    sample of the same workload, one instance per thread at a time."""
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from oracle.oracle import Oracle
    o = Oracle()
    cores = min(16, os.cpu_count() or 1)
    descs = []
    for d, c in zip(w["models"], w["counts"]):
        descs.extend([d] * int(c))
    n_all = len(descs)
    nq = [d.nq for d in descs]
    nv = [d.nv for d in descs]
    qo = np.concatenate([[0], np.cumsum(nq)])
    vo = np.concatenate([[0], np.cumsum(nv)])
    # calibrate on a small slice, then size the sample to about budget_s
    n = min(n_all, 64 * cores)
    steps = 50
    t0 = time.perf_counter()
    integrator = w.get("integrator", 0)
    o.rollout(descs[:n], w["q"][:qo[n]], w["v"][:vo[n]], w["tau"][:vo[n]], w["dt"], steps, integrator=integrator,
              threads=cores)
    dt_cal = max(time.perf_counter() - t0, 1e-6)
    rate = n * steps / dt_cal
    target = rate * budget_s
    n = int(min(n_all, max(cores, target // w["steps"])))
    steps = int(min(w["steps"], max(1, target // n)))
    t0 = time.perf_counter()
    o.rollout(descs[:n], w["q"][:qo[n]], w["v"][:vo[n]], w["tau"][:vo[n]], w["dt"], steps, integrator=integrator,
              threads=cores)
    sec = time.perf_counter() - t0
    algo = ("velocity stepping: CRBA + LU free velocity, interior-point contact impulses"
            if integrator >= abi_vs() else "CRBA+RNEA+LU")
    return dict(value=n * steps / sec, unit="env-steps/s", cores=cores, kind="port",
                sample=f"{n} of {n_all} instances x {steps} steps ({sec:.1f} s), {algo} FP64 oracle (g++ -O3), "
                       f"{cores} threads of {os.cpu_count()} on {cpu_model()}")


def abi_vs():
    from gorilla_physics_amd import abi
    return abi.GP_VELOCITY_STEPPING


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps (0 = the workload's own: arms and mixed "
                                                            "1000, cubes 5000, acrobots 2000)")
    ap.add_argument("--warmup", type=int, default=-1, help="untimed steps (default: one launch's worth)")
    ap.add_argument("--workload", default="arms")
    ap.add_argument("--instances", type=int, default=0)
    ap.add_argument("--steps-per-launch", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--autotune", type=int, default=-1,
                    help="gp_batch_autotune before the warmup (1), or fixed options (0); default: on for the "
                         "multi-bucket workloads (mixed, cubes), off for the arms and acrobots, which run the "
                         "options the tuner picks for them (200 steps per launch, 256-lane workgroups) so that "
                         "a profiler run of the same command sees the same launches")
    ap.add_argument("--lanes-tree", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rates", default="20000")
    ap.add_argument("--serve-chunk", type=int, default=50)
    ap.add_argument("--serve-baseline", type=int, default=100, help="requests replayed with plain FP64 launches")
    ap.add_argument("--no-serve", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: instances are independent, so several "
                                   "GPUs would each run their own share of the instances; pass --gpus 1"}))
        sys.exit(2)
    load_package()
    from gorilla_physics_amd import abi, batch
    serving = None
    if not args.no_serve:
        reqs = serving_requests(args.serve_requests)
        rates = [float(x) for x in args.serve_rates.split(",")]
        sweep = []
        for rate in rates:
            res = serve_run(reqs, rate, chunk=args.serve_chunk)
            log("serve", res)
            sweep.append(res)
        base = None
        if args.serve_baseline:
            base = serve_baseline_plain(reqs[:args.serve_baseline], rates[0])
            log("serve baseline (plain FP64 per request)", base)
        serving = dict(sweep=sweep, plain_fp64_per_request=base,
                       workload="Poisson arrivals; 75% chains 1-12 DOF / floating quadrupeds (dt 1e-3), "
                                "25% free cubes on a plane (dt 1e-4); horizons U{100..2000}; mixed precision")
        if args.workload == "serve":
            print(json.dumps(dict(metric="serving: env-steps/s and p50/p99 rollout latency", serving=serving)))
            return
    w = make_workload(args.workload, args.instances)
    models = [batch.Model(d) for d in w["models"]]
    log(f"workload {args.workload}: {sum(w['counts'])} instances")
    if args.steps <= 0:
        args.steps = int(w["steps"])
    single = args.workload in ("arms", "acrobot")
    is_vs = w.get("integrator", 0) >= abi.GP_VELOCITY_STEPPING
    vstep = is_vs or args.workload == "so101"
    if vstep:  # the runtime-topology kernel in FP64: nothing to tune or compare; every step in one launch
        args.autotune, args.no_compare = 0, True
        args.steps_per_launch = args.steps_per_launch or args.steps
    if args.autotune < 0:
        args.autotune = 0 if single else 1
    if single and not args.autotune:
        args.steps_per_launch = args.steps_per_launch or min(args.steps, 200)
        args.block = args.block or 256
    spl = args.steps_per_launch or min(args.steps, 100)
    if args.warmup < 0:
        args.warmup = spl
    common = dict(block_size=args.block, lanes_per_mech=args.lanes, lanes_per_tree=args.lanes_tree,
                  steps_per_launch=spl, use_graph=args.graph, autotune=bool(args.autotune))
    mixed = run_variant(w, models, args.steps, args.warmup,
                        precision=abi.GP_PREC_F64 if vstep else abi.GP_PREC_MIXED, kernel=abi.GP_KERNEL_AUTO, **common)
    log("mixed", mixed)
    extra = {}
    if not args.no_compare:
        f64 = run_variant(w, models, args.steps, args.warmup, precision=abi.GP_PREC_F64, kernel=abi.GP_KERNEL_AUTO,
                          **common)
        log("fast f64", f64)
        # the same steps (for the cubes, the same free-fall and contact phases)
        plain = run_variant(w, models, args.steps, min(args.warmup, 10), precision=abi.GP_PREC_F64,
                            kernel=abi.GP_KERNEL_PLAIN, use_graph=args.graph)  # runtime topology: nothing to tune
        log("plain f64", plain)
        extra = dict(fast_f64_env_steps_per_s=f64["value"], plain_f64_env_steps_per_s=plain["value"],
                     speedup_mixed_vs_plain_f64=mixed["value"] / plain["value"],
                     speedup_mixed_vs_fast_f64=mixed["value"] / f64["value"])
    # roofline of the dominant kernel: algorithmic flops per launch / launch time
    work = json.load(open(WORK_TABLE))
    flops_per_env_step = work.get(args.workload, {}).get("flops_per_env_step")
    roof = None
    if flops_per_env_step:
        launch_s = mixed["lib_ms"] / 1e3 / max(1, mixed["launches"])
        spl = mixed["opts"]["steps_per_launch"] or spl
        flops = flops_per_env_step * mixed["instances"] * spl
        achieved = flops / launch_s / 1e12
        traffic = work.get(args.workload, {}).get("hbm_bytes_per_launch")
        roof = dict(bound="valu", achieved=achieved, peak=PEAK_TFLOPS["f32"], unit="TFLOP/s",
                    frac=achieved / PEAK_TFLOPS["f32"], traffic=traffic,
                    kernel=work[args.workload].get("kernel"), flops_per_env_step=flops_per_env_step,
                    launch_ms=launch_s * 1e3)
    cpu = None if args.no_cpu_baseline else cpu_baseline(w)
    if cpu:
        log("cpu", cpu)
    line = {
        "metric": "env-steps/sec at fixed FP64-oracle error, 1 GPU; p50/p99 rollout latency",
        "value": mixed["value"],
        "unit": "env-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if vstep else "f32+f64",
        "data": "synthetic (seeded Philox)",
        "config": {"workload": {
            "arms": "65,536 serial 7-DOF revolute arms, dt=1e-3, semi-implicit Euler, FP32 sweeps + FP64 D/q/v",
            "mixed": "32,768 chains 1-12 DOF + floating quadrupeds, dt=1e-3, semi-implicit Euler, mixed precision",
            "cubes": "16,384 free cubes x 8 corner contacts on z=0, dt=1e-4, semi-implicit Euler, mixed precision",
            "acrobot": "1,024 acrobots, dt=1e-3, semi-implicit Euler, mixed precision",
            "thrown_cubes": "16,384 cubes (side 0.1) thrown at z=0, dt=1/120, velocity stepping with continuous "
                            "collision detection (cone-program contact impulses), FP64",
            "so101": "65,536 SO-101 arms (welded base, 6 revolutes), dt=1e-3, RK4, held torques, "
                     "runtime-topology FP64 kernel"}[args.workload],
            "instances": mixed["instances"], "steps_per_launch": mixed["opts"]["steps_per_launch"] or spl,
            "hip_graph": bool(args.graph), "launch_opts": mixed["opts"], "autotuned": bool(args.autotune),
            "error_bound": ("teacher-forced per-step vs FP64 oracle: e(q'), e(v') <= 1e-5 (contact impulses from an "
                            "interior-point method), tests/test_gpu_velocity_stepping.py") if is_vs else
                           ("teacher-forced per-step vs FP64 oracle: e(q'), e(v'), e(qdd) <= 1e-10, "
                            "tests/test_gpu_builders.py") if vstep else
                           ("teacher-forced per-step vs FP64 oracle: e(q') <= 1e-5, e(v'), e(qdd) <= 2e-3, "
                            "p99 e(qdd) <= 1e-4 (FP64 kernels: all <= 1e-10), tests/test_gpu_parity.py")},
        "roofline": roof,
        "cpu_baseline": cpu,
        "faults": mixed["faults"],
    }
    if serving:
        s0 = serving["sweep"][0]
        line["latency_ms"] = {"p50": s0["p50_ms"], "p99": s0["p99_ms"], "rate_req_per_s": s0["rate"],
                              "requests": s0["requests"]}
        line["serving"] = serving
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
