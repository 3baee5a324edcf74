"""Benchmark of the batched rigid-body step on one MI355X.

Default workload: 65,536 serial 7-DOF revolute arms (random links, axes,
masses, inertias, constant torques), dt = 1e-3, semi-implicit Euler. One
"step" is one dt of every instance. The headline `value` is the FP64
compiled-topology kernel, the variant whose per-step error against the FP64
oracle is within 1e-10 (the metric's "fixed FP64-oracle error"); the
mixed-precision kernel (FP32 sweeps, FP64 D solve and q/v accumulation) is
reported beside it with its own measured error bound, and the plain FP64
runtime-topology kernel as the baseline both are compared with. Prints one
JSON line with env-steps/s (all instances and the non-faulted ones), the
dominant kernel's VALU roofline (algorithmic flops from the op count of the
ABA, per-launch duration from HIP events on the batch's stream), the CPU
baseline (the oracle's restatement of the reference algorithm, timed on the
host's cores over a bounded sample), the other configurations (1,024
acrobots, 32,768 mixed chains and quadrupeds, 16,384 cubes on a plane) and
the serving latency.

    python bench.py [--steps K] [--warmup W] [--workload arms|mixed|cubes|acrobot|thrown_cubes|so101|serve]
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
    if name == "trees":
        return workloads.trees(n=n or 16384)
    raise SystemExit(f"unknown workload {name}")


def run_variant(w, models, steps, warmup, autotune=False, **opts):
    """Returns env-steps/s (all instances, and the instances that did not
    fault), ms per step, lib-timed ms, launches, faults and the options the
    batch ran with (after gp_batch_autotune when asked; the tuning runs before
    the warmup, outside the timed region, and restores the state). Every
    hipGraph the warmup and the timed call replay is built by
    gp_step_prepare before either runs, so no capture falls in a timed
    region."""
    from gorilla_physics_amd import batch
    b = batch.Batch(models, w["counts"], **opts)
    b.set_state(w["q"], w["v"])
    b.set_tau(w["tau"])
    tuned_ms = None
    integrator = w.get("integrator", 0)
    if autotune:
        # trials long enough that launch-to-launch noise does not pick the options
        _, tuned_ms = b.autotune(w["dt"], integrator, trial_steps=min(max(steps, 200), 1000))
    o = b.get_opts()
    chosen = dict(block_size=o.block_size, lanes_per_mech=o.lanes_per_mech, lanes_per_tree=o.lanes_per_tree,
                  steps_per_launch=o.steps_per_launch, use_graph=o.use_graph)
    if warmup:
        b.prepare(w["dt"], integrator, warmup)
    b.prepare(w["dt"], integrator, steps)
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
    return dict(value=n * steps / sec, valid_value=(n - faults) * steps / sec, ms_per_step=1e3 * sec / steps,
                lib_ms=lib_ms, launches=launches, faults=faults, instances=n, steps=steps, opts=chosen,
                autotune_trial_ms=tuned_ms)


def roofline(work, variant, fp64_kernel, traffic=None, kernel=None):
    """VALU roofline of one measured variant: algorithmic flops of the timed
    steps / the HIP-event time of the timed launches. The peak is the
    mix-weighted one: the shortest time the flops can take is
    F64 / P64 + F32 / P32 (an FP64 kernel runs both parts at P64)."""
    f32 = float(work["f32_per_batch_step"]) * variant["steps"]
    f64 = float(work["f64_per_batch_step"]) * variant["steps"]
    if fp64_kernel:
        t_min = (f32 + f64) / (PEAK_TFLOPS["f64"] * 1e12)
    else:
        t_min = f32 / (PEAK_TFLOPS["f32"] * 1e12) + f64 / (PEAK_TFLOPS["f64"] * 1e12)
    t = variant["lib_ms"] / 1e3
    achieved = (f32 + f64) / t / 1e12
    peak = (f32 + f64) / t_min / 1e12
    spl = variant["opts"]["steps_per_launch"] or variant["steps"]
    return dict(bound="valu", achieved=achieved, peak=peak, unit="TFLOP/s", frac=achieved / peak, traffic=traffic,
                kernel=kernel, flops_per_env_step=(f32 + f64) / variant["steps"] / variant["instances"],
                f64_share=f64 / (f32 + f64), launch_ms=variant["lib_ms"] / max(1, variant["launches"]),
                launches=variant["launches"], steps_per_launch=spl)


def serving_requests(n, seed=0, n_models=8192):
    """n requests mixing the mixed-workload mechanisms (chains of 1-12 DOF and
    floating quadrupeds, dt 1e-3) and free cubes on a plane (dt 1e-4), 3:1,
    horizons U{100..2000}: n_models distinct mechanisms, each request one of
    them from its own initial state (the mechanism's start velocities plus
    N(0, 0.1))."""
    from gorilla_physics_amd import batch, workloads
    n_cubes = n_models // 4
    mix = workloads.mixed(n=n_models - n_cubes, seed=seed)
    cub = workloads.cubes(n=max(1, n_cubes), seed=seed)
    cube_model = batch.Model(cub["models"][0])
    base, qo, vo = [], 0, 0
    for d in mix["models"]:
        base.append(dict(model=batch.Model(d), q=mix["q"][qo:qo + d.nq], v=mix["v"][vo:vo + d.nv],
                         tau=mix["tau"][vo:vo + d.nv], dt=1e-3))
        qo += d.nq
        vo += d.nv
    for i in range(n_cubes):
        base.append(dict(model=cube_model, q=cub["q"][7 * i:7 * i + 7], v=cub["v"][6 * i:6 * i + 6],
                         tau=np.zeros(6), dt=1e-4))
    rng = np.random.default_rng(seed + 17)
    pick = rng.integers(0, len(base), n)
    horizon = rng.integers(100, 2001, n)
    reqs = []
    for i in range(n):
        b = base[pick[i]]
        reqs.append(dict(model=b["model"], q=b["q"], v=b["v"] + rng.normal(0.0, 0.1, b["v"].size), tau=b["tau"],
                         dt=b["dt"], horizon=int(horizon[i])))
    return reqs


def serve_run(reqs, rate, chunk=50, slots=4096, seed=1):
    """Poisson arrivals at `rate` requests/s into one running server; returns
    throughput and latency percentiles. The server is warm, as a long-running
    one is: the same arrivals are replayed once untimed first, so the pools
    the load needs exist before the measured arrivals. Every request is
    submitted before the first arrival time (the arrivals start after the
    submission loop's expected length), so latency = completion - arrival
    holds no host submission time."""
    from gorilla_physics_amd import abi, serve
    s = serve.Server(precision=abi.GP_PREC_MIXED, chunk_steps=chunk, slots_per_pool=slots)
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, len(reqs))

    def replay():
        t0 = s.clock() + 0.05 + 20e-6 * len(reqs)
        arrivals = t0 + np.cumsum(gaps)
        ids = [s.submit(r["model"], r["q"], r["v"], r["tau"], r["horizon"], r["dt"], float(a))
               for r, a in zip(reqs, arrivals)]
        late = max(0.0, s.clock() - arrivals[0])
        s.run()
        return ids, arrivals, late

    ids, _, _ = replay()  # warm-up: pools for this load
    for i in ids:
        s.result(i)
    st0 = s.stats()
    ids, arrivals, late = replay()
    t_end = s.clock()
    lat = np.array([s.result(i)[2] for i in ids])
    st = s.stats()
    steps = float(sum(r["horizon"] for r in reqs))
    span = t_end - arrivals[0]
    d = {k: st[k] - st0[k] for k in ("rounds", "busy_s", "admit_s", "retire_s", "env_steps")}
    return dict(rate=rate, requests=len(reqs), env_steps_per_s=steps / span, requests_per_s=len(reqs) / span,
                p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                rounds=int(d["rounds"]), pools=int(st["pools"]), gpu_busy_s=d["busy_s"],
                admit_s=d["admit_s"], retire_s=d["retire_s"],
                admit_share=d["admit_s"] / max(1e-12, d["busy_s"] + d["admit_s"]), makespan_s=span,
                max_in_flight_slots=int(slots * st["pools"]), submit_late_s=late)


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


def cpu_baseline(w, budget_s=10.0):
    """The oracle (CRBA + RNEA + LU, the reference's algorithm) over a bounded
    sample of the same workload, one instance per thread at a time; the whole
    workload when it fits the budget."""
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
    whole = " (the whole workload)" if n == n_all and steps == w["steps"] else ""
    return dict(value=n * steps / sec, unit="env-steps/s", cores=cores, kind="port",
                sample=f"{n} of {n_all} instances x {steps} of {w['steps']} steps{whole} ({sec:.1f} s), {algo} "
                       f"FP64 oracle (g++ -O3), {cores} threads of {os.cpu_count()} on {cpu_model()}")


def abi_vs():
    from gorilla_physics_amd import abi
    return abi.GP_VELOCITY_STEPPING


WORKLOAD_TEXT = {
    "arms": "65,536 serial 7-DOF revolute arms, dt=1e-3, semi-implicit Euler",
    "mixed": "32,768 chains of 1-12 revolute/prismatic DOF (75%) + floating-base quadrupeds 6+4x3 DOF (25%), "
             "dt=1e-3, semi-implicit Euler, topology-bucketed grouped launch, autotuned",
    "cubes": "16,384 free unit cubes x 8 corner contacts on z=0 (k=1e4, lambda=1e2, mu=0.5), dt=1e-4, "
             "semi-implicit Euler, autotuned",
    "acrobot": "1,024 acrobots (2-link revolute, l=1 m, m=1 kg, uniform rods), dt=1e-3, semi-implicit Euler, zero torque",
    "thrown_cubes": "16,384 cubes (side 0.1) thrown at z=0, dt=1/120, velocity stepping with continuous "
                    "collision detection (cone-program contact impulses), FP64",
    "so101": "65,536 SO-101 arms (welded base, 6 revolutes), dt=1e-3, RK4, held torques, welded-base compiled kernel",
    "trees": "16,384 trees of 16-40 bodies on a ground plane (hexapods, humanoids = biped + 2 arms, random trees), "
             "dt=1e-4, semi-implicit Euler, one body per lane, grouped launch, autotuned",
}
F64_BOUND = "teacher-forced per-step vs FP64 oracle: e(q'), e(v'), e(qdd) <= 1e-10 (tests/test_gpu_parity.py)"
MIXED_BOUND = ("teacher-forced per-step vs FP64 oracle: e(q') <= 1e-5, e(v'), e(qdd) <= 2e-3, p99 e(qdd) <= 1e-4 "
               "(tests/test_gpu_parity.py)")


def measure_config(name, w, models, steps, warmup, args, compare=True, cpu=True):
    """One configuration: the FP64 compiled-topology kernel (the headline,
    within the FP64-oracle bound), the mixed-precision kernel and the plain
    FP64 kernel on the same inputs, their rooflines and the CPU baseline."""
    from gorilla_physics_amd import abi
    work = json.load(open(WORK_TABLE)).get(name)
    single = name in ("arms", "acrobot")
    autotune = args.autotune if args.autotune >= 0 else (0 if single else 1)
    spl = args.steps_per_launch
    block = args.block
    if single and not autotune:
        spl = spl or min(steps, 200)
        block = block or 256
    spl = spl or min(steps, 100)
    common = dict(block_size=block, lanes_per_mech=args.lanes, lanes_per_tree=args.lanes_tree,
                  steps_per_launch=spl, use_graph=args.graph, autotune=bool(autotune),
                  separate_launches=int(args.separate_launches))
    f64 = run_variant(w, models, steps, warmup, precision=abi.GP_PREC_F64, kernel=abi.GP_KERNEL_AUTO, **common)
    log(name, "fast f64", f64)
    out = dict(workload=WORKLOAD_TEXT[name], instances=f64["instances"], steps=steps, warmup=warmup,
               value=f64["value"], valid_value=f64["valid_value"], ms_per_step=f64["ms_per_step"],
               faults=f64["faults"], dtype="f64", error_bound=F64_BOUND, launch_opts=f64["opts"],
               autotuned=bool(autotune))
    if work:
        kern = work["kernel"]
        out["roofline"] = roofline(work, f64, True, traffic=work["hbm_bytes_per_launch"].get(
            f"f64_spl{f64['opts']['steps_per_launch']}"), kernel=kern["f64"])
    if compare:
        mixed = run_variant(w, models, steps, warmup, precision=abi.GP_PREC_MIXED, kernel=abi.GP_KERNEL_AUTO,
                            **common)
        log(name, "mixed", mixed)
        out["mixed_precision"] = dict(value=mixed["value"], valid_value=mixed["valid_value"],
                                      ms_per_step=mixed["ms_per_step"], faults=mixed["faults"], dtype="f32+f64",
                                      error_bound=MIXED_BOUND, launch_opts=mixed["opts"])
        if work:
            out["mixed_precision"]["roofline"] = roofline(work, mixed, False, traffic=work["hbm_bytes_per_launch"].get(
                f"mixed_spl{mixed['opts']['steps_per_launch']}"), kernel=work["kernel"]["mixed"])
        # the same steps (for the cubes, the same free-fall and contact phases)
        plain = run_variant(w, models, steps, min(warmup, 10), precision=abi.GP_PREC_F64,
                            kernel=abi.GP_KERNEL_PLAIN, use_graph=args.graph)  # runtime topology: nothing to tune
        log(name, "plain f64", plain)
        out["plain_f64"] = dict(value=plain["value"], valid_value=plain["valid_value"],
                                ms_per_step=plain["ms_per_step"], faults=plain["faults"])
        out["speedup_f64_vs_plain_f64"] = f64["value"] / plain["value"]
        out["speedup_mixed_vs_plain_f64"] = mixed["value"] / plain["value"]
        out["speedup_mixed_vs_f64"] = mixed["value"] / f64["value"]
    if cpu:
        out["cpu_baseline"] = cpu_baseline(w)
        log(name, "cpu", out["cpu_baseline"])
    return out


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
    ap.add_argument("--separate-launches", action="store_true",
                    help="one launch per topology bucket instead of one grouped launch (for per-bucket kernel traces)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="leave out the other configurations' lines")
    ap.add_argument("--serve-requests", type=int, default=100000,
                    help="most requests per swept rate (each rate runs min(this, max(2000, rate x 0.5 s)))")
    ap.add_argument("--serve-rates", default="1000,10000,100000,1000000")
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
            k = int(min(len(reqs), max(2000, rate * 0.5)))
            res = serve_run(reqs[:k], rate, chunk=args.serve_chunk)
            log("serve", res)
            sweep.append(res)
        base = None
        if args.serve_baseline:
            base = serve_baseline_plain(reqs[:args.serve_baseline], rates[0])
            log("serve baseline (plain FP64 per request)", base)
        serving = dict(sweep=sweep, plain_fp64_per_request=base,
                       workload="Poisson arrivals; 75% chains 1-12 DOF / floating quadrupeds (dt 1e-3), "
                                "25% free cubes on a plane (dt 1e-4), 8,192 distinct mechanisms; horizons "
                                "U{100..2000}; mixed precision; 4,096 slots per pool")
        if args.workload == "serve":
            print(json.dumps(dict(metric="serving: env-steps/s and p50/p99 rollout latency", serving=serving)))
            return
    w = make_workload(args.workload, args.instances)
    models = [batch.Model(d) for d in w["models"]]
    log(f"workload {args.workload}: {sum(w['counts'])} instances")
    if args.steps <= 0:
        args.steps = int(w["steps"])
    is_vs = w.get("integrator", 0) >= abi.GP_VELOCITY_STEPPING
    vstep = is_vs
    if vstep:  # velocity stepping computes in FP64 only: nothing to tune or compare; every step in one launch
        args.autotune, args.no_compare = 0, True
        args.steps_per_launch = args.steps_per_launch or args.steps
    if args.warmup < 0:
        single = args.workload in ("arms", "acrobot")
        args.warmup = args.steps_per_launch or (min(args.steps, 200) if single else min(args.steps, 100))
    if vstep:
        r = run_variant(w, models, args.steps, args.warmup, precision=abi.GP_PREC_F64, kernel=abi.GP_KERNEL_AUTO,
                        block_size=args.block, steps_per_launch=args.steps_per_launch, use_graph=args.graph)
        log("f64", r)
        head = dict(workload=WORKLOAD_TEXT[args.workload], instances=r["instances"], value=r["value"],
                    valid_value=r["valid_value"], ms_per_step=r["ms_per_step"], faults=r["faults"], dtype="f64",
                    launch_opts=r["opts"], autotuned=False,
                    error_bound=("teacher-forced per-step vs FP64 oracle: e(q'), e(v') <= 1e-5 (contact impulses "
                                 "from an interior-point method), tests/test_gpu_velocity_stepping.py") if is_vs else
                    ("teacher-forced per-step vs FP64 oracle: e(q'), e(v'), e(qdd) <= 1e-10, "
                     "tests/test_gpu_builders.py"), roofline=None)
        if not args.no_cpu_baseline:
            head["cpu_baseline"] = cpu_baseline(w)
            log("cpu", head["cpu_baseline"])
    else:
        head = measure_config(args.workload, w, models, args.steps, args.warmup, args, compare=not args.no_compare,
                              cpu=not args.no_cpu_baseline)
    configs = {}
    if args.workload == "arms" and not args.no_configs and not args.instances:
        # the other configurations, each at its own size and step count
        for name, steps in (("acrobot", 2000), ("mixed", 1000), ("cubes", 5000), ("trees", 1000)):
            wc = make_workload(name, 0)
            mc = [batch.Model(d) for d in wc["models"]]
            log(f"config {name}: {sum(wc['counts'])} instances x {steps} steps")
            configs[name] = measure_config(name, wc, mc, steps, min(steps, 100), args, compare=True,
                                           cpu=not args.no_cpu_baseline)
            del mc
    line = {
        "metric": "env-steps/sec at fixed FP64-oracle error, 1 GPU; p50/p99 rollout latency",
        "value": head["value"],
        "unit": "env-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Philox)",
        "config": {"workload": head["workload"], "instances": head["instances"],
                   "steps_per_launch": head["launch_opts"]["steps_per_launch"], "hip_graph": bool(args.graph),
                   "launch_opts": head["launch_opts"], "autotuned": head["autotuned"],
                   "error_bound": head["error_bound"]},
        "roofline": head.get("roofline"),
        "cpu_baseline": head.get("cpu_baseline"),
        "valid_value": head["valid_value"],
        "faults": head["faults"],
    }
    for k in ("mixed_precision", "plain_f64", "speedup_f64_vs_plain_f64", "speedup_mixed_vs_plain_f64",
              "speedup_mixed_vs_f64"):
        if k in head:
            line[k] = head[k]
    if configs:
        line["configs"] = configs
    if serving:
        line["latency_ms"] = [{"rate_req_per_s": x["rate"], "p50": x["p50_ms"], "p99": x["p99_ms"],
                               "requests": x["requests"], "env_steps_per_s": x["env_steps_per_s"]}
                              for x in serving["sweep"]]
        line["serving"] = serving
    print(json.dumps(line))


if __name__ == "__main__":
    main()
