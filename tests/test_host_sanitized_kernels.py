"""The library's runtime and its kernels compiled for the host CPU with
AddressSanitizer and UBSan (tests/host/: a stand-in for the HIP runtime that
gives every device array a host allocation of its exact size between poisoned
guards, runs every lane of a launch in turn, and range-checks every raw
buffer access against its resource's size). The runtime-topology kernels
(plain dynamics, velocity stepping with and without continuous collision
detection, joint-space terms, energy, transfers) run over the thrown-cube
workload and the velocity-stepping test models; the compiled-topology kernels
that run one instance per lane (chains, welded-base arms, free bodies,
quadrupeds) run one step of SIE, RK2 and RK4 in both precisions and are
compared with the oracle. Any access outside an array, any undefined
conversion (a NaN body index, say) or overflow ends the run with a sanitizer
report. The GPU cannot run sanitizers, so this is where the kernels' indexing
is checked."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gorilla-physics_amd", "csrc")
SOURCES = ["gp_runtime.cpp", "gp_plain.hip", "gp_vstep.hip", "gp_vstep_body.hip", "gp_energy.hip", "gp_terms.hip",
           "gp_xfer.hip"]


FAST_SOURCES = ["gp_fast.hip", "gp_fast_f64.hip", "gp_fast_f32.hip", "gp_fast_rk_f64.hip", "gp_fast_rk_f32.hip"]
FLAGS = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
         "-fsanitize=float-cast-overflow", "-fno-sanitize-recover=all", "-Wno-unused-result"]


def _build(tmp_path_factory, name, sources, defines=()):
    """Compile the sources in parallel (one g++ per file), then link."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    d = tmp_path_factory.mktemp(name)
    inc = ["-I", os.path.join(ROOT, "tests", "host")] + [f"-D{x}" for x in defines]
    objs, procs = [], []
    for src in sources + [os.path.join(ROOT, "tests", "host", "kernels_host_main.cpp")]:
        obj = str(d / (os.path.basename(src) + ".o"))
        path = src if os.path.isabs(src) else os.path.join(CSRC, src)
        procs.append(subprocess.Popen([gxx] + FLAGS + inc + ["-x", "c++", "-c", path, "-o", obj],
                                      stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))
        objs.append(obj)
    for p in procs:
        _, err = p.communicate(timeout=900)
        assert p.returncode == 0, err[-4000:]
    out = str(d / "kernels_host")
    p = subprocess.run([gxx] + FLAGS + objs + ["-o", out], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-4000:]
    return out


@pytest.fixture(scope="module")
def host_binary(tmp_path_factory):
    return _build(tmp_path_factory, "hostbuild", SOURCES)


@pytest.fixture(scope="module")
def host_binary_fast(tmp_path_factory):
    return _build(tmp_path_factory, "hostbuild_fast", SOURCES + FAST_SOURCES, ["GP_HOST_FAST"])


def _write(path, models, counts, q, v, tau, dt, integrator, steps, spl):
    with open(path, "wb") as f:
        f.write(struct.pack("4i", len(models), integrator, steps, spl))
        f.write(struct.pack("d", dt))
        for d in models:
            n = (d.desc.n_bodies, d.desc.n_contacts, d.desc.n_halfspaces, d.desc.n_spring_contacts, d.desc.n_spheres)
            f.write(struct.pack("5i", *n))
            f.write(struct.pack("d", d.desc.gravity))
            for arr, k in ((d._bodies, n[0]), (d._contacts, n[1]), (d._halfspaces, n[2]), (d._springs, n[3]),
                           (d._spheres, n[4])):
                f.write(bytes(arr)[:k * (len(bytes(arr)) // max(1, len(arr)))])
        f.write(np.asarray(counts, dtype=np.int64).tobytes())
        for a in (q, v, tau):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())


def _run(host_binary, tmp_path, out=None, kernel=1, precision=0, **kw):
    path = str(tmp_path / "batch.bin")
    _write(path, **kw)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    args = [host_binary, path] + ([str(out), str(kernel), str(precision)] if out else [])
    p = subprocess.run(args, capture_output=True, text=True, timeout=900, env=env)
    print(p.stdout, p.stderr[-3000:])
    assert p.returncode == 0 and "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr, \
        p.stderr[-4000:]
    return p.stdout


@pytest.mark.parametrize("integrator", [3, 4])
def test_thrown_cubes_velocity_stepping_in_bounds(gp, host_binary, tmp_path, integrator):
    """The benchmark's thrown cubes (collision/ccd/mod.rs:284) over the
    one-second flight and landing, 120 steps in one launch: 192 of them (three
    workgroups) with continuous collision detection, one (a 64-lane pitch,
    512-byte rows) without."""
    from gorilla_physics_amd import workloads
    w = workloads.thrown_cubes(n=192 if integrator == 4 else 1)
    out = _run(host_binary, tmp_path, models=w["models"], counts=w["counts"], q=w["q"], v=w["v"], tau=w["tau"],
               dt=w["dt"], integrator=integrator, steps=120, spl=120)
    assert "faults 0" in out


def test_thrown_cubes_single_body_kernel_in_bounds(gp, host_binary, tmp_path):
    """The benchmark's thrown cubes through the single-body velocity-stepping
    kernel (one lane per cube), CCD, the second's 120 steps in one launch."""
    from gorilla_physics_amd import workloads
    w = workloads.thrown_cubes(n=64)
    out = tmp_path / "out.bin"
    _run(host_binary, tmp_path, out=out, kernel=0, precision=0, models=w["models"], counts=w["counts"], q=w["q"],
         v=w["v"], tau=w["tau"], dt=w["dt"], integrator=4, steps=120, spl=120)
    q1, v1, _, rounds = _read_out(out, w["q"].size, w["v"].size)
    assert np.isfinite(q1).all() and np.isfinite(v1).all() and (rounds >= 1).all()
    z = q1.reshape(-1, 7)[:, 6]
    assert (z > 0.0).all() and (z < 0.2).all()  # resting on the plane, not through it


def test_velocity_stepping_models_in_bounds(gp, host_binary, tmp_path):
    """Spheres, cubes and quadrupeds on a tilted plane (the GPU velocity-
    stepping test's mix: sphere colliders, several contacts per body, trees),
    CCD velocity stepping, 60 steps."""
    import math
    from gorilla_physics_amd import builders
    rng = np.random.default_rng(0)
    normal = np.array([math.sin(0.1), 0.0, math.cos(0.1)])
    states = [builders.build_cube(1.0, 1.0), builders.build_sphere(1.0, 0.5), builders.build_quadruped()]
    for st in states:
        st.add_halfspace(gp.HalfSpace.new(normal, 0.0))
    dc, ds, dq = (st.model_desc() for st in states)
    descs, q, v = [], [], []
    for k in range(96):
        axis = rng.normal(size=3)
        rot = gp.UnitQuaternion.from_axis_angle(axis / np.linalg.norm(axis), rng.uniform(0, 2 * math.pi))
        if k % 3 == 0:
            descs.append(dc)
            q += [*rot.coords(), *rng.uniform(-1, 1, 2), rng.uniform(0.3, 1.0)]
            v += [*rng.normal(size=3), *rng.normal(size=3)]
        elif k % 3 == 1:
            descs.append(ds)
            q += [*rot.coords(), *rng.uniform(-1, 1, 2), rng.uniform(0.45, 0.55)]
            v += [*rng.normal(size=3), *rng.normal(size=3)]
        else:
            descs.append(dq)
            qq = np.zeros(dq.nq)
            qq[3], qq[6] = 1.0, rng.uniform(0.9, 1.3)
            qq[7:] = rng.normal(size=dq.nq - 7) * 0.2
            q += list(qq)
            v += list(rng.normal(size=dq.nv) * 0.3)
    v = np.array(v)
    _run(host_binary, tmp_path, models=descs, counts=[1] * len(descs), q=np.array(q), v=v, tau=np.zeros(v.size),
         dt=1.0 / 60.0, integrator=4, steps=60, spl=60)


def test_plain_kernel_and_terms_in_bounds(gp, host_binary, tmp_path):
    """The plain kernel (RK4, contacts, springs) and the joint-space terms on
    the reference's builder models touching a plane."""
    from test_gpu_builders import _on_ground
    from test_oracle_golden import builder_models
    rng = np.random.default_rng(2)
    descs, qs, vs = [], [], []
    for name in builder_models(gp):
        d, qi, vi = _on_ground(gp, name, rng)
        descs.append(d)
        qs.append(qi)
        vs.append(vi)
    v = np.concatenate(vs)
    _run(host_binary, tmp_path, models=descs, counts=[1] * len(descs), q=np.concatenate(qs), v=v,
         tau=rng.normal(0.0, 0.2, v.size), dt=1e-3, integrator=2, steps=50, spl=25)


def _read_out(path, nq, nv):
    """q, v, vdot and the collision-loop rounds the host driver wrote."""
    raw = open(path, "rb").read()
    d = np.frombuffer(raw[:8 * (nq + 2 * nv)], dtype=np.float64)
    return d[:nq], d[nq:nq + nv], d[nq + nv:], np.frombuffer(raw[8 * (nq + 2 * nv):], dtype=np.int32)


def _vs_mix(gp, rng, n=48):
    """Cubes and spheres near a tilted ground (touching, sliding, in flight)
    and quadrupeds standing on it."""
    import math
    from gorilla_physics_amd import builders
    normal = np.array([math.sin(0.1), 0.0, math.cos(0.1)])
    states = [builders.build_cube(1.0, 1.0), builders.build_sphere(1.0, 0.5), builders.build_quadruped()]
    for st in states:
        st.add_halfspace(gp.HalfSpace.new(normal, 0.0))
    dc, ds, dq = (st.model_desc() for st in states)
    descs, q, v = [], [], []
    for k in range(n):
        axis = rng.normal(size=3)
        rot = gp.UnitQuaternion.from_axis_angle(axis / np.linalg.norm(axis), rng.uniform(0, 2 * math.pi))
        if k % 3 == 0:
            descs.append(dc)
            q += [*rot.coords(), *rng.uniform(-1, 1, 2), rng.uniform(0.3, 1.0)]
            v += [*rng.normal(size=3), *rng.normal(size=3)]
        elif k % 3 == 1:
            descs.append(ds)
            q += [*rot.coords(), *rng.uniform(-1, 1, 2), rng.uniform(0.45, 0.55)]
            v += [*rng.normal(size=3), *rng.normal(size=3)]
        else:
            descs.append(dq)
            qq = np.zeros(dq.nq)
            qq[3], qq[6] = 1.0, rng.uniform(0.9, 1.3)
            qq[7:] = rng.normal(size=dq.nq - 7) * 0.2
            q += list(qq)
            v += list(rng.normal(size=dq.nv) * 0.3)
    return descs, np.array(q), np.array(v)


@pytest.mark.parametrize("integrator", [3, 4])
def test_single_body_velocity_stepping_matches_the_oracle(gp, oracle, host_binary, tmp_path, integrator):
    """The register-resident single-body velocity-stepping kernel (one lane
    per body here: its cross-lane sums are identities), under the
    sanitizers, one step from states along an oracle trajectory of cubes and
    spheres on a tilted plane (quadrupeds beside them run the runtime-topology
    kernel): the same collision-loop rounds as the oracle, q and v within the
    1e-5 per-step tolerance of the GPU velocity-stepping tests."""
    rng = np.random.default_rng(4)
    descs, q, v = _vs_mix(gp, rng)
    qo = np.cumsum([0] + [d.nq for d in descs])
    vo = np.cumsum([0] + [d.nv for d in descs])
    dt, worst = 1e-2, 0.0
    for k in range(4):
        out = tmp_path / f"out{k}.bin"
        _run(host_binary, tmp_path, out=out, kernel=0, precision=0, models=descs, counts=[1] * len(descs), q=q, v=v,
             tau=np.zeros(v.size), dt=dt, integrator=integrator, steps=1, spl=1)
        q1, v1, _, rounds = _read_out(out, q.size, v.size)
        qn, vn = q.copy(), v.copy()
        for i, d in enumerate(descs):
            qs, vs = slice(qo[i], qo[i + 1]), slice(vo[i], vo[i + 1])
            rq, rv, ro = oracle.vs_step(d, q[qs], v[vs], np.zeros(d.nv), dt, integrator == 4)
            assert rounds[i] == ro, (k, i, rounds[i], ro)
            e = max(np.abs(q1[qs] - rq).max() / max(1, np.abs(rq).max()),
                    np.abs(v1[vs] - rv).max() / max(1, np.abs(rv).max()))
            worst = max(worst, e)
            assert e < 1e-5, (k, i, e)
            qn[qs], vn[vs] = rq, rv
        for _ in range(4):  # the next states along the oracle's trajectory
            for i, d in enumerate(descs):
                qs, vs = slice(qo[i], qo[i + 1]), slice(vo[i], vo[i + 1])
                qn[qs], vn[vs], _ = oracle.vs_step(d, qn[qs], vn[vs], np.zeros(d.nv), dt, integrator == 4)
        q, v = qn, vn
    print(f"single-body velocity stepping, integrator {integrator}: worst per-step difference {worst:.3g}")


def _compiled_workload(gp):
    """Chains (arms, acrobots, the mixed workload's chains of 1-12 revolute
    or sprung prismatic joints), quadrupeds on one lane, free cubes with
    their corners near a plane on one lane, and welded-base arms (the SO-101
    and random ones of 1-7 joints)."""
    from gorilla_physics_amd import builders, workloads
    from test_gpu_welded import welded_arm_desc
    rng = np.random.default_rng(8)
    parts = [workloads.arms(n=48), workloads.acrobot(n=16), workloads.mixed(n=96), workloads.cubes(n=24)]
    c = parts[3]
    q = c["q"].reshape(-1, 7).copy()
    q[:, 6] = rng.uniform(0.45, 0.9, q.shape[0])  # some corners inside the plane
    c["q"] = q.reshape(-1)
    descs = [builders.build_so101().model_desc()] * 8 + [welded_arm_desc(rng, k) for k in range(1, 8) for _ in range(4)]
    nq = sum(d.nq for d in descs)
    parts.append(dict(models=descs, counts=[1] * len(descs), q=rng.uniform(-1.5, 1.5, nq), v=rng.normal(0, 0.5, nq),
                      tau=rng.normal(0, 0.05, nq)))
    models, counts, q, v, tau = [], [], [], [], []
    for w in parts:
        models += list(w["models"])
        counts += list(w["counts"])
        q.append(w["q"]), v.append(w["v"]), tau.append(w["tau"])
    return models, counts, np.concatenate(q), np.concatenate(v), np.concatenate(tau)


@pytest.mark.parametrize("integrator", [0, 1, 2])
@pytest.mark.parametrize("precision", [0, 1])
def test_compiled_kernels_in_bounds_and_match_the_oracle(gp, oracle, host_binary_fast, tmp_path, integrator,
                                                        precision):
    """The compiled-topology kernels that run one instance per lane, built
    for the host under the sanitizers: every buffer access inside its array
    (the stand-in for the raw buffer resources ends the run otherwise), and
    one step of every instance against the oracle from the same state: FP64
    within 1e-10, mixed precision within the mixed bounds (e(q') 1e-5, e(v'),
    e(qdd) 2e-3)."""
    from gpu_helpers import step_errors
    models, counts, q, v, tau = _compiled_workload(gp)
    descs = []
    for d, c in zip(models, counts):
        descs += [d] * c
    dt = 1e-3
    out = tmp_path / "out.bin"
    _run(host_binary_fast, tmp_path, out=out, kernel=2, precision=precision, models=models, counts=counts, q=q, v=v,
         tau=tau, dt=dt, integrator=integrator, steps=1, spl=1)
    q1, v1, a1, _ = _read_out(out, q.size, v.size)
    eq, ev, ea = step_errors(oracle, descs, q, v, tau, q1, v1, a1, dt, integrator)
    print(f"host compiled kernels, integrator {integrator}, precision {precision}: e(q') {eq.max():.3g} "
          f"e(v') {ev.max():.3g} e(qdd) {ea.max():.3g}")
    if precision == 0:
        assert max(eq.max(), ev.max(), ea.max()) <= 1e-10
    else:
        assert eq.max() <= 1e-5 and ev.max() <= 2e-3 and ea.max() <= 2e-3
