"""Trees of 13 to 64 bodies on the one-body-per-lane kernel (wave_tree_kernel,
gp_tree_impl.h): hexapods with foot contacts, humanoids (the reference's
biped, builders/biped_builder.rs:12, plus two arms), the biped itself and
random trees of 16-64 bodies with revolute, sprung prismatic and fixed joints
and contact points, all touching or near a ground plane. Every step of every
instance is compared with one oracle step (the reference's CRBA + RNEA + LU,
dynamics.rs:322-364) from the same state. Bounds: FP64 e <= 1e-10 (the FP64
bound of every kernel); mixed precision the mixed bounds of
test_gpu_parity.py. The batches are built with GP_KERNEL_FAST, which fails
unless every bucket runs on a compiled kernel, so the tree kernel is the one
measured."""
import numpy as np
import pytest

from gpu_helpers import step_errors

pytestmark = pytest.mark.gpu

F64_TOL = 1e-10
MIXED_Q_TOL, MIXED_VA_TOL = 1e-5, 2e-3


def _workload(gp, n=96, seed=3):
    from gorilla_physics_amd import builders, workloads
    w = workloads.trees(n=n, seed=seed)
    rng = np.random.default_rng(seed)
    extra = [workloads.random_tree_desc(rng, nb) for nb in (13, 16, 17, 32, 33, 48, 64)]
    biped = builders.build_biped()
    biped.add_halfspace(gp.HalfSpace.new_with_params([0.0, 0.0, 1.0], 0.0, 0.9, 0.5))
    extra.append(biped.model_desc())
    qs, vs, ts = [w["q"]], [w["v"]], [w["tau"]]
    for d in extra:  # joint angles N(0, 0.3), the lowest contact point 2 mm inside the ground
        q = np.zeros(d.nq)
        q[3] = 1.0
        q[7:] = rng.normal(0.0, 0.3, d.nq - 7) * [0.1 if d._bodies[b].jtype == 2 else 1.0
                                                   for b in range(1, d.n_bodies) if d._bodies[b].jtype != 0]
        q[6] = -workloads.lowest_contact_z(d, q) - 0.002
        qs.append(q)
        vs.append(rng.normal(0.0, 0.3, d.nv))
        ts.append(np.concatenate([np.zeros(6), rng.normal(0.0, 0.02, d.nv - 6)]))
    return dict(models=w["models"] + extra, counts=[1] * (n + len(extra)), q=np.concatenate(qs),
                v=np.concatenate(vs), tau=np.concatenate(ts), dt=w["dt"])


def _descs(w):
    out = []
    for d, c in zip(w["models"], w["counts"]):
        out.extend([d] * int(c))
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_tree_kernel_step_parity(gp, oracle, precision):
    from gorilla_physics_amd import abi, batch
    w = _workload(gp)
    descs = _descs(w)
    print(f"case trees precision {precision}: {len(descs)} instances, bodies "
          f"{min(d.n_bodies for d in descs)}-{max(d.n_bodies for d in descs)}", flush=True)
    b, _ = batch.batch_from_workload(w, precision=precision, kernel=abi.GP_KERNEL_FAST, steps_per_launch=1)
    q, v = b.get_state()
    worst = np.zeros(3)
    touched = 0
    for k in range(30):
        b.step(w["dt"], 0, 1)
        q1, v1 = b.get_state()
        a1 = b.get_vdot()
        eq, ev, ea = step_errors(oracle, descs, q, v, w["tau"], q1, v1, a1, w["dt"])
        worst = np.maximum(worst, [eq.max(), ev.max(), ea.max()])
        q, v = q1, v1
    assert not b.faults().any()
    print(f"worst per-step e(q') {worst[0]:.3g} e(v') {worst[1]:.3g} e(qdd) {worst[2]:.3g}", flush=True)
    if precision == abi.GP_PREC_F64:
        assert worst.max() <= F64_TOL
    else:
        assert worst[0] <= MIXED_Q_TOL and worst[1] <= MIXED_VA_TOL and worst[2] <= MIXED_VA_TOL


def test_tree_contacts_engage(gp, oracle):
    """The parity runs are not contact-free: at the start poses some feet and
    contact points are already inside the ground (the oracle's contact
    distances), so contact forces are part of what is compared."""
    w = _workload(gp, n=48)
    qo = np.cumsum([0] + [d.nq for d in _descs(w)])
    inside = 0
    for i, d in enumerate(_descs(w)):
        dist = oracle.contact_distances(d, w["q"][qo[i]:qo[i + 1]])
        inside += int((dist <= 1e-8).sum())
    assert inside > 0


def test_tree_launch_shape_does_not_change_bits(gp):
    """Steps fused per launch, workgroup size and a grouped launch beside
    other buckets change no bit, and an instance alone matches itself in the
    batch."""
    from gorilla_physics_amd import abi, batch
    w = _workload(gp, n=32)
    print("case trees launch shapes", flush=True)
    outs = []
    for spl, block in ((1, 256), (25, 64), (100, 128)):
        b, _ = batch.batch_from_workload(w, precision=0, kernel=abi.GP_KERNEL_FAST, steps_per_launch=spl,
                                         block_size=block)
        b.step(w["dt"], 0, 100)
        outs.append(b.get_state())
    for q, v in outs[1:]:
        assert np.array_equal(q, outs[0][0]) and np.array_equal(v, outs[0][1])
    d = w["models"][5]
    qo = np.cumsum([0] + [m.nq for m in w["models"]])
    vo = np.cumsum([0] + [m.nv for m in w["models"]])
    alone = dict(models=[d], counts=[1], q=w["q"][qo[5]:qo[6]], v=w["v"][vo[5]:vo[6]], tau=w["tau"][vo[5]:vo[6]],
                 dt=w["dt"])
    b, _ = batch.batch_from_workload(alone, precision=0, kernel=abi.GP_KERNEL_FAST, steps_per_launch=100)
    b.step(w["dt"], 0, 100)
    q1, v1 = b.get_state()
    assert np.array_equal(q1, outs[0][0][qo[5]:qo[6]]) and np.array_equal(v1, outs[0][1][vo[5]:vo[6]])
