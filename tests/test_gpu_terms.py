"""GPU parity of the joint-space terms: gp_mass_matrix (mechanism.rs:637-696)
and gp_dynamics_bias (dynamics.rs:233-251) against the oracle's world-frame
CRBA and RNEA (its M and c - tau_c, restating mass_matrix and dynamics_bias),
on the benchmark workloads, cubes with corners in contact, fixed joints and
the reference's model builders. Bound: e <= 1e-10 per instance (the FP64
bound), e(x) = max|x - x_ref| / max(max|x_ref|, 1). Consistency on the GPU
itself: M vdot + c = tau (+ prismatic springs) with gp_dynamics' vdot."""
import numpy as np
import pytest

from gpu_helpers import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _descs(w):
    out = []
    for d, c in zip(w["models"], w["counts"]):
        out.extend([d] * int(c))
    return out


def _check(gp, oracle, w, label):
    from gorilla_physics_amd import batch
    b, _ = batch.batch_from_workload(w)
    print(f"case terms {label}: {b.n_instances} instances", flush=True)
    Ms = b.mass_matrices()
    c = b.dynamics_bias()
    vdot = b.dynamics()
    q, v = b.get_state()
    qo, vo = b.offsets()
    worst_m = worst_c = worst_id = 0.0
    for i, d in enumerate(_descs(w)):
        qi, vi = q[qo[i]:qo[i + 1]], v[vo[i]:vo[i + 1]]
        Mr, cr = oracle.mass_matrix(d, qi, vi)
        worst_m = max(worst_m, rel_err(Ms[i], Mr))
        worst_c = max(worst_c, rel_err(c[vo[i]:vo[i + 1]], cr))
        # M vdot + c = tau + tau_spring (tau = the held torques; springs from the description)
        tau = np.asarray(w["tau"][vo[i]:vo[i + 1]], dtype=float).copy()
        k = 0
        for j in range(d.n_bodies):
            bd = d._bodies[j]
            if bd.jtype == 2 and bd.spring_k != 0.0:
                tau[k] += -bd.spring_k * (qi[_qrow(d, j)] - bd.spring_l)
            k += {0: 0, 1: 1, 2: 1, 3: 6}[bd.jtype]
        lhs = Ms[i] @ vdot[vo[i]:vo[i + 1]] + c[vo[i]:vo[i + 1]]
        scale = max(1.0, np.abs(tau).max(), np.abs(c[vo[i]:vo[i + 1]]).max(),
                    np.abs(Ms[i]).max() * np.abs(vdot[vo[i]:vo[i + 1]]).max())
        worst_id = max(worst_id, np.abs(lhs - tau).max() / scale)
    print(f"  e(M) {worst_m:.3g}  e(c) {worst_c:.3g}  identity {worst_id:.3g}", flush=True)
    assert worst_m <= TOL and worst_c <= TOL and worst_id <= 1e-10
    for M in Ms:
        assert np.array_equal(M, M.T)


def _qrow(d, body):
    k = 0
    for j in range(body):
        k += {0: 0, 1: 1, 2: 1, 3: 7}[d._bodies[j].jtype]
    return k


@pytest.mark.parametrize("name,n", [("acrobot", 64), ("arms", 256), ("mixed", 384)])
def test_terms_on_workloads(gp, oracle, name, n):
    from gorilla_physics_amd import workloads
    _check(gp, oracle, getattr(workloads, name)(n=n), name)


def test_terms_cubes_in_contact(gp, oracle):
    """Cubes lowered so that corners sit inside the plane (and inside the
    1e-8 band, where the force is zero): tau_c enters c."""
    from gorilla_physics_amd import workloads
    w = workloads.cubes(n=128)
    q = w["q"].reshape(-1, 7).copy()
    rng = np.random.default_rng(3)
    for i in range(q.shape[0]):
        # lowest corner at -2 mm .. +1 mm, velocities random
        from gorilla_physics_amd.mechanism import UnitQuaternion
        R = UnitQuaternion(*q[i, :4]).to_matrix()
        corners = np.array([[sx, sy, sz] for sx in (-.5, .5) for sy in (-.5, .5) for sz in (-.5, .5)])
        low = float(np.min(corners @ R.T[:, 2]))
        q[i, 6] = -low + rng.uniform(-2e-3, 1e-3)
    w["q"] = q.reshape(-1)
    w["v"] = rng.normal(0.0, 0.5, w["v"].size)
    _check(gp, oracle, w, "cubes in contact")


def test_terms_builders_with_fixed_joints(gp, oracle):
    """Every model of the reference's builders (the SO-101's welded base, the
    test arm's fixed joint, prismatic springs, spring legs, trees) touching a
    tilted plane with random velocities, in one batch."""
    from gorilla_physics_amd import builders
    from test_gpu_builders import _on_ground
    from test_oracle_golden import builder_models
    rng = np.random.default_rng(5)
    descs, qs, vs = [], [], []
    for name in builder_models(gp):
        d, q, v = _on_ground(gp, name, rng)
        descs.append(d)
        qs.append(q)
        vs.append(v)
    so = builders.build_so101().model_desc()
    descs.append(so)
    qs.append(rng.uniform(-2, 2, 6))
    vs.append(rng.normal(size=6))
    v = np.concatenate(vs)
    w = dict(models=descs, counts=[1] * len(descs), q=np.concatenate(qs), v=v, tau=rng.normal(0.0, 0.5, v.size),
             dt=1e-3)
    _check(gp, oracle, w, "builders")
