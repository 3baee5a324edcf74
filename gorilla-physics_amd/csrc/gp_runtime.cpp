// gp_runtime.cpp — host side of the C ABI (include/gorilla_amd.h).
//
// Model validation and the reference's flattening rules, bucketing of
// instances by topology, structure-of-arrays packing into device memory,
// step dispatch to the HIP kernels (plain / fast), hipGraph capture of
// multi-launch rollouts, the reference's simulate loop, fault flags and
// HIP-event timing. No C++ exception leaves this file: every entry point
// returns a status code and records a thread-local message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <atomic>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/gorilla_amd.h"
#include "gp_internal.h"
#include "gp_runtime.h"

namespace gp {

thread_local std::string g_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return fail(GP_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------------ host math
namespace {

struct M {
  double a[9];
};
M mmul(const M& A, const M& B) {
  M C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C.a[3 * i + j] = A.a[3 * i] * B.a[j] + A.a[3 * i + 1] * B.a[3 + j] + A.a[3 * i + 2] * B.a[6 + j];
  return C;
}
M mT(const M& A) { return M{{A.a[0], A.a[3], A.a[6], A.a[1], A.a[4], A.a[7], A.a[2], A.a[5], A.a[8]}}; }
void mv(const M& A, const double* v, double* o) {
  double t[3];
  for (int i = 0; i < 3; ++i) t[i] = A.a[3 * i] * v[0] + A.a[3 * i + 1] * v[1] + A.a[3 * i + 2] * v[2];
  o[0] = t[0]; o[1] = t[1]; o[2] = t[2];
}
M ident() { return M{{1, 0, 0, 0, 1, 0, 0, 0, 1}}; }
M quat_to_m(const double* q) {  // (x, y, z, w), nalgebra to_rotation_matrix
  double i = q[0], j = q[1], k = q[2], w = q[3];
  double ww = w * w, ii = i * i, jj = j * j, kk = k * k;
  double ij = 2 * i * j, wk = 2 * w * k, wj = 2 * w * j, ik = 2 * i * k, jk = 2 * j * k, wi = 2 * w * i;
  return M{{ww + ii - jj - kk, ij - wk, wj + ik, wk + ij, ww - ii + jj - kk, jk - wi, ik - wj, wi + jk, ww - ii - jj + kk}};
}
// rotation A with A e_z = a: exact signed permutation for +-e_x,y,z
M align_z(const double* a) {
  int nz = (a[0] != 0) + (a[1] != 0) + (a[2] != 0);
  auto cols = [](const double* b1, const double* b2, const double* b3) {
    return M{{b1[0], b2[0], b3[0], b1[1], b2[1], b3[1], b1[2], b2[2], b3[2]}};
  };
  if (nz == 1) {
    if (a[2] != 0) {
      double s = a[2] > 0 ? 1 : -1;
      double b1[3] = {1, 0, 0}, b2[3] = {0, s, 0}, b3[3] = {0, 0, s};
      return cols(b1, b2, b3);
    } else if (a[0] != 0) {
      double s = a[0] > 0 ? 1 : -1;
      double b1[3] = {0, s, 0}, b2[3] = {0, 0, 1}, b3[3] = {s, 0, 0};
      return cols(b1, b2, b3);
    } else {
      double s = a[1] > 0 ? 1 : -1;
      double b1[3] = {0, 0, s}, b2[3] = {1, 0, 0}, b3[3] = {0, s, 0};
      return cols(b1, b2, b3);
    }
  }
  double t[3] = {std::fabs(a[0]) < 0.9 ? 1.0 : 0.0, std::fabs(a[0]) < 0.9 ? 0.0 : 1.0, 0.0};
  double d = t[0] * a[0] + t[1] * a[1] + t[2] * a[2];
  double b1[3] = {t[0] - d * a[0], t[1] - d * a[1], t[2] - d * a[2]};
  double n = std::sqrt(b1[0] * b1[0] + b1[1] * b1[1] + b1[2] * b1[2]);
  for (double& x : b1) x /= n;
  double b2[3] = {a[1] * b1[2] - a[2] * b1[1], a[2] * b1[0] - a[0] * b1[2], a[0] * b1[1] - a[1] * b1[0]};
  return cols(b1, b2, a);
}

uint64_t fnv(uint64_t h, int64_t x) {
  for (int k = 0; k < 8; ++k) {
    h ^= (uint64_t)((x >> (8 * k)) & 0xff);
    h *= 1099511628211ull;
  }
  return h;
}

bool finite_all(const double* p, int n) {
  for (int k = 0; k < n; ++k)
    if (!std::isfinite(p[k])) return false;
  return true;
}

}  // namespace

// ------------------------------------------------------------------ model
int model_create(const gp_model_desc* d, gp_model** out) {
  if (!d || !out) return fail(GP_ERR_INVALID, "null argument");
  if (d->n_bodies < 1 || d->n_bodies > kMaxBodies)
    return fail(GP_ERR_INVALID, "n_bodies %d outside [1, %d]", d->n_bodies, kMaxBodies);
  if (!d->bodies) return fail(GP_ERR_INVALID, "bodies is null");
  if (d->n_contacts < 0 || (d->n_contacts > 0 && !d->contacts)) return fail(GP_ERR_INVALID, "bad contacts");
  if (d->n_halfspaces < 0 || (d->n_halfspaces > 0 && !d->halfspaces)) return fail(GP_ERR_INVALID, "bad half-spaces");
  if (!std::isfinite(d->gravity)) return fail(GP_ERR_INVALID, "gravity not finite");
  std::unique_ptr<gp_model> m(new (std::nothrow) gp_model());
  if (!m) return fail(GP_ERR_NOMEM, "out of host memory");
  static std::atomic<uint64_t> next_serial{1};
  m->serial = next_serial.fetch_add(1);
  const int n = d->n_bodies;
  m->n = n;
  m->g = d->gravity;
  m->parent.resize(n); m->dofcls.resize(n); m->kind.resize(n); m->qoff.resize(n); m->voff.resize(n);
  m->bp.assign((size_t)n * NBF, 0.0);
  std::vector<M> Al(n);
  for (int i = 0; i < n; ++i) {
    const gp_body_desc& b = d->bodies[i];
    // mechanism.rs:91-115 panics on a bad topology; here it is an error code
    if (b.parent < -1 || b.parent >= i)
      return fail(GP_ERR_TOPOLOGY, "body %d: parent %d must be -1 or a lower index", i, b.parent);
    if (b.jtype < GP_JOINT_FIXED || b.jtype > GP_JOINT_FLOATING)
      return fail(GP_ERR_TOPOLOGY, "body %d: unknown joint type %d", i, b.jtype);
    if (!finite_all(b.axis, 3) || !finite_all(b.x_tree_quat, 4) || !finite_all(b.x_tree_trans, 3) ||
        !finite_all(b.moment, 6) || !finite_all(b.mc, 3) || !std::isfinite(b.mass) || !std::isfinite(b.spring_k) ||
        !std::isfinite(b.spring_l))
      return fail(GP_ERR_INVALID, "body %d: non-finite parameter", i);
    if (b.mass < 0) return fail(GP_ERR_INVALID, "body %d: negative mass", i);
    double qn = std::sqrt(b.x_tree_quat[0] * b.x_tree_quat[0] + b.x_tree_quat[1] * b.x_tree_quat[1] +
                          b.x_tree_quat[2] * b.x_tree_quat[2] + b.x_tree_quat[3] * b.x_tree_quat[3]);
    if (std::fabs(qn - 1.0) > 1e-6) return fail(GP_ERR_INVALID, "body %d: x_tree_quat is not a unit quaternion", i);
    m->parent[i] = b.parent;
    int jt = b.jtype;
    m->kind[i] = jt == GP_JOINT_REVOLUTE ? K_REV : jt == GP_JOINT_PRISMATIC ? K_PRI : jt == GP_JOINT_FLOATING ? K_FLOAT : K_FIXED;
    m->dofcls[i] = jt == GP_JOINT_FLOATING ? 6 : jt == GP_JOINT_FIXED ? 0 : 1;
    if (jt == GP_JOINT_PRISMATIC) m->all_rev = false;
    m->qoff[i] = m->nq;
    m->voff[i] = m->nv;
    m->nq += jt == GP_JOINT_FLOATING ? 7 : m->dofcls[i];
    m->nv += m->dofcls[i];
    if (m->dofcls[i] == 1) {
      double an = std::sqrt(b.axis[0] * b.axis[0] + b.axis[1] * b.axis[1] + b.axis[2] * b.axis[2]);
      if (std::fabs(an - 1.0) > 1e-6) return fail(GP_ERR_INVALID, "body %d: joint axis is not a unit vector", i);
      double a[3] = {b.axis[0] / an, b.axis[1] / an, b.axis[2] / an};
      if (an == 1.0) { a[0] = b.axis[0]; a[1] = b.axis[1]; a[2] = b.axis[2]; }
      Al[i] = align_z(a);
    } else {
      Al[i] = ident();
    }
    double qq[4] = {b.x_tree_quat[0] / qn, b.x_tree_quat[1] / qn, b.x_tree_quat[2] / qn, b.x_tree_quat[3] / qn};
    if (qn == 1.0) std::memcpy(qq, b.x_tree_quat, sizeof qq);
    const M Ap = b.parent < 0 ? ident() : Al[b.parent];
    const M Rt = mmul(mmul(mT(Ap), quat_to_m(qq)), Al[i]);
    double pt[3];
    mv(mT(Ap), b.x_tree_trans, pt);
    M J{{b.moment[0], b.moment[3], b.moment[4], b.moment[3], b.moment[1], b.moment[5], b.moment[4], b.moment[5], b.moment[2]}};
    M Ja = mmul(mmul(mT(Al[i]), J), Al[i]);
    double c[3];
    mv(mT(Al[i]), b.mc, c);
    double* P = &m->bp[(size_t)i * NBF];
    for (int k = 0; k < 9; ++k) P[BF_RT + k] = Rt.a[k];
    for (int k = 0; k < 3; ++k) P[BF_PT + k] = pt[k];
    P[BF_J + 0] = Ja.a[0]; P[BF_J + 1] = Ja.a[4]; P[BF_J + 2] = Ja.a[8];
    P[BF_J + 3] = Ja.a[1]; P[BF_J + 4] = Ja.a[2]; P[BF_J + 5] = Ja.a[5];
    for (int k = 0; k < 3; ++k) P[BF_C + k] = c[k];
    P[BF_M] = b.mass;
    P[BF_SK] = jt == GP_JOINT_PRISMATIC ? b.spring_k : 0.0;
    P[BF_SL] = b.spring_l;
  }
  // contacts, stably sorted by body (contact.rs:103-128 visits bodies in order)
  std::vector<int> order(d->n_contacts);
  for (int c = 0; c < d->n_contacts; ++c) {
    const gp_contact_desc& cd = d->contacts[c];
    if (cd.body < 0 || cd.body >= n) return fail(GP_ERR_INVALID, "contact %d: body %d out of range", c, cd.body);
    if (!finite_all(cd.p, 3) || !std::isfinite(cd.k) || cd.k <= 0) return fail(GP_ERR_INVALID, "contact %d: bad parameters", c);
    order[c] = c;
  }
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return d->contacts[x].body < d->contacts[y].body; });
  m->nc = d->n_contacts;
  m->cstart.assign(n + 1, 0);
  m->cp.assign((size_t)m->nc * NCF, 0.0);
  m->contact_order = order;
  for (int s = 0; s < m->nc; ++s) {
    const gp_contact_desc& cd = d->contacts[order[s]];
    m->cstart[cd.body + 1]++;
    double p[3];
    mv(mT(Al[cd.body]), cd.p, p);
    for (int k = 0; k < 3; ++k) m->cp[(size_t)s * NCF + CF_P + k] = p[k];
    m->cp[(size_t)s * NCF + CF_K] = cd.k;
  }
  for (int i = 0; i < n; ++i) m->cstart[i + 1] += m->cstart[i];
  m->nh = d->n_halfspaces;
  m->hp.assign((size_t)m->nh * NHF, 0.0);
  for (int h = 0; h < m->nh; ++h) {
    const gp_halfspace_desc& hd = d->halfspaces[h];
    if (!finite_all(hd.point, 3) || !finite_all(hd.normal, 3) || !std::isfinite(hd.alpha) || !std::isfinite(hd.mu))
      return fail(GP_ERR_INVALID, "half-space %d: non-finite parameter", h);
    double nn = std::sqrt(hd.normal[0] * hd.normal[0] + hd.normal[1] * hd.normal[1] + hd.normal[2] * hd.normal[2]);
    if (std::fabs(nn - 1.0) > 1e-6) return fail(GP_ERR_INVALID, "half-space %d: normal is not a unit vector", h);
    double* H = &m->hp[(size_t)h * NHF];
    for (int k = 0; k < 3; ++k) { H[HF_POINT + k] = hd.point[k]; H[HF_NORMAL + k] = hd.normal[k]; }
    H[HF_ALPHA] = hd.alpha;
    H[HF_MU] = hd.mu;
  }
  // spring contacts (contact.rs:74-94), stably sorted by body, directions in
  // the aligned body frames; A^T kept to store a world direction the same way
  if (d->n_spring_contacts < 0 || (d->n_spring_contacts > 0 && !d->spring_contacts))
    return fail(GP_ERR_INVALID, "bad spring contacts");
  std::vector<int> sorder(d->n_spring_contacts);
  for (int c = 0; c < d->n_spring_contacts; ++c) {
    const gp_spring_contact_desc& sd = d->spring_contacts[c];
    if (sd.body < 0 || sd.body >= n) return fail(GP_ERR_INVALID, "spring contact %d: body %d out of range", c, sd.body);
    if (!finite_all(sd.direction, 3) || !std::isfinite(sd.l_rest) || !std::isfinite(sd.k) || sd.l_rest < 0)
      return fail(GP_ERR_INVALID, "spring contact %d: bad parameters", c);
    const double dn = std::sqrt(sd.direction[0] * sd.direction[0] + sd.direction[1] * sd.direction[1] +
                                sd.direction[2] * sd.direction[2]);
    if (std::fabs(dn - 1.0) > 1e-6) return fail(GP_ERR_INVALID, "spring contact %d: direction is not a unit vector", c);
    sorder[c] = c;
  }
  std::stable_sort(sorder.begin(), sorder.end(),
                   [&](int x, int y) { return d->spring_contacts[x].body < d->spring_contacts[y].body; });
  m->nsp = d->n_spring_contacts;
  m->sstart.assign(n + 1, 0);
  m->sp.assign((size_t)m->nsp * NSF, 0.0);
  m->spring_slot.assign(m->nsp, 0);
  for (int s = 0; s < m->nsp; ++s) {
    const gp_spring_contact_desc& sd = d->spring_contacts[sorder[s]];
    m->spring_slot[sorder[s]] = s;
    m->sstart[sd.body + 1]++;
    double* S = &m->sp[(size_t)s * NSF];
    double dir[3];
    mv(mT(Al[sd.body]), sd.direction, dir);
    for (int k = 0; k < 3; ++k) S[SF_DIR + k] = dir[k];
    S[SF_L] = sd.l_rest;
    S[SF_K] = sd.k;
    const M At = mT(Al[sd.body]);
    for (int k = 0; k < 9; ++k) S[SF_AT + k] = At.a[k];
  }
  for (int i = 0; i < n; ++i) m->sstart[i + 1] += m->sstart[i];
  // sphere colliders (collision/sphere.rs:6-39), at most one per body
  // (RigidBody::collider is an Option), centres in the aligned body frames
  if (d->n_spheres < 0 || (d->n_spheres > 0 && !d->spheres)) return fail(GP_ERR_INVALID, "bad sphere colliders");
  m->sphere_of.assign(n, -1);
  for (int c = 0; c < d->n_spheres; ++c) {
    const gp_sphere_desc& sd = d->spheres[c];
    if (sd.body < 0 || sd.body >= n) return fail(GP_ERR_INVALID, "sphere %d: body %d out of range", c, sd.body);
    if (m->sphere_of[sd.body] >= 0) return fail(GP_ERR_INVALID, "sphere %d: body %d already has a collider", c, sd.body);
    if (!finite_all(sd.center, 3) || !std::isfinite(sd.radius) || sd.radius < 0)
      return fail(GP_ERR_INVALID, "sphere %d: bad parameters", c);
    m->sphere_of[sd.body] = 0;
  }
  m->ns = 0;
  for (int i = 0; i < n; ++i)
    if (m->sphere_of[i] >= 0) m->sphere_of[i] = m->ns++;
  m->shp.assign((size_t)m->ns * NSH, 0.0);
  for (int c = 0; c < d->n_spheres; ++c) {
    const gp_sphere_desc& sd = d->spheres[c];
    double* S = &m->shp[(size_t)m->sphere_of[sd.body] * NSH];
    double ca[3];
    mv(mT(Al[sd.body]), sd.center, ca);
    for (int k = 0; k < 3; ++k) S[SH_C + k] = ca[k];
    S[SH_R] = sd.radius;
    S[SH_ON] = sd.enabled ? 1.0 : 0.0;
  }
  // topology signature: bodies, parents, DOF class, contacts per body, half-spaces
  uint64_t h = 1469598103934665603ull;
  h = fnv(h, n);
  for (int i = 0; i < n; ++i) {
    h = fnv(h, m->parent[i]);
    h = fnv(h, m->dofcls[i]);
    h = fnv(h, m->cstart[i + 1] - m->cstart[i]);
    h = fnv(h, m->sstart[i + 1] - m->sstart[i]);
    h = fnv(h, m->sphere_of[i] >= 0 ? 1 : 0);
  }
  h = fnv(h, m->nh);
  m->sig = h;
  // spring contacts hold per-instance state the compiled-topology kernels do not carry
  m->fast_topo = m->nsp > 0 ? -1 : fast_topology_id(n, m->parent.data(), m->dofcls.data(), m->nc);
  *out = m.release();
  return GP_OK;
}

// ------------------------------------------------------------------ buckets
Bucket::~Bucket() { release(); }

void Bucket::release() {
  void* ptrs[] = {d_arena, d_scratch, d_active, d_flat_off};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  d_arena = nullptr;
  d_flat_off = nullptr;
  d_sstart = nullptr;
  d_sp = d_ss = nullptr;
  d_sphere_of = nullptr;
  d_shp = nullptr;
  scratch_doubles = 0;
  d_parent = d_qoff = d_voff = d_cstart = d_kind = d_fault = d_active = d_rounds = nullptr;
  d_bp = d_cp = d_hp = d_q = d_v = d_tau = d_tau_tmp = d_vdot = d_scratch = nullptr;
  d_bpf = d_cpf = d_hpf = nullptr;
}

// Allocate and fill one bucket from the models of its lanes: every array of
// the bucket lives in one device allocation (256-byte aligned pieces), filled
// by one host-to-device copy. fill_lanes = false (serving pools, whose slots
// are written at admission) skips the per-lane packing: the lanes start zeroed.
int bucket_build(Bucket& b, const std::vector<const gp_model*>& lane_models, bool fill_lanes) {
  const gp_model* m0 = lane_models[0];
  b.nb = m0->n; b.nq = m0->nq; b.nv = m0->nv; b.nc = m0->nc; b.nh = m0->nh; b.nsp = m0->nsp; b.g = m0->g;
  b.ns = m0->ns; b.sphere_of = m0->sphere_of;
  b.parent = m0->parent; b.qoff = m0->qoff; b.voff = m0->voff; b.dofcls = m0->dofcls; b.cstart = m0->cstart;
  b.sstart = m0->sstart; b.spring_slot = m0->spring_slot;
  b.fast_topo = m0->fast_topo;
  b.all_rev = true;
  b.B = (int)lane_models.size();
  b.Bp = (b.B + 63) / 64 * 64;
  const size_t Bp = b.Bp;
  const size_t nq1 = std::max(1, b.nq), nv1 = std::max(1, b.nv);
  // layout
  size_t at = 0;
  auto take = [&](size_t bytes) {
    const size_t o = at;
    at += (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
    return o;
  };
  const size_t o_parent = take(4 * b.nb), o_qoff = take(4 * b.nb), o_voff = take(4 * b.nb),
               o_cstart = take(4 * (b.nb + 1)), o_kind = take(4 * b.nb * Bp), o_bp = take(8 * NBF * b.nb * Bp),
               o_bpf = take(4 * NBF * b.nb * Bp), o_cp = take(8 * NCF * b.nc * Bp), o_cpf = take(4 * NCF * b.nc * Bp),
               o_hp = take(8 * NHF * b.nh * Bp), o_hpf = take(4 * NHF * b.nh * Bp), o_q = take(8 * nq1 * Bp),
               o_v = take(8 * nv1 * Bp), o_tau = take(8 * nv1 * Bp), o_tau_tmp = take(8 * nv1 * Bp),
               o_vdot = take(8 * nv1 * Bp), o_fault = take(4 * Bp), o_sstart = take(4 * (b.nb + 1)),
               o_sp = take(8 * NSF * b.nsp * Bp), o_ss = take(8 * NSS * b.nsp * Bp),
               o_sphere_of = take(4 * b.nb), o_shp = take(8 * NSH * b.ns * Bp), o_rounds = take(4 * Bp);
  std::vector<char> h(at, 0);
  auto put = [&](size_t off, const void* src, size_t bytes) { if (bytes) std::memcpy(h.data() + off, src, bytes); };
  put(o_parent, b.parent.data(), 4 * b.nb);
  put(o_qoff, b.qoff.data(), 4 * b.nb);
  put(o_voff, b.voff.data(), 4 * b.nb);
  put(o_cstart, b.cstart.data(), 4 * (b.nb + 1));
  put(o_sstart, b.sstart.data(), 4 * (b.nb + 1));
  put(o_sphere_of, b.sphere_of.data(), 4 * b.nb);
  double* shp = (double*)(h.data() + o_shp);
  double* sp = (double*)(h.data() + o_sp);
  int* kind = (int*)(h.data() + o_kind);
  double* bp = (double*)(h.data() + o_bp);
  float* bpf = (float*)(h.data() + o_bpf);
  double* cp = (double*)(h.data() + o_cp);
  float* cpf = (float*)(h.data() + o_cpf);
  double* hp = (double*)(h.data() + o_hp);
  float* hpf = (float*)(h.data() + o_hpf);
  for (int l = 0; l < b.B; ++l) {
    const gp_model* m = lane_models[l];
    for (int i = 0; i < b.nb; ++i)
      if (m->kind[i] == K_PRI) b.all_rev = false;
    if (!fill_lanes) continue;
    for (int i = 0; i < b.nb; ++i) {
      kind[i * Bp + l] = m->kind[i];
      for (int f = 0; f < NBF; ++f) {
        const double x = m->bp[(size_t)i * NBF + f];
        bp[((size_t)f * b.nb + i) * Bp + l] = x;
        bpf[((size_t)f * b.nb + i) * Bp + l] = (float)x;
      }
    }
    for (int c = 0; c < b.nc; ++c)
      for (int f = 0; f < NCF; ++f) {
        const double x = m->cp[(size_t)c * NCF + f];
        cp[((size_t)f * b.nc + c) * Bp + l] = x;
        cpf[((size_t)f * b.nc + c) * Bp + l] = (float)x;
      }
    for (int hh = 0; hh < b.nh; ++hh)
      for (int f = 0; f < NHF; ++f) {
        const double x = m->hp[(size_t)hh * NHF + f];
        hp[((size_t)f * b.nh + hh) * Bp + l] = x;
        hpf[((size_t)f * b.nh + hh) * Bp + l] = (float)x;
      }
    for (int s = 0; s < b.nsp; ++s)
      for (int f = 0; f < NSF; ++f) sp[((size_t)f * b.nsp + s) * Bp + l] = m->sp[(size_t)s * NSF + f];
    for (int s = 0; s < b.ns; ++s)
      for (int f = 0; f < NSH; ++f) shp[((size_t)f * b.ns + s) * Bp + l] = m->shp[(size_t)s * NSH + f];
  }
  // zero state (mechanism.rs:68-88): q = 0, floating q = identity pose, v = 0, tau = 0
  double* q0 = (double*)(h.data() + o_q);
  for (int i = 0; i < b.nb; ++i)
    if (b.dofcls[i] == 6)
      for (size_t l = 0; l < Bp; ++l) q0[(size_t)(b.qoff[i] + 3) * Bp + l] = 1.0;
  hipError_t e = hipMalloc(&b.d_arena, at);
  if (e != hipSuccess) return fail(GP_ERR_NOMEM, "hipMalloc(%zu bytes): %s", at, hipGetErrorString(e));
  e = hipMemcpy(b.d_arena, h.data(), at, hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail(GP_ERR_DEVICE, "hipMemcpy H2D: %s", hipGetErrorString(e));
  char* d = (char*)b.d_arena;
  b.d_parent = (int*)(d + o_parent); b.d_qoff = (int*)(d + o_qoff); b.d_voff = (int*)(d + o_voff);
  b.d_cstart = (int*)(d + o_cstart); b.d_kind = (int*)(d + o_kind);
  b.d_bp = (double*)(d + o_bp); b.d_bpf = (float*)(d + o_bpf); b.d_cp = (double*)(d + o_cp);
  b.d_cpf = (float*)(d + o_cpf); b.d_hp = (double*)(d + o_hp); b.d_hpf = (float*)(d + o_hpf);
  b.d_q = (double*)(d + o_q); b.d_v = (double*)(d + o_v); b.d_tau = (double*)(d + o_tau);
  b.d_tau_tmp = (double*)(d + o_tau_tmp); b.d_vdot = (double*)(d + o_vdot); b.d_fault = (int*)(d + o_fault);
  b.d_sstart = (int*)(d + o_sstart); b.d_sp = (double*)(d + o_sp); b.d_ss = (double*)(d + o_ss);
  b.d_sphere_of = (int*)(d + o_sphere_of); b.d_shp = (double*)(d + o_shp); b.d_rounds = (int*)(d + o_rounds);
  b.arena_bytes = at;
  return GP_OK;
}

// Scratch of the runtime-topology kernels: room for the plain kernel (per-body
// fields and Runge-Kutta stages) and for the velocity-stepping kernel.
size_t bucket_scratch_doubles(const Bucket& b) {
  return std::max<size_t>({(size_t)1, plain_scratch_doubles(b.nb, b.nq, b.nv, b.Bp),
                           vs_scratch_doubles(b.nb, b.nq, b.nv, b.nc, b.nh, b.ns, b.Bp)});
}

int bucket_ensure_scratch(Bucket& b) {
  const size_t need = bucket_scratch_doubles(b);
  if (b.d_scratch && b.scratch_doubles >= need) return GP_OK;
  if (b.d_scratch) {
    (void)hipFree(b.d_scratch);
    b.d_scratch = nullptr;
    b.scratch_doubles = 0;
  }
  const size_t bytes = sizeof(double) * need;
  hipError_t e = hipMalloc((void**)&b.d_scratch, bytes);
  if (e != hipSuccess) return fail(GP_ERR_NOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
  b.scratch_doubles = need;
  return GP_OK;
}

int64_t soa_max_lanes(const gp_model& m) {
  const int64_t per_lane = 8 * std::max<int64_t>({(int64_t)NBF * m.n, (int64_t)NCF * m.nc, (int64_t)NHF * m.nh,
                                                   (int64_t)NSF * m.nsp, (int64_t)m.nq, (int64_t)m.nv, 1});
  return std::max<int64_t>(64, ((((int64_t)1 << 31) - 1) / per_lane) / 64 * 64);
}

BucketArgs bucket_args(const Bucket& b, int lanes, const double* tau) {
  BucketArgs a{};
  a.B = b.B; a.Bp = b.Bp; a.nb = b.nb; a.nq = b.nq; a.nv = b.nv; a.nc = b.nc; a.nh = b.nh;
  a.lanes = lanes; a.g = b.g; a.all_rev = b.all_rev ? 1 : 0;
  a.parent = b.d_parent; a.qoff = b.d_qoff; a.voff = b.d_voff; a.cstart = b.d_cstart;
  a.kind = b.d_kind; a.bp = b.d_bp; a.bpf = b.d_bpf; a.cp = b.d_cp; a.cpf = b.d_cpf; a.hp = b.d_hp; a.hpf = b.d_hpf;
  a.q = b.d_q; a.v = b.d_v; a.tau = tau ? tau : b.d_tau; a.vdot = b.d_vdot; a.fault = b.d_fault;
  a.rounds = b.d_rounds;
  a.active = b.d_active; a.scratch = b.d_scratch;
  a.nsp = b.nsp; a.sstart = b.d_sstart; a.sp = b.d_sp; a.ss = b.d_ss;
  a.ns = b.ns; a.sphere_of = b.d_sphere_of; a.shp = b.d_shp; a.no_springs = 0;
  return a;
}

// Time of one step of one wave of a bucket's compiled kernel in microseconds,
// by topology and precision (measured on MI355X with 16,384 instances per
// topology, before the FP64 chains' LDS tier; entries for 5, 7, 9 and 11
// joints interpolated). Per-lane revolute/prismatic kinds add about a fifth.
// Only the ratios matter: they order a grouped launch's buckets and set the
// serving loop's steps per pool.
double bucket_step_cost(const Bucket& b, int precision, int lanes_per_tree) {
  static const double chain_mixed[12] = {0.25, 1.07, 2.13, 3.08, 4.2, 5.27, 6.4, 7.6, 11.5, 15.5, 17.9, 20.2};
  static const double chain_f64[12] = {0.5, 1.35, 2.77, 4.16, 7.3, 10.5, 14.2, 17.9, 33.4, 48.9, 61.7, 74.5};
  const bool mixed = precision == GP_PREC_MIXED;
  const int t = b.fast_topo;
  if (t >= 1 && t <= 12) return (mixed ? chain_mixed : chain_f64)[t - 1] * (b.all_rev ? 1.0 : 1.2);
  if (t > TOPO_WELDED && t <= TOPO_WELDED + kMaxWeldedJoints)  // a welded base costs nothing per step
    return (mixed ? chain_mixed : chain_f64)[t - TOPO_WELDED - 1] * (b.all_rev ? 1.0 : 1.2);
  if (t == TOPO_QUAD && lanes_per_tree == 1) return mixed ? 29.4 : 116.0;
  if (t == TOPO_QUAD) return mixed ? 7.0 : 14.3;   // one lane per leg
  if (t == TOPO_FREE) return mixed ? 1.4 : 1.7;    // four lanes per body, 8 contact points
  return mixed ? 20.0 : 75.0;
}

}  // namespace gp

using namespace gp;

// ------------------------------------------------------------------ batch
static bool is_velocity_stepping(int integrator) {
  return integrator == GP_VELOCITY_STEPPING || integrator == GP_CCD_VELOCITY_STEPPING;
}
static bool bucket_is_fast(const gp_batch_opts& o, const Bucket& b, int integrator) {
  // the one-body-per-lane tree kernel is semi-implicit Euler only
  return o.kernel != GP_KERNEL_PLAIN && b.fast_topo >= 0 && !is_velocity_stepping(integrator) &&
         !(b.fast_topo == TOPO_TREE && integrator != GP_SEMI_IMPLICIT_EULER);
}
static int bucket_lanes(const gp_batch_opts& o, const Bucket& b, int integrator) {
  // the tree-parallel quadruped kernel is semi-implicit Euler only
  const int tree = integrator == GP_SEMI_IMPLICIT_EULER ? o.lanes_per_tree : 1;
  return fast_lanes(b.fast_topo, o.lanes_per_mech, tree, b.nb);
}
static int group_block_size(const gp_batch_opts& o) { return o.block_size > 0 ? o.block_size : 128; }

// Number of FAST buckets that one grouped launch would cover (0 = no grouping).
static int grouped_count(const gp_batch* bt, int integrator) {
  const gp_batch_opts& o = bt->opts;
  if (o.separate_launches || integrator != GP_SEMI_IMPLICIT_EULER) return 0;
  int n = 0;
  for (const Bucket& b : bt->buckets) n += bucket_is_fast(o, b, integrator) ? 1 : 0;
  return (n >= 2 && n <= kMaxGroup) ? n : 0;
}

// Build (or reuse) the device table of the grouped launch. Synchronous: it
// must run before any graph capture.
static int ensure_group(gp_batch* bt, int integrator) {
  const gp_batch_opts& o = bt->opts;
  const int block = group_block_size(o);
  if (bt->group_valid && bt->group_block == block) return GP_OK;
  GroupTable t{};
  int blocks = 0;
  // costliest steps first: workgroups start in index order, so the longest
  // ones do not form the launch's tail
  std::vector<const Bucket*> order;
  for (const Bucket& b : bt->buckets)
    if (bucket_is_fast(o, b, integrator)) order.push_back(&b);
  const int tree = integrator == GP_SEMI_IMPLICIT_EULER ? o.lanes_per_tree : 1;
  std::stable_sort(order.begin(), order.end(), [&](const Bucket* x, const Bucket* y) {
    return bucket_step_cost(*x, o.precision, tree) > bucket_step_cost(*y, o.precision, tree);
  });
  for (const Bucket* bb : order) {
    const Bucket& b = *bb;
    const int lanes = bucket_lanes(o, b, integrator);
    t.args[t.n] = bucket_args(b, lanes, nullptr);
    t.topo[t.n] = b.fast_topo;
    t.block_start[t.n] = blocks;
    blocks += (int)(((int64_t)b.B * lanes + block - 1) / block);
    ++t.n;
  }
  t.block_start[t.n] = blocks;
  if (!bt->d_group) {
    hipError_t e = hipMalloc((void**)&bt->d_group, sizeof(GroupTable));
    if (e != hipSuccess) return fail(GP_ERR_NOMEM, "hipMalloc(group table): %s", hipGetErrorString(e));
  }
  HIP_TRY(hipMemcpy(bt->d_group, &t, sizeof t, hipMemcpyHostToDevice));
  bt->group_blocks = blocks;
  bt->group_block = block;
  bt->group_valid = true;
  return GP_OK;
}

// Contact candidates of a bucket that the single-body velocity-stepping
// kernel can take (one floating body, at most kVsBodyMaxCandidates point /
// sphere x half-space pairs), else -1: such a bucket runs the runtime-
// topology kernel.
static int vs_body_candidate_count(const Bucket& b) {
  if (b.nb != 1 || b.dofcls.empty() || b.dofcls[0] != 6) return -1;
  const int P = b.nc * b.nh + (b.ns > 0 ? b.nh : 0);
  return P <= kVsBodyMaxCandidates ? P : -1;
}

// Every bucket's launches for n_steps steps. scratch_override: where the
// runtime-topology and velocity-stepping kernels keep their per-lane working
// set (nullptr: each bucket's own buffer, allocated on first use; else that
// array, which must hold the largest bucket's, and no grouped launch, whose
// table is allocated on first use).
static int batch_launch_all(gp_batch* bt, double dt, int integrator, int n_steps, hipStream_t s, int64_t* launches,
                            double* scratch_override = nullptr) {
  const gp_batch_opts& o = bt->opts;
  const bool grouped = scratch_override == nullptr && grouped_count(bt, integrator) > 0;
  if (grouped) {
    if (!bt->group_valid) return fail(GP_ERR_INVALID, "group table not prepared");
    int spl = o.steps_per_launch > 0 ? o.steps_per_launch : 100;
    for (int done = 0; done < n_steps; done += spl) {
      int n = std::min(spl, n_steps - done);
      int e = launch_grouped_step(bt->d_group, bt->group_blocks, o.precision, dt, n, bt->group_block, s);
      if (e) return fail(GP_ERR_DEVICE, "grouped kernel launch: %s", hipGetErrorString((hipError_t)e));
      ++*launches;
    }
  }
  for (Bucket& b : bt->buckets) {
    bool fast = bucket_is_fast(o, b, integrator);
    if (o.kernel == GP_KERNEL_FAST && !fast)
      return fail(GP_ERR_UNSUPPORTED, "no FAST kernel for this bucket's topology / integrator");
    if (fast && grouped) continue;
    if (is_velocity_stepping(integrator)) {
      if (o.precision == GP_PREC_MIXED && o.kernel == GP_KERNEL_PLAIN)
        return fail(GP_ERR_UNSUPPORTED, "velocity stepping computes in FP64 only");
      const int ccd = integrator == GP_CCD_VELOCITY_STEPPING ? 1 : 0;
      const int spl = o.steps_per_launch > 0 ? o.steps_per_launch : 100;
      const int cand = vs_body_candidate_count(b);
      if (o.kernel != GP_KERNEL_PLAIN && cand >= 0) {
        // a lone floating body: the register-resident kernel, several lanes per instance
        const int lanes = o.lanes_per_mech == 1 ? 1 : 8;
        BucketArgs a = bucket_args(b, lanes, nullptr);
        for (int done = 0; done < n_steps; done += spl) {
          const int n = std::min(spl, n_steps - done);
          int e = launch_vs_body(a, dt, ccd, n, lanes, cand, o.block_size, s);
          if (e) return fail(GP_ERR_DEVICE, "velocity-stepping launch: %s", hipGetErrorString((hipError_t)e));
          ++*launches;
        }
        continue;
      }
      int st = scratch_override ? GP_OK : bucket_ensure_scratch(b);
      if (st) return st;
      BucketArgs a = bucket_args(b, 1, nullptr);
      if (scratch_override) a.scratch = scratch_override;
      for (int done = 0; done < n_steps; done += spl) {
        const int n = std::min(spl, n_steps - done);
        int e = launch_vs_step(a, dt, ccd, n, s);
        if (e) return fail(GP_ERR_DEVICE, "velocity-stepping launch: %s", hipGetErrorString((hipError_t)e));
        ++*launches;
      }
      continue;
    }
    if (fast) {
      int spl = o.steps_per_launch > 0 ? o.steps_per_launch : 100;
      int lanes = bucket_lanes(o, b, integrator);
      BucketArgs a = bucket_args(b, lanes, nullptr);
      for (int done = 0; done < n_steps; done += spl) {
        int n = std::min(spl, n_steps - done);
        int e = integrator == GP_SEMI_IMPLICIT_EULER
                    ? launch_fast_step(b.fast_topo, o.precision, a, dt, n, o.block_size, s)
                    : launch_fast_rk_step(b.fast_topo, o.precision, integrator, a, dt, n, o.block_size, s);
        if (e) return fail(GP_ERR_DEVICE, "fast kernel launch: %s", hipGetErrorString((hipError_t)e));
        ++*launches;
      }
    } else {
      if (o.precision == GP_PREC_MIXED && o.kernel == GP_KERNEL_PLAIN)
        return fail(GP_ERR_UNSUPPORTED, "the plain kernel computes in FP64 only");
      int st = scratch_override ? GP_OK : bucket_ensure_scratch(b);
      if (st) return st;
      BucketArgs a = bucket_args(b, 1, nullptr);
      if (scratch_override) a.scratch = scratch_override;
      const int spl = o.steps_per_launch > 0 ? o.steps_per_launch : 100;
      for (int done = 0; done < n_steps; done += spl) {
        int e = launch_plain_step(a, dt, integrator, 0, std::min(spl, n_steps - done), s);
        if (e) return fail(GP_ERR_DEVICE, "plain kernel launch: %s", hipGetErrorString((hipError_t)e));
        ++*launches;
      }
    }
  }
  return GP_OK;
}

// ------------------------------------------------------------------ no-allocation step
// gp_step_noalloc: the kernels' scratch (the runtime-topology and
// velocity-stepping kernels' per-lane working set) comes from one file-scope
// device array per GPU instead of the buckets' lazily allocated buffers; no
// graph, no grouped launch table, no allocation of any kind. Every call on a
// GPU shares that GPU's array: calls that overlap on one GPU (other streams,
// other threads) race on it.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of this GPU's `scratch` (host code may not use the
// array's name as an address), looked up once per device.
int noalloc_scratch(double** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(GP_ERR_DEVICE, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(GP_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return GP_OK;
}
}  // namespace

extern "C" {

const char* gp_last_error(void) { return g_error.c_str(); }
int gp_abi_version(void) { return GP_ABI_VERSION; }

int gp_model_create(const gp_model_desc* desc, gp_model** out) {
  try {
    return model_create(desc, out);
  } catch (const std::bad_alloc&) {
    return fail(GP_ERR_NOMEM, "out of host memory");
  } catch (...) {
    return fail(GP_ERR_INVALID, "unexpected exception");
  }
}

void gp_model_destroy(gp_model* model) { delete model; }

int gp_model_dims(const gp_model* m, int32_t* nq, int32_t* nv, int32_t* nb) {
  if (!m) return fail(GP_ERR_INVALID, "null model");
  if (nq) *nq = m->nq;
  if (nv) *nv = m->nv;
  if (nb) *nb = m->n;
  return GP_OK;
}

int gp_model_signature(const gp_model* m, uint64_t* sig) {
  if (!m || !sig) return fail(GP_ERR_INVALID, "null argument");
  *sig = m->sig;
  return GP_OK;
}

static int validate_opts(const gp_batch_opts& o) {
  if (o.precision != GP_PREC_F64 && o.precision != GP_PREC_MIXED) return fail(GP_ERR_INVALID, "unknown precision %d", o.precision);
  if (o.kernel < GP_KERNEL_AUTO || o.kernel > GP_KERNEL_FAST) return fail(GP_ERR_INVALID, "unknown kernel %d", o.kernel);
  if (o.steps_per_launch < 0) return fail(GP_ERR_INVALID, "steps_per_launch < 0");
  if (o.block_size != 0 && (o.block_size < 64 || o.block_size > 1024 || o.block_size % 64))
    return fail(GP_ERR_INVALID, "block_size must be a multiple of 64 in [64, 256]");
  if (o.block_size > 256) return fail(GP_ERR_INVALID, "block_size above the kernels' launch bound of 256");
  if (o.lanes_per_mech != 0 && o.lanes_per_mech != 1 && o.lanes_per_mech != 2 && o.lanes_per_mech != 4 &&
      o.lanes_per_mech != 8)
    return fail(GP_ERR_INVALID, "lanes_per_mech must be 1, 2, 4 or 8");
  if (o.lanes_per_tree != 0 && o.lanes_per_tree != 1 && o.lanes_per_tree != 4)
    return fail(GP_ERR_INVALID, "lanes_per_tree must be 1 or 4");
  return GP_OK;
}

int gp_batch_create(gp_model* const* models, const int64_t* counts, int32_t n_models, const gp_batch_opts* opts,
                    gp_batch** out) {
  if (!models || !counts || !out || n_models < 1) return fail(GP_ERR_INVALID, "null argument or no models");
  try {
    std::unique_ptr<gp_batch> bt(new gp_batch());
    if (opts) bt->opts = *opts;
    int st = validate_opts(bt->opts);
    if (st) return st;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(GP_ERR_DEVICE, "no HIP device");
    // group instances by topology signature (first-seen order)
    std::map<uint64_t, int> sig_bucket;
    std::vector<std::vector<const gp_model*>> lanes;
    std::vector<std::vector<int64_t>> inst_of_lane;
    int64_t inst = 0;
    for (int k = 0; k < n_models; ++k) {
      if (!models[k]) return fail(GP_ERR_INVALID, "model %d is null", k);
      if (counts[k] < 0) return fail(GP_ERR_INVALID, "count %d is negative", k);
      const int64_t cap = soa_max_lanes(*models[k]);
      for (int64_t c = 0; c < counts[k]; ++c) {
        // a full bucket (its largest array would reach 2^31 bytes) starts another of the same signature
        auto it = sig_bucket.find(models[k]->sig);
        int bi;
        if (it == sig_bucket.end() || (int64_t)lanes[it->second].size() >= cap) {
          bi = (int)lanes.size();
          sig_bucket[models[k]->sig] = bi;
          lanes.emplace_back();
          inst_of_lane.emplace_back();
        } else {
          bi = it->second;
        }
        lanes[bi].push_back(models[k]);
        inst_of_lane[bi].push_back(inst++);
      }
    }
    if (inst == 0) return fail(GP_ERR_INVALID, "the batch has no instances");
    if (inst > (int64_t)1 << 31) return fail(GP_ERR_INVALID, "too many instances");
    bt->n_inst = inst;
    bt->inst_bucket.assign(inst, -1);
    bt->inst_lane.assign(inst, -1);
    bt->buckets.resize(lanes.size());
    for (size_t bi = 0; bi < lanes.size(); ++bi) {
      if (lanes[bi].empty()) continue;
      Bucket& b = bt->buckets[bi];
      b.inst = inst_of_lane[bi];
      if ((st = bucket_build(b, lanes[bi]))) return st;
      for (int l = 0; l < b.B; ++l) { bt->inst_bucket[b.inst[l]] = (int)bi; bt->inst_lane[b.inst[l]] = l; }
    }
    // drop empty buckets (all counts 0 for a signature)
    bt->buckets.erase(std::remove_if(bt->buckets.begin(), bt->buckets.end(), [](const Bucket& b) { return b.B == 0; }),
                      bt->buckets.end());
    for (size_t bi = 0; bi < bt->buckets.size(); ++bi)
      for (int l = 0; l < bt->buckets[bi].B; ++l) bt->inst_bucket[bt->buckets[bi].inst[l]] = (int)bi;
    bt->qoff.assign(inst + 1, 0);
    bt->voff.assign(inst + 1, 0);
    for (int64_t i = 0; i < inst; ++i) {
      const Bucket& b = bt->buckets[bt->inst_bucket[i]];
      bt->qoff[i + 1] = bt->qoff[i] + b.nq;
      bt->voff[i + 1] = bt->voff[i] + b.nv;
    }
    // each lane's offsets into the flattened arrays, for the device-array calls
    for (Bucket& b : bt->buckets) {
      std::vector<int64_t> off(2 * (size_t)b.B);
      for (int l = 0; l < b.B; ++l) {
        off[l] = bt->qoff[b.inst[l]];
        off[b.B + l] = bt->voff[b.inst[l]];
      }
      HIP_TRY(hipMalloc((void**)&b.d_flat_off, off.size() * sizeof(int64_t)));
      HIP_TRY(hipMemcpy(b.d_flat_off, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipStreamCreate(&bt->stream));
    HIP_TRY(hipEventCreate(&bt->ev0));
    HIP_TRY(hipEventCreate(&bt->ev1));
    *out = bt.release();
    return GP_OK;
  } catch (const std::bad_alloc&) {
    return fail(GP_ERR_NOMEM, "out of host memory");
  }
}

void gp_batch_destroy(gp_batch* bt) {
  if (!bt) return;
  (void)hipDeviceSynchronize();
  for (auto& kv : bt->graphs) (void)hipGraphExecDestroy(kv.second);
  if (bt->d_group) (void)hipFree(bt->d_group);
  if (bt->stream) (void)hipStreamDestroy(bt->stream);
  if (bt->ev0) (void)hipEventDestroy(bt->ev0);
  if (bt->ev1) (void)hipEventDestroy(bt->ev1);
  delete bt;
}

int gp_batch_info(const gp_batch* bt, int64_t* n, int64_t* nq, int64_t* nv, int32_t* nb) {
  if (!bt) return fail(GP_ERR_INVALID, "null batch");
  if (n) *n = bt->n_inst;
  if (nq) *nq = bt->qoff.back();
  if (nv) *nv = bt->voff.back();
  if (nb) *nb = (int32_t)bt->buckets.size();
  return GP_OK;
}

int gp_batch_offsets(const gp_batch* bt, int64_t* qo, int64_t* vo) {
  if (!bt) return fail(GP_ERR_INVALID, "null batch");
  if (qo) std::copy(bt->qoff.begin(), bt->qoff.end(), qo);
  if (vo) std::copy(bt->voff.begin(), bt->voff.end(), vo);
  return GP_OK;
}

int gp_batch_set_opts(gp_batch* bt, const gp_batch_opts* o) {
  if (!bt || !o) return fail(GP_ERR_INVALID, "null argument");
  int st = validate_opts(*o);
  if (st) return st;
  HIP_TRY(hipStreamSynchronize(bt->stream));
  bt->opts = *o;
  for (auto& kv : bt->graphs) (void)hipGraphExecDestroy(kv.second);
  bt->graphs.clear();
  bt->graph_launches.clear();
  bt->group_valid = false;
  return GP_OK;
}

int gp_batch_get_opts(const gp_batch* bt, gp_batch_opts* o) {
  if (!bt || !o) return fail(GP_ERR_INVALID, "null argument");
  *o = bt->opts;
  return GP_OK;
}

// host flattened (instance-major) <-> bucket SoA
static int transfer(gp_batch* bt, const double* hq, const double* hv, double* oq, double* ov, bool to_dev,
                     int which /*0 state, 1 tau, 2 vdot, 3 tau_tmp*/) {
  HIP_TRY(hipStreamSynchronize(bt->stream));
  for (Bucket& b : bt->buckets) {
    const size_t Bp = b.Bp;
    const int nq = which == 0 ? b.nq : 0;
    const int nv = b.nv;
    std::vector<double> sq((size_t)nq * Bp), sv((size_t)nv * Bp);
    double* dq = b.d_q;
    double* dv = which == 0 ? b.d_v : which == 1 ? b.d_tau : which == 2 ? b.d_vdot : b.d_tau_tmp;
    if (!to_dev) {
      if (nq) HIP_TRY(hipMemcpy(sq.data(), dq, sq.size() * sizeof(double), hipMemcpyDeviceToHost));
      if (nv) HIP_TRY(hipMemcpy(sv.data(), dv, sv.size() * sizeof(double), hipMemcpyDeviceToHost));
    } else if (which == 0 && b.B < (int)Bp) {
      // padding lanes keep a valid zero state
      for (int i = 0; i < b.nb; ++i)
        if (b.dofcls[i] == 6)
          for (size_t l = b.B; l < Bp; ++l) sq[(size_t)(b.qoff[i] + 3) * Bp + l] = 1.0;
    }
    for (int l = 0; l < b.B; ++l) {
      const int64_t i = b.inst[l];
      if (to_dev) {
        for (int k = 0; k < nq; ++k) sq[(size_t)k * Bp + l] = hq[bt->qoff[i] + k];
        for (int k = 0; k < nv; ++k) sv[(size_t)k * Bp + l] = hv ? hv[bt->voff[i] + k] : 0.0;
      } else {
        for (int k = 0; k < nq; ++k) oq[bt->qoff[i] + k] = sq[(size_t)k * Bp + l];
        for (int k = 0; k < nv; ++k) ov[bt->voff[i] + k] = sv[(size_t)k * Bp + l];
      }
    }
    if (to_dev) {
      if (nq) HIP_TRY(hipMemcpy(dq, sq.data(), sq.size() * sizeof(double), hipMemcpyHostToDevice));
      if (nv) HIP_TRY(hipMemcpy(dv, sv.data(), sv.size() * sizeof(double), hipMemcpyHostToDevice));
    }
  }
  return GP_OK;
}

// device arrays <-> bucket SoA on `stream` (which = 0 state, 1 tau, 2 vdot)
static int transfer_dev(gp_batch* bt, const double* q, const double* v, double* oq, double* ov, int which,
                        void* stream) {
  hipStream_t s = stream ? (hipStream_t)stream : bt->stream;
  if (stream) bt->foreign_stream = true;
  const bool to_dev = oq == nullptr && ov == nullptr;
  for (Bucket& b : bt->buckets) {
    int e = 0;
    if (which == 0)
      e = launch_flat_xfer(b.d_q, b.Bp, b.B, b.nq, b.d_flat_off, to_dev ? const_cast<double*>(q) : oq, to_dev, s);
    if (e) return fail(GP_ERR_DEVICE, "transposition launch: %s", hipGetErrorString((hipError_t)e));
    double* dv = which == 0 ? b.d_v : which == 1 ? b.d_tau : b.d_vdot;
    e = launch_flat_xfer(dv, b.Bp, b.B, b.nv, b.d_flat_off + b.B, to_dev ? const_cast<double*>(v) : ov, to_dev, s);
    if (e) return fail(GP_ERR_DEVICE, "transposition launch: %s", hipGetErrorString((hipError_t)e));
  }
  return GP_OK;
}

int gp_batch_set_state_dev(gp_batch* bt, const double* q, const double* v, void* stream) {
  if (!bt || !q || !v) return fail(GP_ERR_INVALID, "null argument");
  return transfer_dev(bt, q, v, nullptr, nullptr, 0, stream);
}

int gp_batch_get_state_dev(gp_batch* bt, double* q, double* v, void* stream) {
  if (!bt || !q || !v) return fail(GP_ERR_INVALID, "null argument");
  return transfer_dev(bt, nullptr, nullptr, q, v, 0, stream);
}

int gp_batch_set_tau_dev(gp_batch* bt, const double* tau, void* stream) {
  if (!bt || !tau) return fail(GP_ERR_INVALID, "null argument");
  return transfer_dev(bt, nullptr, tau, nullptr, nullptr, 1, stream);
}

int gp_batch_get_vdot_dev(gp_batch* bt, double* vdot, void* stream) {
  if (!bt || !vdot) return fail(GP_ERR_INVALID, "null argument");
  return transfer_dev(bt, nullptr, nullptr, nullptr, vdot, 2, stream);
}

int gp_batch_set_spring_contact(gp_batch* bt, int64_t inst, int32_t index, const double* direction, double l_rest) {
  if (!bt || !direction) return fail(GP_ERR_INVALID, "null argument");
  if (inst < 0 || inst >= bt->n_inst) return fail(GP_ERR_INVALID, "instance %lld out of range", (long long)inst);
  Bucket& b = bt->buckets[bt->inst_bucket[inst]];
  if (index < 0 || index >= b.nsp) return fail(GP_ERR_INVALID, "spring contact %d out of range", index);
  if (!finite_all(direction, 3) || !std::isfinite(l_rest) || l_rest < 0) return fail(GP_ERR_INVALID, "bad parameters");
  const int s = b.spring_slot[index], l = bt->inst_lane[inst];
  HIP_TRY(hipStreamSynchronize(bt->stream));
  double at[9];
  for (int k = 0; k < 9; ++k)
    HIP_TRY(hipMemcpy(&at[k], b.d_sp + ((size_t)(SF_AT + k) * b.nsp + s) * b.Bp + l, 8, hipMemcpyDeviceToHost));
  double vals[4];
  for (int r = 0; r < 3; ++r) vals[r] = at[3 * r] * direction[0] + at[3 * r + 1] * direction[1] + at[3 * r + 2] * direction[2];
  vals[3] = l_rest;
  for (int k = 0; k < 4; ++k)
    HIP_TRY(hipMemcpy(b.d_sp + ((size_t)(SF_DIR + k) * b.nsp + s) * b.Bp + l, &vals[k], 8, hipMemcpyHostToDevice));
  return GP_OK;
}

int gp_batch_get_spring_contact(gp_batch* bt, int64_t inst, int32_t index, int32_t* registered, int32_t* halfspace,
                                double* point, double* direction, double* l_rest) {
  if (!bt) return fail(GP_ERR_INVALID, "null batch");
  if (inst < 0 || inst >= bt->n_inst) return fail(GP_ERR_INVALID, "instance %lld out of range", (long long)inst);
  Bucket& b = bt->buckets[bt->inst_bucket[inst]];
  if (index < 0 || index >= b.nsp) return fail(GP_ERR_INVALID, "spring contact %d out of range", index);
  const int s = b.spring_slot[index], l = bt->inst_lane[inst];
  HIP_TRY(hipStreamSynchronize(bt->stream));
  auto sp = [&](int f) {
    double x = 0;
    (void)hipMemcpy(&x, b.d_sp + ((size_t)f * b.nsp + s) * b.Bp + l, 8, hipMemcpyDeviceToHost);
    return x;
  };
  auto ss = [&](int f) {
    double x = 0;
    (void)hipMemcpy(&x, b.d_ss + ((size_t)f * b.nsp + s) * b.Bp + l, 8, hipMemcpyDeviceToHost);
    return x;
  };
  const bool reg = ss(SS_REG) != 0.0;
  if (registered) *registered = reg ? 1 : 0;
  if (halfspace) *halfspace = reg ? (int32_t)ss(SS_H) : -1;
  if (point)
    for (int k = 0; k < 3; ++k) point[k] = ss(SS_P + k);
  if (direction) {  // back to the reference's body frame: A d (A^T stored row-major)
    double at[9], da[3];
    for (int k = 0; k < 9; ++k) at[k] = sp(SF_AT + k);
    for (int k = 0; k < 3; ++k) da[k] = sp(SF_DIR + k);
    for (int r = 0; r < 3; ++r) direction[r] = at[r] * da[0] + at[3 + r] * da[1] + at[6 + r] * da[2];
  }
  if (l_rest) *l_rest = sp(SF_L);
  HIP_TRY(hipGetLastError());
  return GP_OK;
}

int gp_batch_set_state(gp_batch* bt, const double* q, const double* v) {
  if (!bt || !q || !v) return fail(GP_ERR_INVALID, "null argument");
  return transfer(bt, q, v, nullptr, nullptr, true, 0);
}

int gp_batch_get_state(gp_batch* bt, double* q, double* v) {
  if (!bt || !q || !v) return fail(GP_ERR_INVALID, "null argument");
  return transfer(bt, nullptr, nullptr, q, v, false, 0);
}

int gp_batch_set_tau(gp_batch* bt, const double* tau) {
  if (!bt) return fail(GP_ERR_INVALID, "null batch");
  return transfer(bt, nullptr, tau, nullptr, nullptr, true, 1);
}

static int check_step_args(gp_batch* bt, double dt, int32_t integrator, int32_t n_steps) {
  if (!bt) return fail(GP_ERR_INVALID, "null batch");
  if (!std::isfinite(dt)) return fail(GP_ERR_INVALID, "dt is not finite");
  if (integrator < GP_SEMI_IMPLICIT_EULER || integrator > GP_CCD_VELOCITY_STEPPING)
    return fail(GP_ERR_INVALID, "unknown integrator %d", integrator);
  if (n_steps < 0) return fail(GP_ERR_INVALID, "n_steps < 0");
  if (integrator == GP_RUNGE_KUTTA_2 || integrator == GP_RUNGE_KUTTA_4)
    for (const Bucket& b : bt->buckets)
      if (b.nsp > 0)  // simulate.rs:56-71 asserts the same
        return fail(GP_ERR_UNSUPPORTED, "Runge-Kutta on a state with spring contacts");
  return GP_OK;
}

// Steps that one launch of every bucket covers; a rollout is a series of such
// chunks and a shorter last one.
static int chunk_steps(const gp_batch* bt) {
  return bt->opts.steps_per_launch > 0 ? bt->opts.steps_per_launch : 100;
}

// The cached graph of one chunk of `steps` steps (one launch per bucket, or one
// grouped launch), captured and instantiated on first use.
static int chunk_graph(gp_batch* bt, double dt, int integrator, int steps, hipStream_t s, hipGraphExec_t* ex,
                       int64_t* launches) {
  const auto key = std::make_tuple(dt, integrator, steps);
  auto it = bt->graphs.find(key);
  if (it == bt->graphs.end()) {
    // bucket scratch must exist before capture (no allocation inside a graph)
    for (Bucket& b : bt->buckets) {
      int st = bucket_ensure_scratch(b);
      if (st) return st;
    }
    int64_t n = 0;
    hipGraph_t g;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int st = batch_launch_all(bt, dt, integrator, steps, s, &n);
    hipError_t e = hipStreamEndCapture(s, &g);
    if (st) return st;
    if (e != hipSuccess) return fail(GP_ERR_DEVICE, "graph capture: %s", hipGetErrorString(e));
    hipGraphExec_t x;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(GP_ERR_DEVICE, "graph instantiate: %s", hipGetErrorString(e));
    it = bt->graphs.emplace(key, x).first;
    bt->graph_launches[key] = n;
  }
  *ex = it->second;
  *launches = bt->graph_launches[key];
  return GP_OK;
}

// The chunks of a graph-launched rollout of n_steps: steps per graph launch,
// full chunks of steps_per_launch then the rest.
static std::vector<int> chunk_plan(gp_batch* bt, int n_steps) {
  const int spl = chunk_steps(bt);
  std::vector<int> plan;
  for (int done = 0; done < n_steps; done += spl) plan.push_back(std::min(spl, n_steps - done));
  return plan;
}

// Everything a rollout of n_steps needs before its first launch: the grouped
// launch's table and, with graphs on, the chunk graphs.
static int step_prepare(gp_batch* bt, double dt, int integrator, int n_steps, hipStream_t s) {
  if (grouped_count(bt, integrator) > 0) {
    int st = ensure_group(bt, integrator);
    if (st) return st;
  }
  if (!bt->opts.use_graph || n_steps <= 0) return GP_OK;
  hipGraphExec_t ex;
  int64_t l = 0;
  for (int c : chunk_plan(bt, n_steps)) {
    int st = chunk_graph(bt, dt, integrator, c, s, &ex, &l);
    if (st) return st;
  }
  return GP_OK;
}

int gp_step_prepare(gp_batch* bt, double dt, int32_t integrator, int32_t n_steps, void* stream) {
  int st = check_step_args(bt, dt, integrator, n_steps);
  if (st) return st;
  if (stream) bt->foreign_stream = true;
  return step_prepare(bt, dt, integrator, n_steps, stream ? (hipStream_t)stream : bt->stream);
}

int gp_step(gp_batch* bt, double dt, int32_t integrator, int32_t n_steps, void* stream) {
  int st = check_step_args(bt, dt, integrator, n_steps);
  if (st) return st;
  hipStream_t s = stream ? (hipStream_t)stream : bt->stream;
  if (stream) bt->foreign_stream = true;
  if ((st = step_prepare(bt, dt, integrator, n_steps, s))) return st;
  int64_t launches = 0;
  // the timing events bracket the launches only (any preparation is done above)
  HIP_TRY(hipEventRecord(bt->ev0, s));
  if (bt->opts.use_graph && n_steps > 0) {
    for (int c : chunk_plan(bt, n_steps)) {
      hipGraphExec_t ex;
      int64_t l = 0;
      if ((st = chunk_graph(bt, dt, integrator, c, s, &ex, &l))) return st;
      if (l == 1) {
        // a graph of one kernel saves no launches and adds the graph launch's
        // own latency (~7 us on a 0.23 ms launch): launch that kernel directly,
        // with the arguments the capture recorded
        int64_t n = 0;
        if ((st = batch_launch_all(bt, dt, integrator, c, s, &n))) return st;
        launches += n;
      } else {
        HIP_TRY(hipGraphLaunch(ex, s));
        launches += l;
      }
    }
  } else {
    if ((st = batch_launch_all(bt, dt, integrator, n_steps, s, &launches))) return st;
  }
  HIP_TRY(hipEventRecord(bt->ev1, s));
  bt->last_launches = launches;
  bt->timing_valid = true;
  return GP_OK;
}


int gp_step_noalloc(gp_batch* bt, double dt, int32_t integrator, int32_t n_steps, void* stream) {
  int st = check_step_args(bt, dt, integrator, n_steps);
  if (st) return st;
  size_t need = 0;  // the largest working set among the buckets that need one
  for (const Bucket& b : bt->buckets)
    if (!bucket_is_fast(bt->opts, b, integrator)) need = std::max(need, bucket_scratch_doubles(b) * sizeof(double));
  if (need > kNoAllocScratchBytes)
    return fail(GP_ERR_UNSUPPORTED,
                "gp_step_noalloc: a bucket needs %zu bytes of scratch, more than the %zu of the static array; "
                "use gp_step or fewer instances per batch",
                need, kNoAllocScratchBytes);
  double* scr = nullptr;
  if ((st = noalloc_scratch(&scr))) return st;
  int64_t launches = 0;
  hipStream_t s = stream ? (hipStream_t)stream : bt->stream;
  if (stream) bt->foreign_stream = true;
  if ((st = batch_launch_all(bt, dt, integrator, n_steps, s, &launches, scr))) return st;
  bt->timing_valid = false;
  return GP_OK;
}

int gp_batch_sync(gp_batch* bt) {
  if (!bt) return fail(GP_ERR_INVALID, "null batch");
  HIP_TRY(hipStreamSynchronize(bt->stream));
  if (bt->foreign_stream) {
    // work the caller put on streams of its own: wait for the whole device
    HIP_TRY(hipDeviceSynchronize());
    bt->foreign_stream = false;
  }
  return GP_OK;
}

int gp_batch_get_vdot(gp_batch* bt, double* vdot) {
  if (!bt || !vdot) return fail(GP_ERR_INVALID, "null argument");
  return transfer(bt, nullptr, nullptr, nullptr, vdot, false, 2);
}

int gp_dynamics(gp_batch* bt, const double* tau, double* vdot) {
  if (!bt || !vdot) return fail(GP_ERR_INVALID, "null argument");
  int st;
  if (tau && (st = transfer(bt, nullptr, tau, nullptr, nullptr, true, 3))) return st;
  for (Bucket& b : bt->buckets) {
    if ((st = bucket_ensure_scratch(b))) return st;
    BucketArgs a = bucket_args(b, 1, tau ? b.d_tau_tmp : nullptr);
    int e = launch_plain_step(a, 0.0, GP_SEMI_IMPLICIT_EULER, 1, 1, bt->stream);
    if (e) return fail(GP_ERR_DEVICE, "dynamics launch: %s", hipGetErrorString((hipError_t)e));
  }
  return transfer(bt, nullptr, nullptr, nullptr, vdot, false, 2);
}

// M (instance-major, each nv_i x nv_i row-major) and / or c - tau_c (the v
// flattening) of every instance at the batch's state
static int joint_space_terms(gp_batch* bt, double* Mh, double* ch) {
  std::vector<int64_t> moff(bt->n_inst + 1, 0);
  for (int64_t i = 0; i < bt->n_inst; ++i) {
    const int64_t nv = bt->voff[i + 1] - bt->voff[i];
    moff[i + 1] = moff[i] + nv * nv;
  }
  HIP_TRY(hipStreamSynchronize(bt->stream));
  for (Bucket& b : bt->buckets) {
    const size_t rows = (size_t)b.nv * b.nv + b.nv;
    if (b.nv == 0 || b.B == 0) continue;
    int st = bucket_ensure_scratch(b);
    if (st) return st;
    double* d = nullptr;
    hipError_t e = hipMalloc((void**)&d, rows * b.Bp * sizeof(double));
    if (e != hipSuccess) return fail(GP_ERR_NOMEM, "hipMalloc(terms): %s", hipGetErrorString(e));
    std::vector<double> h(rows * b.Bp);
    BucketArgs a = bucket_args(b, 1, nullptr);
    int le = launch_terms(a, d, bt->stream);
    if (!le) le = (int)hipStreamSynchronize(bt->stream);
    if (!le) le = (int)hipMemcpy(h.data(), d, h.size() * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (le) return fail(GP_ERR_DEVICE, "joint-space terms: %s", hipGetErrorString((hipError_t)le));
    const size_t Bp = b.Bp, nv = b.nv;
    for (int l = 0; l < b.B; ++l) {
      const int64_t i = b.inst[l];
      if (Mh)
        for (size_t k = 0; k < nv * nv; ++k) Mh[moff[i] + k] = h[k * Bp + l];
      if (ch)
        for (size_t k = 0; k < nv; ++k) ch[bt->voff[i] + k] = h[(nv * nv + k) * Bp + l];
    }
  }
  return GP_OK;
}

int gp_mass_matrix(gp_batch* bt, double* M) {
  if (!bt || !M) return fail(GP_ERR_INVALID, "null argument");
  return joint_space_terms(bt, M, nullptr);
}

int gp_dynamics_bias(gp_batch* bt, double* c) {
  if (!bt || !c) return fail(GP_ERR_INVALID, "null argument");
  return joint_space_terms(bt, nullptr, c);
}

int gp_simulate(gp_batch* bt, double final_time, double dt, gp_control_fn control_fn, void* user, int32_t integrator,
                double* qs, double* vs, int64_t max_records, int64_t* n_records) {
  if (!bt) return fail(GP_ERR_INVALID, "null batch");
  if (!std::isfinite(final_time) || !std::isfinite(dt)) return fail(GP_ERR_INVALID, "non-finite time");
  if (!(dt > 0) && final_time > 0) return fail(GP_ERR_INVALID, "dt must be positive");
  const int64_t nq = bt->qoff.back(), nv = bt->voff.back();
  std::vector<double> q(std::max<int64_t>(1, nq)), v(std::max<int64_t>(1, nv)), tau(std::max<int64_t>(1, nv), 0.0);
  int st;
  if ((st = gp_batch_get_state(bt, q.data(), v.data()))) return st;
  int64_t rec = 0;
  auto record = [&]() {
    if (rec < max_records) {
      if (qs) std::copy(q.begin(), q.begin() + nq, qs + rec * nq);
      if (vs) std::copy(v.begin(), v.begin() + nv, vs + rec * nv);
    }
    ++rec;
  };
  record();
  if (!control_fn && (st = gp_batch_set_tau(bt, nullptr))) return st;
  double t = 0.0;
  while (t < final_time) {  // simulate.rs:102-109
    if (control_fn) {
      if (control_fn(user, q.data(), v.data(), t, tau.data()) != 0) return fail(GP_ERR_INVALID, "control_fn failed");
      if ((st = gp_batch_set_tau(bt, tau.data()))) return st;
    }
    if ((st = gp_step(bt, dt, integrator, 1, nullptr))) return st;
    if ((st = gp_batch_get_state(bt, q.data(), v.data()))) return st;
    record();
    t += dt;
  }
  if (n_records) *n_records = rec;
  return GP_OK;
}

int gp_batch_energy(gp_batch* bt, double* kinetic, double* gravitational) {
  if (!bt) return fail(GP_ERR_INVALID, "null batch");
  HIP_TRY(hipStreamSynchronize(bt->stream));
  for (Bucket& b : bt->buckets) {
    int st = bucket_ensure_scratch(b);
    if (st) return st;
    double* d = nullptr;
    HIP_TRY(hipMalloc((void**)&d, sizeof(double) * 2 * b.Bp));
    BucketArgs a = bucket_args(b, 1, nullptr);
    int e = launch_energy(a, d, d + b.Bp, bt->stream);
    std::vector<double> h(2 * (size_t)b.Bp);
    hipError_t ce = e ? hipSuccess : hipMemcpyAsync(h.data(), d, sizeof(double) * 2 * b.Bp, hipMemcpyDeviceToHost,
                                                    bt->stream);
    hipError_t se = hipStreamSynchronize(bt->stream);
    (void)hipFree(d);
    if (e) return fail(GP_ERR_DEVICE, "energy launch: %s", hipGetErrorString((hipError_t)e));
    if (ce != hipSuccess || se != hipSuccess)
      return fail(GP_ERR_DEVICE, "energy copy: %s", hipGetErrorString(ce != hipSuccess ? ce : se));
    for (int l = 0; l < b.B; ++l) {
      if (kinetic) kinetic[b.inst[l]] = h[l];
      if (gravitational) gravitational[b.inst[l]] = h[b.Bp + l];
    }
  }
  return GP_OK;
}

int gp_batch_faults(gp_batch* bt, uint8_t* faults) {
  if (!bt || !faults) return fail(GP_ERR_INVALID, "null argument");
  HIP_TRY(hipStreamSynchronize(bt->stream));
  for (Bucket& b : bt->buckets) {
    std::vector<int> f(b.Bp);
    HIP_TRY(hipMemcpy(f.data(), b.d_fault, sizeof(int) * b.Bp, hipMemcpyDeviceToHost));
    for (int l = 0; l < b.B; ++l) faults[b.inst[l]] = f[l] ? 1 : 0;
  }
  return GP_OK;
}

int gp_batch_collision_rounds(gp_batch* bt, int32_t* rounds) {
  if (!bt || !rounds) return fail(GP_ERR_INVALID, "null argument");
  HIP_TRY(hipStreamSynchronize(bt->stream));
  for (Bucket& b : bt->buckets) {
    std::vector<int> r(b.Bp);
    HIP_TRY(hipMemcpy(r.data(), b.d_rounds, sizeof(int) * b.Bp, hipMemcpyDeviceToHost));
    for (int l = 0; l < b.B; ++l) rounds[b.inst[l]] = r[l];
  }
  return GP_OK;
}

int gp_batch_clear_faults(gp_batch* bt) {
  if (!bt) return fail(GP_ERR_INVALID, "null batch");
  HIP_TRY(hipStreamSynchronize(bt->stream));
  for (Bucket& b : bt->buckets) HIP_TRY(hipMemset(b.d_fault, 0, sizeof(int) * b.Bp));
  return GP_OK;
}

int gp_batch_last_timing(const gp_batch* bt, double* ms, int64_t* launches) {
  if (!bt) return fail(GP_ERR_INVALID, "null batch");
  if (!bt->timing_valid) return fail(GP_ERR_INVALID, "no gp_step has run on this batch");
  HIP_TRY(hipEventSynchronize(bt->ev1));
  float f = 0;
  HIP_TRY(hipEventElapsedTime(&f, bt->ev0, bt->ev1));
  if (ms) *ms = f;
  if (launches) *launches = bt->last_launches;
  return GP_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ autotuner
namespace {

// Device copies of every bucket's mutable state, so that trial rollouts leave
// the batch as they found it.
struct StateSnapshot {
  struct Saved { double *q = nullptr, *v = nullptr, *vdot = nullptr; int* fault = nullptr; };
  std::vector<Saved> saved;
  ~StateSnapshot() {
    for (Saved& s : saved) {
      (void)hipFree(s.q); (void)hipFree(s.v); (void)hipFree(s.vdot); (void)hipFree(s.fault);
    }
  }
  static size_t qn(const Bucket& b) { return (size_t)std::max(1, b.nq) * b.Bp; }
  static size_t vn(const Bucket& b) { return (size_t)std::max(1, b.nv) * b.Bp; }
  int take(gp_batch* bt) {
    saved.resize(bt->buckets.size());
    for (size_t i = 0; i < bt->buckets.size(); ++i) {
      const Bucket& b = bt->buckets[i];
      Saved& s = saved[i];
      HIP_TRY(hipMalloc((void**)&s.q, qn(b) * sizeof(double)));
      HIP_TRY(hipMalloc((void**)&s.v, vn(b) * sizeof(double)));
      HIP_TRY(hipMalloc((void**)&s.vdot, vn(b) * sizeof(double)));
      HIP_TRY(hipMalloc((void**)&s.fault, (size_t)b.Bp * sizeof(int)));
    }
    return copy(bt, true);
  }
  int copy(gp_batch* bt, bool save) {
    for (size_t i = 0; i < bt->buckets.size(); ++i) {
      const Bucket& b = bt->buckets[i];
      Saved& s = saved[i];
      auto cp = [&](void* dev, void* keep, size_t bytes) {
        return hipMemcpyAsync(save ? keep : dev, save ? dev : keep, bytes, hipMemcpyDeviceToDevice, bt->stream);
      };
      HIP_TRY(cp(b.d_q, s.q, qn(b) * sizeof(double)));
      HIP_TRY(cp(b.d_v, s.v, vn(b) * sizeof(double)));
      HIP_TRY(cp(b.d_vdot, s.vdot, vn(b) * sizeof(double)));
      HIP_TRY(cp(b.d_fault, s.fault, (size_t)b.Bp * sizeof(int)));
    }
    HIP_TRY(hipStreamSynchronize(bt->stream));
    return GP_OK;
  }
};

// Milliseconds of one trial rollout under the batch's current options (the
// fastest of three runs, each from the saved state).
int trial_ms(gp_batch* bt, StateSnapshot& snap, double dt, int integrator, int steps, double* ms) {
  double best = 1e300;
  for (int rep = 0; rep < 3; ++rep) {
    int st = snap.copy(bt, false);
    if (st) return st;
    if (grouped_count(bt, integrator) > 0 && (st = ensure_group(bt, integrator))) return st;
    int64_t launches = 0;
    HIP_TRY(hipEventRecord(bt->ev0, bt->stream));
    if ((st = batch_launch_all(bt, dt, integrator, steps, bt->stream, &launches))) return st;
    HIP_TRY(hipEventRecord(bt->ev1, bt->stream));
    HIP_TRY(hipEventSynchronize(bt->ev1));
    float f = 0;
    HIP_TRY(hipEventElapsedTime(&f, bt->ev0, bt->ev1));
    best = std::min(best, (double)f);
  }
  *ms = best;
  return GP_OK;
}

}  // namespace

extern "C" int gp_batch_autotune(gp_batch* bt, double dt, int32_t integrator, int32_t trial_steps,
                                 gp_batch_opts* chosen, double* best_ms) {
  if (!bt) return fail(GP_ERR_INVALID, "null batch");
  if (!(dt > 0) || !std::isfinite(dt)) return fail(GP_ERR_INVALID, "dt must be positive");
  if (integrator < GP_SEMI_IMPLICIT_EULER || integrator > GP_RUNGE_KUTTA_4)
    return fail(GP_ERR_INVALID, "unknown integrator %d", integrator);
  if (trial_steps < 0) return fail(GP_ERR_INVALID, "trial_steps < 0");
  const int steps = trial_steps > 0 ? trial_steps : 200;
  HIP_TRY(hipStreamSynchronize(bt->stream));
  bool has_free = false, has_tree = false, has_fast = false;
  for (const Bucket& b : bt->buckets) {
    const bool fast = bucket_is_fast(bt->opts, b, integrator);
    has_fast = has_fast || fast;
    has_free = has_free || (fast && b.fast_topo == TOPO_FREE);
    has_tree = has_tree || (fast && b.fast_topo == TOPO_QUAD);
  }
  const gp_batch_opts user = bt->opts;
  StateSnapshot snap;
  int st = snap.take(bt);
  if (st) return st;
  gp_batch_opts best = user;
  best.use_graph = 0;  // trials time the launches themselves
  // spell out the defaults, so that no candidate re-measures the starting point
  if (best.lanes_per_mech == 0) best.lanes_per_mech = 4;
  if (best.lanes_per_tree == 0) best.lanes_per_tree = 4;
  if (best.steps_per_launch == 0) best.steps_per_launch = 100;
  if (best.block_size == 0) best.block_size = grouped_count(bt, integrator) > 0 ? 128 : 256;
  double best_t = 0;
  auto run = [&](const gp_batch_opts& o, double* t) {
    int e = gp_batch_set_opts(bt, &o);
    return e ? e : trial_ms(bt, snap, dt, integrator, steps, t);
  };
  st = run(best, &best_t);                  // warms code objects and clocks
  if (!st) st = run(best, &best_t);
  // coordinate descent over the options that change speed but not the step:
  // lanes for contact-heavy free bodies, lanes for quadruped trees,
  // workgroup size, steps fused per launch
  struct Dim { int32_t gp_batch_opts::*field; std::vector<int> values; bool used; };
  const Dim dims[] = {
      {&gp_batch_opts::lanes_per_mech, {1, 2, 4, 8}, has_free},
      {&gp_batch_opts::lanes_per_tree, {1, 4}, has_tree && integrator == GP_SEMI_IMPLICIT_EULER},
      {&gp_batch_opts::block_size, {64, 128, 256}, has_fast},
      {&gp_batch_opts::steps_per_launch, {25, 50, 100, 200}, has_fast},
  };
  for (const Dim& d : dims) {
    if (st || !d.used) continue;
    for (int v : d.values) {
      if (best.*d.field == v) continue;
      gp_batch_opts o = best;
      o.*d.field = v;
      double t = 0;
      if ((st = run(o, &t))) break;
      if (t < 0.98 * best_t) { best_t = t; best = o; }  // a change must beat timing noise
    }
  }
  // leave the batch in its saved state with the chosen options (or, after an
  // error, the caller's)
  gp_batch_opts fin = best;
  fin.use_graph = user.use_graph;
  const int st2 = gp_batch_set_opts(bt, st ? &user : &fin);
  const int st3 = snap.copy(bt, false);
  if (st) return st;
  if (st2) return st2;
  if (st3) return st3;
  bt->timing_valid = false;
  if (chosen) *chosen = fin;
  if (best_ms) *best_ms = best_t;
  return GP_OK;
}
