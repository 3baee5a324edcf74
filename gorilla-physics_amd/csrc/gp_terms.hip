// gp_terms.hip — the joint-space terms of M(q) vdot + c(q, v) = tau + tau_c
// at a batch's state: the mass matrix M by the composite-rigid-body algorithm
// (the quantity of mechanism.rs:637-696, RBDA Table 6.2) and the bias
// c(q, v) - tau_c by recursive Newton-Euler with zero joint accelerations
// (the quantity of dynamics.rs:233-251, RBDA Table 5.1). FP64, runtime
// topology, one instance per lane, in the aligned body frames of the plain
// kernel: a 1-DOF joint's motion subspace is e_z (angular for revolute,
// linear for prismatic) and a floating joint's is the identity in the body
// frame, so every entry S_i^T (...) S_j is the reference's, whose floating
// velocity is the body twist in the body frame.
#include "gp_plain_impl.h"

namespace gp {

namespace {

// component k of the motion-subspace column c of body i's joint applied to a
// force-like 6-vector f: S_c^T f
__device__ inline double s_dot(int kind, int c, const sv<double>& f) {
  if (kind == K_REV) return f.w.z;
  if (kind == K_PRI) return f.v.z;
  return c < 3 ? (&f.w.x)[c] : (&f.v.x)[c - 3];
}
__device__ inline int n_dof(int kind) { return kind == K_FLOAT ? 6 : (kind == K_FIXED ? 0 : 1); }

// out: [nv * nv + nv][Bp]: M row-major, then c - tau_c
__global__ void __launch_bounds__(64) terms_kernel(BucketArgs a, double* out) {
  const int lane = blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= a.B) return;
  const int nb = a.nb, nv = a.nv, Bp = a.Bp;
  const Strided q{a.q + lane, Bp}, v{a.v + lane, Bp};
  const Strided M{out + lane, Bp}, c{out + (size_t)nv * nv * Bp + lane, Bp};
  PlainCtx X{a, lane};
  // transforms, velocities, velocity-product accelerations, rigid inertias,
  // bias forces v x* (I v) - f_contact; the spring contacts' state is left alone
  plain_kinematics(a, lane, q, v, true, false);
  // ---- recursive Newton-Euler with vdot = 0: a_i = X_i a_parent + c_i
  // (gravity as the base's acceleration), f_i = I_i a_i + v x* (I v) - f_ext
  for (int i = 0; i < nb; ++i) {
    const int pr = a.parent[i];
    const sv<double> ap = pr < 0 ? sv<double>{{0, 0, 0}, {0, 0, a.g}} : X.ldv(PF_ACC, pr);
    const sv<double> ai = xmotion(X.ldm(PF_E, i), X.ld3(PF_R, i), ap) + X.ldv(PF_CB, i);
    X.stv(PF_ACC, i, ai);
    X.stv(PF_PA, i, X.ldv(PF_PA, i) + mul(X.ldi(i), ai));
  }
  for (int i = nb - 1; i >= 0; --i) {  // leaves to root: tau_i = S_i^T f_i, f_parent += X^T f_i
    const int k = a.kind[i * Bp + lane];
    const sv<double> f = X.ldv(PF_PA, i);
    for (int col = 0; col < n_dof(k); ++col) c[a.voff[i] + col] = s_dot(k, col, f);
    const int pr = a.parent[i];
    if (pr >= 0) X.stv(PF_PA, pr, X.ldv(PF_PA, pr) + xtforce(X.ldm(PF_E, i), X.ld3(PF_R, i), f));
  }
  // ---- composite inertias, leaves to root (the rigid inertias are in PF_IA)
  for (int i = nb - 1; i >= 0; --i) {
    const int pr = a.parent[i];
    if (pr < 0) continue;
    ai6<double> P = X.ldi(pr);
    add_to(P, xtix(X.ldm(PF_E, i), X.ld3(PF_R, i), X.ldi(i)));
    X.sti(pr, P);
  }
  // ---- M: F = Ic_i S_i carried towards the root, M_ij = S_j^T F
  for (int i = 0; i < nb; ++i) {
    const int ki = a.kind[i * Bp + lane];
    const ai6<double> Ic = X.ldi(i);
    for (int ca = 0; ca < n_dof(ki); ++ca) {
      const int row = a.voff[i] + ca;
      const int sidx = ki == K_REV ? 2 : ki == K_PRI ? 5 : ca;
      sv<double> F = col(Ic, sidx);
      for (int cb = 0; cb < n_dof(ki); ++cb) M[row * nv + a.voff[i] + cb] = s_dot(ki, cb, F);
      for (int j = i; a.parent[j] >= 0;) {
        F = xtforce(X.ldm(PF_E, j), X.ld3(PF_R, j), F);
        j = a.parent[j];
        const int kj = a.kind[j * Bp + lane];
        for (int cb = 0; cb < n_dof(kj); ++cb) {
          const int colj = a.voff[j] + cb;
          const double m = s_dot(kj, cb, F);
          M[row * nv + colj] = m;
          M[colj * nv + row] = m;
        }
      }
    }
  }
  // entries of joints that support neither body stay zero
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j) {
      bool related = false;
      for (int k = i; k >= 0 && !related; k = a.parent[k]) related = k == j;
      for (int k = j; k >= 0 && !related; k = a.parent[k]) related = k == i;
      if (related) continue;
      const int ni = n_dof(a.kind[i * Bp + lane]), nj = n_dof(a.kind[j * Bp + lane]);
      for (int r = 0; r < ni; ++r)
        for (int cc = 0; cc < nj; ++cc) M[(a.voff[i] + r) * nv + a.voff[j] + cc] = 0.0;
    }
}

}  // namespace

int launch_terms(const BucketArgs& a, double* out, void* stream) {
  const int block = 64;
  hipLaunchKernelGGL(terms_kernel, dim3((a.B + block - 1) / block), dim3(block), 0, (hipStream_t)stream, a, out);
  return (int)hipGetLastError();
}

}  // namespace gp
