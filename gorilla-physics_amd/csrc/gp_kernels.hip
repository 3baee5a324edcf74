// gp_kernels.hip — batched forward dynamics + time integration for gfx950.
//
// One mechanism instance per lane (or per `L` lanes for contact-heavy
// single bodies). The dynamics is Featherstone's articulated-body algorithm
// (RBDA Table 7.1) in body frames whose z axis is each 1-DOF joint's axis; it
// computes the same joint accelerations as the reference's
// dynamics_continuous (dynamics.rs:322-364: CRBA mass matrix + RNEA bias + LU)
// in O(n) per step. Contact forces follow contact.rs:97-129,260-338 with the
// half-space test of halfspace.rs:39-44; the integrators follow
// integrators.rs:177-319.
//
// Two kernel families:
//  * plain: runtime topology, FP64, one step per launch, per-body
//    intermediates in device scratch (structure-of-arrays, coalesced);
//    supports every joint type, contacts, springs, SIE / RK2 / RK4.
//  * fast: compile-time topology (serial chains of 1-12 one-DOF joints, a free
//    body, a floating-base quadruped tree), per-body state in registers,
//    `n_steps` semi-implicit Euler steps per launch; arithmetic either FP64 or
//    mixed (FP32 sweeps, FP64 D solve and q/v accumulation).
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/gorilla_amd.h"
#include "gp_internal.h"
#include "gp_spatial.h"

namespace gp {

// ---------------------------------------------------------------- shared pieces
// Hunt-Crossley normal force + regularised Coulomb friction, force from the
// half-space on the point (contact.rs:260-302). z^1.5 is z * sqrt(z): NaN for
// z < 0 exactly like powf, and fmax(NaN, 0) = 0 like Rust's f64::max.
template <class R>
__device__ __forceinline__ v3<R> contact_force(R z, v3<R> n, v3<R> vel, R kA, R kB, R alpha, R mu) {
  R z_dot = -dot(vel, n);
  R zn = z * sqrt(z);
  R k = kA * kB / (kA + kB);
  R lam = R(1.5) * alpha * k;
  R pi = fmax(lam * zn * z_dot + k * zn, R(0));
  v3<R> f = pi * n;
  v3<R> vt = vel + z_dot * n;
  R vtn = sqrt(dot(vt, vt));
  if (vtn != R(0)) {
    R s = vtn / R(kSlipTol);
    R mu_eff = s > R(1) ? mu : mu * s;
    f = f - (mu_eff * pi / vtn) * vt;
  }
  return f;
}

// sin/cos of a joint angle held in FP64: exact sincos in FP64; for FP32 the
// angle is first reduced to [-pi, pi] in FP64.
__device__ __forceinline__ void sincos_q(double th, double* s, double* c) { sincos(th, s, c); }
__device__ __forceinline__ void sincos_q(double th, float* s, float* c) {
  const double two_pi = 6.283185307179586476925286766559;
  double k = rint(th * (1.0 / two_pi));
  float r = float(fma(-k, two_pi, th));
  sincosf(r, s, c);
}

// Cholesky of a symmetric positive-definite 6x6 (lower, packed by rows);
// false if a pivot is not positive.
__device__ __forceinline__ bool chol6(const double (&M)[6][6], double (&L)[21]) {
  bool ok = true;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    double s = M[j][j];
#pragma unroll
    for (int k = 0; k < j; ++k) s -= L[j * (j + 1) / 2 + k] * L[j * (j + 1) / 2 + k];
    ok = ok && (s > 0.0);
    double d = sqrt(s);
    L[j * (j + 1) / 2 + j] = d;
    double inv = 1.0 / d;
#pragma unroll
    for (int i = j + 1; i < 6; ++i) {
      double t = M[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) t -= L[i * (i + 1) / 2 + k] * L[j * (j + 1) / 2 + k];
      L[i * (i + 1) / 2 + j] = t * inv;
    }
  }
  return ok;
}
__device__ __forceinline__ void chol6_solve(const double (&L)[21], double (&x)[6]) {
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double s = x[i];
#pragma unroll
    for (int k = 0; k < i; ++k) s -= L[i * (i + 1) / 2 + k] * x[k];
    x[i] = s / L[i * (i + 1) / 2 + i];
  }
#pragma unroll
  for (int i = 5; i >= 0; --i) {
    double s = x[i];
#pragma unroll
    for (int k = i + 1; k < 6; ++k) s -= L[k * (k + 1) / 2 + i] * x[k];
    x[i] = s / L[i * (i + 1) / 2 + i];
  }
}
template <class R> __device__ __forceinline__ void to_dense(const ai6<R>& I, double (&M)[6][6]) {
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    sv<R> c = col(I, j);
    M[0][j] = c.w.x; M[1][j] = c.w.y; M[2][j] = c.w.z; M[3][j] = c.v.x; M[4][j] = c.v.y; M[5][j] = c.v.z;
  }
}

// Floating-joint position update q(k+1) = q(k) (+) dt v (integrators.rs:296-319,
// util.rs:83-102): translation by the OLD rotation, quaternion derivative
// 1/2 q (x) (0, w_body), then renormalisation (UnitQuaternion::from_quaternion).
__device__ __forceinline__ void advance_floating(const double* qi, const double* vi, double dt, double* qo) {
  double x = qi[0], y = qi[1], z = qi[2], w = qi[3];
  double ox = vi[0], oy = vi[1], oz = vi[2];
  double dw = (-x * ox - y * oy - z * oz) / 2.0;
  double dx = (w * ox - z * oy + y * oz) / 2.0;
  double dy = (z * ox + w * oy - x * oz) / 2.0;
  double dz = (-y * ox + x * oy + w * oz) / 2.0;
  v3<double> td = quat_rot(x, y, z, w, v3<double>{vi[3], vi[4], vi[5]});
  double t0 = qi[4] + td.x * dt, t1 = qi[5] + td.y * dt, t2 = qi[6] + td.z * dt;
  double nx = x + dx * dt, ny = y + dy * dt, nz = z + dz * dt, nw = w + dw * dt;
  double inv = 1.0 / sqrt(nx * nx + ny * ny + nz * nz + nw * nw);
  qo[0] = nx * inv; qo[1] = ny * inv; qo[2] = nz * inv; qo[3] = nw * inv;
  qo[4] = t0; qo[5] = t1; qo[6] = t2;
}

// ============================================================ plain kernel
struct Strided {  // a[k] of one lane in a [k][Bp] array
  double* p;
  int Bp;
  __device__ double& operator[](int k) const { return p[(size_t)k * Bp]; }
};

struct PlainCtx {
  const BucketArgs& a;
  int lane;
  __device__ double bpar(int f, int i) const { return a.bp[((size_t)f * a.nb + i) * a.Bp + lane]; }
  __device__ double& s(int f, int i) const { return a.scratch[((size_t)f * a.nb + i) * a.Bp + lane]; }
  __device__ void st3(int f, int i, v3<double> x) const { s(f, i) = x.x; s(f + 1, i) = x.y; s(f + 2, i) = x.z; }
  __device__ v3<double> ld3(int f, int i) const { return {s(f, i), s(f + 1, i), s(f + 2, i)}; }
  __device__ void stv(int f, int i, const sv<double>& x) const { st3(f, i, x.w); st3(f + 3, i, x.v); }
  __device__ sv<double> ldv(int f, int i) const { return {ld3(f, i), ld3(f + 3, i)}; }
  __device__ void stm(int f, int i, const m3<double>& M) const {
    for (int k = 0; k < 9; ++k) s(f + k, i) = M.a[k];
  }
  __device__ m3<double> ldm(int f, int i) const {
    m3<double> M;
    for (int k = 0; k < 9; ++k) M.a[k] = s(f + k, i);
    return M;
  }
  __device__ void sti(int i, const ai6<double>& I) const {
    int f = PF_IA;
    s(f++, i) = I.A.xx; s(f++, i) = I.A.yy; s(f++, i) = I.A.zz; s(f++, i) = I.A.xy; s(f++, i) = I.A.xz; s(f++, i) = I.A.yz;
    for (int k = 0; k < 9; ++k) s(f++, i) = I.B.a[k];
    s(f++, i) = I.C.xx; s(f++, i) = I.C.yy; s(f++, i) = I.C.zz; s(f++, i) = I.C.xy; s(f++, i) = I.C.xz; s(f++, i) = I.C.yz;
  }
  __device__ ai6<double> ldi(int i) const {
    ai6<double> I;
    int f = PF_IA;
    I.A.xx = s(f++, i); I.A.yy = s(f++, i); I.A.zz = s(f++, i); I.A.xy = s(f++, i); I.A.xz = s(f++, i); I.A.yz = s(f++, i);
    for (int k = 0; k < 9; ++k) I.B.a[k] = s(f++, i);
    I.C.xx = s(f++, i); I.C.yy = s(f++, i); I.C.zz = s(f++, i); I.C.xy = s(f++, i); I.C.xz = s(f++, i); I.C.yz = s(f++, i);
    return I;
  }
};

// Articulated-body forward dynamics of one lane; q, v, tau, vdot are the
// lane's columns of [k][Bp] arrays. Returns false on a non-positive D.
__device__ bool plain_dynamics(const BucketArgs& a, int lane, Strided q, Strided v, Strided tau, Strided vdot) {
  PlainCtx X{a, lane};
  const int nb = a.nb, Bp = a.Bp;
  bool ok = true;
  // ---- pass 1: kinematics, velocities, bias forces, contact forces
  for (int i = 0; i < nb; ++i) {
    const int k = a.kind[i * Bp + lane];
    const int pr = a.parent[i];
    const int qo = a.qoff[i], vo = a.voff[i];
    m3<double> Rt;
    for (int f = 0; f < 9; ++f) Rt.a[f] = X.bpar(BF_RT + f, i);
    v3<double> pt{X.bpar(BF_PT, i), X.bpar(BF_PT + 1, i), X.bpar(BF_PT + 2, i)};
    m3<double> Rm = Rt;
    v3<double> p = pt;
    sv<double> vj{{0, 0, 0}, {0, 0, 0}};
    if (k == K_REV) {
      double s, c;
      sincos(q[qo], &s, &c);
      m3<double> Rz{{c, -s, 0.0, s, c, 0.0, 0.0, 0.0, 1.0}};
      Rm = mul(Rt, Rz);
      vj.w.z = v[vo];
    } else if (k == K_PRI) {
      double d = q[qo];
      p = pt + d * v3<double>{Rt.a[2], Rt.a[5], Rt.a[8]};
      vj.v.z = v[vo];
    } else if (k == K_FLOAT) {
      m3<double> Rq = quat_mat(q[qo], q[qo + 1], q[qo + 2], q[qo + 3]);
      Rm = mul(Rt, Rq);
      p = pt + mul(Rt, v3<double>{q[qo + 4], q[qo + 5], q[qo + 6]});
      vj.w = v3<double>{v[vo], v[vo + 1], v[vo + 2]};
      vj.v = v3<double>{v[vo + 3], v[vo + 4], v[vo + 5]};
    }
    m3<double> E{{Rm.a[0], Rm.a[3], Rm.a[6], Rm.a[1], Rm.a[4], Rm.a[7], Rm.a[2], Rm.a[5], Rm.a[8]}};
    X.stm(PF_E, i, E);
    X.st3(PF_R, i, p);
    sv<double> vel = pr < 0 ? vj : xmotion(E, p, X.ldv(PF_VEL, pr)) + vj;
    X.stv(PF_VEL, i, vel);
    X.stv(PF_CB, i, crm(vel, vj));
    s3<double> J{X.bpar(BF_J, i), X.bpar(BF_J + 1, i), X.bpar(BF_J + 2, i), X.bpar(BF_J + 3, i), X.bpar(BF_J + 4, i),
                 X.bpar(BF_J + 5, i)};
    v3<double> cm{X.bpar(BF_C, i), X.bpar(BF_C + 1, i), X.bpar(BF_C + 2, i)};
    ai6<double> I = rigid(J, cm, X.bpar(BF_M, i));
    sv<double> f = crf(vel, mul(I, vel));
    if (a.nc > 0) {
      m3<double> Rw = Rm;
      v3<double> tw = p;
      if (pr >= 0) {
        m3<double> Rwp = X.ldm(PF_RW, pr);
        Rw = mul(Rwp, Rm);
        tw = X.ld3(PF_TW, pr) + mul(Rwp, p);
      }
      X.stm(PF_RW, i, Rw);
      X.st3(PF_TW, i, tw);
      for (int cc = a.cstart[i]; cc < a.cstart[i + 1]; ++cc) {
        auto cpar = [&](int fld) { return a.cp[((size_t)fld * a.nc + cc) * Bp + lane]; };
        v3<double> pb{cpar(CF_P), cpar(CF_P + 1), cpar(CF_P + 2)};
        double kc = cpar(CF_K);
        v3<double> pw = mul(Rw, pb) + tw;
        v3<double> vw = mul(Rw, vel.v + cross(vel.w, pb));
        for (int h = 0; h < a.nh; ++h) {
          auto hpar = [&](int fld) { return a.hp[((size_t)fld * a.nh + h) * Bp + lane]; };
          v3<double> pt_h{hpar(HF_POINT), hpar(HF_POINT + 1), hpar(HF_POINT + 2)};
          v3<double> n{hpar(HF_NORMAL), hpar(HF_NORMAL + 1), hpar(HF_NORMAL + 2)};
          double d = dot(pw - pt_h, n);
          if (d <= kInsideMargin) {
            v3<double> F = contact_force(-d, n, vw, kc, kHalfspaceK, hpar(HF_ALPHA), hpar(HF_MU));
            v3<double> Fb = mulT(Rw, F);
            f.w = f.w - cross(pb, Fb);
            f.v = f.v - Fb;
          }
        }
      }
    }
    X.stv(PF_PA, i, f);
    X.sti(i, I);
  }
  // ---- pass 2: articulated inertias and bias forces, leaves to root
  for (int i = nb - 1; i >= 0; --i) {
    const int k = a.kind[i * Bp + lane];
    const int pr = a.parent[i];
    const int vo = a.voff[i];
    ai6<double> IA = X.ldi(i);
    sv<double> pA = X.ldv(PF_PA, i);
    sv<double> cb = X.ldv(PF_CB, i);
    ai6<double> Ia;
    sv<double> pa;
    if (k == K_REV || k == K_PRI) {
      const int sidx = k == K_REV ? 2 : 5;
      sv<double> U = col(IA, sidx);
      double D = comp(U, sidx);
      double tq = tau[vo];
      double sk = X.bpar(BF_SK, i);
      if (k == K_PRI && sk != 0.0) tq += -sk * (q[a.qoff[i]] - X.bpar(BF_SL, i));  // dynamics.rs:298-315
      double u = tq - comp(pA, sidx);
      ok = ok && (D > 0.0);
      X.stv(PF_U, i, U);
      X.s(PF_D, i) = D;
      X.s(PF_UU, i) = u;
      double inv = 1.0 / D;
      sv<double> uD{inv * U.w, inv * U.v};
      Ia = rank1_down(IA, U, uD);
      pa = pA + mul(Ia, cb);
      pa = pa + sv<double>{u * uD.w, u * uD.v};
    } else if (k == K_FLOAT) {
      sv<double> u{{tau[vo] - pA.w.x, tau[vo + 1] - pA.w.y, tau[vo + 2] - pA.w.z},
                   {tau[vo + 3] - pA.v.x, tau[vo + 4] - pA.v.y, tau[vo + 5] - pA.v.z}};
      X.stv(PF_UU, i, u);
      Ia = ai6<double>{};
      pa = pA + u;  // Ia = 0 and U D^-1 u = u for S = 1
    } else {
      Ia = IA;
      pa = pA + mul(IA, cb);
    }
    if (pr >= 0) {
      m3<double> E = X.ldm(PF_E, i);
      v3<double> r = X.ld3(PF_R, i);
      ai6<double> P = X.ldi(pr);
      add_to(P, xtix(E, r, Ia));
      X.sti(pr, P);
      X.stv(PF_PA, pr, X.ldv(PF_PA, pr) + xtforce(E, r, pa));
    }
  }
  // ---- pass 3: accelerations, root to leaves
  for (int i = 0; i < nb; ++i) {
    const int k = a.kind[i * Bp + lane];
    const int pr = a.parent[i];
    const int vo = a.voff[i];
    m3<double> E = X.ldm(PF_E, i);
    v3<double> r = X.ld3(PF_R, i);
    sv<double> ap_parent = pr < 0 ? sv<double>{{0, 0, 0}, {0, 0, a.g}} : X.ldv(PF_ACC, pr);
    sv<double> ap = xmotion(E, r, ap_parent) + X.ldv(PF_CB, i);
    if (k == K_REV || k == K_PRI) {
      sv<double> U = X.ldv(PF_U, i);
      double qdd = (X.s(PF_UU, i) - dot6(U, ap)) / X.s(PF_D, i);
      vdot[vo] = qdd;
      if (k == K_REV) ap.w.z += qdd; else ap.v.z += qdd;
    } else if (k == K_FLOAT) {
      ai6<double> IA = X.ldi(i);
      double M[6][6], L[21];
      to_dense(IA, M);
      ok = ok && chol6(M, L);
      sv<double> u = X.ldv(PF_UU, i);
      sv<double> Ia_ap = mul(IA, ap);
      double x[6] = {u.w.x - Ia_ap.w.x, u.w.y - Ia_ap.w.y, u.w.z - Ia_ap.w.z,
                     u.v.x - Ia_ap.v.x, u.v.y - Ia_ap.v.y, u.v.z - Ia_ap.v.z};
      chol6_solve(L, x);
      for (int t = 0; t < 6; ++t) vdot[vo + t] = x[t];
      ap = ap + sv<double>{{x[0], x[1], x[2]}, {x[3], x[4], x[5]}};
    }
    X.stv(PF_ACC, i, ap);
  }
  return ok;
}

// q_out = q (+) h v  for every joint of the lane
__device__ void plain_advance(const BucketArgs& a, int lane, Strided q, Strided v, double h, Strided qo) {
  for (int i = 0; i < a.nb; ++i) {
    const int k = a.kind[i * a.Bp + lane];
    const int o = a.qoff[i], w = a.voff[i];
    if (k == K_REV || k == K_PRI) {
      qo[o] = q[o] + v[w] * h;
    } else if (k == K_FLOAT) {
      double qi[7], vi[6], out[7];
      for (int t = 0; t < 7; ++t) qi[t] = q[o + t];
      for (int t = 0; t < 6; ++t) vi[t] = v[w + t];
      advance_floating(qi, vi, h, out);
      for (int t = 0; t < 7; ++t) qo[o + t] = out[t];
    }
  }
}

__device__ bool finite_cols(Strided x, int n) {
  bool ok = true;
  for (int k = 0; k < n; ++k) ok = ok && isfinite(x[k]);
  return ok;
}

// One step (or one dynamics evaluation) per launch; one lane per instance.
__global__ void __launch_bounds__(256) plain_step_kernel(BucketArgs a, double dt, int integrator, int dynamics_only) {
  const int lane = blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= a.B) return;
  if (a.active && a.active[lane] <= 0) return;
  const int Bp = a.Bp;
  Strided q{a.q + lane, Bp}, v{a.v + lane, Bp}, tau{const_cast<double*>(a.tau) + lane, Bp};
  Strided vdot{a.vdot + lane, Bp};
  double* stage = a.scratch + (size_t)NPF * a.nb * Bp;  // RK stages after the per-body fields
  Strided q1{stage + lane, Bp};
  Strided v1{stage + (size_t)a.nq * Bp + lane, Bp};
  Strided acc[4];
  for (int s = 0; s < 4; ++s) acc[s] = Strided{stage + (size_t)(a.nq + a.nv * (1 + s)) * Bp + lane, Bp};
  bool ok = plain_dynamics(a, lane, q, v, tau, vdot);
  if (!dynamics_only) {
    const int nv = a.nv;
    if (integrator == GP_SEMI_IMPLICIT_EULER) {  // integrators.rs:276-292
      for (int k = 0; k < nv; ++k) v[k] = v[k] + vdot[k] * dt;
      plain_advance(a, lane, q, v, dt, q);
    } else {
      // euler_step (integrators.rs:230-271): v1 = v + h a; q1 = q (+) h v
      auto euler = [&](Strided ak, double h, Strided qo, Strided vo_) {
        plain_advance(a, lane, q, v, h, qo);
        for (int k = 0; k < nv; ++k) vo_[k] = v[k] + ak[k] * h;
      };
      for (int k = 0; k < nv; ++k) acc[0][k] = vdot[k];
      if (integrator == GP_RUNGE_KUTTA_2) {  // integrators.rs:177-193
        euler(acc[0], dt / 2.0, q1, v1);
        ok = plain_dynamics(a, lane, q1, v1, tau, acc[1]) && ok;
        euler(acc[1], dt, q1, v1);
      } else {  // RK4, integrators.rs:195-225
        euler(acc[0], dt / 2.0, q1, v1);
        ok = plain_dynamics(a, lane, q1, v1, tau, acc[1]) && ok;
        euler(acc[1], dt / 2.0, q1, v1);
        ok = plain_dynamics(a, lane, q1, v1, tau, acc[2]) && ok;
        euler(acc[2], dt, q1, v1);
        ok = plain_dynamics(a, lane, q1, v1, tau, acc[3]) && ok;
        for (int k = 0; k < nv; ++k) acc[0][k] = (acc[0][k] + acc[1][k] * 2.0 + acc[2][k] * 2.0 + acc[3][k]) / 6.0;
        euler(acc[0], dt, q1, v1);
      }
      for (int k = 0; k < a.nq; ++k) q[k] = q1[k];
      for (int k = 0; k < nv; ++k) v[k] = v1[k];
    }
  }
  ok = ok && finite_cols(vdot, a.nv) && finite_cols(q, a.nq) && finite_cols(v, a.nv);
  if (!ok) a.fault[lane] = 1;
}

size_t plain_scratch_doubles(int nb, int nq, int nv, int Bp) {
  return ((size_t)NPF * nb + nq + 5 * (size_t)nv) * Bp;
}

int launch_plain_step(const BucketArgs& a, double dt, int integrator, int dynamics_only, void* stream) {
  if (a.B == 0) return 0;
  dim3 block(256), grid((a.B + 255) / 256);
  hipLaunchKernelGGL(plain_step_kernel, grid, block, 0, (hipStream_t)stream, a, dt, integrator, dynamics_only);
  return (int)hipGetLastError();
}

// ============================================================ fast kernels
template <int N_>
struct TopoChain {
  static constexpr int N = N_;
  static constexpr __host__ __device__ int parent(int i) { return i - 1; }
  static constexpr __host__ __device__ int dof(int) { return 1; }
  static constexpr __host__ __device__ int qdim(int) { return 1; }
};
struct TopoFree {
  static constexpr int N = 1;
  static constexpr __host__ __device__ int parent(int) { return -1; }
  static constexpr __host__ __device__ int dof(int) { return 6; }
  static constexpr __host__ __device__ int qdim(int) { return 7; }
};
struct TopoQuad {  // floating base + 4 legs of 3 one-DOF joints
  static constexpr int N = 13;
  static constexpr __host__ __device__ int parent(int i) {
    return i == 0 ? -1 : ((i - 1) % 3 == 0 ? 0 : i - 1);
  }
  static constexpr __host__ __device__ int dof(int i) { return i == 0 ? 6 : 1; }
  static constexpr __host__ __device__ int qdim(int i) { return i == 0 ? 7 : 1; }
};
template <class T> constexpr __host__ __device__ int qoff_of(int i) { int s = 0; for (int k = 0; k < i; ++k) s += T::qdim(k); return s; }
template <class T> constexpr __host__ __device__ int voff_of(int i) { int s = 0; for (int k = 0; k < i; ++k) s += T::dof(k); return s; }
template <class T> constexpr int nq_of() { return qoff_of<T>(T::N); }
template <class T> constexpr int nv_of() { return voff_of<T>(T::N); }

template <class R> __device__ __forceinline__ const R* pick(const double* d, const float* f);
template <> __device__ __forceinline__ const double* pick<double>(const double* d, const float*) { return d; }
template <> __device__ __forceinline__ const float* pick<float>(const double*, const float* f) { return f; }

template <class R, int L> __device__ __forceinline__ R lane_sum(R x) {
#pragma unroll
  for (int off = 1; off < L; off <<= 1) x += __shfl_xor(x, off, L);
  return x;
}

// Articulated-body dynamics of one instance with a compile-time topology.
// Sweep arithmetic in R; D, u, the 6x6 solve and the results in FP64.
template <class Topo, class R, int L, bool CONTACTS>
__device__ __forceinline__ bool fast_dynamics(const BucketArgs& a, int lane, int sub, const double* q,
                                              const double* v, const double* tau, double* vdot) {
  constexpr int N = Topo::N;
  const R* bp = pick<R>(a.bp, a.bpf);
  const int Bp = a.Bp;
  auto P = [&](int f, int i) -> R { return bp[(size_t)(f * N + i) * Bp + lane]; };
  m3<R> E[N];
  v3<R> r[N];
  sv<R> vel[N], cb[N], pA[N];
  bool isR[N];
  m3<double> Rw[CONTACTS ? N : 1];
  v3<double> tw[CONTACTS ? N : 1];
  bool ok = true;
  // ---- pass 1
#pragma unroll
  for (int i = 0; i < N; ++i) {
    constexpr int dummy = 0;
    (void)dummy;
    const int pr = Topo::parent(i);
    m3<R> Rt;
#pragma unroll
    for (int f = 0; f < 9; ++f) Rt.a[f] = P(BF_RT + f, i);
    v3<R> pt{P(BF_PT, i), P(BF_PT + 1, i), P(BF_PT + 2, i)};
    m3<R> Rm;
    v3<R> p;
    sv<R> vj{{R(0), R(0), R(0)}, {R(0), R(0), R(0)}};
    m3<double> Rm_d;  // FP64 joint rotation for the world pose (contacts)
    v3<double> p_d;
    if (Topo::dof(i) == 1) {
      isR[i] = a.kind[i * Bp + lane] == K_REV;
      const double qi = q[qoff_of<Topo>(i)];
      R s, c;
      sincos_q(isR[i] ? qi : 0.0, &s, &c);
      const R d = isR[i] ? R(0) : R(qi);
      // Rm = Rt Rz(theta)
      Rm = m3<R>{{c * Rt.a[0] + s * Rt.a[1], -s * Rt.a[0] + c * Rt.a[1], Rt.a[2],
                  c * Rt.a[3] + s * Rt.a[4], -s * Rt.a[3] + c * Rt.a[4], Rt.a[5],
                  c * Rt.a[6] + s * Rt.a[7], -s * Rt.a[6] + c * Rt.a[7], Rt.a[8]}};
      p = pt + d * v3<R>{Rt.a[2], Rt.a[5], Rt.a[8]};
      const R qd = R(v[voff_of<Topo>(i)]);
      vj.w.z = isR[i] ? qd : R(0);
      vj.v.z = isR[i] ? R(0) : qd;
      if (CONTACTS) { Rm_d = cvt<double>(Rm); p_d = cvt<double>(p); }
    } else {
      const int o = qoff_of<Topo>(i), w = voff_of<Topo>(i);
      m3<double> Rq = quat_mat(q[o], q[o + 1], q[o + 2], q[o + 3]);
      v3<double> tq{q[o + 4], q[o + 5], q[o + 6]};
      if (CONTACTS) {
        m3<double> Rtd;
#pragma unroll
        for (int f = 0; f < 9; ++f) Rtd.a[f] = a.bp[(size_t)((BF_RT + f) * N + i) * Bp + lane];
        v3<double> ptd{a.bp[(size_t)((BF_PT) * N + i) * Bp + lane], a.bp[(size_t)((BF_PT + 1) * N + i) * Bp + lane],
                       a.bp[(size_t)((BF_PT + 2) * N + i) * Bp + lane]};
        Rm_d = mul(Rtd, Rq);
        p_d = ptd + mul(Rtd, tq);
        Rm = cvt<R>(Rm_d);
        p = cvt<R>(p_d);
      } else {
        Rm = mul(Rt, cvt<R>(Rq));
        p = pt + mul(Rt, cvt<R>(tq));
      }
      vj.w = v3<R>{R(v[w]), R(v[w + 1]), R(v[w + 2])};
      vj.v = v3<R>{R(v[w + 3]), R(v[w + 4]), R(v[w + 5])};
    }
    E[i] = m3<R>{{Rm.a[0], Rm.a[3], Rm.a[6], Rm.a[1], Rm.a[4], Rm.a[7], Rm.a[2], Rm.a[5], Rm.a[8]}};
    r[i] = p;
    if (pr < 0) vel[i] = vj;
    else vel[i] = xmotion(E[i], r[i], vel[pr < 0 ? 0 : pr]) + vj;
    cb[i] = crm(vel[i], vj);
    s3<R> J{P(BF_J, i), P(BF_J + 1, i), P(BF_J + 2, i), P(BF_J + 3, i), P(BF_J + 4, i), P(BF_J + 5, i)};
    ai6<R> I = rigid(J, v3<R>{P(BF_C, i), P(BF_C + 1, i), P(BF_C + 2, i)}, P(BF_M, i));
    pA[i] = crf(vel[i], mul(I, vel[i]));
    if (CONTACTS) {
      if (pr < 0) { Rw[i] = Rm_d; tw[i] = p_d; }
      else {
        Rw[i] = mul(Rw[pr < 0 ? 0 : pr], Rm_d);
        tw[i] = tw[pr < 0 ? 0 : pr] + mul(Rw[pr < 0 ? 0 : pr], p_d);
      }
      const m3<R> RwR = cvt<R>(Rw[i]);
      sv<R> fe{{R(0), R(0), R(0)}, {R(0), R(0), R(0)}};
      const int c0 = a.cstart[i], c1 = a.cstart[i + 1];
      for (int cc = c0 + sub; cc < c1; cc += L) {
        auto cpar = [&](int fld) { return a.cp[((size_t)fld * a.nc + cc) * Bp + lane]; };
        v3<double> pb{cpar(CF_P), cpar(CF_P + 1), cpar(CF_P + 2)};
        const R kc = R(cpar(CF_K));
        v3<double> pw = mul(Rw[i], pb) + tw[i];
        v3<R> pbR = cvt<R>(pb);
        v3<R> vw = mul(RwR, vel[i].v + cross(vel[i].w, pbR));
        for (int h = 0; h < a.nh; ++h) {
          auto hpar = [&](int fld) { return a.hp[((size_t)fld * a.nh + h) * Bp + lane]; };
          v3<double> n{hpar(HF_NORMAL), hpar(HF_NORMAL + 1), hpar(HF_NORMAL + 2)};
          double dist = dot(pw - v3<double>{hpar(HF_POINT), hpar(HF_POINT + 1), hpar(HF_POINT + 2)}, n);
          if (dist <= kInsideMargin) {
            v3<R> F = contact_force(R(-dist), cvt<R>(n), vw, kc, R(kHalfspaceK), R(hpar(HF_ALPHA)), R(hpar(HF_MU)));
            v3<R> Fb = mulT(RwR, F);
            fe.w = fe.w + cross(pbR, Fb);
            fe.v = fe.v + Fb;
          }
        }
      }
      if (L > 1) {
        fe.w.x = lane_sum<R, L>(fe.w.x); fe.w.y = lane_sum<R, L>(fe.w.y); fe.w.z = lane_sum<R, L>(fe.w.z);
        fe.v.x = lane_sum<R, L>(fe.v.x); fe.v.y = lane_sum<R, L>(fe.v.y); fe.v.z = lane_sum<R, L>(fe.v.z);
      }
      pA[i] = pA[i] - fe;
    }
  }
  // ---- pass 2
  ai6<R> acc[N];
  bool has_acc[N];
#pragma unroll
  for (int i = 0; i < N; ++i) has_acc[i] = false;
  sv<R> U[N];
  double Dd[N], uu[N];
  double Lf[21];
  double D6[6][6];
  double u6[6];
#pragma unroll
  for (int i = N - 1; i >= 0; --i) {
    const int pr = Topo::parent(i);
    s3<R> J{P(BF_J, i), P(BF_J + 1, i), P(BF_J + 2, i), P(BF_J + 3, i), P(BF_J + 4, i), P(BF_J + 5, i)};
    ai6<R> IA = rigid(J, v3<R>{P(BF_C, i), P(BF_C + 1, i), P(BF_C + 2, i)}, P(BF_M, i));
    if (has_acc[i]) add_to(IA, acc[i]);
    if (Topo::dof(i) == 1) {
      U[i] = isR[i] ? col(IA, 2) : col(IA, 5);
      const R D = isR[i] ? IA.A.zz : IA.C.zz;
      Dd[i] = double(D);
      double tq = tau[voff_of<Topo>(i)];
      const R sk = P(BF_SK, i);
      if (!isR[i] && sk != R(0)) tq += -double(sk) * (q[qoff_of<Topo>(i)] - double(P(BF_SL, i)));
      uu[i] = tq - double(isR[i] ? pA[i].w.z : pA[i].v.z);
      ok = ok && (Dd[i] > 0.0);
      if (pr >= 0) {
        const R inv = R(1.0 / Dd[i]);
        sv<R> uD{inv * U[i].w, inv * U[i].v};
        ai6<R> Ia = rank1_down(IA, U[i], uD);
        const R uR = R(uu[i]);
        sv<R> pa = pA[i] + mul(Ia, cb[i]) + sv<R>{uR * uD.w, uR * uD.v};
        ai6<R> add = xtix(E[i], r[i], Ia);
        constexpr int dummy2 = 0;
        (void)dummy2;
        const int pp = pr < 0 ? 0 : pr;
        if (has_acc[pp]) add_to(acc[pp], add);
        else { acc[pp] = add; has_acc[pp] = true; }
        pA[pp] = pA[pp] + xtforce(E[i], r[i], pa);
      }
    } else {
      // floating joint (root only in the compiled topologies)
      to_dense(IA, D6);
      const int w = voff_of<Topo>(i);
      u6[0] = tau[w] - double(pA[i].w.x); u6[1] = tau[w + 1] - double(pA[i].w.y); u6[2] = tau[w + 2] - double(pA[i].w.z);
      u6[3] = tau[w + 3] - double(pA[i].v.x); u6[4] = tau[w + 4] - double(pA[i].v.y); u6[5] = tau[w + 5] - double(pA[i].v.z);
      ok = chol6(D6, Lf) && ok;
    }
  }
  // ---- pass 3
  sv<R> accel[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int pr = Topo::parent(i);
    const sv<R> a0{{R(0), R(0), R(0)}, {R(0), R(0), R(a.g)}};
    sv<R> ap = xmotion(E[i], r[i], pr < 0 ? a0 : accel[pr < 0 ? 0 : pr]) + cb[i];
    if (Topo::dof(i) == 1) {
      const double qdd = (uu[i] - double(dot6(U[i], ap))) / Dd[i];
      vdot[voff_of<Topo>(i)] = qdd;
      if (isR[i]) ap.w.z += R(qdd); else ap.v.z += R(qdd);
    } else {
      const int w = voff_of<Topo>(i);
      double apd[6] = {double(ap.w.x), double(ap.w.y), double(ap.w.z), double(ap.v.x), double(ap.v.y), double(ap.v.z)};
      double x[6];
#pragma unroll
      for (int rr = 0; rr < 6; ++rr) {
        double s = u6[rr];
#pragma unroll
        for (int cc = 0; cc < 6; ++cc) s -= D6[rr][cc] * apd[cc];
        x[rr] = s;
      }
      chol6_solve(Lf, x);
#pragma unroll
      for (int t = 0; t < 6; ++t) vdot[w + t] = x[t];
      ap = ap + sv<R>{{R(x[0]), R(x[1]), R(x[2])}, {R(x[3]), R(x[4]), R(x[5])}};
    }
    accel[i] = ap;
  }
  return ok;
}

// n_steps semi-implicit Euler steps per launch (integrators.rs:276-319) with
// the state held in registers; q, v accumulate in FP64.
template <class Topo, class R, int L, bool CONTACTS>
__global__ void __launch_bounds__(256) fast_sie_kernel(BucketArgs a, double dt, int n_steps) {
  constexpr int NQ = nq_of<Topo>(), NV = nv_of<Topo>();
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int inst = tid / L;
  const int sub = tid % L;
  if (inst >= a.B) return;
  const int Bp = a.Bp;
  int my_steps = n_steps;
  if (a.active) my_steps = min(my_steps, a.active[inst]);
  double q[NQ], v[NV], tau[NV], vd[NV];
#pragma unroll
  for (int k = 0; k < NQ; ++k) q[k] = a.q[(size_t)k * Bp + inst];
#pragma unroll
  for (int k = 0; k < NV; ++k) { v[k] = a.v[(size_t)k * Bp + inst]; tau[k] = a.tau[(size_t)k * Bp + inst]; vd[k] = 0.0; }
  bool ok = true;
  for (int s = 0; s < my_steps; ++s) {
    ok = fast_dynamics<Topo, R, L, CONTACTS>(a, inst, sub, q, v, tau, vd) && ok;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = v[k] + vd[k] * dt;
#pragma unroll
    for (int i = 0; i < Topo::N; ++i) {
      if (Topo::dof(i) == 1) {
        q[qoff_of<Topo>(i)] += v[voff_of<Topo>(i)] * dt;
      } else {
        advance_floating(&q[qoff_of<Topo>(i)], &v[voff_of<Topo>(i)], dt, &q[qoff_of<Topo>(i)]);
      }
    }
  }
  if (sub != 0) return;
#pragma unroll
  for (int k = 0; k < NQ; ++k) { a.q[(size_t)k * Bp + inst] = q[k]; ok = ok && isfinite(q[k]); }
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    a.v[(size_t)k * Bp + inst] = v[k];
    a.vdot[(size_t)k * Bp + inst] = vd[k];
    ok = ok && isfinite(v[k]) && isfinite(vd[k]);
  }
  if (!ok) a.fault[inst] = 1;
}

enum { TOPO_FREE = 100, TOPO_QUAD = 101 };

int fast_topology_id(int nb, const int* parent, const int* dofcls, int nc) {
  bool chain = nb >= 1 && nb <= 12 && nc == 0;
  for (int i = 0; i < nb && chain; ++i) chain = parent[i] == i - 1 && dofcls[i] == 1;
  if (chain) return nb;
  if (nb == 1 && parent[0] == -1 && dofcls[0] == 6) return TOPO_FREE;
  bool quad = nb == 13 && nc == 0;
  for (int i = 0; i < nb && quad; ++i) quad = parent[i] == TopoQuad::parent(i) && dofcls[i] == TopoQuad::dof(i);
  if (quad) return TOPO_QUAD;
  return -1;
}

template <class Topo, class R, int L, bool C>
static int launch_t(const BucketArgs& a, double dt, int n_steps, int block, hipStream_t s) {
  long threads = (long)a.B * L;
  dim3 blk(block), grid((unsigned)((threads + block - 1) / block));
  hipLaunchKernelGGL((fast_sie_kernel<Topo, R, L, C>), grid, blk, 0, s, a, dt, n_steps);
  return (int)hipGetLastError();
}

template <class R>
static int launch_prec(int topo, const BucketArgs& a, double dt, int n, int block, hipStream_t s) {
  switch (topo) {
    case 1: return launch_t<TopoChain<1>, R, 1, false>(a, dt, n, block, s);
    case 2: return launch_t<TopoChain<2>, R, 1, false>(a, dt, n, block, s);
    case 3: return launch_t<TopoChain<3>, R, 1, false>(a, dt, n, block, s);
    case 4: return launch_t<TopoChain<4>, R, 1, false>(a, dt, n, block, s);
    case 5: return launch_t<TopoChain<5>, R, 1, false>(a, dt, n, block, s);
    case 6: return launch_t<TopoChain<6>, R, 1, false>(a, dt, n, block, s);
    case 7: return launch_t<TopoChain<7>, R, 1, false>(a, dt, n, block, s);
    case 8: return launch_t<TopoChain<8>, R, 1, false>(a, dt, n, block, s);
    case 9: return launch_t<TopoChain<9>, R, 1, false>(a, dt, n, block, s);
    case 10: return launch_t<TopoChain<10>, R, 1, false>(a, dt, n, block, s);
    case 11: return launch_t<TopoChain<11>, R, 1, false>(a, dt, n, block, s);
    case 12: return launch_t<TopoChain<12>, R, 1, false>(a, dt, n, block, s);
    case TOPO_QUAD: return launch_t<TopoQuad, R, 1, false>(a, dt, n, block, s);
    case TOPO_FREE:
      switch (a.lanes) {
        case 2: return launch_t<TopoFree, R, 2, true>(a, dt, n, block, s);
        case 4: return launch_t<TopoFree, R, 4, true>(a, dt, n, block, s);
        case 8: return launch_t<TopoFree, R, 8, true>(a, dt, n, block, s);
        default: return launch_t<TopoFree, R, 1, true>(a, dt, n, block, s);
      }
    default: return (int)hipErrorInvalidValue;
  }
}

int launch_fast_step(int topo_id, int precision, const BucketArgs& a, double dt, int n_steps, int block,
                     void* stream) {
  if (a.B == 0 || n_steps <= 0) return 0;
  if (block <= 0) block = 256;
  hipStream_t s = (hipStream_t)stream;
  if (precision == GP_PREC_MIXED) return launch_prec<float>(topo_id, a, dt, n_steps, block, s);
  return launch_prec<double>(topo_id, a, dt, n_steps, block, s);
}

}  // namespace gp
