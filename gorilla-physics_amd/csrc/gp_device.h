// gp_device.h — device helpers shared by the plain and fast kernels:
// contact force law, joint-angle sin/cos, the 6x6 Cholesky of a floating
// joint's D and the floating-joint position update.
#pragma once

#include <hip/hip_runtime.h>

#include "gp_internal.h"
#include "gp_spatial.h"

namespace gp {

// Cross-lane moves within aligned groups of up to 8 lanes by DPP (VALU moves;
// a shuffle, ds_bpermute, would be a round trip through the LDS unit): an
// 8-lane group is two 4-lane banks of one 16-lane DPP row, so a quad
// permutation moves within each bank and a row shift by 4, written to one
// bank of the pair only (bank_mask), moves between the pair. `old` is what
// the lanes the bank mask leaves out keep.
constexpr int kDppRowShl4 = 0x104, kDppRowShr4 = 0x114;  // lane i takes lane i + 4 / i - 4 of its row
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E;           // quad permutations [1,0,3,2], [2,3,0,1]
template <int CTRL, int BANKS> __device__ __forceinline__ int dpp32(int old, int x) {
  return __builtin_amdgcn_update_dpp(old, x, CTRL, 0xF, BANKS, false);
}
template <int CTRL, int BANKS, class T> __device__ __forceinline__ T dpp_move(T old, T x) {
  if constexpr (sizeof(T) == 8) {
    const long long o = __builtin_bit_cast(long long, old), v = __builtin_bit_cast(long long, x);
    const int lo = dpp32<CTRL, BANKS>((int)o, (int)v);
    const int hi = dpp32<CTRL, BANKS>((int)(o >> 32), (int)(v >> 32));
    return __builtin_bit_cast(T, ((long long)hi << 32) | (long long)(unsigned)lo);
  } else {
    return __builtin_bit_cast(T, dpp32<CTRL, BANKS>(__builtin_bit_cast(int, old), __builtin_bit_cast(int, x)));
  }
}
// x of the lane O apart (O = 1, 2, 4) in its aligned group
template <int O, class T> __device__ __forceinline__ T lane_xor(T x) {
  static_assert(O == 1 || O == 2 || O == 4, "DPP lane xor within 8 lanes");
  if constexpr (O == 1) return dpp_move<kDppXor1, 0xF>(x, x);
  else if constexpr (O == 2) return dpp_move<kDppXor2, 0xF>(x, x);
  else return dpp_move<kDppRowShl4, 0x5>(dpp_move<kDppRowShr4, 0xA>(x, x), x);
}

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

// sin/cos of a joint angle held in FP64. Quadrant reduction th = k pi/2 + r,
// |r| <= pi/4, with pi/2 split into three parts (the first two of 33
// significant bits, so k C1 and k C2 are exact for |k| < 2^20; Cody & Waite
// 1980), then the Taylor polynomials of sin and cos in r^2 (truncated after
// r^17 / r^18: below 6e-20 relative on |r| <= pi/4). About 30 operations.
// Angles beyond 2^19 quarter turns first lose the nearest multiple n of 2 pi,
// held as a head and a tail (2 pi - head - tail = -6e-33): n head is split
// into its rounded value and its exact error, th - round(n head) is exact
// (Sterbenz), so the remainder is off by a few 1e-16 for any |th| < 1e15; an
// infinite or NaN angle gives NaN. The library's sincos would hold its large-
// argument reduction's registers live in every caller.
__device__ __forceinline__ void sincos_q(double th, double* s, double* c) {
  double k = rint(th * 0.6366197723675814);  // th * 2 / pi
  if (!(fabs(k) < 524288.0)) {
    const double n = rint(th * 0.15915494309189535);  // th / (2 pi)
    const double p = n * 6.283185307179586;
    const double e = fma(n, 6.283185307179586, -p);
    th = fma(-n, 2.4492935982947064e-16, (th - p) - e);
    k = rint(th * 0.6366197723675814);
  }
  const double r = fma(-k, 2.0222662487959506e-21, fma(-k, 6.077100506303966e-11, fma(-k, 1.5707963267341256, th)));
  const double x2 = r * r;
  double sn = fma(x2, 2.8114572543455206e-15, -7.647163731819816e-13);
  sn = fma(x2, sn, 1.6059043836821613e-10);
  sn = fma(x2, sn, -2.505210838544172e-08);
  sn = fma(x2, sn, 2.7557319223985893e-06);
  sn = fma(x2, sn, -0.0001984126984126984);
  sn = fma(x2, sn, 0.008333333333333333);
  sn = fma(x2, sn, -0.16666666666666666);
  sn = fma(x2 * r, sn, r);
  double cs = fma(x2, -1.5619206968586225e-16, 4.779477332387385e-14);
  cs = fma(x2, cs, -1.1470745597729725e-11);
  cs = fma(x2, cs, 2.08767569878681e-09);
  cs = fma(x2, cs, -2.755731922398589e-07);
  cs = fma(x2, cs, 2.48015873015873e-05);
  cs = fma(x2, cs, -0.001388888888888889);
  cs = fma(x2, cs, 0.041666666666666664);
  cs = fma(x2, cs, -0.5);
  cs = fma(x2, cs, 1.0);
  const int q = int(k) & 3;  // quadrant: (s, c) -> (c, -s) per quarter turn
  const double s1 = (q & 1) ? cs : sn;
  const double c1 = (q & 1) ? sn : cs;
  *s = (q & 2) ? -s1 : s1;
  *c = ((q + 1) & 2) ? -c1 : c1;
}
// FP32 path: quadrant reduction th = k pi/2 + r, |r| <= pi/4, in FP64 (two
// FMAs with pi/2 split into a head and a tail), then the Taylor polynomials of
// sin and cos in FP32. Truncation after x^9 / x^10 is
// below 2e-9 on |r| <= pi/4, under FP32 rounding.
__device__ __forceinline__ void sincos_q(double th, float* s, float* c) {
  const double pio2_hi = 1.5707963267948966192313216916398;  // fl(pi / 2)
  const double pio2_lo = 6.123233995736766035868820147292e-17;  // pi / 2 - fl(pi / 2)
  const double k = rint(th * 0.63661977236758134307553505349006);  // th * 2 / pi
  const float r = float(fma(-k, pio2_lo, fma(-k, pio2_hi, th)));
  const float x2 = r * r;
  float sn = fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, 1.0f / 362880.0f, -1.0f / 5040.0f), 1.0f / 120.0f), -1.0f / 6.0f), 1.0f);
  sn *= r;
  float cs = fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, -1.0f / 3628800.0f, 1.0f / 40320.0f), -1.0f / 720.0f),
                                    1.0f / 24.0f), -0.5f), 1.0f);
  const int q = int(k) & 3;  // quadrant: (s, c) -> (c, -s) per quarter turn
  const float s1 = (q & 1) ? cs : sn;
  const float c1 = (q & 1) ? sn : cs;
  *s = (q & 2) ? -s1 : s1;
  *c = ((q + 1) & 2) ? -c1 : c1;
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

}  // namespace gp
