// gp_fast_impl.h — templates of the register-resident fast kernels, shared by
// the per-precision translation units (gp_fast_f32/f64.hip, gp_group_f32/f64.hip).
#pragma once

// One instance per lane (or per L lanes for contact-heavy free bodies), the
// whole per-instance state in registers, n_steps semi-implicit Euler steps
// (integrators.rs:276-319) per launch. The dynamics is Featherstone's
// articulated-body algorithm (RBDA Table 7.1) in body frames whose z axis is
// each 1-DOF joint's axis; it returns the joint accelerations the reference's
// dynamics_continuous computes with CRBA + RNEA + LU (dynamics.rs:322-364).
//
// Arithmetic: R = double (FP64 everywhere) or R = float (FP32 kinematics,
// bias forces, contact forces and both sweeps; the joint-space inertia D,
// the bias u, the 6x6 floating-base solve and the q/v accumulation in FP64).
//
// Register budget: per 1-DOF body only (cos, sin, displacement), the bias
// acceleration's four non-zero components, p^A, U, D and u are kept across
// the sweeps; the Plucker transform E = (R_t R_z(q))^T and r are rebuilt from
// the tree parameters when a sweep needs them. Torques are read where used and
// accelerations are folded into v and q in the outward pass, so only q and v
// persist between steps. Every per-bucket array is read and written through a
// raw buffer resource (one shared 32-bit lane offset instead of a 64-bit
// address per access), and per-lane revolute/prismatic choices select values,
// never addresses: with both, no FP32 chain of up to 12 joints spills.
#include <hip/hip_runtime.h>

#include <utility>

#include "../../include/gorilla_amd.h"
#include "gp_device.h"
#include "gp_internal.h"
#include "gp_spatial.h"

namespace gp {

template <int N_>
struct TopoChain {
  static constexpr int N = N_;
  static constexpr int first = 0;  // first body the sweeps visit
  static constexpr __host__ __device__ int parent(int i) { return i - 1; }
  static constexpr __host__ __device__ int dof(int) { return 1; }
  static constexpr __host__ __device__ int qdim(int) { return 1; }
};
struct TopoFree {
  static constexpr int N = 1;
  static constexpr int first = 0;
  static constexpr __host__ __device__ int parent(int) { return -1; }
  static constexpr __host__ __device__ int dof(int) { return 6; }
  static constexpr __host__ __device__ int qdim(int) { return 7; }
};
struct TopoQuad {  // floating base + 4 legs of 3 one-DOF joints
  static constexpr int N = 13;
  static constexpr int first = 0;
  static constexpr __host__ __device__ int parent(int i) { return i == 0 ? -1 : ((i - 1) % 3 == 0 ? 0 : i - 1); }
  static constexpr __host__ __device__ int dof(int i) { return i == 0 ? 6 : 1; }
  static constexpr __host__ __device__ int qdim(int i) { return i == 0 ? 7 : 1; }
};
// A body welded to the world (a fixed joint, no coordinates) and a chain of
// N - 1 one-DOF joints behind it: a fixed-base manipulator (the SO-101 of
// builders/mod.rs:252). Body 0 never moves, so the sweeps start at body 1 as
// the root of a chain whose base acceleration is gravity seen in body 0's
// frame, formed once per launch.
template <int N_>
struct TopoWelded {
  static constexpr int N = N_;
  static constexpr int first = 1;
  static constexpr __host__ __device__ int parent(int i) { return i - 1; }
  static constexpr __host__ __device__ int dof(int i) { return i == 0 ? 0 : 1; }
  static constexpr __host__ __device__ int qdim(int i) { return i == 0 ? 0 : 1; }
};
// the parent a sweep sees: the first swept body is a root
template <class T> constexpr __host__ __device__ int sweep_parent(int i) { return i == T::first ? -1 : T::parent(i); }
template <class T> constexpr __host__ __device__ int qoff_of(int i) {
  int s = 0;
  for (int k = 0; k < i; ++k) s += T::qdim(k);
  return s;
}
template <class T> constexpr __host__ __device__ int voff_of(int i) {
  int s = 0;
  for (int k = 0; k < i; ++k) s += T::dof(k);
  return s;
}
template <class T> constexpr int nq_of() { return qoff_of<T>(T::N); }
template <class T> constexpr int nv_of() { return voff_of<T>(T::N); }

template <class R> __device__ __forceinline__ const R* pick(const double* d, const float* f);
template <> __device__ __forceinline__ const double* pick<double>(const double* d, const float*) { return d; }
template <> __device__ __forceinline__ const float* pick<float>(const double*, const float* f) { return f; }

// One structure-of-arrays block ([row][Bp lanes]) as a raw buffer resource:
// a load or store names its row with a uniform byte offset (row * Bp * size,
// scalar registers) and its lane with one 32-bit VGPR offset shared by every
// access, instead of a 64-bit per-lane address per access. The hardware range
// check drops accesses beyond the block (loads return 0). The host keeps every
// block below 2^31 bytes.
constexpr int kRawBufferWord3 = 0x00020000;  // CDNA raw buffer: 32-bit data format
template <class T>
struct Soa {
  __amdgpu_buffer_rsrc_t r;
  int pitch;  // bytes per row
  __device__ __forceinline__ Soa(const T* base, int rows, int Bp)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(base), (short)0, rows * Bp * (int)sizeof(T),
                                            kRawBufferWord3)),
        pitch(Bp * (int)sizeof(T)) {}
  __device__ __forceinline__ T ld(int row, int lane) const {
    if constexpr (sizeof(T) == 8)
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, lane * 8, row * pitch, 0));
    else
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, lane * 4, row * pitch, 0));
  }
  // row + xrow with xrow varying across lanes (a per-lane part must go into the
  // VGPR offset: a divergent scalar offset would need a loop over its values)
  __device__ __forceinline__ T ldx(int row, int lane, int xrow) const {
    const int voff = lane * (int)sizeof(T) + xrow * pitch;
    if constexpr (sizeof(T) == 8)
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, row * pitch, 0));
    else
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, row * pitch, 0));
  }
  __device__ __forceinline__ void stx(int row, int lane, int xrow, T x) const {
    const int voff = lane * (int)sizeof(T) + xrow * pitch;
    if constexpr (sizeof(T) == 8) {
      using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, x), r, voff, row * pitch, 0);
    } else {
      using V = decltype(__builtin_amdgcn_raw_buffer_load_b32(r, 0, 0, 0));
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(V, x), r, voff, row * pitch, 0);
    }
  }
  __device__ __forceinline__ void st(int row, int lane, T x) const {
    if constexpr (sizeof(T) == 8) {
      using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, x), r, lane * 8, row * pitch, 0);
    } else {
      using V = decltype(__builtin_amdgcn_raw_buffer_load_b32(r, 0, 0, 0));
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(V, x), r, lane * 4, row * pitch, 0);
    }
  }
};

// A per-lane choice between two values. Taking them by value keeps the
// compiler from selecting between two addresses inside a register-resident
// aggregate, which would force that aggregate into scratch memory.
template <class T> __device__ __forceinline__ T sel(bool c, T x, T y) { return c ? x : y; }

template <class R, int L> __device__ __forceinline__ R lane_sum(R x) {
#pragma unroll
  for (int off = 1; off < L; off <<= 1) x += __shfl_xor(x, off, L);
  return x;
}

// E = (R_t R_z)^T and r = p_t + d R_t e_z for an aligned 1-DOF joint
template <class R>
__device__ __forceinline__ void joint_frame(const m3<R>& Rt, v3<R> pt, R c, R s, R d, m3<R>& E, v3<R>& r) {
  E = m3<R>{{c * Rt.a[0] + s * Rt.a[1], c * Rt.a[3] + s * Rt.a[4], c * Rt.a[6] + s * Rt.a[7],
             -s * Rt.a[0] + c * Rt.a[1], -s * Rt.a[3] + c * Rt.a[4], -s * Rt.a[6] + c * Rt.a[7],
             Rt.a[2], Rt.a[5], Rt.a[8]}};
  r = pt + d * v3<R>{Rt.a[2], Rt.a[5], Rt.a[8]};
}

template <class R>
struct Body1 {  // what a 1-DOF body keeps between the sweeps
  R c, s, d;       // joint angle cos / sin (1, 0 for prismatic), displacement (0 for revolute)
  R cw0, cw1;      // bias acceleration v x_m (S qd): angular x, y (z is 0)
  R cv0, cv1;      // linear x, y (z is 0)
  sv<R> pA;        // articulated bias force
  sv<R> U;         // I^A S
  double Dinv, u;  // 1 / (S^T I^A S) and tau - S^T p^A
  int rev;  // 1 revolute, 0 prismatic
};

// Lane-private LDS tier of the tree parameters that do not stay in
// registers. Per body (from the first uncached one on): its tree rotation as a
// unit quaternion and its translation, 7 values, the rotation rebuilt where a
// sweep uses it (24 operations in place of 9 loads); then, while room is left,
// the rigid inertias (10 values) of the first of those bodies. One row of
// `pitch` values per lane (lane-major, so every access is the lane's base
// register plus an immediate offset; an odd pitch spreads a wave's 64-bit
// accesses over all banks), written once per launch; a lane reads and writes
// only its own row, so no barrier is needed. Budget: kLdsLaneBytes per lane,
// the CU's 160 KB over its 256 lanes at one wave per SIMD (the occupancy of
// every kernel that has uncached bodies). Without it the FP64 7-joint chain
// re-reads every body's 22 values in every sweep from beyond L2 (~110 MB per
// step at 65,536 arms).
constexpr int kLdsLaneBytes = 640;
template <class Topo, class R, int CN> struct LdsTier {
  static constexpr int lo = CN > Topo::first ? CN : Topo::first;  // first body in the tier
  static constexpr int U = Topo::N - lo;
  static constexpr int cap = kLdsLaneBytes / (int)sizeof(R) - 1;    // room for an odd pitch
  static constexpr bool on = Topo::dof(Topo::first) == 1 && U > 0;  // chains and welded chains
  static constexpr int NT = on ? (U < cap / 7 ? U : cap / 7) : 0;   // bodies [lo, lo + NT): transforms
  static constexpr int NI0 = NT == U ? (cap - 7 * NT) / 10 : 0;
  static constexpr int NI = NI0 < U ? NI0 : U;                     // bodies [lo, lo + NI): inertias
  static constexpr int slots = 7 * NT + 10 * NI;
  static constexpr int pitch = slots > 0 ? (slots | 1) : 0;
  // the inertias of up to 4 more bodies [lo + NI, lo + NI + NR) in registers
  // (a 7-joint FP64 chain then reads nothing from memory inside a step)
  static constexpr int NR = (NT == U && U - NI <= 4) ? U - NI : 0;
};

// Unit quaternion (w, x, y, z) of a rotation matrix (row-major), by
// Shepperd's method: the largest of 1 + trace and 1 + 2 R_kk - trace under
// the square root. The case is picked by selects, not branches (the same
// arithmetic in each case): a launch's prologue converts every cached body's
// rotation, and branches there keep the compiler from issuing the next body's
// parameter loads before this one's conversion is done.
template <class T>
__device__ __forceinline__ void rot_to_quat(const m3<T>& M, T* o) {
  const T m00 = M.a[0], m01 = M.a[1], m02 = M.a[2], m10 = M.a[3], m11 = M.a[4], m12 = M.a[5], m20 = M.a[6],
          m21 = M.a[7], m22 = M.a[8];
  const T tr = m00 + m11 + m22;
  const bool c0 = tr >= m00 && tr >= m11 && tr >= m22;     // w largest
  const bool c1 = !c0 && m00 >= m11 && m00 >= m22;         // x largest
  const bool c2 = !c0 && !c1 && m11 >= m22;                // y largest (else z)
  const T s0 = T(1) + tr, s1 = T(1) + m00 - m11 - m22, s2 = T(1) + m11 - m00 - m22, s3 = T(1) + m22 - m00 - m11;
  const T r = sqrt(c0 ? s0 : c1 ? s1 : c2 ? s2 : s3), h = T(0.5) / r;
  const T big = T(0.5) * r;
  const T a = (m21 - m12) * h, b = (m02 - m20) * h, c = (m10 - m01) * h;
  const T d = (m01 + m10) * h, e = (m02 + m20) * h, f = (m12 + m21) * h;
  o[0] = c0 ? big : c1 ? a : c2 ? b : c;
  o[1] = c0 ? a : c1 ? big : c2 ? d : e;
  o[2] = c0 ? b : c1 ? d : c2 ? big : f;
  o[3] = c0 ? c : c1 ? e : c2 ? f : big;
}

template <class T>
__device__ __forceinline__ m3<T> quat_to_rot(T w, T x, T y, T z) {
  const T x2 = x + x, y2 = y + y, z2 = z + z;
  const T xx = x * x2, yy = y * y2, zz = z * z2, xy = x * y2, xz = x * z2, yz = y * z2;
  const T wx = w * x2, wy = w * y2, wz = w * z2;
  return m3<T>{{T(1) - (yy + zz), xy - wz, xz + wy, xy + wz, T(1) - (xx + zz), yz - wx, xz - wy, yz + wx,
                T(1) - (xx + yy)}};
}

// Per-body tree parameters of one instance ([NBF][N][Bp] in the bucket).
// The first CN bodies' are loaded into registers once per launch (loads whose
// values are never used, e.g. spring constants of an all-revolute chain, are
// dropped); with LDS, the next bodies' transforms and inertias are kept in the
// lane's LDS column (LdsTier); the others are read from the bucket's array
// where used (an L1/L2 hit at best). Caching pays while the ~22 values per
// body fit beside the sweep state (see cached_bodies).
template <class Topo, class R, int CN, bool LDS> struct TierOf {
  static constexpr int NT = 0, NI = 0, NR = 0, lo = 0, pitch = 0;
};
template <class Topo, class R, int CN> struct TierOf<Topo, R, CN, true> {
  using T = LdsTier<Topo, R, CN>;
  static constexpr int NT = T::NT, NI = T::NI, NR = T::NR, lo = T::lo, pitch = T::pitch;
};

template <class Topo, class R, int CN = 0, bool LDS = false>
struct FastParams {
  using Tier = TierOf<Topo, R, CN, LDS>;
  static constexpr int NT = Tier::NT, NI = Tier::NI, NR = Tier::NR, lo = Tier::lo, pitch = Tier::pitch;
  Soa<R> bp;
  int lane;
  int xbody;  // per-lane shift of the body index (a lane that runs one leg of a tree)
  R* lds;     // this lane's LDS row (LdsTier): slot k at lds[k]
  R c[CN > 0 ? NBF : 1][CN > 0 ? CN : 1];
  R ri[NR > 0 ? NR : 1][10];  // inertias of bodies lo + NI + j (J, m c, m)
  __device__ __forceinline__ FastParams(const R* p, int Bp, int lane_, int xbody_ = 0, R* lds_base = nullptr)
      : bp(p, NBF * Topo::N, Bp), lane(lane_), xbody(xbody_),
        lds(lds_base ? lds_base + (int)threadIdx.x * pitch : nullptr) {
    if constexpr (CN > 0) {
#pragma unroll
      for (int f = 0; f < NBF; ++f)
#pragma unroll
        for (int i = 0; i < CN; ++i) c[f][i] = bp.ldx(f * Topo::N + i, lane, xbody);
    }
    if constexpr (NT > 0) {
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        m3<R> M;
#pragma unroll
        for (int k = 0; k < 9; ++k) M.a[k] = glob(BF_RT + k, lo + j);
        R qt[4];
        rot_to_quat(M, qt);
#pragma unroll
        for (int k = 0; k < 4; ++k) lds[7 * j + k] = qt[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) lds[7 * j + 4 + k] = glob(BF_PT + k, lo + j);
      }
#pragma unroll
      for (int j = 0; j < NI; ++j)
#pragma unroll
        for (int k = 0; k < 10; ++k) lds[7 * NT + 10 * j + k] = glob(BF_J + k, lo + j);
#pragma unroll
      for (int j = 0; j < NR; ++j)
#pragma unroll
        for (int k = 0; k < 10; ++k) ri[j][k] = glob(BF_J + k, lo + NI + j);
    }
  }
  __device__ __forceinline__ R glob(int f, int i) const { return bp.ldx(f * Topo::N + i, lane, xbody); }
  // a compiler-only memory barrier: LDS values read after it are loaded again
  __device__ __forceinline__ void sweep_fence() const {
    if constexpr (NT > 0) asm volatile("" ::: "memory");
  }
  __device__ __forceinline__ R operator()(int f, int i) const {
    if (i < CN) return c[f][i < CN ? i : 0];
    return glob(f, i);
  }
  __device__ __forceinline__ m3<R> Rt(int i) const {
    if (NT > 0 && i >= lo && i < lo + NT && i >= CN) {
      const int j = i - lo;
      return quat_to_rot(lds[7 * j], lds[7 * j + 1], lds[7 * j + 2], lds[7 * j + 3]);
    }
    m3<R> M;
#pragma unroll
    for (int k = 0; k < 9; ++k) M.a[k] = (*this)(BF_RT + k, i);
    return M;
  }
  __device__ __forceinline__ v3<R> pt(int i) const {
    if (NT > 0 && i >= lo && i < lo + NT && i >= CN) {
      const int j = i - lo;
      return {lds[7 * j + 4], lds[7 * j + 5], lds[7 * j + 6]};
    }
    return {(*this)(BF_PT, i), (*this)(BF_PT + 1, i), (*this)(BF_PT + 2, i)};
  }
  __device__ __forceinline__ ai6<R> inertia(int i) const {
    if (NI > 0 && i >= lo && i < lo + NI && i >= CN) {
      const int b = 7 * NT + 10 * (i - lo);
      auto at = [&](int k) { return lds[b + k]; };
      return rigid(s3<R>{at(0), at(1), at(2), at(3), at(4), at(5)}, v3<R>{at(6), at(7), at(8)}, at(9));
    }
    if (NR > 0 && i >= lo + NI && i < lo + NI + NR) {
      const R* e = ri[NR > 0 ? i - lo - NI : 0];
      return rigid(s3<R>{e[0], e[1], e[2], e[3], e[4], e[5]}, v3<R>{e[6], e[7], e[8]}, e[9]);
    }
    s3<R> J{(*this)(BF_J, i), (*this)(BF_J + 1, i), (*this)(BF_J + 2, i), (*this)(BF_J + 3, i), (*this)(BF_J + 4, i),
            (*this)(BF_J + 5, i)};
    return rigid(J, v3<R>{(*this)(BF_C, i), (*this)(BF_C + 1, i), (*this)(BF_C + 2, i)}, (*this)(BF_M, i));
  }
};

// Dynamic LDS of a launch (the tier's columns); GP_DYN_LDS names it.
#ifndef GP_DYN_LDS
extern __shared__ __align__(16) unsigned char gp_dyn_lds[];
#define GP_DYN_LDS gp_dyn_lds
#endif

// Bodies whose parameters stay in registers for a whole launch without
// spilling: every body of a chain of up to 9 joints (FP32 sweeps) or 4 (FP64),
// one fewer with per-lane joint kinds; the first few of a longer chain; a lone
// free body's one.
#ifndef GP_PCACHE_F32
#define GP_PCACHE_F32 9
#endif
#ifndef GP_PCACHE_F64
#define GP_PCACHE_F64 4
#endif
#ifndef GP_PCACHE_LONG
#define GP_PCACHE_LONG 6
#endif
template <class Topo, class R, bool ALLREV> constexpr int cached_bodies() {
  constexpr int n_max = (sizeof(R) == 4 ? GP_PCACHE_F32 : GP_PCACHE_F64) - (ALLREV ? 0 : 1);
  if (Topo::N == 1) return 1;
  if (Topo::dof(Topo::first) != 1) return 0;
  if (Topo::N <= n_max) return Topo::N;
  return sizeof(R) == 4 ? (ALLREV ? GP_PCACHE_LONG : GP_PCACHE_LONG / 2) : 0;
}

// Bytes of dynamic LDS a launch of `block` lanes needs for the tier of a
// kernel with this topology and precision.
template <class Topo, class R, bool ALLREV> constexpr int lds_tier_bytes(int block) {
  return LdsTier<Topo, R, cached_bodies<Topo, R, ALLREV>()>::pitch * (int)sizeof(R) * block;
}

// A lone floating body: its articulated inertia is its rigid inertia, the same
// every step, so its inverse is formed once per launch.
template <class Topo> constexpr bool lone_free_body() { return Topo::N == 1 && Topo::dof(0) == 6; }

// index of (r, c) in a packed symmetric 6x6 (lower triangle, row-major)
constexpr __host__ __device__ int sym6(int r, int c) { return r >= c ? r * (r + 1) / 2 + c : c * (c + 1) / 2 + r; }

// I^-1 (packed symmetric) of a rigid-body inertia; false if I is not SPD.
template <class R>
__device__ __forceinline__ bool inverse_inertia(const ai6<R>& I, double (&Iinv)[21]) {
  double D6[6][6], Lf[21];
  to_dense(I, D6);
  const bool ok = chol6(D6, Lf);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double e[6] = {0, 0, 0, 0, 0, 0};
    e[k] = 1.0;
    chol6_solve(Lf, e);
#pragma unroll
    for (int r = k; r < 6; ++r) Iinv[sym6(r, k)] = e[r];
  }
  return ok;
}

// Per-launch constants in registers. TAU: the held torques (topologies whose
// parameters are all cached). A lone free body: its inverse inertia, its
// joint's tree transform (FP64), and (with contacts) this lane's contact points
// (p, k) and up to MAXH half-spaces when the bucket's counts fit (`cached`;
// else they are read from memory every step).
template <class Topo, int L, bool CONTACTS, bool TAU>
struct LaunchConst {
  static constexpr bool FREE = lone_free_body<Topo>();
  static constexpr bool WELDED = Topo::first > 0;
  double abase[WELDED ? 3 : 1];  // welded base: linear acceleration of body 0 (gravity in its frame)
  static constexpr bool TAU_REG = TAU || FREE;
  static constexpr int MAXC = (FREE && CONTACTS) ? (8 + L - 1) / L : 1;  // points per lane
  static constexpr int MAXH = (FREE && CONTACTS) ? 2 : 1;
  double Iinv[FREE ? 21 : 1];
  double rt[FREE ? 12 : 1];   // R_tree (9), p_tree (3)
  double tau[TAU_REG ? nv_of<Topo>() : 1];
  double cp[MAXC][NCF];
  double hp[MAXH][NHF];
  int ncl = 0, nh = 0;
  bool cached = false;
};

template <class Topo, class R, int L, bool CONTACTS, class FP, bool TAU>
__device__ __forceinline__ bool load_launch_const(const BucketArgs& a, const FP& P, int lane,
                                                  int sub, LaunchConst<Topo, L, CONTACTS, TAU>& lc) {
  using LC = LaunchConst<Topo, L, CONTACTS, TAU>;
  bool ok = true;
  if constexpr (LC::WELDED) {
    // body 0 (row f * N + 0 of field f): E = R_t^T, so E (0, 0, g) = g times
    // the third row of R_t
    const Soa<double> bpS(a.bp, NBF * Topo::N, a.Bp);
#pragma unroll
    for (int k = 0; k < 3; ++k) lc.abase[k] = a.g * bpS.ld((BF_RT + 6 + k) * Topo::N, lane);
  }
  if constexpr (LC::TAU_REG) {
    const Soa<double> tauS(a.tau, nv_of<Topo>(), a.Bp);
#pragma unroll
    for (int k = 0; k < nv_of<Topo>(); ++k) lc.tau[k] = tauS.ld(k, lane);
  }
  if constexpr (LC::FREE) {
    ok = inverse_inertia(P.inertia(0), lc.Iinv);
    const Soa<double> bpS(a.bp, NBF, a.Bp);
#pragma unroll
    for (int k = 0; k < 9; ++k) lc.rt[k] = bpS.ld(BF_RT + k, lane);
#pragma unroll
    for (int k = 0; k < 3; ++k) lc.rt[9 + k] = bpS.ld(BF_PT + k, lane);
    if constexpr (CONTACTS) {
      const int c0 = a.cstart[0], c1 = a.cstart[1];
      lc.cached = c1 - c0 <= LC::MAXC * L && a.nh <= LC::MAXH;  // uniform over the bucket
      if (lc.cached) {
        const Soa<double> cpS(a.cp, NCF * a.nc, a.Bp), hpS(a.hp, NHF * a.nh, a.Bp);
        lc.ncl = c1 - c0 > sub ? (c1 - c0 - sub + L - 1) / L : 0;
        lc.nh = a.nh;
#pragma unroll
        for (int m = 0; m < LC::MAXC; ++m)
#pragma unroll
          for (int f = 0; f < NCF; ++f) lc.cp[m][f] = m < lc.ncl ? cpS.ld(f * a.nc + c0 + sub + m * L, lane) : 0.0;
#pragma unroll
        for (int h = 0; h < LC::MAXH; ++h)
#pragma unroll
          for (int f = 0; f < NHF; ++f) lc.hp[h][f] = h < lc.nh ? hpS.ld(f * a.nh + h, lane) : 0.0;
      }
    }
  }
  return ok;
}

// Articulated-body dynamics of one instance with a compile-time topology.
// SIE = true: one semi-implicit Euler step (integrators.rs:276-319), each
// joint's v and q advanced as soon as its acceleration is known in the outward
// pass (nothing later in the step reads them), so no acceleration array is
// held; on the last step the accelerations are stored to vdot. SIE = false:
// the accelerations at (q, v) go to qdd_out and q, v are left alone (one stage of
// RK2 / RK4). Torques are read from memory where used. ALLREV: every 1-DOF
// joint of the bucket is revolute (no per-lane select). Iinv: the inverse
// inertia of a lone free body (lone_free_body), else unused.
template <class Topo, class R, int L, bool CONTACTS, bool ALLREV, bool SIE, class FP, class LC>
__device__ __forceinline__ bool fast_step(const BucketArgs& a, const FP& P, int lane, int sub,
                                          double* q, double* v, double dt, bool last, const LC& lc, double* qdd_out) {
  constexpr int N = Topo::N;
  const int Bp = a.Bp;
  const Soa<double> tauS(a.tau, a.nv, Bp), vdotS(a.vdot, a.nv, Bp), bpS(a.bp, NBF * N, Bp);
  const Soa<int> kindS(a.kind, a.nb, Bp);
  const Soa<double> cpS(a.cp, NCF * a.nc, Bp), hpS(a.hp, NHF * a.nh, Bp);
  auto tau = [&](int k) { return LC::TAU_REG ? lc.tau[LC::TAU_REG ? k : 0] : tauS.ld(k, lane); };
  Body1<R> b[N];
  sv<R> vel[N];
  // the floating root (if any) keeps its full transform
  m3<R> E0;
  v3<R> r0;
  sv<R> cb0, pA0;
  m3<double> Rw[CONTACTS ? N : 1];
  v3<double> tw[CONTACTS ? N : 1];
  bool ok = true;
  // the tier's values are re-read in every sweep, not kept in registers
  // across sweeps (the compiler would otherwise hold every body's rebuilt
  // rotation from pass 1 to pass 3 and spill)
  P.sweep_fence();
  // ---- pass 1: kinematics, velocities, bias terms (root to leaves)
#pragma unroll
  for (int i = Topo::first; i < N; ++i) {
    const int pr = sweep_parent<Topo>(i);
    const int pp = pr < 0 ? 0 : pr;
    m3<R> E;
    v3<R> r;
    sv<R> vj{{R(0), R(0), R(0)}, {R(0), R(0), R(0)}};
    m3<double> Rm_d;
    v3<double> p_d;
    if (Topo::dof(i) == 1) {
      const bool rev = ALLREV || kindS.ld(i, lane) == K_REV;
      b[i].rev = rev;
      const double qi = q[qoff_of<Topo>(i)];
      sincos_q(rev ? qi : 0.0, &b[i].s, &b[i].c);
      b[i].d = rev ? R(0) : R(qi);
      joint_frame(P.Rt(i), P.pt(i), b[i].c, b[i].s, b[i].d, E, r);
      const R qd = R(v[voff_of<Topo>(i)]);
      vj.w.z = rev ? qd : R(0);
      vj.v.z = rev ? R(0) : qd;
      if (CONTACTS) {
        m3<double> Ed = cvt<double>(E);
        Rm_d = m3<double>{{Ed.a[0], Ed.a[3], Ed.a[6], Ed.a[1], Ed.a[4], Ed.a[7], Ed.a[2], Ed.a[5], Ed.a[8]}};
        p_d = cvt<double>(r);
      }
    } else {
      const int o = qoff_of<Topo>(i), w = voff_of<Topo>(i);
      m3<double> Rq = quat_mat(q[o], q[o + 1], q[o + 2], q[o + 3]);
      v3<double> tq{q[o + 4], q[o + 5], q[o + 6]};
      m3<double> Rtd;
      v3<double> ptd;
#pragma unroll
      for (int k = 0; k < 9; ++k) Rtd.a[k] = lone_free_body<Topo>() ? lc.rt[k] : bpS.ld((BF_RT + k) * N + i, lane);
#pragma unroll
      for (int k = 0; k < 3; ++k)
        (&ptd.x)[k] = lone_free_body<Topo>() ? lc.rt[9 + k] : bpS.ld((BF_PT + k) * N + i, lane);
      Rm_d = mul(Rtd, Rq);
      p_d = ptd + mul(Rtd, tq);
      const m3<R> Rm = cvt<R>(Rm_d);
      E = m3<R>{{Rm.a[0], Rm.a[3], Rm.a[6], Rm.a[1], Rm.a[4], Rm.a[7], Rm.a[2], Rm.a[5], Rm.a[8]}};
      r = cvt<R>(p_d);
      vj.w = v3<R>{R(v[w]), R(v[w + 1]), R(v[w + 2])};
      vj.v = v3<R>{R(v[w + 3]), R(v[w + 4]), R(v[w + 5])};
      E0 = E;
      r0 = r;
    }
    vel[i] = pr < 0 ? vj : xmotion(E, r, vel[pp]) + vj;
    const sv<R>& V = vel[i];
    ai6<R> I = P.inertia(i);
    sv<R> pA = crf(V, mul(I, V));
    if (Topo::dof(i) == 1) {
      const R qd = sel(b[i].rev, vj.w.z, vj.v.z);
      // v x_m (S qd) with S = e_z angular (revolute) or linear (prismatic)
      b[i].cw0 = b[i].rev ? qd * V.w.y : R(0);
      b[i].cw1 = b[i].rev ? -qd * V.w.x : R(0);
      b[i].cv0 = qd * sel(b[i].rev, V.v.y, V.w.y);
      b[i].cv1 = -qd * sel(b[i].rev, V.v.x, V.w.x);
    } else {
      cb0 = crm(V, vj);
    }
    if (CONTACTS) {
      if (pr < 0) { Rw[i] = Rm_d; tw[i] = p_d; }
      else { Rw[i] = mul(Rw[pp], Rm_d); tw[i] = tw[pp] + mul(Rw[pp], p_d); }
      const m3<R> RwR = cvt<R>(Rw[i]);
      sv<R> fe{{R(0), R(0), R(0)}, {R(0), R(0), R(0)}};
      const int c0 = a.cstart[i], c1 = a.cstart[i + 1];
      // half-spaces outside, this lane's contact points inside: per half-space
      // the signed distance of a world point R p + t is (R^T n) . p + n . (t - point),
      // so each point costs one FP64 dot product; velocity and force only in contact
      auto touch = [&](const v3<double>& pb, double kc, double dist, const v3<double>& n, R alpha, R mu) {
        const v3<R> pbR = cvt<R>(pb);
        const v3<R> vw = mul(RwR, V.v + cross(V.w, pbR));
        const v3<R> F = contact_force(R(-dist), cvt<R>(n), vw, R(kc), R(kHalfspaceK), alpha, mu);
        const v3<R> Fb = mulT(RwR, F);
        fe.w = fe.w + cross(pbR, Fb);
        fe.v = fe.v + Fb;
      };
      if (lc.cached) {
#pragma unroll
        for (int h = 0; h < LC::MAXH; ++h) {
          if (h >= lc.nh) break;
          const v3<double> n{lc.hp[h][HF_NORMAL], lc.hp[h][HF_NORMAL + 1], lc.hp[h][HF_NORMAL + 2]};
          const v3<double> nb = mulT(Rw[i], n);
          const double off = dot(tw[i] - v3<double>{lc.hp[h][HF_POINT], lc.hp[h][HF_POINT + 1], lc.hp[h][HF_POINT + 2]}, n);
          const R alpha = R(lc.hp[h][HF_ALPHA]), mu = R(lc.hp[h][HF_MU]);
#pragma unroll
          for (int m = 0; m < LC::MAXC; ++m) {
            if (m >= lc.ncl) break;
            const v3<double> pb{lc.cp[m][CF_P], lc.cp[m][CF_P + 1], lc.cp[m][CF_P + 2]};
            const double dist = dot(nb, pb) + off;
            if (dist <= kInsideMargin) touch(pb, lc.cp[m][CF_K], dist, n, alpha, mu);
          }
        }
      } else
      for (int h = 0; h < a.nh && c0 + sub < c1; ++h) {
        auto hpar = [&](int fld) { return hpS.ld(fld * a.nh + h, lane); };
        const v3<double> n{hpar(HF_NORMAL), hpar(HF_NORMAL + 1), hpar(HF_NORMAL + 2)};
        const v3<double> nb = mulT(Rw[i], n);
        const double off = dot(tw[i] - v3<double>{hpar(HF_POINT), hpar(HF_POINT + 1), hpar(HF_POINT + 2)}, n);
        const R alpha = R(hpar(HF_ALPHA)), mu = R(hpar(HF_MU));
        for (int cc = c0 + sub; cc < c1; cc += L) {
          auto cpar = [&](int fld) { return cpS.ld(fld * a.nc + cc, lane); };
          const v3<double> pb{cpar(CF_P), cpar(CF_P + 1), cpar(CF_P + 2)};
          const double dist = dot(nb, pb) + off;
          if (dist <= kInsideMargin) touch(pb, cpar(CF_K), dist, n, alpha, mu);
        }
      }
      if (L > 1) {
        fe.w.x = lane_sum<R, L>(fe.w.x); fe.w.y = lane_sum<R, L>(fe.w.y); fe.w.z = lane_sum<R, L>(fe.w.z);
        fe.v.x = lane_sum<R, L>(fe.v.x); fe.v.y = lane_sum<R, L>(fe.v.y); fe.v.z = lane_sum<R, L>(fe.v.z);
      }
      pA = pA - fe;
    }
    if (Topo::dof(i) == 1) b[i].pA = pA; else pA0 = pA;
  }
  P.sweep_fence();
  // ---- pass 2: articulated inertias (leaves to root)
  ai6<R> acc[N];
  bool has_acc[N];
#pragma unroll
  for (int i = 0; i < N; ++i) has_acc[i] = false;
  double Lf[21], D6[6][6], u6[6];
#pragma unroll
  for (int i = N - 1; i >= Topo::first; --i) {
    const int pr = sweep_parent<Topo>(i);
    const int pp = pr < 0 ? 0 : pr;
    ai6<R> IA = P.inertia(i);
    if (has_acc[i]) add_to(IA, acc[i]);
    if (Topo::dof(i) == 1) {
      Body1<R>& B = b[i];
      B.U = B.rev ? col(IA, 2) : col(IA, 5);
      B.Dinv = 1.0 / double(sel(B.rev, IA.A.zz, IA.C.zz));
      double tq = tau(voff_of<Topo>(i));
      if (!ALLREV) {
        // prismatic spring, branch-free: a revolute joint or k = 0 adds -0
        const double sk = B.rev ? 0.0 : double(P(BF_SK, i));
        tq += -sk * (q[qoff_of<Topo>(i)] - double(P(BF_SL, i)));
      }
      B.u = tq - double(sel(B.rev, B.pA.w.z, B.pA.v.z));
      ok = ok && (B.Dinv > 0.0);
      if (pr >= 0) {
        const R inv = R(B.Dinv);
        const sv<R> uD{inv * B.U.w, inv * B.U.v};
        const ai6<R> Ia = rank1_down(IA, B.U, uD);
        const R uR = R(B.u);
        const sv<R> cbv{{B.cw0, B.cw1, R(0)}, {B.cv0, B.cv1, R(0)}};
        const sv<R> pa = B.pA + mul(Ia, cbv) + sv<R>{uR * uD.w, uR * uD.v};
        m3<R> E;
        v3<R> r;
        joint_frame(P.Rt(i), P.pt(i), B.c, B.s, B.d, E, r);
        const ai6<R> add = xtix(E, r, Ia);
        if (has_acc[pp]) add_to(acc[pp], add);
        else { acc[pp] = add; has_acc[pp] = true; }
        if (Topo::dof(pp) == 1) b[pp].pA = b[pp].pA + xtforce(E, r, pa);
        else pA0 = pA0 + xtforce(E, r, pa);
      }
    } else {
      // floating joint (the root in every compiled topology): D = I^A (6x6)
      const int w = voff_of<Topo>(i);
      u6[0] = tau(w) - double(pA0.w.x); u6[1] = tau(w + 1) - double(pA0.w.y); u6[2] = tau(w + 2) - double(pA0.w.z);
      u6[3] = tau(w + 3) - double(pA0.v.x); u6[4] = tau(w + 4) - double(pA0.v.y); u6[5] = tau(w + 5) - double(pA0.v.z);
      if constexpr (!lone_free_body<Topo>()) {
        to_dense(IA, D6);
        ok = chol6(D6, Lf) && ok;
      }
    }
  }
  P.sweep_fence();
  // ---- pass 3: accelerations (root to leaves)
  sv<R> accel[N];
  sv<R> a0{{R(0), R(0), R(0)}, {R(0), R(0), R(a.g)}};
  if constexpr (LC::WELDED) a0.v = v3<R>{R(lc.abase[0]), R(lc.abase[1]), R(lc.abase[2])};
#pragma unroll
  for (int i = Topo::first; i < N; ++i) {
    const int pr = sweep_parent<Topo>(i);
    const int pp = pr < 0 ? 0 : pr;
    if (Topo::dof(i) == 1) {
      const Body1<R>& B = b[i];
      m3<R> E;
      v3<R> r;
      joint_frame(P.Rt(i), P.pt(i), B.c, B.s, B.d, E, r);
      sv<R> ap = xmotion(E, r, pr < 0 ? a0 : accel[pp]);
      ap.w.x += B.cw0; ap.w.y += B.cw1; ap.v.x += B.cv0; ap.v.y += B.cv1;
      const double qdd = (B.u - double(dot6(B.U, ap))) * B.Dinv;
      const int vo = voff_of<Topo>(i);
      ok = ok && isfinite(qdd);
      if constexpr (SIE) {
        if (last && sub == 0) vdotS.st(vo, lane, qdd);
        v[vo] = v[vo] + qdd * dt;
        q[qoff_of<Topo>(i)] += v[vo] * dt;
      } else {
        qdd_out[vo] = qdd;
      }
      ap.w.z = B.rev ? ap.w.z + R(qdd) : ap.w.z;
      ap.v.z = B.rev ? ap.v.z : ap.v.z + R(qdd);
      accel[i] = ap;
    } else {
      sv<R> ap = xmotion(E0, r0, pr < 0 ? a0 : accel[pp]) + cb0;
      const int w = voff_of<Topo>(i);
      const double apd[6] = {double(ap.w.x), double(ap.w.y), double(ap.w.z), double(ap.v.x), double(ap.v.y), double(ap.v.z)};
      double x[6];
      if constexpr (lone_free_body<Topo>()) {
        // x = I^-1 (u - I a_p) = I^-1 u - a_p
#pragma unroll
        for (int rr = 0; rr < 6; ++rr) {
          double s = 0.0;
#pragma unroll
          for (int cc = 0; cc < 6; ++cc) s += lc.Iinv[sym6(rr, cc)] * u6[cc];
          x[rr] = s - apd[rr];
        }
      } else {
#pragma unroll
        for (int rr = 0; rr < 6; ++rr) {
          double s = u6[rr];
#pragma unroll
          for (int cc = 0; cc < 6; ++cc) s -= D6[rr][cc] * apd[cc];
          x[rr] = s;
        }
        chol6_solve(Lf, x);
      }
#pragma unroll
      for (int t = 0; t < 6; ++t) {
        ok = ok && isfinite(x[t]);
        if constexpr (SIE) {
          if (last && sub == 0) vdotS.st(w + t, lane, x[t]);
          v[w + t] = v[w + t] + x[t] * dt;
        } else {
          qdd_out[w + t] = x[t];
        }
      }
      if constexpr (SIE) advance_floating(&q[qoff_of<Topo>(i)], &v[w], dt, &q[qoff_of<Topo>(i)]);
      accel[i] = ap + sv<R>{{R(x[0]), R(x[1]), R(x[2])}, {R(x[3]), R(x[4]), R(x[5])}};
    }
  }
  return ok;
}

// n_steps semi-implicit Euler steps per launch with the state in registers;
// q and v accumulate in FP64. With `active`, lane l runs min(n_steps,
// active[l]) steps (serving: instances with different horizons).
template <class Topo, class R, int L, bool CONTACTS, bool ALLREV, bool LDS = false>
__device__ __forceinline__ void fast_sie_body(const BucketArgs& a, int tid, double dt, int n_steps) {
  constexpr int NQ = nq_of<Topo>(), NV = nv_of<Topo>();
  const int inst = tid / L;
  const int sub = tid % L;
  if (inst >= a.B) return;
  const int Bp = a.Bp;
  int my_steps = n_steps;
  if (a.active) {
    // serving: an idle slot does nothing; a busy one runs what is left of its
    // horizon in this chunk and counts it down
    const int left = a.active[inst];
    if (left <= 0) return;
    my_steps = min(my_steps, left);
    if (sub == 0) a.active[inst] = left - my_steps;
  }
  const Soa<double> qS(a.q, NQ, Bp), vS(a.v, NV, Bp);
  double q[NQ], v[NV];
#pragma unroll
  for (int k = 0; k < NQ; ++k) q[k] = qS.ld(k, inst);
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = vS.ld(k, inst);
  bool ok = true;
  constexpr int PC = cached_bodies<Topo, R, ALLREV>();
  const FastParams<Topo, R, PC, LDS> P(pick<R>(a.bp, a.bpf), Bp, inst, 0, LDS ? (R*)GP_DYN_LDS : nullptr);
  // the held torques in registers for the launch, except in the FP64 chains
  // whose per-body state fills the register file (12 joints, or 9-12 with
  // per-lane joint kinds): they read them in each step, from L1
  constexpr bool long64 = sizeof(R) == 8 && (Topo::N >= 12 || (!ALLREV && Topo::N >= 9));
  constexpr bool tau_reg = (PC == Topo::N || FastParams<Topo, R, PC, LDS>::NT > 0) && !long64;
  LaunchConst<Topo, L, CONTACTS, tau_reg> lc;
  ok = load_launch_const<Topo, R, L, CONTACTS>(a, P, inst, sub, lc);
  for (int s = 0; s < my_steps; ++s)
    ok = fast_step<Topo, R, L, CONTACTS, ALLREV, true>(a, P, inst, sub, q, v, dt, s + 1 == my_steps, lc,
                                                      nullptr) && ok;
  if (sub != 0) return;
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    qS.st(k, inst, q[k]);
    ok = ok && isfinite(q[k]);
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    vS.st(k, inst, v[k]);
    ok = ok && isfinite(v[k]);
  }
  if (!ok) a.fault[inst] = 1;
}

template <class Topo, class R, int L, bool CONTACTS, bool ALLREV>
__global__ void __launch_bounds__(256) fast_sie_kernel(BucketArgs a, double dt, int n_steps) {
  fast_sie_body<Topo, R, L, CONTACTS, ALLREV, true>(a, blockIdx.x * blockDim.x + threadIdx.x, dt, n_steps);
}

// euler_step (integrators.rs:230-271): qo = q (+) h v, vo = v + h acc. In place
// (qo = q, vo = v) is allowed: q is advanced before v changes.
template <class Topo>
__device__ __forceinline__ void euler_stage(const double* q, const double* v, const double* acc, double h, double* qo,
                                            double* vo) {
#pragma unroll
  for (int i = 0; i < Topo::N; ++i) {
    if (Topo::dof(i) == 1) qo[qoff_of<Topo>(i)] = q[qoff_of<Topo>(i)] + v[voff_of<Topo>(i)] * h;
    else if (Topo::dof(i) == 6) advance_floating(&q[qoff_of<Topo>(i)], &v[voff_of<Topo>(i)], h, &qo[qoff_of<Topo>(i)]);
  }
#pragma unroll
  for (int k = 0; k < nv_of<Topo>(); ++k) vo[k] = v[k] + acc[k] * h;
}

// n_steps RK2 or RK4 steps per launch (integrators.rs:177-225) with the state
// in registers. The reference's stage rule: every stage state is an
// euler_step from the step's start (q, v) with the previous stage's
// accelerations (h = dt/2, dt/2, dt), and RK4 combines
// (a1 + 2 a2 + 2 a3 + a4) / 6 in that order. One copy of the dynamics serves
// every stage (a stage loop that is not unrolled). vdot receives the last
// step's first-stage accelerations, as the plain kernel leaves it.
template <class Topo, class R, int L, bool CONTACTS, bool ALLREV, bool LDS = false>
__device__ __forceinline__ void fast_rk_body(const BucketArgs& a, int tid, double dt, int n_steps, int integrator) {
  constexpr int NQ = nq_of<Topo>(), NV = nv_of<Topo>();
  const int inst = tid / L;
  const int sub = tid % L;
  if (inst >= a.B) return;
  const int Bp = a.Bp;
  int my_steps = n_steps;
  if (a.active) {
    const int left = a.active[inst];
    if (left <= 0) return;
    my_steps = min(my_steps, left);
    if (sub == 0) a.active[inst] = left - my_steps;
  }
  const Soa<double> qS(a.q, NQ, Bp), vS(a.v, NV, Bp), vdotS(a.vdot, NV, Bp);
  double q[NQ], v[NV];
#pragma unroll
  for (int k = 0; k < NQ; ++k) q[k] = qS.ld(k, inst);
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = vS.ld(k, inst);
  bool ok = true;
  constexpr int PC = cached_bodies<Topo, R, ALLREV>();
  const FastParams<Topo, R, PC, LDS> P(pick<R>(a.bp, a.bpf), Bp, inst, 0, LDS ? (R*)GP_DYN_LDS : nullptr);
  constexpr bool tau_reg = PC == Topo::N || FastParams<Topo, R, PC, LDS>::NT > 0;
  LaunchConst<Topo, L, CONTACTS, tau_reg> lc;
  ok = load_launch_const<Topo, R, L, CONTACTS>(a, P, inst, sub, lc);
  const bool rk4 = integrator == GP_RUNGE_KUTTA_4;
  const int stages = rk4 ? 4 : 2;
  double qs[NQ], vs[NV], ak[NV], sum[NV];
  for (int s = 0; s < my_steps; ++s) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) qs[k] = q[k];
#pragma unroll
    for (int k = 0; k < NV; ++k) vs[k] = v[k];
#pragma unroll 1
    for (int st = 0; st < stages; ++st) {
      ok = fast_step<Topo, R, L, CONTACTS, ALLREV, false>(a, P, inst, sub, qs, vs, 0.0, false, lc, ak) && ok;
      if (st == 0) {
        if (s + 1 == my_steps && sub == 0) {
#pragma unroll
          for (int k = 0; k < NV; ++k) vdotS.st(k, inst, ak[k]);
        }
#pragma unroll
        for (int k = 0; k < NV; ++k) sum[k] = ak[k];
      } else if (st < 3) {
#pragma unroll
        for (int k = 0; k < NV; ++k) sum[k] = sum[k] + ak[k] * 2.0;
      } else {
#pragma unroll
        for (int k = 0; k < NV; ++k) sum[k] = (sum[k] + ak[k]) / 6.0;
      }
      if (st + 1 < stages) euler_stage<Topo>(q, v, ak, st == 2 ? dt : dt / 2.0, qs, vs);
      else euler_stage<Topo>(q, v, rk4 ? sum : ak, dt, q, v);
    }
  }
  if (sub != 0) return;
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    qS.st(k, inst, q[k]);
    ok = ok && isfinite(q[k]);
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    vS.st(k, inst, v[k]);
    ok = ok && isfinite(v[k]);
  }
  if (!ok) a.fault[inst] = 1;
}

template <class Topo, class R, int L, bool CONTACTS, bool ALLREV>
__global__ void __launch_bounds__(256) fast_rk_kernel(BucketArgs a, double dt, int n_steps, int integrator) {
  fast_rk_body<Topo, R, L, CONTACTS, ALLREV, true>(a, blockIdx.x * blockDim.x + threadIdx.x, dt, n_steps,
                                                   integrator);
}

// ---------------------------------------------------------------- trees, one lane per leg
// A floating base with LEGS legs, each a chain of DEPTH one-DOF joints (body
// 1 + l * DEPTH + k is joint k of leg l; TopoQuad is LEGS = 4, DEPTH = 3).
// LEGS consecutive lanes hold one instance: each runs the base redundantly and
// its own leg; the legs' contributions to the base's articulated inertia and
// bias force are summed across the lanes with xor butterflies (every lane gets
// the same bits), and every lane solves the base's 6x6 system in FP64.
template <int LEGS, int DEPTH>
struct TopoTree {
  static constexpr int N = 1 + LEGS * DEPTH;
};

template <class R, int LEGS>
__device__ __forceinline__ R legs_sum(R x) {
#pragma unroll
  for (int off = 1; off < LEGS; off <<= 1) x += __shfl_xor(x, off, LEGS);
  return x;
}
template <class R, int LEGS>
__device__ __forceinline__ void legs_sum(ai6<R>& I, sv<R>& f) {
  I.A.xx = legs_sum<R, LEGS>(I.A.xx); I.A.yy = legs_sum<R, LEGS>(I.A.yy); I.A.zz = legs_sum<R, LEGS>(I.A.zz);
  I.A.xy = legs_sum<R, LEGS>(I.A.xy); I.A.xz = legs_sum<R, LEGS>(I.A.xz); I.A.yz = legs_sum<R, LEGS>(I.A.yz);
#pragma unroll
  for (int k = 0; k < 9; ++k) I.B.a[k] = legs_sum<R, LEGS>(I.B.a[k]);
  I.C.xx = legs_sum<R, LEGS>(I.C.xx); I.C.yy = legs_sum<R, LEGS>(I.C.yy); I.C.zz = legs_sum<R, LEGS>(I.C.zz);
  I.C.xy = legs_sum<R, LEGS>(I.C.xy); I.C.xz = legs_sum<R, LEGS>(I.C.xz); I.C.yz = legs_sum<R, LEGS>(I.C.yz);
  f.w.x = legs_sum<R, LEGS>(f.w.x); f.w.y = legs_sum<R, LEGS>(f.w.y); f.w.z = legs_sum<R, LEGS>(f.w.z);
  f.v.x = legs_sum<R, LEGS>(f.v.x); f.v.y = legs_sum<R, LEGS>(f.v.y); f.v.z = legs_sum<R, LEGS>(f.v.z);
}

template <class R, int LEGS, int DEPTH, bool ALLREV>
__device__ __forceinline__ void tree_sie_body(const BucketArgs& a, int tid, double dt, int n_steps) {
  constexpr int N = TopoTree<LEGS, DEPTH>::N;
  const int inst = tid / LEGS;
  const int leg = tid % LEGS;
  if (inst >= a.B) return;
  const int Bp = a.Bp;
  int my_steps = n_steps;
  if (a.active) {
    const int left = a.active[inst];
    if (left <= 0) return;
    my_steps = min(my_steps, left);
    if (leg == 0) a.active[inst] = left - my_steps;
  }
  // this lane's leg: bodies 1 + legrow + k, q rows 7 + legrow + k, v rows 6 + legrow + k
  const int legrow = leg * DEPTH;
  const FastParams<TopoTree<LEGS, DEPTH>, R> P(pick<R>(a.bp, a.bpf), Bp, inst);
  const FastParams<TopoTree<LEGS, DEPTH>, R> PL(pick<R>(a.bp, a.bpf), Bp, inst, legrow);
  const Soa<double> qS(a.q, 7 + LEGS * DEPTH, Bp), vS(a.v, 6 + LEGS * DEPTH, Bp), tauS(a.tau, 6 + LEGS * DEPTH, Bp),
      vdotS(a.vdot, 6 + LEGS * DEPTH, Bp), bpS(a.bp, NBF * N, Bp);
  const Soa<int> kindS(a.kind, N, Bp);
  // state: base (q 7, v 6) in every lane, this leg's joints
  double qb[7], vb[6], tb[6], ab[6];
  double ql[DEPTH], vl[DEPTH], tl[DEPTH], al[DEPTH];
#pragma unroll
  for (int k = 0; k < 7; ++k) qb[k] = qS.ld(k, inst);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    vb[k] = vS.ld(k, inst);
    tb[k] = tauS.ld(k, inst);
    ab[k] = vdotS.ld(k, inst);
  }
#pragma unroll
  for (int k = 0; k < DEPTH; ++k) {
    ql[k] = qS.ldx(7 + k, inst, legrow);
    vl[k] = vS.ldx(6 + k, inst, legrow);
    tl[k] = tauS.ldx(6 + k, inst, legrow);
    al[k] = vdotS.ldx(6 + k, inst, legrow);
  }
  bool ok = true;
  for (int s = 0; s < my_steps; ++s) {
    // ---- pass 1: base, then this leg
    m3<double> Rq = quat_mat(qb[0], qb[1], qb[2], qb[3]);
    m3<double> Rtd;
#pragma unroll
    for (int k = 0; k < 9; ++k) Rtd.a[k] = bpS.ld(((BF_RT + k) * N), inst);
    v3<double> ptd{bpS.ld((BF_PT * N), inst), bpS.ld(((BF_PT + 1) * N), inst),
                   bpS.ld(((BF_PT + 2) * N), inst)};
    const m3<R> Rm = cvt<R>(mul(Rtd, Rq));
    const m3<R> E0{{Rm.a[0], Rm.a[3], Rm.a[6], Rm.a[1], Rm.a[4], Rm.a[7], Rm.a[2], Rm.a[5], Rm.a[8]}};
    const v3<R> r0 = cvt<R>(ptd + mul(Rtd, v3<double>{qb[4], qb[5], qb[6]}));
    const sv<R> V0{{R(vb[0]), R(vb[1]), R(vb[2])}, {R(vb[3]), R(vb[4]), R(vb[5])}};
    const sv<R> cb0 = crm(V0, V0);
    const ai6<R> I0 = P.inertia(0);
    sv<R> pA0 = crf(V0, mul(I0, V0));
    Body1<R> b[DEPTH];
    sv<R> vel = V0;
#pragma unroll
    for (int k = 0; k < DEPTH; ++k) {
      const int i = 1 + k;  // in PL, shifted to this leg
      const bool rev = ALLREV || kindS.ldx(i, inst, legrow) == K_REV;
      b[k].rev = rev;
      sincos_q(rev ? ql[k] : 0.0, &b[k].s, &b[k].c);
      b[k].d = rev ? R(0) : R(ql[k]);
      m3<R> E;
      v3<R> r;
      joint_frame(PL.Rt(i), PL.pt(i), b[k].c, b[k].s, b[k].d, E, r);
      const R qd = R(vl[k]);
      sv<R> vj{{R(0), R(0), rev ? qd : R(0)}, {R(0), R(0), rev ? R(0) : qd}};
      vel = xmotion(E, r, vel) + vj;
      b[k].pA = crf(vel, mul(PL.inertia(i), vel));
      b[k].cw0 = rev ? qd * vel.w.y : R(0);
      b[k].cw1 = rev ? -qd * vel.w.x : R(0);
      b[k].cv0 = qd * sel(rev, vel.v.y, vel.w.y);
      b[k].cv1 = -qd * sel(rev, vel.v.x, vel.w.x);
    }
    // ---- pass 2: this leg, leaves to its first joint, then the base
    ai6<R> acc;
    ai6<R> cIA;
    sv<R> cpA;
#pragma unroll
    for (int k = DEPTH - 1; k >= 0; --k) {
      const int i = 1 + k;  // in PL, shifted to this leg
      Body1<R>& B = b[k];
      ai6<R> IA = PL.inertia(i);
      if (k < DEPTH - 1) add_to(IA, acc);
      B.U = B.rev ? col(IA, 2) : col(IA, 5);
      B.Dinv = 1.0 / double(sel(B.rev, IA.A.zz, IA.C.zz));
      double tq = tl[k];
      if (!ALLREV) {
        // prismatic spring, branch-free: a revolute joint or k = 0 adds -0
        const double sk = B.rev ? 0.0 : double(PL(BF_SK, i));
        tq += -sk * (ql[k] - double(PL(BF_SL, i)));
      }
      B.u = tq - double(sel(B.rev, B.pA.w.z, B.pA.v.z));
      ok = ok && (B.Dinv > 0.0);
      const R inv = R(B.Dinv);
      const sv<R> uD{inv * B.U.w, inv * B.U.v};
      const ai6<R> Ia = rank1_down(IA, B.U, uD);
      const R uR = R(B.u);
      const sv<R> cbv{{B.cw0, B.cw1, R(0)}, {B.cv0, B.cv1, R(0)}};
      const sv<R> pa = B.pA + mul(Ia, cbv) + sv<R>{uR * uD.w, uR * uD.v};
      m3<R> E;
      v3<R> r;
      joint_frame(PL.Rt(i), PL.pt(i), B.c, B.s, B.d, E, r);
      if (k > 0) {
        acc = xtix(E, r, Ia);
        b[k - 1].pA = b[k - 1].pA + xtforce(E, r, pa);
      } else {
        cIA = xtix(E, r, Ia);
        cpA = xtforce(E, r, pa);
      }
    }
    legs_sum<R, LEGS>(cIA, cpA);
    ai6<R> IA0 = I0;
    add_to(IA0, cIA);
    pA0 = pA0 + cpA;
    double D6[6][6], Lf[21], u6[6];
    to_dense(IA0, D6);
    ok = chol6(D6, Lf) && ok;
    u6[0] = tb[0] - double(pA0.w.x); u6[1] = tb[1] - double(pA0.w.y); u6[2] = tb[2] - double(pA0.w.z);
    u6[3] = tb[3] - double(pA0.v.x); u6[4] = tb[4] - double(pA0.v.y); u6[5] = tb[5] - double(pA0.v.z);
    // ---- pass 3: base, then this leg
    const sv<R> ag{{R(0), R(0), R(0)}, {R(0), R(0), R(a.g)}};
    sv<R> ap0 = xmotion(E0, r0, ag) + cb0;
    const double apd[6] = {double(ap0.w.x), double(ap0.w.y), double(ap0.w.z),
                           double(ap0.v.x), double(ap0.v.y), double(ap0.v.z)};
    double x[6];
#pragma unroll
    for (int rr = 0; rr < 6; ++rr) {
      double t = u6[rr];
#pragma unroll
      for (int cc = 0; cc < 6; ++cc) t -= D6[rr][cc] * apd[cc];
      x[rr] = t;
    }
    chol6_solve(Lf, x);
#pragma unroll
    for (int t = 0; t < 6; ++t) ab[t] = x[t];
    sv<R> accel = ap0 + sv<R>{{R(x[0]), R(x[1]), R(x[2])}, {R(x[3]), R(x[4]), R(x[5])}};
#pragma unroll
    for (int k = 0; k < DEPTH; ++k) {
      const int i = 1 + k;  // in PL, shifted to this leg
      const Body1<R>& B = b[k];
      m3<R> E;
      v3<R> r;
      joint_frame(PL.Rt(i), PL.pt(i), B.c, B.s, B.d, E, r);
      sv<R> ap = xmotion(E, r, accel);
      ap.w.x += B.cw0; ap.w.y += B.cw1; ap.v.x += B.cv0; ap.v.y += B.cv1;
      const double qdd = (B.u - double(dot6(B.U, ap))) * B.Dinv;
      al[k] = qdd;
      ap.w.z = B.rev ? ap.w.z + R(qdd) : ap.w.z;
      ap.v.z = B.rev ? ap.v.z : ap.v.z + R(qdd);
      accel = ap;
    }
    // ---- semi-implicit Euler (integrators.rs:276-319)
#pragma unroll
    for (int k = 0; k < 6; ++k) vb[k] = vb[k] + ab[k] * dt;
    advance_floating(qb, vb, dt, qb);
#pragma unroll
    for (int k = 0; k < DEPTH; ++k) {
      vl[k] = vl[k] + al[k] * dt;
      ql[k] += vl[k] * dt;
    }
  }
  if (leg == 0) {
#pragma unroll
    for (int k = 0; k < 7; ++k) { qS.st(k, inst, qb[k]); ok = ok && isfinite(qb[k]); }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      vS.st(k, inst, vb[k]);
      vdotS.st(k, inst, ab[k]);
      ok = ok && isfinite(vb[k]) && isfinite(ab[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < DEPTH; ++k) {
    qS.stx(7 + k, inst, legrow, ql[k]);
    vS.stx(6 + k, inst, legrow, vl[k]);
    vdotS.stx(6 + k, inst, legrow, al[k]);
    ok = ok && isfinite(ql[k]) && isfinite(vl[k]) && isfinite(al[k]);
  }
  if (!ok) a.fault[inst] = 1;
}

template <class R, int LEGS, int DEPTH, bool ALLREV>
__global__ void __launch_bounds__(256) tree_sie_kernel(BucketArgs a, double dt, int n_steps) {
  tree_sie_body<R, LEGS, DEPTH, ALLREV>(a, blockIdx.x * blockDim.x + threadIdx.x, dt, n_steps);
}

// ---------------------------------------------------------------- trees, one body per lane
}  // namespace gp
#include "gp_tree_impl.h"
namespace gp {

// ---------------------------------------------------------------- grouped launch
// Every bucket of a batch in one launch: bucket b owns workgroups
// [block_start[b], block_start[b + 1]) and each workgroup runs its bucket's
// compile-time topology. Buckets of different sizes and depths then share the
// chip in one dispatch instead of a series of small launches.
template <class R>
__device__ __forceinline__ void grouped_dispatch(const BucketArgs& a, int topo, int tid, double dt, int n) {
  const bool ar = a.all_rev != 0;
  switch (topo) {
#define GP_CHAIN_CASE(K)                                                               \
  case K:                                                                              \
    if (ar) fast_sie_body<TopoChain<K>, R, 1, false, true, true>(a, tid, dt, n);       \
    else fast_sie_body<TopoChain<K>, R, 1, false, false, true>(a, tid, dt, n);         \
    break;
    GP_CHAIN_CASE(1) GP_CHAIN_CASE(2) GP_CHAIN_CASE(3) GP_CHAIN_CASE(4) GP_CHAIN_CASE(5) GP_CHAIN_CASE(6)
    GP_CHAIN_CASE(7) GP_CHAIN_CASE(8) GP_CHAIN_CASE(9) GP_CHAIN_CASE(10) GP_CHAIN_CASE(11) GP_CHAIN_CASE(12)
#undef GP_CHAIN_CASE
#define GP_WELDED_CASE(K)                                                              \
  case TOPO_WELDED + K:                                                                \
    if (ar) fast_sie_body<TopoWelded<K + 1>, R, 1, false, true, true>(a, tid, dt, n);  \
    else fast_sie_body<TopoWelded<K + 1>, R, 1, false, false, true>(a, tid, dt, n);    \
    break;
    GP_WELDED_CASE(1) GP_WELDED_CASE(2) GP_WELDED_CASE(3) GP_WELDED_CASE(4) GP_WELDED_CASE(5) GP_WELDED_CASE(6)
    GP_WELDED_CASE(7)
#undef GP_WELDED_CASE
    case TOPO_QUAD:
      if (a.lanes == 4) {
        if (ar) tree_sie_body<R, 4, 3, true>(a, tid, dt, n);
        else tree_sie_body<R, 4, 3, false>(a, tid, dt, n);
      } else {
        if (ar) fast_sie_body<TopoQuad, R, 1, false, true>(a, tid, dt, n);
        else fast_sie_body<TopoQuad, R, 1, false, false>(a, tid, dt, n);
      }
      break;
    case TOPO_TREE:
      switch (a.lanes) {
        case 16: wave_tree_body<R, 16>(a, tid, dt, n); break;
        case 32: wave_tree_body<R, 32>(a, tid, dt, n); break;
        case 64: wave_tree_body<R, 64>(a, tid, dt, n); break;
        default: break;
      }
      break;
    case TOPO_FREE:
      switch (a.lanes) {
        case 2: fast_sie_body<TopoFree, R, 2, true, false>(a, tid, dt, n); break;
        case 4: fast_sie_body<TopoFree, R, 4, true, false>(a, tid, dt, n); break;
        case 8: fast_sie_body<TopoFree, R, 8, true, false>(a, tid, dt, n); break;
        default: fast_sie_body<TopoFree, R, 1, true, false>(a, tid, dt, n); break;
      }
      break;
    default: break;
  }
}

template <class R>
__global__ void __launch_bounds__(256) grouped_sie_kernel(const GroupTable* __restrict__ t, double dt, int n_steps) {
  int b = 0;
  while (b + 1 < t->n && (int)blockIdx.x >= t->block_start[b + 1]) ++b;  // uniform per workgroup
  const BucketArgs a = t->args[b];
  const int tid = ((int)blockIdx.x - t->block_start[b]) * (int)blockDim.x + (int)threadIdx.x;
  grouped_dispatch<R>(a, t->topo[b], tid, a.dt > 0.0 ? a.dt : dt, a.n_steps > 0 ? a.n_steps : n_steps);
}


// The largest tier row (bytes per lane) of any chain or welded-chain kernel
// in this precision: a grouped launch reserves it for every lane, whichever
// buckets the launch holds (its kernel runs at one wave per SIMD anyway).
template <class R, bool AR, int... K>
constexpr int max_tier_lane_bytes(std::integer_sequence<int, K...>) {
  int m = 0;
  ((m = m > lds_tier_bytes<TopoChain<K + 1>, R, AR>(1) ? m : lds_tier_bytes<TopoChain<K + 1>, R, AR>(1)), ...);
  ((m = m > lds_tier_bytes<TopoWelded<K + 1>, R, AR>(1) ? m : lds_tier_bytes<TopoWelded<K + 1>, R, AR>(1)), ...);
  return m;
}
template <class R> constexpr int grouped_tier_lane_bytes() {
  constexpr int a = max_tier_lane_bytes<R, true>(std::make_integer_sequence<int, 12>{});
  constexpr int b = max_tier_lane_bytes<R, false>(std::make_integer_sequence<int, 12>{});
  constexpr int m = a > b ? a : b;
  // the tree kernel's exchange slots share the same LDS in a grouped launch
  static_assert(m >= kTreeSlotBytes, "grouped LDS row smaller than a tree exchange slot");
  return m;
}

// Workgroup size that fits the tier's LDS in what the device allows a
// workgroup (the requested size, halved while it does not fit; 64 lanes need
// at most 40 KB).
template <class Topo, class R, bool AR>
static int tier_block(int block) {
  static int max_lds = -1;
  if (max_lds < 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      v = 64 << 10;
    max_lds = v;
  }
  while (block > 64 && lds_tier_bytes<Topo, R, AR>(block) > max_lds) block /= 2;
  return block;
}

template <class Topo, class R, int L, bool C, bool AR>
static int launch_t(const BucketArgs& a, double dt, int n_steps, int block, hipStream_t s) {
  block = tier_block<Topo, R, AR>(block);
  long threads = (long)a.B * L;
  dim3 blk(block), grid((unsigned)((threads + block - 1) / block));
  const unsigned lds = (unsigned)lds_tier_bytes<Topo, R, AR>(block);
  hipLaunchKernelGGL((fast_sie_kernel<Topo, R, L, C, AR>), grid, blk, lds, s, a, dt, n_steps);
  return (int)hipGetLastError();
}

template <class Topo, class R>
static int launch_chain(const BucketArgs& a, double dt, int n, int block, hipStream_t s) {
  if (a.all_rev) return launch_t<Topo, R, 1, false, true>(a, dt, n, block, s);
  return launch_t<Topo, R, 1, false, false>(a, dt, n, block, s);
}

template <class R, int LEGS, int DEPTH>
static int launch_tree(const BucketArgs& a, double dt, int n_steps, int block, hipStream_t s) {
  long threads = (long)a.B * LEGS;
  dim3 blk(block), grid((unsigned)((threads + block - 1) / block));
  if (a.all_rev) hipLaunchKernelGGL((tree_sie_kernel<R, LEGS, DEPTH, true>), grid, blk, 0, s, a, dt, n_steps);
  else hipLaunchKernelGGL((tree_sie_kernel<R, LEGS, DEPTH, false>), grid, blk, 0, s, a, dt, n_steps);
  return (int)hipGetLastError();
}

template <class Topo, class R, int L, bool C, bool AR>
static int launch_rk_t(const BucketArgs& a, double dt, int n_steps, int integrator, int block, hipStream_t s) {
  block = tier_block<Topo, R, AR>(block);
  long threads = (long)a.B * L;
  dim3 blk(block), grid((unsigned)((threads + block - 1) / block));
  const unsigned lds = (unsigned)lds_tier_bytes<Topo, R, AR>(block);
  hipLaunchKernelGGL((fast_rk_kernel<Topo, R, L, C, AR>), grid, blk, lds, s, a, dt, n_steps, integrator);
  return (int)hipGetLastError();
}

template <class Topo, class R>
static int launch_rk_chain(const BucketArgs& a, double dt, int n, int integ, int block, hipStream_t s) {
  if (a.all_rev) return launch_rk_t<Topo, R, 1, false, true>(a, dt, n, integ, block, s);
  return launch_rk_t<Topo, R, 1, false, false>(a, dt, n, integ, block, s);
}

// RK2 / RK4 launches (one bucket per launch; quadrupeds one lane per instance)
template <class R>
static int launch_rk_prec(int topo, const BucketArgs& a, double dt, int n, int integ, int block, hipStream_t s) {
  switch (topo) {
#define GP_RK_CASE(K) \
    case K: return launch_rk_chain<TopoChain<K>, R>(a, dt, n, integ, block, s);
    GP_RK_CASE(1) GP_RK_CASE(2) GP_RK_CASE(3) GP_RK_CASE(4) GP_RK_CASE(5) GP_RK_CASE(6)
    GP_RK_CASE(7) GP_RK_CASE(8) GP_RK_CASE(9) GP_RK_CASE(10) GP_RK_CASE(11) GP_RK_CASE(12)
#undef GP_RK_CASE
#define GP_RK_WELDED_CASE(K) \
    case TOPO_WELDED + K: return launch_rk_chain<TopoWelded<K + 1>, R>(a, dt, n, integ, block, s);
    GP_RK_WELDED_CASE(1) GP_RK_WELDED_CASE(2) GP_RK_WELDED_CASE(3) GP_RK_WELDED_CASE(4) GP_RK_WELDED_CASE(5)
    GP_RK_WELDED_CASE(6) GP_RK_WELDED_CASE(7)
#undef GP_RK_WELDED_CASE
    case TOPO_QUAD: return launch_rk_chain<TopoQuad, R>(a, dt, n, integ, block, s);
    case TOPO_FREE:
      switch (a.lanes) {
        case 2: return launch_rk_t<TopoFree, R, 2, true, false>(a, dt, n, integ, block, s);
        case 4: return launch_rk_t<TopoFree, R, 4, true, false>(a, dt, n, integ, block, s);
        case 8: return launch_rk_t<TopoFree, R, 8, true, false>(a, dt, n, integ, block, s);
        default: return launch_rk_t<TopoFree, R, 1, true, false>(a, dt, n, integ, block, s);
      }
    default: return (int)hipErrorInvalidValue;
  }
}

template <class R>
static int launch_prec(int topo, const BucketArgs& a, double dt, int n, int block, hipStream_t s) {
  switch (topo) {
    case 1: return launch_chain<TopoChain<1>, R>(a, dt, n, block, s);
    case 2: return launch_chain<TopoChain<2>, R>(a, dt, n, block, s);
    case 3: return launch_chain<TopoChain<3>, R>(a, dt, n, block, s);
    case 4: return launch_chain<TopoChain<4>, R>(a, dt, n, block, s);
    case 5: return launch_chain<TopoChain<5>, R>(a, dt, n, block, s);
    case 6: return launch_chain<TopoChain<6>, R>(a, dt, n, block, s);
    case 7: return launch_chain<TopoChain<7>, R>(a, dt, n, block, s);
    case 8: return launch_chain<TopoChain<8>, R>(a, dt, n, block, s);
    case 9: return launch_chain<TopoChain<9>, R>(a, dt, n, block, s);
    case 10: return launch_chain<TopoChain<10>, R>(a, dt, n, block, s);
    case 11: return launch_chain<TopoChain<11>, R>(a, dt, n, block, s);
    case 12: return launch_chain<TopoChain<12>, R>(a, dt, n, block, s);
    case TOPO_WELDED + 1: return launch_chain<TopoWelded<2>, R>(a, dt, n, block, s);
    case TOPO_WELDED + 2: return launch_chain<TopoWelded<3>, R>(a, dt, n, block, s);
    case TOPO_WELDED + 3: return launch_chain<TopoWelded<4>, R>(a, dt, n, block, s);
    case TOPO_WELDED + 4: return launch_chain<TopoWelded<5>, R>(a, dt, n, block, s);
    case TOPO_WELDED + 5: return launch_chain<TopoWelded<6>, R>(a, dt, n, block, s);
    case TOPO_WELDED + 6: return launch_chain<TopoWelded<7>, R>(a, dt, n, block, s);
    case TOPO_WELDED + 7: return launch_chain<TopoWelded<8>, R>(a, dt, n, block, s);
    case TOPO_QUAD:
      if (a.lanes == 4) return launch_tree<R, 4, 3>(a, dt, n, block, s);
      return launch_chain<TopoQuad, R>(a, dt, n, block, s);
    case TOPO_TREE: return launch_wave_tree<R>(a, dt, n, block, s);
    case TOPO_FREE:
      switch (a.lanes) {
        case 2: return launch_t<TopoFree, R, 2, true, false>(a, dt, n, block, s);
        case 4: return launch_t<TopoFree, R, 4, true, false>(a, dt, n, block, s);
        case 8: return launch_t<TopoFree, R, 8, true, false>(a, dt, n, block, s);
        default: return launch_t<TopoFree, R, 1, true, false>(a, dt, n, block, s);
      }
    default: return (int)hipErrorInvalidValue;
  }
}

}  // namespace gp
