// gp_vstep_body.hip — velocity stepping with and without continuous collision
// detection (integrators.rs:42-175, dynamics.rs:366-742) for a lone floating
// body with at most 8 contact candidates (a contact point or its sphere
// against one half-space): a cube on a plane, a sphere, any rigid body with
// contact points and a sphere collider against a few half-spaces.
//
// The step is the runtime-topology kernel's (gp_vstep.hip), operation for
// operation: the free velocity without contact forces, the contacts, their
// Jacobian rows, Y = M^-1 J^T and G = J Y, g = J v_free, the impulse program of
// dynamics.rs:668-704 (minimise 1/2 x'Gx + g'x with D x in the product of
// second-order cones, D = diag(0.95, 1, 1) per contact) by the same
// primal-dual interior point (Nesterov-Todd scaling, Mehrotra's
// predictor-corrector, the Cholesky factor of G + D W^-2 D, the same start,
// stopping and best-iterate rules), v_next = v_free + Y x, and the same
// collision loop. What differs is where it runs: L lanes share an instance,
// each holding the rows of T = 8 / L candidates (R = 3 T rows of the 24-row
// program; an inactive candidate's rows are identity rows with a zero right-
// hand side, which leaves the active rows' arithmetic unchanged), every
// working array of the program in registers except G's rows, which sit in a
// lane-private LDS row (written once per collision round, read each
// iteration). The Cholesky factorisation is row-distributed: at column j the
// lane owning row j broadcasts it, every lane updates its own rows; the
// triangular solves broadcast each solved entry. Sums over rows of different
// lanes are xor butterflies (commutative adds: every lane holds the same
// bits and takes the same branches). At 8 lanes, each collision round puts
// the touching candidates on an instance's first lanes and the wave skips the
// program's columns past 3 x its largest touching count; broadcasts and sums
// are DPP moves; the instance's long-lived state waits in LDS while the
// interior point runs. M is the body's spatial inertia in its frame (the
// joint-space mass matrix of a floating joint).
#include <hip/hip_runtime.h>

#include "gp_cone.h"
#include "gp_device.h"
#include "gp_internal.h"
#include "gp_spatial.h"

#ifndef GP_DYN_LDS
extern __shared__ __align__(16) unsigned char gp_dyn_lds[];
#define GP_DYN_LDS gp_dyn_lds
#endif

namespace gp {

namespace {

constexpr int kBodyCand = kVsBodyMaxCandidates;  // candidates per instance
constexpr int kBodyRows = 3 * kBodyCand;  // rows of the padded program
// doubles of an instance's parked state: q, v (the step's start), the
// current q and v, v_free, the world rotation and translation
constexpr int kStash = 7 + 6 + 7 + 6 + 6 + 9 + 3;

// x of the lane `O` apart (O = 1, 2, 4) in the instance's 8-lane group
template <int O, class T> __device__ __forceinline__ T xor8(T x) { return lane_xor<O>(x); }
template <int CTRL, int BANKS> __device__ __forceinline__ double dpp64(double old, double x) {
  return dpp_move<CTRL, BANKS>(old, x);
}
// lane S of the group: a quad broadcast of lane S % 4 in each bank, then the
// other bank of the pair takes it
template <int S> __device__ __forceinline__ double bcast8(double x) {
  constexpr int q = S & 3, perm = q | (q << 2) | (q << 4) | (q << 6);
  const double t = dpp64<perm, 0xF>(x, x);
  if constexpr (S < 4) return dpp64<kDppRowShr4, 0xA>(t, t);
  else return dpp64<kDppRowShl4, 0x5>(t, t);
}

template <int L> __device__ __forceinline__ double gsum(double x) {
  if constexpr (L == 8) {
    x += xor8<1>(x);
    x += xor8<2>(x);
    return x + xor8<4>(x);
  } else {
#pragma unroll
    for (int o = 1; o < L; o <<= 1) x += __shfl_xor(x, o, 64);
    return x;
  }
}
template <int L> __device__ __forceinline__ double gmax(double x) {
  if constexpr (L == 8) {
    x = fmax(x, xor8<1>(x));
    x = fmax(x, xor8<2>(x));
    return fmax(x, xor8<4>(x));
  } else {
#pragma unroll
    for (int o = 1; o < L; o <<= 1) x = fmax(x, __shfl_xor(x, o, 64));
    return x;
  }
}
template <int L> __device__ __forceinline__ double gmin(double x) {
  if constexpr (L == 8) {
    x = fmin(x, xor8<1>(x));
    x = fmin(x, xor8<2>(x));
    return fmin(x, xor8<4>(x));
  } else {
#pragma unroll
    for (int o = 1; o < L; o <<= 1) x = fmin(x, __shfl_xor(x, o, 64));
    return x;
  }
}
// lane `src` of this instance's group: its value of x (src is a constant
// wherever the loops are unrolled, and the switch folds)
template <int L> __device__ __forceinline__ double bc(double x, int src) {
  if constexpr (L == 1) return x;
  else if constexpr (L == 8) {
    switch (src) {
      case 0: return bcast8<0>(x);
      case 1: return bcast8<1>(x);
      case 2: return bcast8<2>(x);
      case 3: return bcast8<3>(x);
      case 4: return bcast8<4>(x);
      case 5: return bcast8<5>(x);
      case 6: return bcast8<6>(x);
      default: return bcast8<7>(x);
    }
  } else return __shfl(x, src, L);
}
// the smallest (t, key) over the lanes, the smaller key on a tie
template <int L, int O> __device__ __forceinline__ void argmin_step(double& t, int& key) {
  double t2;
  int k2;
  if constexpr (L == 8) {
    t2 = xor8<O>(t);
    k2 = xor8<O>(key);
  } else {
    t2 = __shfl_xor(t, O, 64);
    k2 = __shfl_xor(key, O, 64);
  }
  if (t2 < t || (t2 == t && k2 < key)) { t = t2; key = k2; }
}
template <int L> __device__ __forceinline__ void gargmin(double& t, int& key) {
  if constexpr (L >= 2) argmin_step<L, 1>(t, key);
  if constexpr (L >= 4) argmin_step<L, 2>(t, key);
  if constexpr (L >= 8) argmin_step<L, 4>(t, key);
}

__device__ __forceinline__ double dmu(int row) { return row % 3 == 0 ? kConeMu : 1.0; }

// One contact candidate: a contact point (body frame) or the body's sphere
// against one half-space.
struct Cand {
  int kind;  // -1 none, 0 point, 1 sphere
  int hh;    // half-space index
  int id;    // candidate index (contact point c x half-space h: c nh + h; spheres after)
  double pb[3];
  double r;  // sphere radius
  double hpt[3], hn[3];
};

template <int L>
__device__ void vs_body(const BucketArgs& a, int tid, double dt, int ccd, int n_steps, double* lds) {
  constexpr int T = kBodyCand / L, R = 3 * T, m = kBodyRows;
  constexpr int pitch = (R * m) | 1;
  const int inst = tid / L, sub = tid % L;
  if (inst >= a.B) return;
  double* Gs = lds + (size_t)threadIdx.x * pitch;  // this lane's rows of G: Gs[i * m + k]
  const int Bp = a.Bp, nb = a.nb, nc = a.nc, nh = a.nh;
  auto bpar = [&](int f) { return a.bp[(size_t)f * nb * Bp + inst]; };
  // ---- the contact candidates (the body's parameters are re-read, from
  // L1/L2, in every collision round: registers go to the impulse program)
  const int sl = a.ns > 0 ? a.sphere_of[0] : -1;
  const bool sphere_on = sl >= 0 && a.shp[((size_t)SH_ON * a.ns + sl) * Bp + inst] != 0.0;
  const int np = nc * nh, P = np + (sphere_on ? nh : 0);
  Cand cd[T];
#pragma unroll
  for (int t = 0; t < T; ++t) {
    const int c = sub * T + t;
    Cand& C = cd[t];
    C.kind = c < np ? 0 : (c < P ? 1 : -1);
    C.hh = C.kind == 0 ? c % nh : (C.kind == 1 ? c - np : 0);
    C.id = c;
    C.r = C.kind == 1 ? a.shp[((size_t)SH_R * a.ns + sl) * Bp + inst] : 0.0;
    const int cc = C.kind == 0 ? c / nh : 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      C.pb[k] = C.kind == 0 ? a.cp[((size_t)(CF_P + k) * nc + cc) * Bp + inst]
                : C.kind == 1 ? a.shp[((size_t)(SH_C + k) * a.ns + sl) * Bp + inst] : 0.0;
      C.hpt[k] = C.kind >= 0 ? a.hp[((size_t)(HF_POINT + k) * nh + C.hh) * Bp + inst] : 0.0;
      C.hn[k] = C.kind >= 0 ? a.hp[((size_t)(HF_NORMAL + k) * nh + C.hh) * Bp + inst] : 1.0;
    }
  }
  double q[7], v[6];
#pragma unroll
  for (int k = 0; k < 7; ++k) q[k] = a.q[(size_t)k * Bp + inst];
#pragma unroll
  for (int k = 0; k < 6; ++k) v[k] = a.v[(size_t)k * Bp + inst];

  int last_rounds = 0;
  for (int step = 0; step < n_steps; ++step) {
    double qc[7], vc[6], vn[6];
#pragma unroll
    for (int k = 0; k < 7; ++k) qc[k] = q[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) { vc[k] = v[k]; vn[k] = v[k]; }
    bool ok = true, done = false;
    double t_remain = dt;
    int forced_h = -1, rounds = 0;
    for (int round = 0; round < kCcdMaxRounds && !done; ++round) {
      rounds = round + 1;
      const double h = t_remain;
      m3<double> Rt;
#pragma unroll
      for (int k = 0; k < 9; ++k) Rt.a[k] = bpar(BF_RT + k);
      const v3<double> pt{bpar(BF_PT), bpar(BF_PT + 1), bpar(BF_PT + 2)};
      const ai6<double> I = rigid(s3<double>{bpar(BF_J), bpar(BF_J + 1), bpar(BF_J + 2), bpar(BF_J + 3),
                                             bpar(BF_J + 4), bpar(BF_J + 5)},
                                  v3<double>{bpar(BF_C), bpar(BF_C + 1), bpar(BF_C + 2)}, bpar(BF_M));
      double LM[21];  // M = I in the body's twist coordinates, its Cholesky factor
      {
        double M6[6][6];
        to_dense(I, M6);
        ok = chol6(M6, LM) && ok;
      }
      // 1. free velocity, no contact forces: vdot = M^-1 (tau - v x* M v) - X a_g
      m3<double> Rw = mul(Rt, quat_mat(qc[0], qc[1], qc[2], qc[3]));
      v3<double> tw = pt + mul(Rt, v3<double>{qc[4], qc[5], qc[6]});
      double vfree[6];
      {
        const sv<double> V{{vc[0], vc[1], vc[2]}, {vc[3], vc[4], vc[5]}};
        const sv<double> pA = crf(V, mul(I, V));
        auto tau = [&](int k) { return a.tau[(size_t)k * Bp + inst]; };
        double acc[6] = {tau(0) - pA.w.x, tau(1) - pA.w.y, tau(2) - pA.w.z,
                         tau(3) - pA.v.x, tau(4) - pA.v.y, tau(5) - pA.v.z};
        chol6_solve(LM, acc);
        // the world's base acceleration +g z in the body frame: g R^T e_z
        acc[3] -= Rw.a[6] * a.g; acc[4] -= Rw.a[7] * a.g; acc[5] -= Rw.a[8] * a.g;
#pragma unroll
        for (int k = 0; k < 6; ++k) vfree[k] = vc[k] + h * acc[k];
      }
      // This round's candidates, cw: for L = 8 the group's touching ones on
      // its first lanes in candidate order, the others after them (cd keeps
      // the candidate order, so the layout depends on this round's contacts
      // only, not on how the steps are split into launches). The program's
      // active rows then come first, and the wave works on 3 x (its largest
      // touching count) = jmax rows. A candidate that does not touch has
      // identity rows, zero columns and a zero right-hand side: the columns
      // past jmax would only add and subtract exact zeros, so skipping them
      // changes no value (the sums over lanes do change order with the
      // lanes' candidates).
      Cand cw[T];
#pragma unroll
      for (int t = 0; t < T; ++t) cw[t] = cd[t];
      int mw = kBodyCand;
      if constexpr (L == 8) {
        bool touch;
        {
          const Cand& C = cd[0];
          const v3<double> pw = mul(Rw, v3<double>{C.pb[0], C.pb[1], C.pb[2]}) + tw;
          const v3<double> n{C.hn[0], C.hn[1], C.hn[2]};
          const double dist = dot(pw - v3<double>{C.hpt[0], C.hpt[1], C.hpt[2]}, n);
          touch = C.kind == 0 ? dist <= kInsideMargin : (C.kind == 1 ? (dist <= C.r || C.hh == forced_h) : false);
        }
        const unsigned long long wm = __ballot(touch);
        const unsigned gmask = (unsigned)(wm >> (8 * ((threadIdx.x & 63) >> 3))) & 0xFFu;
        const int na = __popc(gmask);
        int src = 0, ca = 0, ci = 0;  // the lane whose candidate this lane takes
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          if ((gmask >> b) & 1u) {
            if (ca == sub) src = b;
            ++ca;
          } else {
            if (na + ci == sub) src = b;
            ++ci;
          }
        }
        Cand& C = cw[0];
        C.kind = __shfl(C.kind, src, 8);
        C.hh = __shfl(C.hh, src, 8);
        C.id = __shfl(C.id, src, 8);
        C.r = __shfl(C.r, src, 8);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          C.pb[k] = __shfl(C.pb[k], src, 8);
          C.hpt[k] = __shfl(C.hpt[k], src, 8);
          C.hn[k] = __shfl(C.hn[k], src, 8);
        }
        mw = 0;
#pragma unroll
        for (int g8 = 0; g8 < 8; ++g8) mw = max(mw, __popc((unsigned)(wm >> (8 * g8)) & 0xFFu));
      }
      const int jmax = 3 * mw;
      // 2. contacts
      bool act[T];
      double nact = 0.0;
      double Y[T][6][3], g[R];
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const Cand& C = cw[t];
        const v3<double> pw = mul(Rw, v3<double>{C.pb[0], C.pb[1], C.pb[2]}) + tw;
        const v3<double> n{C.hn[0], C.hn[1], C.hn[2]};
        const double dist = dot(pw - v3<double>{C.hpt[0], C.hpt[1], C.hpt[2]}, n);
        act[t] = C.kind == 0 ? dist <= kInsideMargin : (C.kind == 1 ? (dist <= C.r || C.hh == forced_h) : false);
        nact += act[t] ? 1.0 : 0.0;
        // 3. Jacobian rows [n; t; b] . (velocity of the contact point) in the
        // body's twist coordinates (angular R^T (rel x d), linear R^T d), then
        // Y = M^-1 J^T and g = J v_free
        const v3<double> p = C.kind == 1 ? pw - dist * n : pw;
        v3<double> ta = cross(n, v3<double>{1.0, 0.0, 0.0});
        double tn = sqrt(dot(ta, ta));
        if (tn == 0.0) {
          ta = cross(n, v3<double>{0.0, 1.0, 0.0});
          tn = sqrt(dot(ta, ta));
        }
        ta = v3<double>{ta.x / tn, ta.y / tn, ta.z / tn};
        v3<double> tb = cross(n, ta);
        const double bn = sqrt(dot(tb, tb));
        tb = v3<double>{tb.x / bn, tb.y / bn, tb.z / bn};
        const v3<double> rel = p - tw;
        const v3<double> dirs[3] = {n, ta, tb};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          const v3<double> ang = mulT(Rw, cross(rel, dirs[r])), lin = mulT(Rw, dirs[r]);
          double Jr[6] = {ang.x, ang.y, ang.z, lin.x, lin.y, lin.z};
          double s = 0.0;
#pragma unroll
          for (int k = 0; k < 6; ++k) s += Jr[k] * vfree[k];
          g[3 * t + r] = act[t] ? s : 0.0;
          chol6_solve(LM, Jr);
#pragma unroll
          for (int k = 0; k < 6; ++k) Y[t][k][r] = act[t] ? Jr[k] : 0.0;
        }
      }
      const double nct = gsum<L>(nact);
      if (nct == 0.0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) vn[k] = vfree[k];
      } else {
        // G = J Y, this lane's rows, into its LDS row (an inactive
        // candidate's rows and columns are zero)
        {
          double Jrow[R][6];
#pragma unroll
          for (int t = 0; t < T; ++t) {
            // J_c = (M Y_c)^T: recomputed from Y so J need not stay live
#pragma unroll
            for (int r = 0; r < 3; ++r) {
              const sv<double> y{{Y[t][0][r], Y[t][1][r], Y[t][2][r]}, {Y[t][3][r], Y[t][4][r], Y[t][5][r]}};
              const sv<double> jr = mul(I, y);
              Jrow[3 * t + r][0] = jr.w.x; Jrow[3 * t + r][1] = jr.w.y; Jrow[3 * t + r][2] = jr.w.z;
              Jrow[3 * t + r][3] = jr.v.x; Jrow[3 * t + r][4] = jr.v.y; Jrow[3 * t + r][5] = jr.v.z;
            }
          }
#pragma unroll
          for (int c2 = 0; c2 < kBodyCand; ++c2) {
            if (c2 >= mw) continue;  // columns past jmax are never read
            double Yc[6][3];
#pragma unroll
            for (int k = 0; k < 6; ++k)
#pragma unroll
              for (int r = 0; r < 3; ++r) Yc[k][r] = bc<L>(Y[c2 % T][k][r], c2 / T);
#pragma unroll
            for (int i = 0; i < R; ++i)
#pragma unroll
              for (int r = 0; r < 3; ++r) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < 6; ++k) s += Jrow[i][k] * Yc[k][r];
                Gs[i * m + 3 * c2 + r] = s;
              }
          }
        }
        // The instance's state that outlives the interior point waits in LDS
        // meanwhile, its registers free for the program: every lane of the
        // instance holds the same bits, its first lane writes, all read back.
        double* const stash = lds + (size_t)blockDim.x * pitch + (size_t)(threadIdx.x / L) * kStash;
        auto park = [&](bool back) {
          auto mv = [&](double& x, int k) {
            if (back) x = stash[k];
            else if (sub == 0) stash[k] = x;
          };
#pragma unroll
          for (int k = 0; k < 7; ++k) { mv(q[k], k); mv(qc[k], 7 + k); }
#pragma unroll
          for (int k = 0; k < 6; ++k) { mv(v[k], 14 + k); mv(vc[k], 20 + k); mv(vfree[k], 26 + k); }
#pragma unroll
          for (int k = 0; k < 9; ++k) mv(Rw.a[k], 32 + k);
          mv(tw.x, 41); mv(tw.y, 42); mv(tw.z, 43);
          asm volatile("" ::: "memory");
        };
        park(false);
        // 5. the impulse program (the runtime-topology kernel's cone_program)
        // dinv: 1 / (this lane's rows' factor diagonal), one division per
        // column of the factorisation and none in the solves
        double x[R], z[R], xb[R], rd[R], dx[R], dz[R], xi[R], lt[R], W[T][9], Wi[T][9], Lc[R][m], xall[m], dinv[R];
#pragma unroll
        for (int i = 0; i < R; ++i) dinv[i] = 1.0;
#pragma unroll
        for (int k = 0; k < m; ++k) xall[k] = 0.0;  // entries past jmax stay zero
        double gl = 0.0;
#pragma unroll
        for (int t = 0; t < T; ++t)
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const int i = 3 * t + r;
            x[i] = act[t] && r == 0 ? 1.0 / kConeMu : 0.0;
            z[i] = act[t] && r == 0 ? 1.0 : 0.0;
            xb[i] = x[i];
            gl = fmax(gl, fabs(g[i]));
          }
        const double gnorm = gmax<L>(gl);
        double best = 1e300;
        int since_best = 0;
        auto gather = [&](const double (&loc)[R], double (&all)[m]) {
#pragma unroll
          for (int k = 0; k < m; ++k) {
            if (k >= jmax) continue;
            all[k] = bc<L>(loc[k % R], k / R);
          }
        };
        // Newton direction for a right-hand side xi of the scaled complementarity
        auto solve = [&]() {
          double rhs[R];
#pragma unroll
          for (int t = 0; t < T; ++t)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              double s = 0.0;
#pragma unroll
              for (int j = 0; j < 3; ++j) s += Wi[t][3 * i + j] * xi[3 * t + j];
              rhs[3 * t + i] = act[t] ? -rd[3 * t + i] + dmu(i) * s : 0.0;
            }
          // L y = rhs: the owner of row j forms its numerator, everyone learns y_j
#pragma unroll
          for (int j = 0; j < m; ++j) {
            if (j >= jmax) continue;
            double s = rhs[j % R];
#pragma unroll
            for (int k = 0; k < j; ++k) s -= Lc[j % R][k] * xall[k];
            xall[j] = bc<L>(s * dinv[j % R], j / R);  // only the owner's product is kept
          }
          // L^T dx = y: column j of L is spread over the rows' owners
#pragma unroll
          for (int i = 0; i < R; ++i) dx[i] = 0.0;
#pragma unroll
          for (int j = m - 1; j >= 0; --j) {
            if (j >= jmax) continue;
            double part = 0.0;
#pragma unroll
            for (int i = 0; i < R; ++i)
              if (sub * R + i > j) part += Lc[i][j] * dx[i];
            const double num = xall[j] - gsum<L>(part);
            xall[j] = bc<L>(num * dinv[j % R], j / R);
            if (j / R == sub) dx[j % R] = xall[j];
          }
#pragma unroll
          for (int t = 0; t < T; ++t) {
            double a3[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              a3[i] = 0.0;
#pragma unroll
              for (int j = 0; j < 3; ++j) a3[i] += Wi[t][3 * i + j] * dmu(j) * dx[3 * t + j];
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              double s = 0.0;
#pragma unroll
              for (int j = 0; j < 3; ++j) s += Wi[t][3 * i + j] * (xi[3 * t + j] - a3[j]);
              dz[3 * t + i] = act[t] ? s : 0.0;
            }
          }
        };
        auto max_step = [&](double scale) {
          double alpha = 1.0;
#pragma unroll
          for (int t = 0; t < T; ++t) {
            if (!act[t]) continue;
            const double sc[3] = {dmu(0) * x[3 * t], x[3 * t + 1], x[3 * t + 2]};
            const double dsc[3] = {dmu(0) * dx[3 * t], dx[3 * t + 1], dx[3 * t + 2]};
            alpha = fmin(alpha, scale * fmin(soc_max_step(sc, dsc), soc_max_step(&z[3 * t], &dz[3 * t])));
          }
          return gmin<L>(alpha);
        };
        for (int it = 0; it < kConeMaxIter; ++it) {
          // G is read from LDS in each iteration: without this compiler-only
          // barrier the compiler keeps the rows it wrote in registers, and
          // spills them
          asm volatile("" ::: "memory");
          // rd = G x + g - D z
          gather(x, xall);
          double gap_l = 0.0, rmax_l = 0.0, xmax_l = 0.0;
#pragma unroll
          for (int i = 0; i < R; ++i) {
            double s = g[i] - dmu(i) * z[i];
#pragma unroll
            for (int k = 0; k < m; ++k) {
              if (k >= jmax) continue;
              s += Gs[i * m + k] * xall[k];
            }
            rd[i] = s;
            if (act[i / 3]) {
              rmax_l = fmax(rmax_l, fabs(s));
              gap_l += dmu(i) * x[i] * z[i];
              xmax_l = fmax(xmax_l, fabs(x[i]));
            }
          }
          const double gap = gsum<L>(gap_l), rmax = gmax<L>(rmax_l), xmax = gmax<L>(xmax_l);
          if (!isfinite(gap) || !isfinite(rmax)) break;
          const double gap_rel = gap / ((1.0 + gnorm) * (1.0 + xmax)), res_rel = rmax / (1.0 + gnorm);
          const double merit = fmax(gap_rel, res_rel);
          if (merit < best) {
            best = merit;
            since_best = 0;
#pragma unroll
            for (int i = 0; i < R; ++i) xb[i] = x[i];
          }
          if (gap_rel <= kConeGapTol && res_rel <= kConeResTol) break;
          if (++since_best > kConeStall) break;
          const double mug = gap / nct;
          // scalings and K = G + D W^-2 D (inactive rows: identity)
#pragma unroll
          for (int t = 0; t < T; ++t) {
            if (act[t]) {
              const double sc[3] = {dmu(0) * x[3 * t], x[3 * t + 1], x[3 * t + 2]};
              nt_scaling<true>(sc, &z[3 * t], W[t], Wi[t]);
            } else {
#pragma unroll
              for (int k = 0; k < 9; ++k) { W[t][k] = 0.0; Wi[t][k] = 0.0; }
            }
#pragma unroll
            for (int i = 0; i < 3; ++i)
              lt[3 * t + i] = W[t][3 * i] * z[3 * t] + W[t][3 * i + 1] * z[3 * t + 1] + W[t][3 * i + 2] * z[3 * t + 2];
          }
#pragma unroll
          for (int i = 0; i < R; ++i) {
            const int gi = sub * R + i;
            const bool ai = act[i / 3];
#pragma unroll
            for (int k = 0; k < m; ++k) Lc[i][k] = ai ? (k <= gi ? Gs[i * m + k] : 0.0) : (k == gi ? 1.0 : 0.0);
          }
          // + D W^-2 D on the active candidates' diagonal blocks (global
          // columns 3 c + j of candidate c = sub T + t: every block position
          // is visited, the lane's own one selected)
#pragma unroll
          for (int t = 0; t < T; ++t) {
            if (!act[t]) continue;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
              for (int j = 0; j <= i; ++j) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < 3; ++k) s += Wi[t][3 * i + k] * Wi[t][3 * k + j];
                const double e = dmu(i) * s * dmu(j);
#pragma unroll
                for (int cb = 0; cb < kBodyCand; ++cb)
                  if (cb == sub * T + t) Lc[3 * t + i][3 * cb + j] += e;
              }
          }
          // Cholesky, lower triangle in place: at column j the owner of row j
          // broadcasts its entries, every lane updates its rows below it
          bool pd = true;
#pragma unroll
          for (int j = 0; j < m; ++j) {
            if (j >= jmax) continue;  // identity rows past jmax: already their factor
            double sjj = bc<L>(Lc[j % R][j], j / R);
            double u[R];
#pragma unroll
            for (int i = 0; i < R; ++i) u[i] = Lc[i][j];
#pragma unroll
            for (int k = 0; k < j; ++k) {
              const double ljk = bc<L>(Lc[j % R][k], j / R);
              sjj -= ljk * ljk;
#pragma unroll
              for (int i = 0; i < R; ++i) u[i] -= Lc[i][k] * ljk;
            }
            pd = pd && sjj > 0.0;
            const double ljj = sqrt(sjj), rj = 1.0 / ljj;
#pragma unroll
            for (int i = 0; i < R; ++i) {
              const int gi = sub * R + i;
              if (gi == j) {
                Lc[i][j] = ljj;
                dinv[i] = rj;
              } else if (gi > j) {
                Lc[i][j] = u[i] * rj;
              }
            }
          }
          if (!pd) break;
          // predictor
#pragma unroll
          for (int i = 0; i < R; ++i) xi[i] = -lt[i];
          solve();
          const double alpha_a = max_step(1.0);
          double sz_l = 0.0;
#pragma unroll
          for (int i = 0; i < R; ++i)
            if (act[i / 3]) sz_l += (dmu(i) * x[i] + alpha_a * dmu(i) * dx[i]) * (z[i] + alpha_a * dz[i]);
          const double ratio = gsum<L>(sz_l) / gap, sigma = ratio * ratio * ratio;
          // corrector: xi = lt \ (sigma mu e - lt o lt - (W^-1 ds) o (W dz))
#pragma unroll
          for (int t = 0; t < T; ++t) {
            if (!act[t]) {
#pragma unroll
              for (int i = 0; i < 3; ++i) xi[3 * t + i] = 0.0;
              continue;
            }
            const double ds[3] = {dmu(0) * dx[3 * t], dx[3 * t + 1], dx[3 * t + 2]};
            double a3[3], b3[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              a3[i] = Wi[t][3 * i] * ds[0] + Wi[t][3 * i + 1] * ds[1] + Wi[t][3 * i + 2] * ds[2];
              b3[i] = W[t][3 * i] * dz[3 * t] + W[t][3 * i + 1] * dz[3 * t + 1] + W[t][3 * i + 2] * dz[3 * t + 2];
            }
            const double l0 = lt[3 * t], l1 = lt[3 * t + 1], l2 = lt[3 * t + 2];
            const double w0 = sigma * mug - (l0 * l0 + l1 * l1 + l2 * l2) - (a3[0] * b3[0] + a3[1] * b3[1] + a3[2] * b3[2]);
            const double w1 = -(2.0 * l0 * l1) - (a3[0] * b3[1] + b3[0] * a3[1]);
            const double w2 = -(2.0 * l0 * l2) - (a3[0] * b3[2] + b3[0] * a3[2]);
            const double y0 = (l0 * w0 - l1 * w1 - l2 * w2) / soc_det(l0, l1, l2);
            xi[3 * t] = y0;
            xi[3 * t + 1] = (w1 - y0 * l1) / l0;
            xi[3 * t + 2] = (w2 - y0 * l2) / l0;
          }
          solve();
          const double alpha = max_step(0.99);
#pragma unroll
          for (int i = 0; i < R; ++i)
            if (act[i / 3]) {
              x[i] += alpha * dx[i];
              z[i] += alpha * dz[i];
            }
        }
        double bad = best < 1e300 ? 0.0 : 1.0;
#pragma unroll
        for (int i = 0; i < R; ++i) {
          x[i] = xb[i];
          if (act[i / 3] && !isfinite(x[i])) bad = 1.0;
        }
        ok = ok && gmax<L>(bad) == 0.0;
        park(true);
        // 6. v_next = v_free + Y x
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          double s = 0.0;
#pragma unroll
          for (int t = 0; t < T; ++t)
#pragma unroll
            for (int r = 0; r < 3; ++r) s += Y[t][k][r] * x[3 * t + r];
          vn[k] = vfree[k] + gsum<L>(s);
        }
      }
      // the collision loop: the earliest time a separated contact point, or
      // the sphere of a body without contacts, reaches a half-space at the
      // new velocities (integrators.rs:62-158)
      double earliest = 1e300;
      int key = 0x7fffffff;
      if (ccd) {
        const v3<double> w = mul(Rw, v3<double>{vn[0], vn[1], vn[2]});
        const v3<double> v0 = mul(Rw, v3<double>{vn[3], vn[4], vn[5]}) + cross(tw, w);
#pragma unroll
        for (int t = 0; t < T; ++t) {
          const Cand& C = cd[t];
          if (C.kind < 0 || (C.kind == 1 && nct != 0.0)) continue;
          const v3<double> p = mul(Rw, v3<double>{C.pb[0], C.pb[1], C.pb[2]}) + tw;
          const v3<double> n{C.hn[0], C.hn[1], C.hn[2]};
          const double vdn = dot(v0 + cross(w, p), n);
          if (vdn == 0.0) continue;
          const double dist = dot(p - v3<double>{C.hpt[0], C.hpt[1], C.hpt[2]}, n) - C.r;
          if (dist <= 0.0) continue;
          const double tt = dist / -vdn;
          if (tt > 0.0 && tt < earliest) {
            earliest = tt;
            key = C.id;
          }
        }
        gargmin<L>(earliest, key);
      }
      if (earliest < t_remain) {
        advance_floating(qc, vn, earliest, qc);
#pragma unroll
        for (int k = 0; k < 6; ++k) vc[k] = vn[k];
        t_remain -= earliest;
        forced_h = key >= np ? key - np : -1;  // a sphere's hit: that half-space touches next round
      } else {
        advance_floating(qc, vn, t_remain, qc);
        done = true;
      }
    }
    ok = ok && done;  // the collision loop gave up: a fault, the state still advances
    if (!done) advance_floating(qc, vn, t_remain, qc);
    // the state, vdot and the rounds leave the registers once, after the
    // launch's last step (a fault flag when it happens)
    bool fin = true;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double vd = (vn[k] - v[k]) / dt;
      if (sub == 0 && step + 1 == n_steps) a.vdot[(size_t)k * Bp + inst] = vd;
      v[k] = vn[k];
      fin = fin && isfinite(v[k]);
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      q[k] = qc[k];
      fin = fin && isfinite(q[k]);
    }
    if (sub == 0 && !(ok && fin)) a.fault[inst] = 1;
    last_rounds = rounds;
  }
  if (sub == 0) {
#pragma unroll
    for (int k = 0; k < 7; ++k) a.q[(size_t)k * Bp + inst] = q[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) a.v[(size_t)k * Bp + inst] = v[k];
    a.rounds[inst] = last_rounds;
  }
}

template <int L> constexpr int body_lds_pitch() { return ((3 * kBodyCand / L) * kBodyRows) | 1; }

template <int L>
__global__ void __launch_bounds__(256) vs_body_kernel(BucketArgs a, double dt, int ccd, int n_steps) {
  vs_body<L>(a, blockIdx.x * blockDim.x + threadIdx.x, dt, ccd, n_steps, (double*)GP_DYN_LDS);
}

template <int L>
int launch_l(const BucketArgs& a, double dt, int ccd, int n_steps, int block, hipStream_t s) {
  const long threads = (long)a.B * L;
  dim3 blk(block), grid((unsigned)((threads + block - 1) / block));
  const unsigned lds = (unsigned)((body_lds_pitch<L>() * block + block / L * kStash) * sizeof(double));
  hipLaunchKernelGGL((vs_body_kernel<L>), grid, blk, lds, s, a, dt, ccd, n_steps);
  return (int)hipGetLastError();
}

}  // namespace

int launch_vs_body(const BucketArgs& a, double dt, int ccd, int n_steps, int lanes, int candidates, int block,
                   void* stream) {
  if (a.B == 0 || n_steps <= 0) return 0;
  if (candidates > kBodyCand) return (int)hipErrorInvalidValue;
  // each lane's LDS row is 584 bytes at 8 lanes per body: at most 256 lanes
  // per workgroup; whole waves only
  if (block <= 0) block = kVsBodyBlock;
  if (block > 256 || block % 64 != 0) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream;
  // one lane per body: its LDS row holds all 24 rows of G (4.6 KB), so 16
  // lanes per workgroup (74 KB)
  if (lanes == 1) return launch_l<1>(a, dt, ccd, n_steps, 16, s);
  return launch_l<8>(a, dt, ccd, n_steps, block, s);
}

}  // namespace gp
