// gp_tree_impl.h — trees of many bodies, one instance per L lanes (L = 16, 32
// or 64, at least the body count), lane j of an instance holding body j and
// its joint. The articulated-body algorithm (RBDA Table 7.1) runs its three
// sweeps level by level: every body of one depth at once, a body's parent
// quantities (velocity, world pose, acceleration) taken from the parent's
// lane by a cross-lane shuffle on the way out, and on the way in each parent
// summing its children's articulated inertias and bias forces by reading them
// from the children's lanes. The lanes exchange through LDS: each lane has a
// slot of kTreeSlot doubles, written by the lane when its level is done and
// read by its children (pass 1 and 3) or its parent (pass 2) one level later;
// all lanes of a segment are in one wave, whose LDS instructions execute in
// order, so no barrier is needed, only a compiler fence between levels. Runtime
// topology: any tree of up to 64 bodies with at most kTreeMaxChildren
// children per body, every joint type, contact points, prismatic springs;
// semi-implicit Euler (integrators.rs:276-319), n_steps steps per launch with
// each lane's joint state, parameters and per-body quantities in registers.
// The joint accelerations are those of the reference's dynamics_continuous
// (dynamics.rs:322-364). R = double: FP64 throughout; R = float: FP32 sweeps,
// FP64 joint-space solves, contact distances and q/v (as the other mixed
// kernels).
// Included by gp_fast_impl.h (it uses Soa, pick and joint_frame from there).
#pragma once

namespace gp {

constexpr int kTreeMaxChildren = 8;
// doubles per lane in the exchange slot (an articulated inertia and a force,
// 21 + 6; odd, so a wave's 64-bit accesses spread over the banks)
constexpr int kTreeSlot = 27;
constexpr int kTreeSlotBytes = kTreeSlot * 8;

// slot <-> spatial vector / articulated inertia, in the sweeps' precision R
// from the slot's start (the slot is a row of doubles)
template <class R> __device__ __forceinline__ void put_sv(R* p, const sv<R>& x) {
  p[0] = x.w.x; p[1] = x.w.y; p[2] = x.w.z; p[3] = x.v.x; p[4] = x.v.y; p[5] = x.v.z;
}
template <class R> __device__ __forceinline__ sv<R> get_sv(const R* p) {
  return {{p[0], p[1], p[2]}, {p[3], p[4], p[5]}};
}
template <class R> __device__ __forceinline__ void put_ai6(R* p, const ai6<R>& I) {
  p[0] = I.A.xx; p[1] = I.A.yy; p[2] = I.A.zz; p[3] = I.A.xy; p[4] = I.A.xz; p[5] = I.A.yz;
#pragma unroll
  for (int k = 0; k < 9; ++k) p[6 + k] = I.B.a[k];
  p[15] = I.C.xx; p[16] = I.C.yy; p[17] = I.C.zz; p[18] = I.C.xy; p[19] = I.C.xz; p[20] = I.C.yz;
}
template <class R> __device__ __forceinline__ ai6<R> get_ai6(const R* p) {
  ai6<R> o;
  o.A = s3<R>{p[0], p[1], p[2], p[3], p[4], p[5]};
#pragma unroll
  for (int k = 0; k < 9; ++k) o.B.a[k] = p[6 + k];
  o.C = s3<R>{p[15], p[16], p[17], p[18], p[19], p[20]};
  return o;
}
// a compiler-only barrier between levels: slot reads are not moved above the
// writes of the level before (the hardware keeps one wave's LDS order)
__device__ __forceinline__ void level_fence() { asm volatile("" ::: "memory"); }
template <int L> __device__ __forceinline__ int seg_max(int x) {
#pragma unroll
  for (int off = 1; off < L; off <<= 1) x = max(x, __shfl_xor(x, off, L));
  return x;
}

// Whether a bucket's topology runs on this kernel: more bodies than the
// compiled chains, at most 64, no body with more than kTreeMaxChildren children.
inline bool tree_topology_ok(int nb, const int* parent) {
  if (nb < 13 || nb > 64) return false;
  for (int i = 0; i < nb; ++i) {
    int c = 0;
    for (int k = i + 1; k < nb; ++k) c += parent[k] == i ? 1 : 0;
    if (c > kTreeMaxChildren) return false;
  }
  return true;
}
inline int tree_lanes(int nb) { return nb <= 16 ? 16 : nb <= 32 ? 32 : 64; }

template <class R, int L>
__device__ __forceinline__ void wave_tree_body(const BucketArgs& a, int tid, double dt, int n_steps) {
  const int inst = tid / L, j = tid % L;
  if (inst >= a.B) return;  // the whole L-lane segment leaves together
  const int nb = a.nb, Bp = a.Bp;
  int my_steps = n_steps;
  if (a.active) {  // serving: the instance's remaining horizon
    const int left = a.active[inst];
    if (left <= 0) return;
    my_steps = min(my_steps, left);
    if (j == 0) a.active[inst] = left - my_steps;
  }
  const bool live = j < nb;
  // ---- topology of this lane's body
  const int par = live ? a.parent[j] : -1;
  // exchange slots: the segment's (its lane 0 is block lane threadIdx.x - j,
  // blockDim being a multiple of L) and this lane's
  double* const seg = (double*)GP_DYN_LDS + (size_t)((int)threadIdx.x - j) * kTreeSlot;
  double* const mine = seg + j * kTreeSlot;
  R* const mineR = (R*)mine;
  int depth = -1;
  if (live) {
    depth = 0;
    for (int k = par; k >= 0; k = a.parent[k]) ++depth;
  }
  const int maxd = seg_max<L>(depth);
  int child[kTreeMaxChildren];
  int nch = 0;
#pragma unroll
  for (int k = 0; k < kTreeMaxChildren; ++k) child[k] = j;
  if (live)
    for (int k = j + 1; k < nb; ++k)
      if (a.parent[k] == j && nch < kTreeMaxChildren) child[nch++] = k;
  const int maxc = seg_max<L>(nch);
  const Soa<int> kindS(a.kind, nb, Bp);
  const int kind = live ? kindS.ld(j, inst) : K_FIXED;
  const int qo = live ? a.qoff[j] : 0, vo = live ? a.voff[j] : 0;
  const int nqj = kind == K_FLOAT ? 7 : (kind == K_FIXED ? 0 : 1);
  const int nvj = kind == K_FLOAT ? 6 : (kind == K_FIXED ? 0 : 1);
  // ---- parameters and state of this lane's body and joint, for the launch
  const Soa<double> bpS(a.bp, NBF * nb, Bp);
  const Soa<R> bpR(pick<R>(a.bp, a.bpf), NBF * nb, Bp);
  const int jb = live ? j : 0;
  m3<double> Rt;
  v3<double> pt;
#pragma unroll
  for (int k = 0; k < 9; ++k) Rt.a[k] = bpS.ld((BF_RT + k) * nb + jb, inst);
  pt = v3<double>{bpS.ld(BF_PT * nb + jb, inst), bpS.ld((BF_PT + 1) * nb + jb, inst), bpS.ld((BF_PT + 2) * nb + jb, inst)};
  const m3<R> RtR = cvt<R>(Rt);
  const v3<R> ptR = cvt<R>(pt);
  const s3<R> J{bpR.ld(BF_J * nb + jb, inst), bpR.ld((BF_J + 1) * nb + jb, inst), bpR.ld((BF_J + 2) * nb + jb, inst),
                bpR.ld((BF_J + 3) * nb + jb, inst), bpR.ld((BF_J + 4) * nb + jb, inst),
                bpR.ld((BF_J + 5) * nb + jb, inst)};
  const v3<R> mc{bpR.ld(BF_C * nb + jb, inst), bpR.ld((BF_C + 1) * nb + jb, inst), bpR.ld((BF_C + 2) * nb + jb, inst)};
  const R mass = bpR.ld(BF_M * nb + jb, inst);
  const double sk = kind == K_PRI ? bpS.ld(BF_SK * nb + jb, inst) : 0.0;
  const double sl = kind == K_PRI ? bpS.ld(BF_SL * nb + jb, inst) : 0.0;
  const Soa<double> qS(a.q, a.nq, Bp), vS(a.v, a.nv, Bp), tauS(a.tau, a.nv, Bp), vdotS(a.vdot, a.nv, Bp);
  const Soa<double> cpS(a.cp, NCF * a.nc, Bp), hpS(a.hp, NHF * a.nh, Bp);
  double q[7], v[6], tau[6], vd[6];
#pragma unroll
  for (int k = 0; k < 7; ++k) q[k] = k < nqj ? qS.ldx(0, inst, qo + k) : 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    v[k] = k < nvj ? vS.ldx(0, inst, vo + k) : 0.0;
    tau[k] = k < nvj ? tauS.ldx(0, inst, vo + k) : 0.0;
    vd[k] = 0.0;
  }
  const bool contacts = a.nc > 0;
  const int c0 = live ? a.cstart[j] : 0, c1 = live ? a.cstart[j + 1] : 0;
  bool ok = true;
  const ai6<R> I = rigid(J, mc, mass);
  for (int s = 0; s < my_steps; ++s) {
    // ---- this joint's transform (E, r: parent -> body) and velocity S qd
    m3<R> E;
    v3<R> r;
    m3<double> Rm_d;  // body -> parent rotation and translation, FP64 (contacts, floating pose)
    v3<double> p_d;
    sv<R> vj{{R(0), R(0), R(0)}, {R(0), R(0), R(0)}};
    if (kind == K_FLOAT) {
      Rm_d = mul(Rt, quat_mat(q[0], q[1], q[2], q[3]));
      p_d = pt + mul(Rt, v3<double>{q[4], q[5], q[6]});
      const m3<R> Rm = cvt<R>(Rm_d);
      E = m3<R>{{Rm.a[0], Rm.a[3], Rm.a[6], Rm.a[1], Rm.a[4], Rm.a[7], Rm.a[2], Rm.a[5], Rm.a[8]}};
      r = cvt<R>(p_d);
      vj = sv<R>{{R(v[0]), R(v[1]), R(v[2])}, {R(v[3]), R(v[4]), R(v[5])}};
    } else {
      R cs = R(1), sn = R(0), dd = R(0);
      if (kind == K_REV) sincos_q(q[0], &sn, &cs);
      if (kind == K_PRI) dd = R(q[0]);
      joint_frame(RtR, ptR, cs, sn, dd, E, r);
      if (kind == K_REV) vj.w.z = R(v[0]);
      if (kind == K_PRI) vj.v.z = R(v[0]);
      if (contacts) {
        const m3<double> Ed = cvt<double>(E);
        Rm_d = m3<double>{{Ed.a[0], Ed.a[3], Ed.a[6], Ed.a[1], Ed.a[4], Ed.a[7], Ed.a[2], Ed.a[5], Ed.a[8]}};
        p_d = cvt<double>(r);
      }
    }
    // ---- pass 1, root to leaves by depth: velocities (and world poses);
    // slot: velocity (R) from the start, world rotation doubles [6, 15),
    // world translation doubles [15, 18)
    sv<R> vel = vj;
    m3<double> Rw = Rm_d;
    v3<double> tw = p_d;
    auto put_pose = [&]() {
      put_sv(mineR, vel);
      if (contacts) {
#pragma unroll
        for (int k = 0; k < 9; ++k) mine[6 + k] = Rw.a[k];
        mine[15] = tw.x; mine[16] = tw.y; mine[17] = tw.z;
      }
    };
    level_fence();  // the previous step's pass-3 reads come first
    if (depth == 0) put_pose();
    for (int d = 1; d <= maxd; ++d) {
      level_fence();
      if (depth == d) {
        const double* ps = seg + par * kTreeSlot;
        vel = xmotion(E, r, get_sv<R>((const R*)ps)) + vj;
        if (contacts) {
          m3<double> pR;
#pragma unroll
          for (int k = 0; k < 9; ++k) pR.a[k] = ps[6 + k];
          tw = v3<double>{ps[15], ps[16], ps[17]} + mul(pR, p_d);
          Rw = mul(pR, Rm_d);
        }
        if (d < maxd) put_pose();
      }
    }
    // velocity-product acceleration, rigid-body bias force, contact forces
    const sv<R> cb = crm(vel, vj);
    sv<R> pA = crf(vel, mul(I, vel));
    if (contacts && c0 < c1) {
      const m3<R> RwR = cvt<R>(Rw);
      sv<R> fe{{R(0), R(0), R(0)}, {R(0), R(0), R(0)}};
      for (int h = 0; h < a.nh; ++h) {
        auto hpar = [&](int fld) { return hpS.ld(fld * a.nh + h, inst); };
        const v3<double> n{hpar(HF_NORMAL), hpar(HF_NORMAL + 1), hpar(HF_NORMAL + 2)};
        const v3<double> nb_ = mulT(Rw, n);
        const double off = dot(tw - v3<double>{hpar(HF_POINT), hpar(HF_POINT + 1), hpar(HF_POINT + 2)}, n);
        const R alpha = R(hpar(HF_ALPHA)), mu = R(hpar(HF_MU));
        for (int cc = c0; cc < c1; ++cc) {
          auto cpar = [&](int fld) { return cpS.ld(fld * a.nc + cc, inst); };
          const v3<double> pb{cpar(CF_P), cpar(CF_P + 1), cpar(CF_P + 2)};
          const double dist = dot(nb_, pb) + off;
          if (dist <= kInsideMargin) {
            const v3<R> pbR = cvt<R>(pb);
            const v3<R> vw = mul(RwR, vel.v + cross(vel.w, pbR));
            const v3<R> F = contact_force(R(-dist), cvt<R>(n), vw, R(cpar(CF_K)), R(kHalfspaceK), alpha, mu);
            const v3<R> Fb = mulT(RwR, F);
            fe.w = fe.w + cross(pbR, Fb);
            fe.v = fe.v + Fb;
          }
        }
      }
      pA = pA - fe;
    }
    // ---- pass 2, leaves to root by depth: articulated inertias and bias forces
    ai6<R> IA = I;
    sv<R> U{{R(0), R(0), R(0)}, {R(0), R(0), R(0)}};
    double Dinv = 0.0, u = 0.0, u6[6] = {0, 0, 0, 0, 0, 0};
    for (int d = maxd; d >= 0; --d) {
      ai6<R> cI{};
      sv<R> cf{{R(0), R(0), R(0)}, {R(0), R(0), R(0)}};
      if (depth == d) {
        ai6<R> Ia = IA;
        sv<R> pa = pA + mul(IA, cb);
        if (kind == K_REV || kind == K_PRI) {
          const bool rev = kind == K_REV;
          U = rev ? col(IA, 2) : col(IA, 5);
          Dinv = 1.0 / double(rev ? IA.A.zz : IA.C.zz);
          const double tq = tau[0] - sk * (q[0] - sl);  // prismatic spring (dynamics.rs:298-315); 0 otherwise
          u = tq - double(rev ? pA.w.z : pA.v.z);
          ok = ok && (Dinv > 0.0);
          const R inv = R(Dinv);
          const sv<R> uD{inv * U.w, inv * U.v};
          Ia = rank1_down(IA, U, uD);
          const R uR = R(u);
          pa = pA + mul(Ia, cb) + sv<R>{uR * uD.w, uR * uD.v};
        } else if (kind == K_FLOAT) {
          u6[0] = tau[0] - double(pA.w.x); u6[1] = tau[1] - double(pA.w.y); u6[2] = tau[2] - double(pA.w.z);
          u6[3] = tau[3] - double(pA.v.x); u6[4] = tau[4] - double(pA.v.y); u6[5] = tau[5] - double(pA.v.z);
          Ia = ai6<R>{};  // I^A - I^A (I^A)^-1 I^A = 0
          pa = pA + sv<R>{{R(u6[0]), R(u6[1]), R(u6[2])}, {R(u6[3]), R(u6[4]), R(u6[5])}};
        }
        cI = xtix(E, r, Ia);
        cf = xtforce(E, r, pa);
        if (d > 0) {  // slot (R): articulated inertia [0, 21), bias force [21, 27)
          put_ai6(mineR, cI);
          put_sv(mineR + 21, cf);
        }
      }
      if (d == 0) break;
      // parents (depth d - 1) take their children's parts, in child order
      level_fence();
      if (depth == d - 1) {
#pragma unroll
        for (int k = 0; k < kTreeMaxChildren; ++k) {
          if (k < nch) {
            const R* cs = (const R*)(seg + child[k] * kTreeSlot);
            add_to(IA, get_ai6<R>(cs));
            pA = pA + get_sv<R>(cs + 21);
          }
        }
      }
      level_fence();  // the slots read here are written again only in pass 3
    }
    // ---- pass 3, root to leaves by depth: accelerations, then the joint's step
    sv<R> acc{{R(0), R(0), R(0)}, {R(0), R(0), R(a.g)}};  // gravity as the base's acceleration
    for (int d = 0; d <= maxd; ++d) {
      level_fence();
      if (depth == d) {
        const sv<R> pacc = par < 0 ? sv<R>{{R(0), R(0), R(0)}, {R(0), R(0), R(a.g)}} : get_sv<R>((const R*)(seg + par * kTreeSlot));
        sv<R> ap = xmotion(E, r, pacc) + cb;
        if (kind == K_REV || kind == K_PRI) {
          const double qdd = (u - double(dot6(U, ap))) * Dinv;
          vd[0] = qdd;
          if (kind == K_REV) ap.w.z += R(qdd);
          else ap.v.z += R(qdd);
        } else if (kind == K_FLOAT) {
          double D6[6][6], Lf[21];
          to_dense(IA, D6);
          ok = chol6(D6, Lf) && ok;
          const double apd[6] = {double(ap.w.x), double(ap.w.y), double(ap.w.z),
                                 double(ap.v.x), double(ap.v.y), double(ap.v.z)};
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
          for (int t = 0; t < 6; ++t) vd[t] = x[t];
          ap = ap + sv<R>{{R(x[0]), R(x[1]), R(x[2])}, {R(x[3]), R(x[4]), R(x[5])}};
        }
        acc = ap;
        if (d < maxd) put_sv(mineR, acc);  // slot (R): acceleration from the start
      }
    }
    // ---- semi-implicit Euler of this lane's joint (integrators.rs:276-319)
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if (k < nvj) ok = ok && isfinite(vd[k]);
      v[k] = v[k] + vd[k] * dt;
    }
    if (kind == K_FLOAT) advance_floating(q, v, dt, q);
    else q[0] += v[0] * dt;
  }
  if (live) {
#pragma unroll
    for (int k = 0; k < 7; ++k)
      if (k < nqj) {
        qS.stx(0, inst, qo + k, q[k]);
        ok = ok && isfinite(q[k]);
      }
#pragma unroll
    for (int k = 0; k < 6; ++k)
      if (k < nvj) {
        vS.stx(0, inst, vo + k, v[k]);
        vdotS.stx(0, inst, vo + k, vd[k]);
        ok = ok && isfinite(v[k]);
      }
  }
  // any lane's fault is the instance's
  int bad = ok ? 0 : 1;
#pragma unroll
  for (int off = 1; off < L; off <<= 1) bad |= __shfl_xor(bad, off, L);
  if (j == 0 && bad) a.fault[inst] = 1;
}

template <class R, int L>
__global__ void __launch_bounds__(256) wave_tree_kernel(BucketArgs a, double dt, int n_steps) {
  wave_tree_body<R, L>(a, blockIdx.x * blockDim.x + threadIdx.x, dt, n_steps);
}

template <class R>
static int launch_wave_tree(const BucketArgs& a, double dt, int n_steps, int block, hipStream_t s) {
  const long threads = (long)a.B * a.lanes;
  dim3 blk(block), grid((unsigned)((threads + block - 1) / block));
  const unsigned lds = (unsigned)(kTreeSlotBytes * block);
  switch (a.lanes) {
    case 16: hipLaunchKernelGGL((wave_tree_kernel<R, 16>), grid, blk, lds, s, a, dt, n_steps); break;
    case 32: hipLaunchKernelGGL((wave_tree_kernel<R, 32>), grid, blk, lds, s, a, dt, n_steps); break;
    case 64: hipLaunchKernelGGL((wave_tree_kernel<R, 64>), grid, blk, lds, s, a, dt, n_steps); break;
    default: return (int)hipErrorInvalidValue;
  }
  return (int)hipGetLastError();
}

}  // namespace gp
