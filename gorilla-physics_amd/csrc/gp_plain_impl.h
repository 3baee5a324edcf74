// gp_plain_impl.h — the runtime-topology FP64 dynamics shared by the plain
// kernel (gp_plain.hip) and the velocity-stepping kernel (gp_vstep.hip).
//
// plain_dynamics is the articulated-body algorithm (RBDA Table 7.1) for one
// lane in body frames whose z axis is each 1-DOF joint's axis: the joint
// accelerations of the reference's dynamics_continuous (dynamics.rs:322-364:
// CRBA + RNEA + LU) in O(n). Contact forces follow
// contact.rs:97-129,260-338 with the half-space test of halfspace.rs:39-44.
// Per-body intermediates live in device scratch as structure-of-arrays
// ([field][body][lane]), so every access is coalesced.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/gorilla_amd.h"
#include "gp_device.h"
#include "gp_internal.h"
#include "gp_spatial.h"

namespace gp {

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

// Pass 1 of the dynamics for one lane, root to leaves: each body's joint
// transform (PF_E = E, PF_R = r), velocity (PF_VEL), velocity-product
// acceleration v x_m S qd (PF_CB), rigid-body inertia (PF_IA) and bias force
// v x* (I v) - f_ext (PF_PA) with the contact and spring-contact forces, in
// the aligned body frame. contacts = false leaves the forces out;
// update_springs = false evaluates the spring contacts without changing their
// registration (a query: a spring registering in this evaluation, or
// detaching, exerts no force either way). Returns false where a spring
// contact pulls into its half-space.
__device__ inline bool plain_kinematics(const BucketArgs& a, int lane, Strided q, Strided v, bool contacts = true,
                                        bool update_springs = true) {
  PlainCtx X{a, lane};
  const int nb = a.nb, Bp = a.Bp;
  bool ok = true;
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
    if (contacts && (a.nc > 0 || a.nsp > 0)) {
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
      // spring contacts (contact.rs:131-186); their registration changes here,
      // once per dynamics evaluation, as in the reference
      for (int sc = a.sstart[i]; sc < a.sstart[i + 1]; ++sc) {
        auto spar = [&](int fld) -> double& { return a.sp[((size_t)fld * a.nsp + sc) * Bp + lane]; };
        auto sst = [&](int fld) -> double& { return a.ss[((size_t)fld * a.nsp + sc) * Bp + lane]; };
        auto hpar = [&](int h, int fld) { return a.hp[((size_t)fld * a.nh + h) * Bp + lane]; };
        if (sst(SS_REG) == 0.0) {
          if (!update_springs) continue;  // registering exerts no force in this evaluation
          // the leg's tip: body origin + R direction l_rest; the first half-space holding it registers it
          const v3<double> tip = tw + spar(SF_L) * mul(Rw, v3<double>{spar(SF_DIR), spar(SF_DIR + 1), spar(SF_DIR + 2)});
          for (int h = 0; h < a.nh; ++h) {
            const v3<double> pt_h{hpar(h, HF_POINT), hpar(h, HF_POINT + 1), hpar(h, HF_POINT + 2)};
            const v3<double> n{hpar(h, HF_NORMAL), hpar(h, HF_NORMAL + 1), hpar(h, HF_NORMAL + 2)};
            if (!(dot(tip - pt_h, n) <= kInsideMargin)) continue;
            sst(SS_REG) = 1.0;
            sst(SS_P) = tip.x; sst(SS_P + 1) = tip.y; sst(SS_P + 2) = tip.z;
            sst(SS_H) = (double)h;
            break;
          }
        } else {
          const v3<double> diff = v3<double>{sst(SS_P), sst(SS_P + 1), sst(SS_P + 2)} - tw;
          const double dist = sqrt(dot(diff, diff));
          const v3<double> sdir{diff.x / dist, diff.y / dist, diff.z / dist};
          const int h = (int)sst(SS_H);
          const v3<double> n{hpar(h, HF_NORMAL), hpar(h, HF_NORMAL + 1), hpar(h, HF_NORMAL + 2)};
          if (dot(sdir, n) > 0.0) ok = false;  // a pull into the half-space: the reference panics
          if (dist < spar(SF_L)) {
            const double force = -spar(SF_K) * (dist - spar(SF_L));  // energy_control.rs:13
            const v3<double> F = force * (-sdir);                    // at the body's origin: no torque about it
            f.v = f.v - mulT(Rw, F);
          } else if (update_springs) {
            // detached: the leg keeps its last direction, a world vector stored as
            // the body-frame direction (contact.rs:176-181), here in the aligned frame
            sst(SS_REG) = 0.0;
            const double sn = sqrt(dot(sdir, sdir));
            const v3<double> w{sdir.x / sn, sdir.y / sn, sdir.z / sn};
            for (int r = 0; r < 3; ++r)
              spar(SF_DIR + r) = spar(SF_AT + 3 * r) * w.x + spar(SF_AT + 3 * r + 1) * w.y + spar(SF_AT + 3 * r + 2) * w.z;
          }
        }
      }
    }
    X.stv(PF_PA, i, f);
    X.sti(i, I);
  }
  return ok;
}

// Articulated-body forward dynamics of one lane; q, v, tau, vdot are the
// lane's columns of [k][Bp] arrays. contacts = false leaves out the contact
// and spring-contact forces (the velocity-stepping path). Returns false on a
// non-positive D.
__device__ inline bool plain_dynamics(const BucketArgs& a, int lane, Strided q, Strided v, Strided tau, Strided vdot,
                                      bool contacts = true) {
  PlainCtx X{a, lane};
  const int nb = a.nb, Bp = a.Bp;
  // ---- pass 1: kinematics, velocities, bias forces, contact forces
  bool ok = plain_kinematics(a, lane, q, v, contacts);
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
      if (k == K_PRI && sk != 0.0 && !a.no_springs) tq += -sk * (q[a.qoff[i]] - X.bpar(BF_SL, i));  // dynamics.rs:298-315
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
__device__ inline void plain_advance(const BucketArgs& a, int lane, Strided q, Strided v, double h, Strided qo) {
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

__device__ inline bool finite_cols(Strided x, int n) {
  bool ok = true;
  for (int k = 0; k < n; ++k) ok = ok && isfinite(x[k]);
  return ok;
}

}  // namespace gp
