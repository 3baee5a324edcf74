// gp_vstep.hip — velocity stepping with and without continuous collision
// detection (integrators.rs:42-175, dynamics.rs:366-742): one instance per
// lane, FP64, runtime topology, one step per launch.
//
// Per lane and step:
//  1. free velocity v_free = v + dt M^-1 (tau + springs - c(q, v)) by the
//     articulated-body algorithm without contact forces (dynamics.rs:375-395);
//  2. contacts: contact points inside a half-space and sphere colliders
//     touching one (collision_detection, dynamics.rs:546-606);
//  3. contact Jacobian rows [n; t; b] (v_o + w x p) over the joints from the
//     contact's body to the root (compose_contact_jacobian, dynamics.rs:747-820);
//  4. Y = M^-1 J^T, one articulated-body solve per row of J (zero velocity,
//     zero gravity, no springs: the solve returns M^-1 x);
//  5. G = J Y, g = J v_free and the impulse program of dynamics.rs:668-704:
//     minimise 1/2 x'Gx + g'x with (0.95 x_n, x_t, x_b) in the second-order
//     cone per contact, by a primal-dual interior-point method with
//     Nesterov-Todd scaling and Mehrotra's predictor-corrector;
//  6. v_next = v_free + Y x, q advanced with v_next.
// With continuous collision detection the step is cut at the earliest time
// a separated contact point, or the sphere of a body without contacts, would
// reach a half-space with v_next, and the rest of the step is stepped again;
// a sphere that ended a cut touches that half-space in the next round.
//
// Every per-lane array is a [k][Bp] column block of the bucket's scratch
// (vs_layout), so each access of a wavefront is one coalesced 512-byte line.
#include "gp_cone.h"
#include "gp_plain_impl.h"

namespace gp {

namespace {

struct VsLane {
  const BucketArgs& a;
  int lane;
  VsLayout L;
  double* base;  // a.scratch + lane
  __device__ Strided col(size_t off) const { return Strided{base + off * (size_t)a.Bp, a.Bp}; }
};

__device__ inline v3<double> mcol(const m3<double>& R, int k) { return {R.a[k], R.a[3 + k], R.a[6 + k]}; }

// world frames of every body from the parent-relative transforms the last
// plain_dynamics call left in scratch (PF_E = R^T, PF_R)
__device__ void world_frames(const BucketArgs& a, int lane) {
  PlainCtx X{a, lane};
  for (int i = 0; i < a.nb; ++i) {
    const m3<double> E = X.ldm(PF_E, i);
    const m3<double> Rm{{E.a[0], E.a[3], E.a[6], E.a[1], E.a[4], E.a[7], E.a[2], E.a[5], E.a[8]}};
    const v3<double> p = X.ld3(PF_R, i);
    const int pr = a.parent[i];
    if (pr < 0) {
      X.stm(PF_RW, i, Rm);
      X.st3(PF_TW, i, p);
    } else {
      const m3<double> Rp = X.ldm(PF_RW, pr);
      X.stm(PF_RW, i, mul(Rp, Rm));
      X.st3(PF_TW, i, X.ld3(PF_TW, pr) + mul(Rp, p));
    }
  }
}

// The impulse program for nc contacts (G, g in scratch) -> x. Returns false
// when no finite iterate was produced.
__device__ bool cone_program(const VsLane& V, int nc) {
  const int m = 3 * nc;
  const VsLayout& L = V.L;
  Strided G = V.col(L.g), g = V.col(L.gv), Lc = V.col(L.l), x = V.col(L.x), z = V.col(L.z), s = V.col(L.s),
          rd = V.col(L.rd), lt = V.col(L.lt), xb = V.col(L.xb), dx = V.col(L.dx), dz = V.col(L.dz), ds = V.col(L.ds),
          xi = V.col(L.xi), rhs = V.col(L.rhs), Wc = V.col(L.w), Wic = V.col(L.wi);
  auto d = [](int k) { return k % 3 == 0 ? kConeMu : 1.0; };
  for (int c = 0; c < nc; ++c) {
    x[3 * c] = 1.0 / kConeMu; x[3 * c + 1] = 0.0; x[3 * c + 2] = 0.0;
    z[3 * c] = 1.0; z[3 * c + 1] = 0.0; z[3 * c + 2] = 0.0;
  }
  double gnorm = 0.0;
  for (int k = 0; k < m; ++k) gnorm = fmax(gnorm, fabs(g[k]));
  for (int k = 0; k < m; ++k) xb[k] = x[k];
  double best = 1e300;
  int since_best = 0;
  // Newton direction for a right-hand side xi of the scaled complementarity
  auto solve = [&]() {
    for (int c = 0; c < nc; ++c)
      for (int i = 0; i < 3; ++i) {
        double t = 0.0;
        for (int j = 0; j < 3; ++j) t += Wic[9 * c + 3 * i + j] * xi[3 * c + j];
        rhs[3 * c + i] = -rd[3 * c + i] + d(i) * t;
      }
    for (int r = 0; r < m; ++r) {
      double t = rhs[r];
      for (int k = 0; k < r; ++k) t -= Lc[r * m + k] * dx[k];
      dx[r] = t / Lc[r * m + r];
    }
    for (int r = m - 1; r >= 0; --r) {
      double t = dx[r];
      for (int k = r + 1; k < m; ++k) t -= Lc[k * m + r] * dx[k];
      dx[r] = t / Lc[r * m + r];
    }
    for (int c = 0; c < nc; ++c) {
      double a3[3];
      for (int i = 0; i < 3; ++i) {
        a3[i] = 0.0;
        for (int j = 0; j < 3; ++j) a3[i] += Wic[9 * c + 3 * i + j] * d(j) * dx[3 * c + j];
      }
      for (int i = 0; i < 3; ++i) {
        double t = 0.0;
        for (int j = 0; j < 3; ++j) t += Wic[9 * c + 3 * i + j] * (xi[3 * c + j] - a3[j]);
        dz[3 * c + i] = t;
      }
    }
  };
  auto max_step = [&](double scale) {
    double alpha = 1.0;
    for (int c = 0; c < nc; ++c) {
      const double sc[3] = {s[3 * c], s[3 * c + 1], s[3 * c + 2]}, dsc[3] = {ds[3 * c], ds[3 * c + 1], ds[3 * c + 2]};
      const double zc[3] = {z[3 * c], z[3 * c + 1], z[3 * c + 2]}, dzc[3] = {dz[3 * c], dz[3 * c + 1], dz[3 * c + 2]};
      alpha = fmin(alpha, scale * fmin(soc_max_step(sc, dsc), soc_max_step(zc, dzc)));
    }
    return alpha;
  };
  for (int it = 0; it < kConeMaxIter; ++it) {
    double gap = 0.0, rmax = 0.0, xmax = 0.0;
    for (int k = 0; k < m; ++k) s[k] = d(k) * x[k];
    for (int r = 0; r < m; ++r) {
      double t = g[r] - d(r) * z[r];
      for (int k = 0; k < m; ++k) t += G[r * m + k] * x[k];
      rd[r] = t;
      rmax = fmax(rmax, fabs(t));
      gap += s[r] * z[r];
      xmax = fmax(xmax, fabs(x[r]));
    }
    if (!isfinite(gap) || !isfinite(rmax)) break;
    const double gap_rel = gap / ((1.0 + gnorm) * (1.0 + xmax)), res_rel = rmax / (1.0 + gnorm);
    const double merit = fmax(gap_rel, res_rel);
    if (merit < best) {
      best = merit;
      since_best = 0;
      for (int k = 0; k < m; ++k) xb[k] = x[k];
    }
    if (gap_rel <= kConeGapTol && res_rel <= kConeResTol) break;
    if (++since_best > kConeStall) break;
    const double mug = gap / nc;
    for (int c = 0; c < nc; ++c) {
      const double sc[3] = {s[3 * c], s[3 * c + 1], s[3 * c + 2]}, zc[3] = {z[3 * c], z[3 * c + 1], z[3 * c + 2]};
      double W[9], Wi[9];
      nt_scaling(sc, zc, W, Wi);
      for (int k = 0; k < 9; ++k) { Wc[9 * c + k] = W[k]; Wic[9 * c + k] = Wi[k]; }
      for (int i = 0; i < 3; ++i) lt[3 * c + i] = W[3 * i] * zc[0] + W[3 * i + 1] * zc[1] + W[3 * i + 2] * zc[2];
    }
    // K = G + D W^-2 D, Cholesky in place in Lc (lower triangle)
    for (int r = 0; r < m; ++r)
      for (int k = 0; k <= r; ++k) Lc[r * m + k] = G[r * m + k];
    for (int c = 0; c < nc; ++c)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j <= i; ++j) {
          double t = 0.0;
          for (int k = 0; k < 3; ++k) t += Wic[9 * c + 3 * i + k] * Wic[9 * c + 3 * k + j];
          Lc[(3 * c + i) * m + 3 * c + j] += d(i) * t * d(j);
        }
    bool pd = true;
    for (int j = 0; j < m && pd; ++j) {
      double t = Lc[j * m + j];
      for (int k = 0; k < j; ++k) t -= Lc[j * m + k] * Lc[j * m + k];
      if (!(t > 0.0)) { pd = false; break; }
      const double ljj = sqrt(t);
      Lc[j * m + j] = ljj;
      for (int r = j + 1; r < m; ++r) {
        double u = Lc[r * m + j];
        for (int k = 0; k < j; ++k) u -= Lc[r * m + k] * Lc[j * m + k];
        Lc[r * m + j] = u / ljj;
      }
    }
    if (!pd) break;
    // predictor
    for (int k = 0; k < m; ++k) xi[k] = -lt[k];
    solve();
    for (int k = 0; k < m; ++k) ds[k] = d(k) * dx[k];
    const double alpha_a = max_step(1.0);
    double sz = 0.0;
    for (int k = 0; k < m; ++k) sz += (s[k] + alpha_a * ds[k]) * (z[k] + alpha_a * dz[k]);
    const double ratio = sz / gap, sigma = ratio * ratio * ratio;
    // corrector: xi = lt \ (sigma mu e - lt o lt - (W^-1 ds) o (W dz))
    for (int c = 0; c < nc; ++c) {
      double a3[3], b3[3];
      for (int i = 0; i < 3; ++i) {
        a3[i] = Wic[9 * c + 3 * i] * ds[3 * c] + Wic[9 * c + 3 * i + 1] * ds[3 * c + 1] + Wic[9 * c + 3 * i + 2] * ds[3 * c + 2];
        b3[i] = Wc[9 * c + 3 * i] * dz[3 * c] + Wc[9 * c + 3 * i + 1] * dz[3 * c + 1] + Wc[9 * c + 3 * i + 2] * dz[3 * c + 2];
      }
      const double l0 = lt[3 * c], l1 = lt[3 * c + 1], l2 = lt[3 * c + 2];
      const double w0 = sigma * mug - (l0 * l0 + l1 * l1 + l2 * l2) - (a3[0] * b3[0] + a3[1] * b3[1] + a3[2] * b3[2]);
      const double w1 = -(2.0 * l0 * l1) - (a3[0] * b3[1] + b3[0] * a3[1]);
      const double w2 = -(2.0 * l0 * l2) - (a3[0] * b3[2] + b3[0] * a3[2]);
      const double y0 = (l0 * w0 - l1 * w1 - l2 * w2) / soc_det(l0, l1, l2);
      xi[3 * c] = y0;
      xi[3 * c + 1] = (w1 - y0 * l1) / l0;
      xi[3 * c + 2] = (w2 - y0 * l2) / l0;
    }
    solve();
    for (int k = 0; k < m; ++k) ds[k] = d(k) * dx[k];
    const double alpha = max_step(0.99);
    for (int k = 0; k < m; ++k) {
      x[k] += alpha * dx[k];
      z[k] += alpha * dz[k];
    }
  }
  bool ok = best < 1e300;
  for (int k = 0; k < m; ++k) {
    x[k] = xb[k];
    ok = ok && isfinite(x[k]);
  }
  return ok;
}

// motion subspace column k of body i's joint in world Plücker coordinates
// (angular w, linear velocity of the point at the world origin)
__device__ inline void world_column(int kind, const m3<double>& Rw, v3<double> o, int k, v3<double>& w,
                                    v3<double>& v0) {
  if (kind == K_REV) {
    w = mcol(Rw, 2);
    v0 = cross(o, w);
  } else if (kind == K_PRI) {
    w = v3<double>{0, 0, 0};
    v0 = mcol(Rw, 2);
  } else if (k < 3) {
    w = mcol(Rw, k);
    v0 = cross(o, w);
  } else {
    w = v3<double>{0, 0, 0};
    v0 = mcol(Rw, k - 3);
  }
}

// Mass matrix by the composite-rigid-body algorithm in world coordinates
// (the quantity of mechanism.rs:637-696), then its Cholesky factor, lower
// triangle in place (row-major nv x nv). World frames of the current q must be
// in scratch. Composite inertias (J about the world origin, m c, m) are kept
// in the first ten articulated-inertia fields, free at this point.
__device__ bool mass_matrix_cholesky(const VsLane& V, Strided Mm) {
  const BucketArgs& a = V.a;
  const int lane = V.lane, nb = a.nb, nv = a.nv, Bp = a.Bp;
  PlainCtx X{a, lane};
  for (int i = 0; i < nb; ++i) {  // world inertias (inertia.rs:106-134)
    const m3<double> R = X.ldm(PF_RW, i);
    const v3<double> p = X.ld3(PF_TW, i);
    const double J[9] = {X.bpar(BF_J, i), X.bpar(BF_J + 3, i), X.bpar(BF_J + 4, i),
                         X.bpar(BF_J + 3, i), X.bpar(BF_J + 1, i), X.bpar(BF_J + 5, i),
                         X.bpar(BF_J + 4, i), X.bpar(BF_J + 5, i), X.bpar(BF_J + 2, i)};
    const double mass = X.bpar(BF_M, i);
    const v3<double> Rmc = mul(R, v3<double>{X.bpar(BF_C, i), X.bpar(BF_C + 1, i), X.bpar(BF_C + 2, i)});
    const v3<double> mp = mass * p;
    double RJ[9], Jw[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) RJ[3 * r + c] = R.a[3 * r] * J[c] + R.a[3 * r + 1] * J[3 + c] + R.a[3 * r + 2] * J[6 + c];
    const double rm[3] = {Rmc.x, Rmc.y, Rmc.z}, pp[3] = {p.x, p.y, p.z}, mpp[3] = {mp.x, mp.y, mp.z};
    double Y[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) Y[3 * r + c] = rm[r] * pp[c] + rm[c] * pp[r] + mpp[r] * pp[c];
    const double trY = Y[0] + Y[4] + Y[8];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        Jw[3 * r + c] = RJ[3 * r] * R.a[3 * c] + RJ[3 * r + 1] * R.a[3 * c + 1] + RJ[3 * r + 2] * R.a[3 * c + 2] -
                        Y[3 * r + c] + (r == c ? trY : 0.0);
    X.s(PF_IA, i) = Jw[0]; X.s(PF_IA + 1, i) = Jw[4]; X.s(PF_IA + 2, i) = Jw[8];
    X.s(PF_IA + 3, i) = Jw[1]; X.s(PF_IA + 4, i) = Jw[2]; X.s(PF_IA + 5, i) = Jw[5];
    X.st3(PF_IA + 6, i, Rmc + mp);
    X.s(PF_IA + 9, i) = mass;
  }
  for (int i = nb - 1; i > 0; --i) {  // composite inertias, leaves to root
    const int pr = a.parent[i];
    if (pr < 0) continue;
    for (int f = 0; f < 10; ++f) X.s(PF_IA + f, pr) += X.s(PF_IA + f, i);
  }
  for (int r = 0; r < nv; ++r)
    for (int c = 0; c <= r; ++c) Mm[r * nv + c] = 0.0;
  for (int i = 0; i < nb; ++i) {
    const int ki = a.kind[i * Bp + lane], ni = ki == K_FLOAT ? 6 : (ki == K_FIXED ? 0 : 1);
    if (ni == 0) continue;
    const double Jxx = X.s(PF_IA, i), Jyy = X.s(PF_IA + 1, i), Jzz = X.s(PF_IA + 2, i);
    const double Jxy = X.s(PF_IA + 3, i), Jxz = X.s(PF_IA + 4, i), Jyz = X.s(PF_IA + 5, i);
    const v3<double> c = X.ld3(PF_IA + 6, i);
    const double mass = X.s(PF_IA + 9, i);
    const m3<double> Ri = X.ldm(PF_RW, i);
    const v3<double> oi = X.ld3(PF_TW, i);
    for (int ca = 0; ca < ni; ++ca) {
      v3<double> sw, sv;
      world_column(ki, Ri, oi, ca, sw, sv);
      // F = Ic s (momentum.rs:17-32)
      const v3<double> Jw{Jxx * sw.x + Jxy * sw.y + Jxz * sw.z, Jxy * sw.x + Jyy * sw.y + Jyz * sw.z,
                          Jxz * sw.x + Jyz * sw.y + Jzz * sw.z};
      const v3<double> Fa = Jw + cross(c, sv), Fl = mass * sv - cross(c, sw);
      const int row = a.voff[i] + ca;
      for (int j = i; j >= 0; j = a.parent[j]) {
        const int kj = a.kind[j * Bp + lane], nj = kj == K_FLOAT ? 6 : (kj == K_FIXED ? 0 : 1);
        const m3<double> Rj = X.ldm(PF_RW, j);
        const v3<double> oj = X.ld3(PF_TW, j);
        for (int cb2 = 0; cb2 < nj; ++cb2) {
          const int col = a.voff[j] + cb2;
          if (col > row) continue;
          v3<double> tw, tv;
          world_column(kj, Rj, oj, cb2, tw, tv);
          Mm[row * nv + col] = dot(Fa, tw) + dot(Fl, tv);
        }
      }
    }
  }
  bool pd = true;
  for (int j = 0; j < nv; ++j) {
    double t = Mm[j * nv + j];
    for (int k = 0; k < j; ++k) t -= Mm[j * nv + k] * Mm[j * nv + k];
    if (!(t > 0.0)) {
      pd = false;
      t = 1.0;
    }
    const double ljj = sqrt(t);
    Mm[j * nv + j] = ljj;
    for (int r = j + 1; r < nv; ++r) {
      double u = Mm[r * nv + j];
      for (int k = 0; k < j; ++k) u -= Mm[r * nv + k] * Mm[j * nv + k];
      Mm[r * nv + j] = u / ljj;
    }
  }
  return pd;
}

// dynamics_discrete for one lane at (q, v) with step h: new velocities in
// L.vn. Returns false on a non-positive D or a failed impulse program.
// *n_contacts = the touching contacts (their bodies in L.cb). forced_body /
// forced_h: a sphere the collision loop has just moved onto half-space
// forced_h is in contact with it (its distance is zero in exact arithmetic;
// rounding may leave it some ulps above the plane).
__device__ bool vs_dynamics(const VsLane& V, Strided q, Strided v, Strided tau, double h, int* n_contacts,
                            int forced_body, int forced_h) {
  const BucketArgs& a = V.a;
  const int lane = V.lane, nv = a.nv, nb = a.nb, Bp = a.Bp;
  const VsLayout& L = V.L;
  PlainCtx X{a, lane};
  Strided vfree = V.col(L.vfree), zero = V.col(L.zero), colv = V.col(L.col), out = V.col(L.out), vn = V.col(L.vn);
  Strided J = V.col(L.j), Y = V.col(L.y), G = V.col(L.g), gv = V.col(L.gv), x = V.col(L.x);
  Strided cn = V.col(L.cn), cpn = V.col(L.cp), cb = V.col(L.cb);
  // 1. free velocity: no contact forces on this path (dynamics.rs:377-383)
  bool ok = plain_dynamics(a, lane, q, v, tau, out, false);
  for (int k = 0; k < nv; ++k) vfree[k] = v[k] + h * out[k];
  world_frames(a, lane);
  // 2. contacts: points inside a half-space, then touching spheres
  int nc = 0;
  auto add = [&](v3<double> n, v3<double> p, int body) {
    cn[3 * nc] = n.x; cn[3 * nc + 1] = n.y; cn[3 * nc + 2] = n.z;
    cpn[3 * nc] = p.x; cpn[3 * nc + 1] = p.y; cpn[3 * nc + 2] = p.z;
    cb[nc] = (double)body;
    ++nc;
  };
  for (int i = 0; i < nb; ++i) {
    const m3<double> Rw = X.ldm(PF_RW, i);
    const v3<double> tw = X.ld3(PF_TW, i);
    for (int cc = a.cstart[i]; cc < a.cstart[i + 1]; ++cc) {
      auto cpar = [&](int f) { return a.cp[((size_t)f * a.nc + cc) * Bp + lane]; };
      const v3<double> pw = mul(Rw, v3<double>{cpar(CF_P), cpar(CF_P + 1), cpar(CF_P + 2)}) + tw;
      for (int hh = 0; hh < a.nh; ++hh) {
        auto hpar = [&](int f) { return a.hp[((size_t)f * a.nh + hh) * Bp + lane]; };
        const v3<double> pt{hpar(HF_POINT), hpar(HF_POINT + 1), hpar(HF_POINT + 2)};
        const v3<double> n{hpar(HF_NORMAL), hpar(HF_NORMAL + 1), hpar(HF_NORMAL + 2)};
        if (dot(pw - pt, n) <= kInsideMargin) add(n, pw, i);
      }
    }
  }
  for (int i = 0; i < nb; ++i) {
    const int sl = a.sphere_of[i];
    if (sl < 0) continue;
    auto spar = [&](int f) { return a.shp[((size_t)f * a.ns + sl) * Bp + lane]; };
    if (spar(SH_ON) == 0.0) continue;
    const m3<double> Rw = X.ldm(PF_RW, i);
    const v3<double> center = mul(Rw, v3<double>{spar(SH_C), spar(SH_C + 1), spar(SH_C + 2)}) + X.ld3(PF_TW, i);
    const double r = spar(SH_R);
    for (int hh = 0; hh < a.nh; ++hh) {
      auto hpar = [&](int f) { return a.hp[((size_t)f * a.nh + hh) * Bp + lane]; };
      const v3<double> pt{hpar(HF_POINT), hpar(HF_POINT + 1), hpar(HF_POINT + 2)};
      const v3<double> n{hpar(HF_NORMAL), hpar(HF_NORMAL + 1), hpar(HF_NORMAL + 2)};
      const double dist = dot(center - pt, n);
      if (dist <= r || (i == forced_body && hh == forced_h)) add(n, center - dist * n, i);
    }
  }
  *n_contacts = nc;
  if (nc == 0) {
    for (int k = 0; k < nv; ++k) vn[k] = vfree[k];
    return ok;
  }
  const int m = 3 * nc;
  // 3. Jacobian rows
  for (int ci = 0; ci < nc; ++ci) {
    const v3<double> n{cn[3 * ci], cn[3 * ci + 1], cn[3 * ci + 2]};
    const v3<double> p{cpn[3 * ci], cpn[3 * ci + 1], cpn[3 * ci + 2]};
    v3<double> t = cross(n, v3<double>{1.0, 0.0, 0.0});
    double tn = sqrt(dot(t, t));
    if (tn == 0.0) {
      t = cross(n, v3<double>{0.0, 1.0, 0.0});
      tn = sqrt(dot(t, t));
    }
    t = v3<double>{t.x / tn, t.y / tn, t.z / tn};
    v3<double> bb = cross(n, t);
    const double bn = sqrt(dot(bb, bb));
    bb = v3<double>{bb.x / bn, bb.y / bn, bb.z / bn};
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < nv; ++k) J[(3 * ci + r) * nv + k] = 0.0;
    auto put = [&](int colk, v3<double> vel) {
      J[(3 * ci) * nv + colk] = dot(n, vel);
      J[(3 * ci + 1) * nv + colk] = dot(t, vel);
      J[(3 * ci + 2) * nv + colk] = dot(bb, vel);
    };
    for (int jb = (int)cb[ci]; jb >= 0; jb = a.parent[jb]) {
      const int kd = a.kind[jb * Bp + lane], vo = a.voff[jb];
      const m3<double> Rw = X.ldm(PF_RW, jb);
      const v3<double> rel = p - X.ld3(PF_TW, jb);
      if (kd == K_REV) {
        put(vo, cross(mcol(Rw, 2), rel));
      } else if (kd == K_PRI) {
        put(vo, mcol(Rw, 2));
      } else if (kd == K_FLOAT) {
        for (int k = 0; k < 3; ++k) put(vo + k, cross(mcol(Rw, k), rel));
        for (int k = 0; k < 3; ++k) put(vo + 3 + k, mcol(Rw, k));
      }
    }
  }
  // 4. Y = M^-1 J^T with M from the composite-rigid-body algorithm and its
  // Cholesky factor (M is symmetric positive definite)
  Strided Mm = V.col(L.mm);
  ok = mass_matrix_cholesky(V, Mm) && ok;
  for (int r = 0; r < m; ++r) {
    for (int k = 0; k < nv; ++k) {  // L y = J_r
      double t = J[r * nv + k];
      for (int l = 0; l < k; ++l) t -= Mm[k * nv + l] * out[l];
      out[k] = t / Mm[k * nv + k];
    }
    for (int k = nv - 1; k >= 0; --k) {  // L^T y = .
      double t = out[k];
      for (int l = k + 1; l < nv; ++l) t -= Mm[l * nv + k] * Y[l * m + r];
      Y[k * m + r] = t / Mm[k * nv + k];
    }
  }
  // 5. G = J Y, g = J v_free, impulses
  for (int r = 0; r < m; ++r) {
    for (int c = 0; c <= r; ++c) {
      double t = 0.0;
      for (int k = 0; k < nv; ++k) t += J[r * nv + k] * Y[k * m + c];
      G[r * m + c] = t;
      G[c * m + r] = t;
    }
    double t = 0.0;
    for (int k = 0; k < nv; ++k) t += J[r * nv + k] * vfree[k];
    gv[r] = t;
  }
  ok = cone_program(V, nc) && ok;
  // 6. v_next = v_free + M^-1 J^T x
  for (int k = 0; k < nv; ++k) {
    double t = vfree[k];
    for (int r = 0; r < m; ++r) t += Y[k * m + r] * x[r];
    vn[k] = t;
  }
  return ok;
}

// earliest time within which a separated contact point, or the sphere of a
// body without contacts, reaches a half-space at the new velocities
// (integrators.rs:62-158); world frames of the current q are in scratch.
// *hit_body / *hit_h: the sphere and half-space of that time (-1: a point).
__device__ double earliest_collision(const VsLane& V, int nc, int* hit_body, int* hit_h) {
  const BucketArgs& a = V.a;
  const int lane = V.lane, Bp = a.Bp;
  PlainCtx X{a, lane};
  Strided vn = V.col(V.L.vn), cb = V.col(V.L.cb);
  double earliest = 1e300;
  *hit_body = -1;
  *hit_h = -1;
  auto hit = [&](v3<double> p, v3<double> vel, double radius, int sphere_body) {
    for (int hh = 0; hh < a.nh; ++hh) {
      auto hpar = [&](int f) { return a.hp[((size_t)f * a.nh + hh) * Bp + lane]; };
      const v3<double> pt{hpar(HF_POINT), hpar(HF_POINT + 1), hpar(HF_POINT + 2)};
      const v3<double> n{hpar(HF_NORMAL), hpar(HF_NORMAL + 1), hpar(HF_NORMAL + 2)};
      const double vdn = dot(vel, n);
      if (vdn == 0.0) continue;
      const double dist = dot(p - pt, n) - radius;
      if (dist <= 0.0) continue;
      const double t = dist / -vdn;
      if (t < earliest && t > 0.0) {
        earliest = t;
        *hit_body = sphere_body;
        *hit_h = hh;
      }
    }
  };
  for (int i = 0; i < a.nb; ++i) {
    // world twist (w, velocity of the point at the origin) from the new velocities
    const int kd = a.kind[i * Bp + lane], vo = a.voff[i], pr = a.parent[i];
    const m3<double> Rw = X.ldm(PF_RW, i);
    const v3<double> o = X.ld3(PF_TW, i);
    v3<double> w = pr < 0 ? v3<double>{0, 0, 0} : X.ld3(PF_VEL, pr);
    v3<double> v0 = pr < 0 ? v3<double>{0, 0, 0} : X.ld3(PF_VEL + 3, pr);
    if (kd == K_REV) {
      const v3<double> ax = vn[vo] * mcol(Rw, 2);
      w = w + ax;
      v0 = v0 + cross(o, ax);
    } else if (kd == K_PRI) {
      v0 = v0 + vn[vo] * mcol(Rw, 2);
    } else if (kd == K_FLOAT) {
      const v3<double> ww = mul(Rw, v3<double>{vn[vo], vn[vo + 1], vn[vo + 2]});
      w = w + ww;
      v0 = v0 + mul(Rw, v3<double>{vn[vo + 3], vn[vo + 4], vn[vo + 5]}) + cross(o, ww);
    }
    X.st3(PF_VEL, i, w);
    X.st3(PF_VEL + 3, i, v0);
    for (int cc = a.cstart[i]; cc < a.cstart[i + 1]; ++cc) {
      auto cpar = [&](int f) { return a.cp[((size_t)f * a.nc + cc) * Bp + lane]; };
      const v3<double> pw = mul(Rw, v3<double>{cpar(CF_P), cpar(CF_P + 1), cpar(CF_P + 2)}) + o;
      hit(pw, v0 + cross(w, pw), 0.0, -1);
    }
  }
  for (int i = 0; i < a.nb; ++i) {
    const int sl = a.sphere_of[i];
    if (sl < 0) continue;
    auto spar = [&](int f) { return a.shp[((size_t)f * a.ns + sl) * Bp + lane]; };
    if (spar(SH_ON) == 0.0) continue;
    bool touching = false;
    for (int c = 0; c < nc; ++c) touching = touching || (int)cb[c] == i;
    if (touching) continue;
    const v3<double> center = mul(X.ldm(PF_RW, i), v3<double>{spar(SH_C), spar(SH_C + 1), spar(SH_C + 2)}) + X.ld3(PF_TW, i);
    hit(center, X.ld3(PF_VEL + 3, i) + cross(X.ld3(PF_VEL, i), center), spar(SH_R), i);
  }
  return earliest;
}

}  // namespace

// n_steps steps of each lane in one launch: a lane goes on to its next step
// as soon as its own step is done, so a launch lasts as long as the slowest
// lane's sum of steps rather than the sum of each step's slowest lane. Every
// step reads q and v back from the state arrays it wrote, and the lane's
// arithmetic is the same as in n_steps launches of one step.
__global__ void __launch_bounds__(64) vs_step_kernel(BucketArgs a, double dt, int ccd, int n_steps) {
  const int lane = blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= a.B) return;
  const int Bp = a.Bp, nv = a.nv, nq = a.nq;
  const VsLane V{a, lane, vs_layout(a.nb, a.nq, a.nv, a.nc, a.nh, a.ns), a.scratch + lane};
  Strided q{a.q + lane, Bp}, v{a.v + lane, Bp}, tau{const_cast<double*>(a.tau) + lane, Bp};
  Strided vdot{a.vdot + lane, Bp};
  for (int step = 0; step < n_steps; ++step) {
    Strided qc = V.col(V.L.qc), vc = V.col(V.L.vc), vn = V.col(V.L.vn), zero = V.col(V.L.zero);
    for (int k = 0; k < nv; ++k) zero[k] = 0.0;
    for (int k = 0; k < nq; ++k) qc[k] = q[k];
    for (int k = 0; k < nv; ++k) vc[k] = v[k];
    bool ok = true, done = false;
    double t_remain = dt;
    int forced_body = -1, forced_h = -1, rounds = 0;
    for (int round = 0; round < kCcdMaxRounds && !done; ++round) {
      int nc = 0, hit_body = -1, hit_h = -1;
      rounds = round + 1;
      ok = vs_dynamics(V, qc, vc, tau, t_remain, &nc, forced_body, forced_h) && ok;
      const double earliest = ccd ? earliest_collision(V, nc, &hit_body, &hit_h) : 1e300;
      if (earliest < t_remain) {
        plain_advance(a, lane, qc, vn, earliest, qc);
        for (int k = 0; k < nv; ++k) vc[k] = vn[k];
        t_remain -= earliest;
        forced_body = hit_body;
        forced_h = hit_body >= 0 ? hit_h : -1;
      } else {
        plain_advance(a, lane, qc, vn, t_remain, qc);
        done = true;
      }
    }
    ok = ok && done;  // the collision loop gave up: a fault, the state still advances
    if (!done) plain_advance(a, lane, qc, vn, t_remain, qc);
    for (int k = 0; k < nv; ++k) {
      vdot[k] = (vn[k] - v[k]) / dt;
      v[k] = vn[k];
    }
    for (int k = 0; k < nq; ++k) q[k] = qc[k];
    ok = ok && finite_cols(q, nq) && finite_cols(v, nv);
    if (!ok) a.fault[lane] = 1;
    a.rounds[lane] = rounds;
  }
}

size_t vs_scratch_doubles(int nb, int nq, int nv, int nc, int nh, int ns, int Bp) {
  return vs_layout(nb, nq, nv, nc, nh, ns).total * (size_t)Bp;
}

// 64-lane workgroups: one instance per lane is a long serial program, so the
// launch spreads its waves over every compute unit rather than packing four
// to a workgroup on a quarter of them
int launch_vs_step(const BucketArgs& a, double dt, int ccd, int n_steps, void* stream) {
  if (a.B == 0 || n_steps <= 0) return 0;
  dim3 block(64), grid((a.B + 63) / 64);
  hipLaunchKernelGGL(vs_step_kernel, grid, block, 0, (hipStream_t)stream, a, dt, ccd, n_steps);
  return (int)hipGetLastError();
}

}  // namespace gp
