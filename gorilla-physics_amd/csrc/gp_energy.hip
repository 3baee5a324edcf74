// gp_energy.hip — kinetic and gravitational energy of every instance
// (MechanismState::kinetic_energy / gravitational_energy, mechanism.rs:334-362),
// one instance per lane, FP64, runtime topology.
//
// Per body: world frame of the aligned body frame from q (the same joint
// transforms as the plain kernel), world twist (angular velocity w and the
// velocity v0 of the point at the world origin) from v, world spatial inertia
// (inertia.rs:106-134); kinetic energy 1/2 (w.J w + v0.(m v0 + 2 w x c))
// (inertia.rs:182-202); gravitational energy m g z of the body frame's origin,
// as the reference computes it (not of the centre of mass).
#include "gp_plain_impl.h"

namespace gp {

__global__ void __launch_bounds__(64) energy_kernel(BucketArgs a, double* ke, double* pe) {
  const int lane = blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= a.B) return;
  const int Bp = a.Bp;
  PlainCtx X{a, lane};
  double K = 0.0, P = 0.0;
  for (int i = 0; i < a.nb; ++i) {
    const int k = a.kind[i * Bp + lane];
    const int pr = a.parent[i];
    const int qo = a.qoff[i], vo = a.voff[i];
    m3<double> Rt;
    for (int f = 0; f < 9; ++f) Rt.a[f] = X.bpar(BF_RT + f, i);
    const v3<double> pt{X.bpar(BF_PT, i), X.bpar(BF_PT + 1, i), X.bpar(BF_PT + 2, i)};
    m3<double> Rm = Rt;
    v3<double> p = pt;
    if (k == K_REV) {
      double s, c;
      sincos(a.q[(size_t)qo * Bp + lane], &s, &c);
      const m3<double> Rz{{c, -s, 0.0, s, c, 0.0, 0.0, 0.0, 1.0}};
      Rm = mul(Rt, Rz);
    } else if (k == K_PRI) {
      p = pt + a.q[(size_t)qo * Bp + lane] * v3<double>{Rt.a[2], Rt.a[5], Rt.a[8]};
    } else if (k == K_FLOAT) {
      auto qq = [&](int t) { return a.q[(size_t)(qo + t) * Bp + lane]; };
      Rm = mul(Rt, quat_mat(qq(0), qq(1), qq(2), qq(3)));
      p = pt + mul(Rt, v3<double>{qq(4), qq(5), qq(6)});
    }
    m3<double> Rw = Rm;
    v3<double> o = p;
    v3<double> w{0, 0, 0}, v0{0, 0, 0};
    if (pr >= 0) {
      const m3<double> Rp = X.ldm(PF_RW, pr);
      Rw = mul(Rp, Rm);
      o = X.ld3(PF_TW, pr) + mul(Rp, p);
      w = X.ld3(PF_VEL, pr);
      v0 = X.ld3(PF_VEL + 3, pr);
    }
    auto vv = [&](int t) { return a.v[(size_t)(vo + t) * Bp + lane]; };
    if (k == K_REV) {
      const v3<double> ax = vv(0) * v3<double>{Rw.a[2], Rw.a[5], Rw.a[8]};
      w = w + ax;
      v0 = v0 + cross(o, ax);
    } else if (k == K_PRI) {
      v0 = v0 + vv(0) * v3<double>{Rw.a[2], Rw.a[5], Rw.a[8]};
    } else if (k == K_FLOAT) {
      const v3<double> ww = mul(Rw, v3<double>{vv(0), vv(1), vv(2)});
      w = w + ww;
      v0 = v0 + mul(Rw, v3<double>{vv(3), vv(4), vv(5)}) + cross(o, ww);
    }
    X.stm(PF_RW, i, Rw);
    X.st3(PF_TW, i, o);
    X.st3(PF_VEL, i, w);
    X.st3(PF_VEL + 3, i, v0);
    // world spatial inertia: J about the world origin, m c, m
    const double J[9] = {X.bpar(BF_J, i), X.bpar(BF_J + 3, i), X.bpar(BF_J + 4, i),
                         X.bpar(BF_J + 3, i), X.bpar(BF_J + 1, i), X.bpar(BF_J + 5, i),
                         X.bpar(BF_J + 4, i), X.bpar(BF_J + 5, i), X.bpar(BF_J + 2, i)};
    const double mass = X.bpar(BF_M, i);
    const v3<double> Rmc = mul(Rw, v3<double>{X.bpar(BF_C, i), X.bpar(BF_C + 1, i), X.bpar(BF_C + 2, i)});
    const v3<double> mp = mass * o;
    const double rm[3] = {Rmc.x, Rmc.y, Rmc.z}, pp[3] = {o.x, o.y, o.z}, mpp[3] = {mp.x, mp.y, mp.z};
    double RJ[9], Jw[9], Y[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        RJ[3 * r + c] = Rw.a[3 * r] * J[c] + Rw.a[3 * r + 1] * J[3 + c] + Rw.a[3 * r + 2] * J[6 + c];
        Y[3 * r + c] = rm[r] * pp[c] + rm[c] * pp[r] + mpp[r] * pp[c];
      }
    const double trY = Y[0] + Y[4] + Y[8];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        Jw[3 * r + c] = RJ[3 * r] * Rw.a[3 * c] + RJ[3 * r + 1] * Rw.a[3 * c + 1] + RJ[3 * r + 2] * Rw.a[3 * c + 2] -
                        Y[3 * r + c] + (r == c ? trY : 0.0);
    const v3<double> cw = Rmc + mp;
    const v3<double> Jww{Jw[0] * w.x + Jw[1] * w.y + Jw[2] * w.z, Jw[3] * w.x + Jw[4] * w.y + Jw[5] * w.z,
                         Jw[6] * w.x + Jw[7] * w.y + Jw[8] * w.z};
    K += (dot(w, Jww) + dot(v0, mass * v0 + 2.0 * cross(w, cw))) / 2.0;
    P += mass * a.g * o.z;
  }
  ke[lane] = K;
  pe[lane] = P;
}

int launch_energy(const BucketArgs& a, double* ke, double* pe, void* stream) {
  if (a.B == 0) return 0;
  dim3 block(64), grid((a.B + 63) / 64);
  hipLaunchKernelGGL(energy_kernel, grid, block, 0, (hipStream_t)stream, a, ke, pe);
  return (int)hipGetLastError();
}

}  // namespace gp
