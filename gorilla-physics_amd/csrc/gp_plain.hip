// gp_plain.hip — runtime-topology batched step (the plain FP64 kernel).
//
// One mechanism instance per lane, one step (or one dynamics evaluation) per
// launch, every joint type, contacts, prismatic springs and the SIE / RK2 /
// RK4 integrators (integrators.rs:177-319). The dynamics is Featherstone's
// articulated-body algorithm (RBDA Table 7.1) in body frames whose z axis is
// each 1-DOF joint's axis: the joint accelerations of the reference's
// dynamics_continuous (dynamics.rs:322-364: CRBA + RNEA + LU) in O(n). Contact
// forces follow contact.rs:97-129,260-338 with the half-space test of
// halfspace.rs:39-44. Per-body intermediates live in device scratch as
// structure-of-arrays ([field][body][lane]), so every access is coalesced.
#include "gp_plain_impl.h"

namespace gp {

// n_steps steps (or one dynamics evaluation) per launch; one lane per
// instance. Each step reads q and v back from the state arrays it wrote, so a
// lane's arithmetic is the same as in n_steps launches of one step.
__global__ void __launch_bounds__(256) plain_step_kernel(BucketArgs a, double dt, int integrator, int dynamics_only,
                                                         int n_steps) {
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
  for (int step = 0; step < n_steps; ++step) {
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
}

size_t plain_scratch_doubles(int nb, int nq, int nv, int Bp) {
  return ((size_t)NPF * nb + nq + 5 * (size_t)nv) * Bp;
}

int launch_plain_step(const BucketArgs& a, double dt, int integrator, int dynamics_only, int n_steps, void* stream) {
  if (a.B == 0 || n_steps <= 0) return 0;
  dim3 block(256), grid((a.B + 255) / 256);
  hipLaunchKernelGGL(plain_step_kernel, grid, block, 0, (hipStream_t)stream, a, dt, integrator, dynamics_only,
                     dynamics_only ? 1 : n_steps);
  return (int)hipGetLastError();
}

}  // namespace gp
