// gp_fast_rk_f32.hip — RK2 / RK4 fast kernels, float sweeps (a translation unit of
// their own so they compile beside the semi-implicit Euler kernels).
#include "gp_fast_impl.h"

namespace gp {

int launch_fast_rk_f32(int topo, const BucketArgs& a, double dt, int n_steps, int integrator, int block, hipStream_t s) {
  return launch_rk_prec<float>(topo, a, dt, n_steps, integrator, block, s);
}

}  // namespace gp
