// gp_fast_rk_f64.hip — RK2 / RK4 fast kernels, double sweeps (a translation unit of
// their own so they compile beside the semi-implicit Euler kernels).
#include "gp_fast_impl.h"

namespace gp {

int launch_fast_rk_f64(int topo, const BucketArgs& a, double dt, int n_steps, int integrator, int block, hipStream_t s) {
  return launch_rk_prec<double>(topo, a, dt, n_steps, integrator, block, s);
}

}  // namespace gp
