// gp_fast_f64.hip — per-bucket fast kernels, double sweeps (one translation unit
// per precision so the kernels compile in parallel).
#include "gp_fast_impl.h"

namespace gp {

int launch_fast_f64(int topo, const BucketArgs& a, double dt, int n_steps, int block, hipStream_t s) {
  return launch_prec<double>(topo, a, dt, n_steps, block, s);
}

}  // namespace gp
