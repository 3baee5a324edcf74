// gp_fast_f32.hip — per-bucket fast kernels, float sweeps (one translation unit
// per precision so the kernels compile in parallel).
#include "gp_fast_impl.h"

namespace gp {

int launch_fast_f32(int topo, const BucketArgs& a, double dt, int n_steps, int block, hipStream_t s) {
  return launch_prec<float>(topo, a, dt, n_steps, block, s);
}

}  // namespace gp
