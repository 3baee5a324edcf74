// gp_group_f32.hip — grouped launch over all fast buckets of a batch, float sweeps.
#include "gp_fast_impl.h"

namespace gp {

int launch_group_f32(const GroupTable* d_table, int n_blocks, double dt, int n_steps, int block, hipStream_t s) {
  // the chain kernels' LDS tier, at its largest row for every lane
  const unsigned lds = (unsigned)(grouped_tier_lane_bytes<float>() * block);
  hipLaunchKernelGGL(grouped_sie_kernel<float>, dim3(n_blocks), dim3(block), lds, s, d_table, dt, n_steps);
  return (int)hipGetLastError();
}

}  // namespace gp
