// gp_group_f64.hip — grouped launch over all fast buckets of a batch, double sweeps.
#include "gp_fast_impl.h"

namespace gp {

int launch_group_f64(const GroupTable* d_table, int n_blocks, double dt, int n_steps, int block, hipStream_t s) {
  // the chain kernels' LDS tier, at its largest row for every lane
  const unsigned lds = (unsigned)(grouped_tier_lane_bytes<double>() * block);
  hipLaunchKernelGGL(grouped_sie_kernel<double>, dim3(n_blocks), dim3(block), lds, s, d_table, dt, n_steps);
  return (int)hipGetLastError();
}

}  // namespace gp
