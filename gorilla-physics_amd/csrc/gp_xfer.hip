// gp_xfer.hip — transposition between a caller's device arrays (instance-major
// flattening, the layout of the host set/get calls) and a bucket's
// structure-of-arrays rows, for callers whose state or torques live in HBM.
#include <hip/hip_runtime.h>

#include "gp_internal.h"

namespace gp {

// Thread t handles (row k, lane l) with t = k * B + l: consecutive threads
// touch consecutive lanes of one SoA row (coalesced on the bucket side).
__global__ void __launch_bounds__(256) flat_xfer_kernel(double* soa, int Bp, int B, int rows, const int64_t* flat_off,
                                                        double* flat, int to_soa) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)B * rows) return;
  const int k = (int)(t / B), l = (int)(t % B);
  double* s = soa + (size_t)k * Bp + l;
  double* f = flat + flat_off[l] + k;
  if (to_soa) *s = *f;
  else *f = *s;
}

int launch_flat_xfer(double* soa, int Bp, int B, int rows, const int64_t* flat_off, double* flat, int to_soa,
                     void* stream) {
  if (B == 0 || rows == 0) return 0;
  const int64_t n = (int64_t)B * rows;
  hipLaunchKernelGGL(flat_xfer_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, soa, Bp,
                     B, rows, flat_off, flat, to_soa);
  return (int)hipGetLastError();
}

}  // namespace gp
