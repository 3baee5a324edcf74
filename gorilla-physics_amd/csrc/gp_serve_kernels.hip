// gp_serve_kernels.hip — slot admission and retirement for the serving loop.
//
// admit: K requests, packed on the host one after another, are written into
// their slots (lanes) of a pool's structure-of-arrays state and parameters in
// one launch. gather: the final state of K finished slots is packed into a
// compact buffer for one device-to-host copy.
#include <hip/hip_runtime.h>

#include "gp_internal.h"
#include "gp_serve.h"

namespace gp {

__global__ void __launch_bounds__(64) admit_kernel(PoolPtrs p, AdmitBatch x) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= x.K) return;
  const size_t Bp = p.Bp;
  const int l = x.lanes[k];
  const double* bp = x.bp + (size_t)k * p.nb * NBF;
  for (int i = 0; i < p.nb; ++i) {
    p.kind[i * Bp + l] = x.kind[k * p.nb + i];
    for (int f = 0; f < NBF; ++f) {
      const double val = bp[i * NBF + f];
      p.bp[((size_t)f * p.nb + i) * Bp + l] = val;
      p.bpf[((size_t)f * p.nb + i) * Bp + l] = float(val);
    }
  }
  const double* cp = x.cp + (size_t)k * p.nc * NCF;
  for (int c = 0; c < p.nc; ++c)
    for (int f = 0; f < NCF; ++f) {
      p.cp[((size_t)f * p.nc + c) * Bp + l] = cp[c * NCF + f];
      p.cpf[((size_t)f * p.nc + c) * Bp + l] = float(cp[c * NCF + f]);
    }
  const double* hp = x.hp + (size_t)k * p.nh * NHF;
  for (int h = 0; h < p.nh; ++h)
    for (int f = 0; f < NHF; ++f) {
      p.hp[((size_t)f * p.nh + h) * Bp + l] = hp[h * NHF + f];
      p.hpf[((size_t)f * p.nh + h) * Bp + l] = float(hp[h * NHF + f]);
    }
  for (int j = 0; j < p.nq; ++j) p.q[j * Bp + l] = x.q[(size_t)k * p.nq + j];
  for (int j = 0; j < p.nv; ++j) {
    p.v[j * Bp + l] = x.v[(size_t)k * p.nv + j];
    p.tau[j * Bp + l] = x.tau[(size_t)k * p.nv + j];
    p.vdot[j * Bp + l] = 0.0;
  }
  p.fault[l] = 0;
  p.active[l] = x.steps[k];
}

__global__ void __launch_bounds__(64) gather_kernel(PoolPtrs p, GatherBatch g) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.K) return;
  const size_t Bp = p.Bp;
  const int l = g.lanes[k];
  for (int j = 0; j < p.nq; ++j) g.q[(size_t)k * p.nq + j] = p.q[j * Bp + l];
  for (int j = 0; j < p.nv; ++j) g.v[(size_t)k * p.nv + j] = p.v[j * Bp + l];
  g.fault[k] = p.fault[l];
}

int launch_admit(const PoolPtrs& p, const AdmitBatch& x, void* stream) {
  if (x.K <= 0) return 0;
  hipLaunchKernelGGL(admit_kernel, dim3((x.K + 63) / 64), dim3(64), 0, (hipStream_t)stream, p, x);
  return (int)hipGetLastError();
}

int launch_gather(const PoolPtrs& p, const GatherBatch& g, void* stream) {
  if (g.K <= 0) return 0;
  hipLaunchKernelGGL(gather_kernel, dim3((g.K + 63) / 64), dim3(64), 0, (hipStream_t)stream, p, g);
  return (int)hipGetLastError();
}

}  // namespace gp
