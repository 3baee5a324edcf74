// gp_serve_kernels.hip — slot admission and retirement for the serving loop.
//
// admit: the requests of every pool admitted in one round (their state packed
// on the host one after another, their models' parameters from the device
// copies made when the models were first submitted) are written into their
// slots (lanes) of the pools' structure-of-arrays state and parameters by one
// launch. gather: the final
// state of every slot that finished in the round is packed into a compact
// buffer for one device-to-host copy, by one launch.
#include <hip/hip_runtime.h>

#include "gp_internal.h"
#include "gp_serve.h"

namespace gp {

// the table entry holding request r (uniform per workgroup)
template <class T> __device__ __forceinline__ int entry_of(const T* t, int r) {
  int e = 0;
  while (e + 1 < t->n && r >= t->start[e + 1]) ++e;
  return e;
}

// Workgroup r writes request r into its slot: kind, body / contact /
// half-space parameters (both precisions, from its model's device copy), q,
// v, tau, a zero vdot and fault, and the horizon. Lane j handles elements
// j, j + 64, ... of the request.
__global__ void __launch_bounds__(64) admit_kernel(const AdmitTable* __restrict__ t) {
  const int r = blockIdx.x;
  const int e = entry_of(t, r);
  const PoolPtrs& p = t->p[e];
  const AdmitBatch& x = t->x[e];
  const int k = r - t->start[e];
  const size_t Bp = p.Bp;
  const int l = x.lanes[k];
  const ModelDev md = t->models[x.model[k]];
  const int nbp = p.nb * NBF, ncp = p.nc * NCF, nhp = p.nh * NHF;
  const int n_elem = p.nb + nbp + ncp + nhp + p.nq + 2 * p.nv;
  for (int j = threadIdx.x; j < n_elem; j += blockDim.x) {
    int m = j;
    if (m < p.nb) { p.kind[m * Bp + l] = md.kind[m]; continue; }
    m -= p.nb;
    if (m < nbp) {  // packed [nb][NBF] -> SoA [NBF][nb]
      const int i = m / NBF, f = m % NBF;
      const double val = md.bp[m];
      p.bp[((size_t)f * p.nb + i) * Bp + l] = val;
      p.bpf[((size_t)f * p.nb + i) * Bp + l] = float(val);
      continue;
    }
    m -= nbp;
    if (m < ncp) {
      const int c = m / NCF, f = m % NCF;
      const double val = md.cp[m];
      p.cp[((size_t)f * p.nc + c) * Bp + l] = val;
      p.cpf[((size_t)f * p.nc + c) * Bp + l] = float(val);
      continue;
    }
    m -= ncp;
    if (m < nhp) {
      const int h = m / NHF, f = m % NHF;
      const double val = md.hp[m];
      p.hp[((size_t)f * p.nh + h) * Bp + l] = val;
      p.hpf[((size_t)f * p.nh + h) * Bp + l] = float(val);
      continue;
    }
    m -= nhp;
    if (m < p.nq) { p.q[m * Bp + l] = x.q[(size_t)k * p.nq + m]; continue; }
    m -= p.nq;
    if (m < p.nv) {
      p.v[m * Bp + l] = x.v[(size_t)k * p.nv + m];
      p.vdot[m * Bp + l] = 0.0;
      continue;
    }
    m -= p.nv;
    p.tau[m * Bp + l] = x.tau[(size_t)k * p.nv + m];
  }
  if (threadIdx.x == 0) {
    p.fault[l] = 0;
    p.active[l] = x.steps[k];
  }
}

// Workgroup r packs finished slot r's q, v and fault.
__global__ void __launch_bounds__(64) gather_kernel(const GatherTable* __restrict__ t) {
  const int r = blockIdx.x;
  const int e = entry_of(t, r);
  const PoolPtrs& p = t->p[e];
  const GatherBatch& g = t->g[e];
  const int k = r - t->start[e];
  const size_t Bp = p.Bp;
  const int l = g.lanes[k];
  for (int j = threadIdx.x; j < p.nq + p.nv; j += blockDim.x) {
    if (j < p.nq) g.q[(size_t)k * p.nq + j] = p.q[j * Bp + l];
    else g.v[(size_t)k * p.nv + (j - p.nq)] = p.v[(j - p.nq) * Bp + l];
  }
  if (threadIdx.x == 0) g.fault[k] = p.fault[l];
}

int launch_admit(const AdmitTable* d_table, int n_requests, void* stream) {
  if (n_requests <= 0) return 0;
  hipLaunchKernelGGL(admit_kernel, dim3(n_requests), dim3(64), 0, (hipStream_t)stream, d_table);
  return (int)hipGetLastError();
}

int launch_gather(const GatherTable* d_table, int n_requests, void* stream) {
  if (n_requests <= 0) return 0;
  hipLaunchKernelGGL(gather_kernel, dim3(n_requests), dim3(64), 0, (hipStream_t)stream, d_table);
  return (int)hipGetLastError();
}

}  // namespace gp
