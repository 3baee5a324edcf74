// Host stand-in for <hip/hip_runtime.h>, used only by the CPU build of the
// library's runtime and kernels under AddressSanitizer / UBSan
// (tests/test_host_sanitized_kernels.py). Device memory is host memory
// allocated with its exact size (so the sanitizer sees every access past an
// array), a kernel launch runs every (workgroup, lane) in turn on the calling
// thread, and streams, events and graphs are no-ops. Kernels that exchange
// values between lanes (shuffles, shared LDS, barriers) are not built this
// way; lane-private LDS columns are.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) (&(x))

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct gp_host_idx { unsigned x = 0, y = 0, z = 0; };
inline thread_local gp_host_idx blockIdx, threadIdx, blockDim, gridDim;

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorNotSupported = 801 };
typedef struct gp_host_stream* hipStream_t;
typedef struct gp_host_event* hipEvent_t;
typedef struct gp_host_graph* hipGraph_t;
typedef struct gp_host_graph_exec* hipGraphExec_t;
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyHostToHost };
enum hipStreamCaptureMode { hipStreamCaptureModeThreadLocal };

// Every allocation sits between two poisoned 1 MiB guards, so an index that
// overshoots a [k][Bp] block by many rows still lands in poisoned memory.
#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/asan_interface.h>
#define GP_HOST_POISON(p, n) __asan_poison_memory_region((p), (n))
#define GP_HOST_UNPOISON(p, n) __asan_unpoison_memory_region((p), (n))
#else
#define GP_HOST_POISON(p, n) ((void)0)
#define GP_HOST_UNPOISON(p, n) ((void)0)
#endif
constexpr size_t kGpHostGuard = (size_t)1 << 20;
inline hipError_t hipMalloc(void** p, size_t n) {
  const size_t body = (n + 15) / 16 * 16;
  char* raw = (char*)std::calloc(body + 2 * kGpHostGuard + 16, 1);
  if (!raw) return hipErrorOutOfMemory;
  std::memcpy(raw, &body, sizeof body);
  GP_HOST_POISON(raw + 16, kGpHostGuard - 16);
  GP_HOST_POISON(raw + kGpHostGuard + n, body - n + kGpHostGuard + 16);
  *p = raw + kGpHostGuard;
  return hipSuccess;
}
inline hipError_t hipFree(void* p) {
  if (!p) return hipSuccess;
  char* raw = (char*)p - kGpHostGuard;
  size_t body;
  std::memcpy(&body, raw, sizeof body);
  GP_HOST_UNPOISON(raw, body + 2 * kGpHostGuard + 16);
  std::free(raw);
  return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { std::memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) { return hipMemcpy(d, s, n, k); }
inline hipError_t hipMemset(void* d, int v, size_t n) { std::memset(d, v, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { return hipMemset(d, v, n); }
inline hipError_t hipStreamCreate(hipStream_t* s) { *s = nullptr; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
template <class T> inline hipError_t hipGetSymbolAddress(void** p, T* symbol) { *p = (void*)symbol; return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t) { return "host build"; }
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = nullptr; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0.0f; return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*) { return hipErrorNotSupported; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, void*, size_t) { return hipErrorNotSupported; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
inline hipError_t hipMemGetInfo(size_t* f, size_t* t) { *f = *t = (size_t)1 << 34; return hipSuccess; }


inline void sincos(double x, double* s, double* c) { *s = std::sin(x); *c = std::cos(x); }
inline void sincosf(float x, float* s, float* c) { *s = std::sin(x); *c = std::cos(x); }
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
using std::fmin;
using std::fmax;
using std::fabs;
using std::sqrt;
using std::isfinite;

// Raw buffer resources (the compiled-topology kernels' array accesses): a base
// pointer and a byte range. An access outside the range, which the hardware
// would drop (loads read 0), ends the run here: the kernels are meant to stay
// inside their arrays.
struct gp_host_rsrc {
  char* base;
  int bytes;
};
typedef gp_host_rsrc __amdgpu_buffer_rsrc_t;
inline gp_host_rsrc gp_host_make_rsrc(void* p, short, int bytes, int) { return {(char*)p, bytes}; }
inline char* gp_host_rsrc_at(const gp_host_rsrc& r, int voff, int soff, int size) {
  const long long at = (long long)voff + soff;
  if (at < 0 || at + size > r.bytes) {
    std::fprintf(stderr, "buffer access out of range: offset %lld size %d range %d\n", at, size, r.bytes);
    std::abort();
  }
  return r.base + at;
}
inline uint64_t gp_host_load_b64(gp_host_rsrc r, int voff, int soff, int) {
  uint64_t x;
  std::memcpy(&x, gp_host_rsrc_at(r, voff, soff, 8), 8);
  return x;
}
inline uint32_t gp_host_load_b32(gp_host_rsrc r, int voff, int soff, int) {
  uint32_t x;
  std::memcpy(&x, gp_host_rsrc_at(r, voff, soff, 4), 4);
  return x;
}
inline void gp_host_store_b64(uint64_t x, gp_host_rsrc r, int voff, int soff, int) {
  std::memcpy(gp_host_rsrc_at(r, voff, soff, 8), &x, 8);
}
inline void gp_host_store_b32(uint32_t x, gp_host_rsrc r, int voff, int soff, int) {
  std::memcpy(gp_host_rsrc_at(r, voff, soff, 4), &x, 4);
}
#define __builtin_amdgcn_make_buffer_rsrc gp_host_make_rsrc
#define __builtin_amdgcn_raw_buffer_load_b64 gp_host_load_b64
#define __builtin_amdgcn_raw_buffer_load_b32 gp_host_load_b32
#define __builtin_amdgcn_raw_buffer_store_b64 gp_host_store_b64
#define __builtin_amdgcn_raw_buffer_store_b32 gp_host_store_b32

// Lanes run one after another here, so nothing can be exchanged between
// them: a kernel that shuffles (free bodies on several lanes, trees, the
// leg-parallel quadruped) must not be launched in this build.
template <class T> inline T gp_host_no_shuffle() {
  std::fprintf(stderr, "a cross-lane shuffle ran in the host build\n");
  std::abort();
}
template <class T> inline T __shfl_xor(T, int, int = 64) { return gp_host_no_shuffle<T>(); }
template <class T> inline T __shfl(T, int, int = 64) { return gp_host_no_shuffle<T>(); }
inline unsigned long long __ballot(int) { return gp_host_no_shuffle<unsigned long long>(); }
#define __builtin_amdgcn_update_dpp(old, x, ctrl, row_mask, bank_mask, bound_ctrl) gp_host_no_shuffle<int>()
inline int __popc(unsigned x) { return __builtin_popcount(x); }

// Dynamic LDS of the running launch: one guarded allocation of the launch's
// size, shared by its workgroups in turn (the kernels that use it keep
// lane-private columns, so one workgroup's values never reach another's lanes
// before being written).
inline thread_local unsigned char* gp_host_dyn_lds = nullptr;
#define GP_DYN_LDS gp_host_dyn_lds
enum hipDeviceAttribute_t { hipDeviceAttributeMaxSharedMemoryPerBlock };
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) { *v = 160 << 10; return hipSuccess; }

// every (workgroup, lane) of the grid in turn
#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...)          \
  do {                                                                       \
    const dim3 gp_g_ = dim3(grid), gp_b_ = dim3(block);                      \
    gridDim.x = gp_g_.x;                                                     \
    blockDim.x = gp_b_.x;                                                    \
    const size_t gp_sh_ = (size_t)(shmem);                                   \
    void* gp_lds_ = nullptr;                                                 \
    if (gp_sh_) (void)hipMalloc(&gp_lds_, gp_sh_);                           \
    gp_host_dyn_lds = (unsigned char*)gp_lds_;                               \
    for (unsigned gp_bx = 0; gp_bx < gp_g_.x; ++gp_bx)                       \
      for (unsigned gp_tx = 0; gp_tx < gp_b_.x; ++gp_tx) {                   \
        blockIdx.x = gp_bx;                                                  \
        threadIdx.x = gp_tx;                                                 \
        kernel(__VA_ARGS__);                                                 \
      }                                                                      \
    gp_host_dyn_lds = nullptr;                                               \
    if (gp_lds_) (void)hipFree(gp_lds_);                                     \
  } while (0)
