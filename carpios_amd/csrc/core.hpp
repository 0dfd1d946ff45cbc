// Shared internals of libcpgpu: context, error plumbing, kernel timers and
// the deterministic reductions used by every dot product.
//
// Reductions never use device-scope fences or atomics: on gfx950 each of the
// eight XCDs has its own L2, and an agent-scope release in every block writes
// that L2 back.  Instead every block of a kernel stores its partial sums with
// plain stores; the kernel boundary makes them visible, and the consumer (the
// next kernel) reduces all partials in a fixed order in its prologue.  Every
// block of the consumer computes the identical total, so the result is
// deterministic and needs no broadcast.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/cpgpu.h"

struct cpgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // device address of this GPU's static scratch (cpgpu_poisson2d_solve_static)
  void* scratch = nullptr;
  // kernel timers: per family a list of (start, stop) event pairs
  static constexpr int kFamilies = 10;
  bool timer_on[kFamilies] = {};
  std::vector<hipEvent_t> ev[kFamilies];
  size_t ev_used[kFamilies] = {};
};

namespace cpgpu {

int set_error(cpgpu_ctx* ctx, int code, const char* fmt, ...);

// Device address of the static scratch array on `device` (current), looked
// up once per device (poisson.hip).
int static_scratch_setup(int device, void** out);

#define CPGPU_HIP(ctx, call)                                                  \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      return ::cpgpu::set_error((ctx), CPGPU_EHIP, "%s:%d %s: %s", __FILE__,  \
                                __LINE__, #call, hipGetErrorString(e_));      \
  } while (0)

// Start/stop events of one timed launch of a kernel family (both null when
// timing is off for the family or ctx is null).
struct EvPair {
  hipEvent_t a = nullptr, b = nullptr;
};
EvPair timer_events(cpgpu_ctx* ctx, int fam);

// Launch on a stream; with timer events through hipExtLaunchKernelGGL,
// whose events take the dispatch's own start and end timestamps (no extra
// packets between kernels, so the figure is the kernel's execution time).
#define CPGPU_LAUNCH(ev, kernel, grid, block, shmem, st, ...)                  \
  do {                                                                         \
    if ((ev).a)                                                                \
      hipExtLaunchKernelGGL(kernel, grid, block, shmem, st, (ev).a, (ev).b, 0, \
                            __VA_ARGS__);                                      \
    else                                                                       \
      hipLaunchKernelGGL(kernel, grid, block, shmem, st, __VA_ARGS__);         \
  } while (0)

constexpr int kThreads = 256;  // threads per block of every reduction kernel
constexpr int kMaxBlocks = 1024;  // upper bound of partials per reduction

// Wavefront (64 lanes) sum through cross-lane shuffles in a fixed order.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum of NV values per thread; every thread receives the totals.
// Needs kThreads threads; lds holds NV*4 values.
template <typename T, int NV>
__device__ __forceinline__ void block_allsum(T (&v)[NV], T* lds) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int lane = tid & 63, wid = tid >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();  // lds may still be read from a previous use
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[k * 4 + wid] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k)
    v[k] = ((lds[k * 4] + lds[k * 4 + 1]) + lds[k * 4 + 2]) + lds[k * 4 + 3];
}

// Block partials: partials[k*G + bid] = block sum of v[k].
template <typename T, int NV>
__device__ __forceinline__ void store_partials(T (&v)[NV], T* partials, int G,
                                               int bid, T* lds) {
  block_allsum<T, NV>(v, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[k * G + bid] = v[k];
  }
}

// Totals of the partials written by the previous kernel (G blocks), same
// value in every thread of every block.
template <typename T, int NV>
__device__ __forceinline__ void reduce_partials(const T* __restrict__ partials,
                                                int G, T (&tot)[NV], T* lds) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    T s = T(0);
    for (int b = tid; b < G; b += kThreads) s += partials[k * G + b];
    tot[k] = s;
  }
  block_allsum<T, NV>(tot, lds);
}

}  // namespace cpgpu
