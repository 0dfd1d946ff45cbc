// Shared internals of libcpgpu: context, error plumbing, kernel timers and
// the deterministic block/grid reduction used by every dot product.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/cpgpu.h"

struct cpgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // kernel timers: per family a list of (start, stop) event pairs
  bool timer_on[4] = {false, false, false, false};
  std::vector<hipEvent_t> ev[4];
  size_t ev_used[4] = {0, 0, 0, 0};
};

namespace cpgpu {

int set_error(cpgpu_ctx* ctx, int code, const char* fmt, ...);

#define CPGPU_HIP(ctx, call)                                                  \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      return ::cpgpu::set_error((ctx), CPGPU_EHIP, "%s:%d %s: %s", __FILE__,  \
                                __LINE__, #call, hipGetErrorString(e_));      \
  } while (0)

// Brackets kernel launches of one family with events when timing is on.
struct TimerScope {
  cpgpu_ctx* ctx;
  int fam;
  hipEvent_t stop = nullptr;
  TimerScope(cpgpu_ctx* c, int f);
  ~TimerScope();
};

constexpr int kBlock = 256;  // threads per block for 1-D kernels

// Wavefront (64 lanes) sum through cross-lane shuffles; the order of the
// additions is fixed, so the result is deterministic.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum of NV values per thread (blockDim.x*blockDim.y threads, a
// multiple of 64).  Result valid in thread 0.
template <typename T, int NV>
__device__ __forceinline__ void block_sum(T (&v)[NV], T* lds /* >= NV*16 */) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int nthreads = blockDim.x * blockDim.y;
  const int lane = tid & 63, wid = tid >> 6, nw = nthreads >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[k * 16 + wid] = v[k];
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      T s = lds[k * 16];
      for (int w = 1; w < nw; ++w) s += lds[k * 16 + w];
      v[k] = s;
    }
  }
}

// Two-pass grid reduction.  Every block writes its NV partials; the last
// block to arrive (ticket counter) sums all partials in index order, so the
// result does not depend on block scheduling.  Returns true in the block
// that holds the totals (valid in thread 0 of that block, in tot[]).
template <typename T, int NV>
__device__ __forceinline__ bool grid_sum(T (&v)[NV], T* partials,
                                         unsigned int* ticket, int nblocks,
                                         T (&tot)[NV]) {
  __shared__ T lds[NV * 16 > 64 ? NV * 16 : 64];
  __shared__ bool is_last;
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int nthreads = blockDim.x * blockDim.y;
  const int bid = blockIdx.x + blockIdx.y * gridDim.x;
  block_sum<T, NV>(v, lds);
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[k * nblocks + bid] = v[k];
    __threadfence();
    unsigned int t = atomicAdd(ticket, 1u);
    is_last = (t == (unsigned int)(nblocks - 1));
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  T acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    T s = T(0);
    for (int b = tid; b < nblocks; b += nthreads)
      s += __hip_atomic_load(&partials[k * nblocks + b], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    acc[k] = s;
  }
  __syncthreads();
  block_sum<T, NV>(acc, lds);
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) tot[k] = acc[k];
    *ticket = 0u;
  }
  return true;
}

}  // namespace cpgpu
