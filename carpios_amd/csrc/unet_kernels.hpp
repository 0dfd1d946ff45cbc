// BF16 MFMA kernels of the U-Net preconditioner (no reference counterpart).
//
// Activations are NHWC BF16 (channels contiguous; a pixel m = i*W + j of an
// H x W level), gradients of activations are NHWC FP32, weights are BF16
// copies of FP32 master parameters.  Accumulation is FP32 throughout.
//
// MFMA fragment layout on gfx950 (checked on the device): for
// v_mfma_f32_16x16x{16,32}_bf16 lane l holds A[l%16][KPL*(l/16) + t] and
// B[KPL*(l/16) + t][l%16] (KPL = K/4 values per lane) and returns
// D[4*(l/16) + t][l%16], t = 0..3.
//
// Implicit GEMM: one wavefront owns 16*MP output rows (pixels) x 16*NB
// output columns (channels); K runs over (tap, input channel) in chunks of
// the MFMA depth.  Every operand fragment is one 8- or 16-byte load of
// contiguous channels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cpgpu {
namespace unet {

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4_t;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;

typedef uint16_t bf16;  // storage: raw bits

__device__ __host__ __forceinline__ float bf2f(bf16 b) {
  union {
    uint32_t u;
    float f;
  } x;
  x.u = (uint32_t)b << 16;
  return x.f;
}
// round to nearest even; NaN stays NaN
__device__ __host__ __forceinline__ bf16 f2bf(float f) {
  union {
    uint32_t u;
    float f;
  } x;
  x.f = f;
  if ((x.u & 0x7f800000u) == 0x7f800000u && (x.u & 0x7fffffu))
    return (bf16)((x.u >> 16) | 0x40);
  const uint32_t r = x.u + 0x7fffu + ((x.u >> 16) & 1u);
  return (bf16)(r >> 16);
}

// A tensor view: element (m, c) at p[m*cs + co + c]
struct View {
  bf16* p;
  int H, W, cs, co;
};
struct GView {  // FP32 gradient view
  float* p;
  int H, W, cs, co;
};

template <int KC>
struct Frag;
template <>
struct Frag<16> {
  typedef bf16x4_t T;
  static constexpr int KPL = 4;
  __device__ static f32x4_t mfma(T a, T b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  }
  __device__ static T load(const bf16* p) {
    T v;
    __builtin_memcpy(&v, p, 8);
    return v;
  }
  __device__ static T zero() {
    T v = {0, 0, 0, 0};
    return v;
  }
  __device__ static T from_f32(const float* g, const bf16* y, bool mask) {
    T v;
    short s[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      float x = g[t];
      if (mask && !(bf2f(y[t]) > 0.f)) x = 0.f;
      s[t] = (short)f2bf(x);
    }
    __builtin_memcpy(&v, s, 8);
    return v;
  }
};
template <>
struct Frag<32> {
  typedef bf16x8_t T;
  static constexpr int KPL = 8;
  __device__ static f32x4_t mfma(T a, T b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  __device__ static T load(const bf16* p) {
    T v;
    __builtin_memcpy(&v, p, 16);
    return v;
  }
  __device__ static T zero() {
    T v;
    const uint32_t z[4] = {0, 0, 0, 0};
    __builtin_memcpy(&v, z, 16);
    return v;
  }
  __device__ static T from_f32(const float* g, const bf16* y, bool mask) {
    T v;
    bf16 s[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      float x = g[t];
      if (mask && !(bf2f(y[t]) > 0.f)) x = 0.f;
      s[t] = f2bf(x);
    }
    __builtin_memcpy(&v, s, 16);
    return v;
  }
};



// ---------------------------------------------------------------------
// conv 3x3, pad 1: forward.  y = relu(bias + sum_tap W[tap] x[m + off]).
// W layout [tap][co][ci].  grid: x = pixel tiles of 4 waves * 16*MP,
// y = column groups of 16*NB channels.
template <int CIN, int COUT, int NB, int MP>
__global__ __launch_bounds__(256) void k_conv3_fwd(View x, View y,
                                                   const bf16* __restrict__ w,
                                                   const float* __restrict__ bias,
                                                   int relu) {
  constexpr int KC = (CIN % 32 == 0) ? 32 : 16;
  typedef Frag<KC> F;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int M = x.H * x.W;
  const int m0 = (blockIdx.x * 4 + wave) * 16 * MP;
  const int n0 = blockIdx.y * 16 * NB;
  if (m0 >= M) return;
  f32x4_t acc[MP][NB];
#pragma unroll
  for (int a = 0; a < MP; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  int pi[MP], pj[MP];
#pragma unroll
  for (int a = 0; a < MP; ++a) {
    const int m = m0 + a * 16 + (lane & 15);
    pi[a] = m / x.W;
    pj[a] = m - pi[a] * x.W;
    if (m >= M) pi[a] = -1000000;  // never in range
  }
  const int kq = F::KPL * (lane >> 4);
#pragma unroll 1
  for (int tap = 0; tap < 9; ++tap) {
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
#pragma unroll
    for (int c0 = 0; c0 < CIN; c0 += KC) {
      typename F::T bf[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int co = n0 + b * 16 + (lane & 15);
        bf[b] = F::load(w + ((size_t)(tap * COUT + co) * CIN + c0 + kq));
      }
#pragma unroll
      for (int a = 0; a < MP; ++a) {
        const int si = pi[a] + dy, sj = pj[a] + dx;
        typename F::T af = F::zero();
        if (si >= 0 && si < x.H && sj >= 0 && sj < x.W)
          af = F::load(x.p + (size_t)(si * x.W + sj) * x.cs + x.co + c0 + kq);
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = F::mfma(af, bf[b], acc[a][b]);
      }
    }
  }
#pragma unroll
  for (int a = 0; a < MP; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int co = n0 + b * 16 + (lane & 15);
      const float bb = bias[co];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int m = m0 + a * 16 + 4 * (lane >> 4) + t;
        if (m >= M) continue;
        float v = acc[a][b][t] + bb;
        if (relu) v = v > 0.f ? v : 0.f;
        y.p[(size_t)m * y.cs + y.co + co] = f2bf(v);
      }
    }
}

// conv 3x3 input gradient: dx[m][ci] = sum_tap sum_co dyp[m - off][co]
// WT[tap][ci][co], dyp = dy masked by the layer's ReLU output y.
template <int CIN, int COUT, int NB, int MP>
__global__ __launch_bounds__(256) void k_conv3_dgrad(GView dy, View y, GView dx,
                                                     const bf16* __restrict__ wt,
                                                     int relu) {
  constexpr int KC = (COUT % 32 == 0) ? 32 : 16;
  typedef Frag<KC> F;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int M = dy.H * dy.W;
  const int m0 = (blockIdx.x * 4 + wave) * 16 * MP;
  const int n0 = blockIdx.y * 16 * NB;
  if (m0 >= M) return;
  f32x4_t acc[MP][NB];
#pragma unroll
  for (int a = 0; a < MP; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  int pi[MP], pj[MP];
#pragma unroll
  for (int a = 0; a < MP; ++a) {
    const int m = m0 + a * 16 + (lane & 15);
    pi[a] = m / dy.W;
    pj[a] = m - pi[a] * dy.W;
    if (m >= M) pi[a] = -1000000;
  }
  const int kq = F::KPL * (lane >> 4);
#pragma unroll 1
  for (int tap = 0; tap < 9; ++tap) {
    const int dyy = tap / 3 - 1, dxx = tap % 3 - 1;
#pragma unroll
    for (int c0 = 0; c0 < COUT; c0 += KC) {
      typename F::T bf[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int ci = n0 + b * 16 + (lane & 15);
        bf[b] = F::load(wt + ((size_t)(tap * CIN + ci) * COUT + c0 + kq));
      }
#pragma unroll
      for (int a = 0; a < MP; ++a) {
        const int si = pi[a] - dyy, sj = pj[a] - dxx;
        typename F::T af = F::zero();
        if (si >= 0 && si < dy.H && sj >= 0 && sj < dy.W) {
          const size_t sm = (size_t)(si * dy.W + sj);
          af = F::from_f32(dy.p + sm * dy.cs + dy.co + c0 + kq,
                           y.p + sm * y.cs + y.co + c0 + kq, relu != 0);
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = F::mfma(af, bf[b], acc[a][b]);
      }
    }
  }
#pragma unroll
  for (int a = 0; a < MP; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int ci = n0 + b * 16 + (lane & 15);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int m = m0 + a * 16 + 4 * (lane >> 4) + t;
        if (m < M) dx.p[(size_t)m * dx.cs + dx.co + ci] = acc[a][b][t];
      }
    }
}

// up-convolution 4x4 stride 4 (transposed conv): out[4I+a][4J+b][co] =
// bias[co] + sum_ci x[I][J][ci] Wu[ab][co][ci].  GEMM rows = coarse pixels,
// columns = (ab, co).
template <int CIN, int COUT, int NB, int MP>
__global__ __launch_bounds__(256) void k_up_fwd(View x, View y,
                                                const bf16* __restrict__ w,
                                                const float* __restrict__ bias) {
  constexpr int KC = (CIN % 32 == 0) ? 32 : 16;
  typedef Frag<KC> F;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int M = x.H * x.W;
  const int m0 = (blockIdx.x * 4 + wave) * 16 * MP;
  const int n0 = blockIdx.y * 16 * NB;  // over 16*COUT columns
  if (m0 >= M) return;
  f32x4_t acc[MP][NB];
#pragma unroll
  for (int a = 0; a < MP; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const int kq = F::KPL * (lane >> 4);
#pragma unroll
  for (int c0 = 0; c0 < CIN; c0 += KC) {
    typename F::T bf[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int n = n0 + b * 16 + (lane & 15);  // = ab*COUT + co
      bf[b] = F::load(w + ((size_t)n * CIN + c0 + kq));
    }
#pragma unroll
    for (int a = 0; a < MP; ++a) {
      const int m = m0 + a * 16 + (lane & 15);
      typename F::T af = F::zero();
      if (m < M) af = F::load(x.p + (size_t)m * x.cs + x.co + c0 + kq);
#pragma unroll
      for (int b = 0; b < NB; ++b) acc[a][b] = F::mfma(af, bf[b], acc[a][b]);
    }
  }
#pragma unroll
  for (int a = 0; a < MP; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int n = n0 + b * 16 + (lane & 15);
      const int ab = n / COUT, co = n - ab * COUT;
      const int ai = ab >> 2, bj = ab & 3;
      const float bb = bias[co];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int m = m0 + a * 16 + 4 * (lane >> 4) + t;
        if (m >= M) continue;
        const int I = m / x.W, J = m - I * x.W;
        const size_t f = (size_t)(4 * I + ai) * y.W + (4 * J + bj);
        y.p[f * y.cs + y.co + co] = f2bf(acc[a][b][t] + bb);
      }
    }
}

// up-convolution input gradient: dx[m][ci] = sum_{ab,co} dy[fine(m,ab)][co]
// WuT[ab][ci][co]
template <int CIN, int COUT, int NB, int MP>
__global__ __launch_bounds__(256) void k_up_dgrad(GView dy, GView dx,
                                                  const bf16* __restrict__ wt) {
  constexpr int KC = (COUT % 32 == 0) ? 32 : 16;
  typedef Frag<KC> F;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int M = dx.H * dx.W;
  const int m0 = (blockIdx.x * 4 + wave) * 16 * MP;
  const int n0 = blockIdx.y * 16 * NB;
  if (m0 >= M) return;
  f32x4_t acc[MP][NB];
#pragma unroll
  for (int a = 0; a < MP; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const int kq = F::KPL * (lane >> 4);
#pragma unroll 1
  for (int ab = 0; ab < 16; ++ab) {
    const int ai = ab >> 2, bj = ab & 3;
#pragma unroll
    for (int c0 = 0; c0 < COUT; c0 += KC) {
      typename F::T bf[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int ci = n0 + b * 16 + (lane & 15);
        bf[b] = F::load(wt + ((size_t)(ab * CIN + ci) * COUT + c0 + kq));
      }
#pragma unroll
      for (int a = 0; a < MP; ++a) {
        const int m = m0 + a * 16 + (lane & 15);
        typename F::T af = F::zero();
        if (m < M) {
          const int I = m / dx.W, J = m - I * dx.W;
          const size_t f = (size_t)(4 * I + ai) * dy.W + (4 * J + bj);
          af = F::from_f32(dy.p + f * dy.cs + dy.co + c0 + kq, nullptr, false);
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = F::mfma(af, bf[b], acc[a][b]);
      }
    }
  }
#pragma unroll
  for (int a = 0; a < MP; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int ci = n0 + b * 16 + (lane & 15);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int m = m0 + a * 16 + 4 * (lane >> 4) + t;
        if (m < M) dx.p[(size_t)m * dx.cs + dx.co + ci] = acc[a][b][t];
      }
    }
}

// ---------------------------------------------------------------------
// Weight gradients on MFMA.  dW[tap][co][ci] = sum_p g(p, tap, co) x(p, tap,
// ci) with K = pixels: each block takes one tap and a chunk of pixels,
// stages 64-pixel slices of both operands in LDS transposed (channel-major,
// pixels contiguous) so every fragment is one 16-byte LDS read, and writes
// FP32 partial sums per chunk (reduced in a fixed order afterwards).
// conv: g = dy[p] * relu'(y[p]), x = X[p + off(tap)], taps 0..8
// up:   g = dy[fine(p, tap)],     x = X[p],           taps 0..15
template <int CIN, int COUT, bool UP>
__global__ __launch_bounds__(256) void k_wgrad(GView dy, View y, View x,
                                               int relu, int pix_per_chunk,
                                               float* __restrict__ part) {
  constexpr int PS = 64;          // pixels per LDS slice
  constexpr int PITCH = PS + 8;   // bf16 per LDS row (16-byte aligned rows)
  __shared__ __attribute__((aligned(16))) bf16 sg[COUT * PITCH];
  __shared__ __attribute__((aligned(16))) bf16 sx[CIN * PITCH];
  typedef Frag<32> F;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tap = blockIdx.y;
  const int M = x.H * x.W;  // pixels of the GEMM's K dimension
  const int p0 = blockIdx.x * pix_per_chunk;
  const int p1 = min(M, p0 + pix_per_chunk);
  constexpr int TCO = COUT / 16, TCI = CIN / 16, TT = TCO * TCI;
  constexpr int TPW = (TT + 3) / 4;  // tiles per wave
  f32x4_t acc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) acc[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const int dyy = UP ? 0 : tap / 3 - 1, dxx = UP ? 0 : tap % 3 - 1;
  const int ai = tap >> 2, bj = tap & 3;
  for (int ps = p0; ps < p1; ps += PS) {
    __syncthreads();
    // stage g (COUT x PS) and x (CIN x PS), zero beyond p1
    for (int e = threadIdx.x; e < PS * COUT; e += 256) {
      const int pp = e / COUT, c = e - pp * COUT;
      const int p = ps + pp;
      float v = 0.f;
      if (p < p1) {
        const int I = p / x.W, J = p - I * x.W;
        size_t f;
        if (UP) {
          f = (size_t)(4 * I + ai) * dy.W + (4 * J + bj);
          v = dy.p[f * dy.cs + dy.co + c];
        } else {
          f = (size_t)p;
          v = dy.p[f * dy.cs + dy.co + c];
          if (relu && !(bf2f(y.p[f * y.cs + y.co + c]) > 0.f)) v = 0.f;
        }
      }
      sg[c * PITCH + pp] = f2bf(v);
    }
    for (int e = threadIdx.x; e < PS * CIN; e += 256) {
      const int pp = e / CIN, c = e - pp * CIN;
      const int p = ps + pp;
      bf16 v = 0;
      if (p < p1) {
        const int I = p / x.W, J = p - I * x.W;
        const int si = I + dyy, sj = J + dxx;
        if (si >= 0 && si < x.H && sj >= 0 && sj < x.W)
          v = x.p[(size_t)(si * x.W + sj) * x.cs + x.co + c];
      }
      sx[c * PITCH + pp] = v;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const int tile = wave + 4 * t;
      if (tile >= TT) break;
      const int tco = tile / TCI, tci = tile - tco * TCI;
#pragma unroll
      for (int k0 = 0; k0 < PS; k0 += 32) {
        const typename F::T af =
            F::load(&sg[(tco * 16 + (lane & 15)) * PITCH + k0 + 8 * (lane >> 4)]);
        const typename F::T bfr =
            F::load(&sx[(tci * 16 + (lane & 15)) * PITCH + k0 + 8 * (lane >> 4)]);
        acc[t] = F::mfma(af, bfr, acc[t]);
      }
    }
  }
  // partial sums: part[((chunk * NT + tap) * COUT + co) * CIN + ci]
  const int ntap = UP ? 16 : 9;
  float* out = part + ((size_t)blockIdx.x * ntap + tap) * COUT * CIN;
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int tile = wave + 4 * t;
    if (tile >= TT) break;
    const int tco = tile / TCI, tci = tile - tco * TCI;
    const int ci = tci * 16 + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int co = tco * 16 + 4 * (lane >> 4) + r;
      out[(size_t)co * CIN + ci] = acc[t][r];
    }
  }
}

}  // namespace unet
}  // namespace cpgpu
