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
// round to nearest even (the hardware conversion, v_cvt_pk_bf16_f32; two
// conversions side by side become one instruction): branch-free, where the
// bit-level formula compiled to a branch per value for its NaN case
__device__ __forceinline__ bf16 f2bf(float f) { return __builtin_bit_cast(bf16, (__bf16)f); }

// two values rounded as f2bf, packed lo | hi << 16: one v_cvt_pk_bf16_f32
// (two scalar conversions side by side are not always combined)
__device__ __forceinline__ uint32_t bf16x2_bits(float lo, float hi) {
  typedef __bf16 bf2_t __attribute__((ext_vector_type(2)));
  typedef float f2_t __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f2_t{lo, hi}), bf2_t));
}
// the FP32 value of half `hi` (0 low, 1 high) of a packed pair
__device__ __forceinline__ float bf16x2_half(uint32_t bits, int hi) {
  return __uint_as_float(hi ? (bits & 0xffff0000u) : (bits << 16));
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
  constexpr int NC = CIN / KC;
  // per tap, every load of the tap (all channel chunks, weights and
  // activations) is issued before its MFMAs, the activations from clamped
  // valid addresses and zeroed when outside the image: one memory round trip
  // per tap instead of one per chunk
#pragma unroll 1
  for (int tap = 0; tap < 9; ++tap) {
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    typename F::T bf[NC][NB], af[NC][MP];
    bool ok[MP];
    size_t off[MP];
#pragma unroll
    for (int a = 0; a < MP; ++a) {
      const int si = pi[a] + dy, sj = pj[a] + dx;
      ok[a] = si >= 0 && si < x.H && sj >= 0 && sj < x.W;
      const int ci = min(max(si, 0), x.H - 1), cj = min(max(sj, 0), x.W - 1);
      off[a] = (size_t)(ci * x.W + cj) * x.cs + x.co + kq;
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int co = n0 + b * 16 + (lane & 15);
        bf[c][b] = F::load(w + ((size_t)(tap * COUT + co) * CIN + c * KC + kq));
      }
#pragma unroll
      for (int a = 0; a < MP; ++a) af[c][a] = F::load(x.p + off[a] + c * KC);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int a = 0; a < MP; ++a) {
        const typename F::T av = ok[a] ? af[c][a] : F::zero();
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = F::mfma(bf[c][b], av, acc[a][b]);
      }
  }
  // D[channel][pixel]: four consecutive channels of one pixel per lane
#pragma unroll
  for (int a = 0; a < MP; ++a) {
    const int m = m0 + a * 16 + (lane & 15);
    if (m >= M) continue;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int co = n0 + b * 16 + 4 * (lane >> 4);
      uint32_t h[2];
#pragma unroll
      for (int t = 0; t < 4; t += 2) {
        float v0 = acc[a][b][t] + bias[co + t];
        float v1 = acc[a][b][t + 1] + bias[co + t + 1];
        if (relu) {
          v0 = v0 > 0.f ? v0 : 0.f;
          v1 = v1 > 0.f ? v1 : 0.f;
        }
        h[t >> 1] = bf16x2_bits(v0, v1);
      }
      *reinterpret_cast<uint2*>(y.p + (size_t)m * y.cs + y.co + co) = uint2{h[0], h[1]};
    }
  }
}

// conv 3x3 forward for small levels (a few thousand pixels or fewer), the
// nine taps split over the block's four waves: wave w sums taps
// [kT0[w], kT0[w + 1]) = {0,1,2}, {3,4}, {5,6}, {7,8} for the block's
// 16 MP pixels x 16 NB output channels, every load of its taps issued before
// its MFMAs (one memory round trip per wave, where a wave walking all nine
// taps makes nine), and wave 0 adds the four partial sums in wave order from
// LDS and applies the epilogue.  grid: x = groups of MP 16-pixel groups,
// y = groups of NB 16-channel groups; a block's weight fragments serve its
// MP pixel groups and its activation fragments its NB channel groups.
// BLK4 (sides multiples of 4): a 16-pixel group is a 4 x 4 pixel block (lane
// l % 16 at row (l % 16) / 4, column l % 4 of block g), and with pout.p set
// the epilogue also writes the block's average (the 4 x 4 pooling of the
// BF16 output, summed by lane shuffles: the sum of 16 non-negative BF16
// values is exact in FP32, as in k_pool_fwd).
template <int CIN, int COUT, bool BLK4 = false, int MP = 1, int NB = 1>
__global__ __launch_bounds__(256) void k_conv3_fwd_ks(View x, View y, const bf16* __restrict__ w,
                                                      const float* __restrict__ bias, int relu,
                                                      View pout) {
  constexpr int KC = (CIN % 32 == 0) ? 32 : 16;
  constexpr int NC = CIN / KC;
  typedef Frag<KC> F;
  __shared__ float red[3][MP * NB * 256];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int M = x.H * x.W;
  int m[MP], pi[MP], pj[MP];
#pragma unroll
  for (int a = 0; a < MP; ++a) {
    const int g = blockIdx.x * MP + a;
    if constexpr (BLK4) {
      const int nbw = x.W >> 2, bi = g / nbw, bj = g - bi * nbw;
      pi[a] = 4 * bi + ((lane & 15) >> 2);
      pj[a] = 4 * bj + (lane & 3);
      m[a] = pi[a] < x.H ? pi[a] * x.W + pj[a] : M;
    } else {
      m[a] = g * 16 + (lane & 15);
      pi[a] = m[a] < M ? m[a] / x.W : -1000000;
      pj[a] = m[a] < M ? m[a] - (m[a] / x.W) * x.W : 0;
    }
  }
  const int n0 = blockIdx.y * 16 * NB;
  const int kq = F::KPL * (lane >> 4);
  const int t0 = wave == 0 ? 0 : 1 + 2 * wave, nt = wave == 0 ? 3 : 2;
  typename F::T bf[3][NC][NB], af[3][NC][MP];
  bool ok[3][MP];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int tap = min(t0 + k, 8);
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int b = 0; b < NB; ++b)
        bf[k][c][b] =
            F::load(w + ((size_t)(tap * COUT + n0 + b * 16 + (lane & 15)) * CIN + c * KC + kq));
#pragma unroll
    for (int a = 0; a < MP; ++a) {
      const int si = pi[a] + tap / 3 - 1, sj = pj[a] + tap % 3 - 1;
      ok[k][a] = k < nt && si >= 0 && si < x.H && sj >= 0 && sj < x.W;
      const int ci = min(max(si, 0), x.H - 1), cj = min(max(sj, 0), x.W - 1);
      const size_t off = (size_t)(ci * x.W + cj) * x.cs + x.co + kq;
#pragma unroll
      for (int c = 0; c < NC; ++c) af[k][c][a] = F::load(x.p + off + c * KC);
    }
  }
  f32x4_t acc[MP][NB];
#pragma unroll
  for (int a = 0; a < MP; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (k < nt) {
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int a = 0; a < MP; ++a)
#pragma unroll
          for (int b = 0; b < NB; ++b)
            acc[a][b] = F::mfma(bf[k][c][b], ok[k][a] ? af[k][c][a] : F::zero(), acc[a][b]);
    }
  }
  const int e = (4 * (lane >> 4)) * 16 + (lane & 15);  // + 16 t: D[4 (lane / 16) + t][lane % 16]
  if (wave > 0) {
#pragma unroll
    for (int a = 0; a < MP; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int t = 0; t < 4; ++t) red[wave - 1][(a * NB + b) * 256 + e + 16 * t] = acc[a][b][t];
  }
  __syncthreads();
  if (wave > 0) return;
#pragma unroll
  for (int a = 0; a < MP; ++a) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int co = n0 + b * 16 + 4 * (lane >> 4);
      uint32_t h[2];
#pragma unroll
      for (int t = 0; t < 4; t += 2) {
        float v[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int ix = (a * NB + b) * 256 + e + 16 * (t + u);
          float s = acc[a][b][t + u];
          s += red[0][ix];
          s += red[1][ix];
          s += red[2][ix];
          s += bias[co + t + u];
          v[u] = relu ? (s > 0.f ? s : 0.f) : s;
        }
        h[t >> 1] = bf16x2_bits(v[0], v[1]);
      }
      if (m[a] < M)
        *reinterpret_cast<uint2*>(y.p + (size_t)m[a] * y.cs + y.co + co) = uint2{h[0], h[1]};
      if (BLK4 && pout.p) {
        float ps[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          float v = bf16x2_half(h[u >> 1], u & 1);
          v += __shfl_xor(v, 1, 64);
          v += __shfl_xor(v, 2, 64);
          v += __shfl_xor(v, 4, 64);
          v += __shfl_xor(v, 8, 64);
          ps[u] = v * (1.f / 16.f);
        }
        if ((lane & 15) == 0 && m[a] < M)
          *reinterpret_cast<uint2*>(pout.p + (size_t)(blockIdx.x * MP + a) * pout.cs + pout.co +
                                    co) = uint2{bf16x2_bits(ps[0], ps[1]),
                                                bf16x2_bits(ps[2], ps[3])};
      }
    }
  }
}

// conv 3x3 input gradient: dx[m][ci] = sum_tap sum_co g[m - off][co]
// WT[tap][ci][co], g = the layer's output gradient masked by its ReLU and
// rounded to BF16 once (k_mask_grad), read like an activation.
template <int CIN, int COUT, int NB, int MP>
__global__ __launch_bounds__(256) void k_conv3_dgrad(View dy, GView dx,
                                                     const bf16* __restrict__ wt) {
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
        if (si >= 0 && si < dy.H && sj >= 0 && sj < dy.W)
          af = F::load(dy.p + (size_t)(si * dy.W + sj) * dy.cs + dy.co + c0 + kq);
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = F::mfma(bf[b], af, acc[a][b]);
      }
    }
  }
  // D[channel][pixel]: four consecutive channels of one pixel per lane
#pragma unroll
  for (int a = 0; a < MP; ++a) {
    const int m = m0 + a * 16 + (lane & 15);
    if (m >= M) continue;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int ci = n0 + b * 16 + 4 * (lane >> 4);
      *reinterpret_cast<float4*>(dx.p + (size_t)m * dx.cs + dx.co + ci) =
          float4{acc[a][b][0], acc[a][b][1], acc[a][b][2], acc[a][b][3]};
    }
  }
}

// 3x3 convolution (pad 1) of one 16 x 16 pixel tile per block, the input
// tile and its halo staged once in LDS (pixel-major, KCH channels
// contiguous, 16-byte chunks), for levels whose sides are multiples of 16.
// Every tap's A fragment is then an aligned 16-byte LDS read.
//   FWD:  y[p][n]  = relu(bias[n] + sum_tap sum_k w[tap][n][k] x[p + off][k])
//         (BF16 out; off = (tap / 3 - 1, tap % 3 - 1))
//   !FWD: dx[p][n] = sum_tap sum_k w[tap][n][k] x[p - off][k]
//         (the input gradient: x = masked BF16 output gradient, w = the
//         transposed weights [tap][ci][co], FP32 out)
// Wave w computes tile rows 4w..4w+3 (16 pixels each) x NO channels.  The
// product is formed as D[channel][pixel] (weights as the A operand): a lane
// then holds four consecutive output channels of one pixel, stored as one
// 8-byte (BF16) or 16-byte (FP32) write, and the 16 lanes of a group cover 16
// consecutive pixels, so a wave's store is one contiguous row segment.
// With pout.p set (FWD) the epilogue also writes the 4 x 4 average pooling
// of the BF16 output into pout, as the 16-channel tiles' pooling epilogue
// forms it (wave w's four rows are one pooling band).
template <int KCH, int NO, bool FWD>
__global__ __launch_bounds__(256) void k_conv3_tile(View x, View yb, GView yf,
                                                    const bf16* __restrict__ w,
                                                    const float* __restrict__ bias,
                                                    int relu, View pout) {
  constexpr int KC = (KCH % 32 == 0) ? 32 : 16;
  constexpr int NB = NO / 16;
  constexpr int NCK = KCH / KC;  // K steps per tap
  constexpr int TP = 18;       // tile side with halo
  constexpr int CPP = KCH / 8;  // 16-byte chunks per pixel
  typedef Frag<KC> F;
  // at most two K = 32 steps per tap and two column blocks (forward): all
  // nine taps' weight fragments fit in registers and are loaded before the
  // staging, so their latency overlaps it (these levels run one block per
  // compute unit, with nothing else to hide a load behind)
  constexpr bool kPre = (KCH == 32 || (FWD && KCH == 64)) && NB <= 2;
  __shared__ __attribute__((aligned(16))) bf16 sx[TP * TP * KCH];
  // tile element (pixel, channel): with 32 channels (four 16-byte chunks
  // per pixel) the chunks are stored in the order chunk ^ ((column >> 1) & 3);
  // a 128-bit LDS read serves 16 consecutive lanes per cycle, and 16
  // consecutive pixels at the plain 64-byte pitch meet 2-way bank
  // conflicts, in this order none (measured on gfx950).  The 64-channel
  // tiles (one block per CU at their level, latency-bound) measured slower
  // with the key's arithmetic than with the conflicts, and keep the plain
  // order.
  auto sxo = [](int pix, int col, int ch) {
    const int key = CPP == 4 ? (col >> 1) & 3 : 0;
    return pix * KCH + (((ch >> 3) ^ key) << 3) + (ch & 7);
  };
  const int tj0 = blockIdx.x * 16, ti0 = blockIdx.y * 16;
  const int H = x.H, W = x.W;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int px = lane & 15, kq = F::KPL * (lane >> 4);
  typename F::T wpre[kPre ? 9 : 1][kPre ? NCK : 1][NB];
  if constexpr (kPre) {
#pragma unroll
    for (int tap = 0; tap < 9; ++tap)
#pragma unroll
      for (int c = 0; c < NCK; ++c)
#pragma unroll
        for (int b = 0; b < NB; ++b)
          wpre[tap][c][b] = F::load(w + ((size_t)(tap * NO + b * 16 + px) * KCH + c * KC + kq));
  }
  // staging: every load of the tile first, from clamped valid addresses
  // (masked when stored), so a thread's loads are in flight together
  constexpr int NE = TP * TP * CPP, NPT = (NE + 255) / 256;
  uint4 pre[NPT];
  bool pok[NPT];
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const int e = min((int)threadIdx.x + 256 * k, NE - 1);
    const int pix = e / CPP, ch = e - pix * CPP;
    const int r = pix / TP, c = pix - r * TP;
    const int gi = ti0 + r - 1, gj = tj0 + c - 1;
    pok[k] = gi >= 0 && gi < H && gj >= 0 && gj < W;
    const int ci = min(max(gi, 0), H - 1), cj = min(max(gj, 0), W - 1);
    pre[k] = *reinterpret_cast<const uint4*>(x.p + ((size_t)ci * W + cj) * x.cs + x.co + ch * 8);
  }
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const int e = threadIdx.x + 256 * k;
    if (e < NE) {
      const int pix = e / CPP, ch = e - pix * CPP;
      const int c = pix - (pix / TP) * TP;
      *reinterpret_cast<uint4*>(sx + sxo(pix, c, ch * 8)) = pok[k] ? pre[k] : uint4{0u, 0u, 0u, 0u};
    }
  }
  __syncthreads();
  f32x4_t acc[4][NB];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  if constexpr (kPre) {
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int dy = FWD ? tap / 3 - 1 : 1 - tap / 3;
      const int dx = FWD ? tap % 3 - 1 : 1 - tap % 3;
#pragma unroll
      for (int c = 0; c < NCK; ++c) {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const int r = wave * 4 + a;
          const typename F::T af =
              F::load(sx + sxo((r + 1 + dy) * TP + px + 1 + dx, px + 1 + dx, c * KC + kq));
#pragma unroll
          for (int b = 0; b < NB; ++b) acc[a][b] = F::mfma(wpre[tap][c][b], af, acc[a][b]);
        }
      }
    }
  }
#pragma unroll 1
  for (int tap = 0; tap < (kPre ? 0 : 9); ++tap) {
    const int dy = FWD ? tap / 3 - 1 : 1 - tap / 3;
    const int dx = FWD ? tap % 3 - 1 : 1 - tap % 3;
#pragma unroll
    for (int c0 = 0; c0 < KCH; c0 += KC) {
      typename F::T bf[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b)
        bf[b] = F::load(w + ((size_t)(tap * NO + b * 16 + px) * KCH + c0 + kq));
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int r = wave * 4 + a;
        const typename F::T af =
            F::load(sx + sxo((r + 1 + dy) * TP + px + 1 + dx, px + 1 + dx, c0 + kq));
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = F::mfma(bf[b], af, acc[a][b]);
      }
    }
  }
  const int cq = 4 * (lane >> 4);  // first of the lane's four channels
  const bool pool = FWD && pout.p;
  float ps[NB][4];  // pooling: the column sums of the lane's four rows
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int u = 0; u < 4; ++u) ps[b][u] = 0.f;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const size_t m = (size_t)(ti0 + wave * 4 + a) * W + tj0 + px;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int n = b * 16 + cq;
      if (FWD) {
        uint32_t h[2];
#pragma unroll
        for (int t = 0; t < 4; t += 2) {
          float v0 = acc[a][b][t] + bias[n + t];
          float v1 = acc[a][b][t + 1] + bias[n + t + 1];
          if (relu) {
            v0 = v0 > 0.f ? v0 : 0.f;
            v1 = v1 > 0.f ? v1 : 0.f;
          }
          h[t >> 1] = bf16x2_bits(v0, v1);
        }
        if (pool) {
#pragma unroll
          for (int u = 0; u < 4; ++u) ps[b][u] += bf16x2_half(h[u >> 1], u & 1);
        }
        *reinterpret_cast<uint2*>(yb.p + m * yb.cs + yb.co + n) = uint2{h[0], h[1]};
      } else {
        *reinterpret_cast<float4*>(yf.p + m * yf.cs + yf.co + n) =
            float4{acc[a][b][0], acc[a][b][1], acc[a][b][2], acc[a][b][3]};
      }
    }
  }
  if (pool) {
    // the four columns of a window by lane shuffles, then / 16 and rounded
    // (the sum of 16 non-negative BF16 values is exact in FP32)
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        float v = ps[b][u];
        v += __shfl_xor(v, 1, 64);
        v += __shfl_xor(v, 2, 64);
        ps[b][u] = v;
      }
    if ((px & 3) == 0) {
      const int mp = (ti0 / 4 + wave) * pout.W + tj0 / 4 + px / 4;
#pragma unroll
      for (int b = 0; b < NB; ++b)
        *reinterpret_cast<uint2*>(pout.p + (size_t)mp * pout.cs + pout.co + b * 16 + cq) =
            uint2{bf16x2_bits(ps[b][0] * (1.f / 16.f), ps[b][1] * (1.f / 16.f)),
                  bf16x2_bits(ps[b][2] * (1.f / 16.f), ps[b][3] * (1.f / 16.f))};
    }
  }
}

// k_conv3_tile for 16 input channels: the nine taps are paired into five
// K = 32 MFMAs (lanes 0-31 take the pair's first tap, 32-63 its second; the
// fifth pair's second half is zero), and all weight fragments are loaded
// before the tile is staged, so their latency overlaps the staging.  A block
// covers 4*ROWS x 16 pixels (wave w: rows ROWS*w .. ROWS*w + ROWS - 1).
//
// With pout.p set (forward, ROWS = 4: wave w's rows are one 4-row pooling
// band) the epilogue also writes the 4 x 4 average pooling of the BF16
// output into pout: rows summed in the lane, the four columns of a window by
// lane shuffles, then / 16 and rounded (the sum of 16 non-negative BF16
// values is exact in FP32, as in k_pool_fwd).
template <int NO, bool FWD, int ROWS>
__global__ __launch_bounds__(256) void k_conv3_tile16(View x, View yb, GView yf,
                                                      const bf16* __restrict__ w,
                                                      const float* __restrict__ bias,
                                                      int relu, View pout) {
  constexpr int KCH = 16, NB = NO / 16, TP = 18, TR = 4 * ROWS + 2, CPP = 2;
  typedef Frag<32> F;
  __shared__ __attribute__((aligned(16))) bf16 sx[TR * TP * KCH];
  const int tj0 = blockIdx.x * 16, ti0 = blockIdx.y * 4 * ROWS;
  const int H = x.H, W = x.W;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int px = lane & 15, half = lane >> 5, kc = 8 * ((lane >> 4) & 1);
  typename F::T wf[5][NB];
#pragma unroll
  for (int pr = 0; pr < 5; ++pr) {
    const int tap = 2 * pr + half;
#pragma unroll
    for (int b = 0; b < NB; ++b)
      wf[pr][b] = tap < 9 ? F::load(w + ((size_t)(tap * NO + b * 16 + px) * KCH + kc))
                          : F::zero();
  }
  for (int e = threadIdx.x; e < TR * TP * CPP; e += 256) {
    const int pix = e / CPP, ch = e - pix * CPP;
    const int r = pix / TP, c = pix - r * TP;
    const int gi = ti0 + r - 1, gj = tj0 + c - 1;
    uint4 v = {0u, 0u, 0u, 0u};
    if (gi >= 0 && gi < H && gj >= 0 && gj < W)
      v = *reinterpret_cast<const uint4*>(x.p + ((size_t)gi * W + gj) * x.cs + x.co + ch * 8);
    *reinterpret_cast<uint4*>(sx + pix * KCH + ch * 8) = v;
  }
  __syncthreads();
  f32x4_t acc[ROWS][NB];
#pragma unroll
  for (int a = 0; a < ROWS; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int pr = 0; pr < 5; ++pr) {
    const int tap = 2 * pr + half;
    const int tp = tap < 9 ? tap : 4;  // padded lanes: any in-tile position, zeroed below
    const int dy = FWD ? tp / 3 - 1 : 1 - tp / 3;
    const int dx = FWD ? tp % 3 - 1 : 1 - tp % 3;
#pragma unroll
    for (int a = 0; a < ROWS; ++a) {
      const int r = wave * ROWS + a;
      typename F::T xf = F::load(sx + ((r + 1 + dy) * TP + px + 1 + dx) * KCH + kc);
      if (tap >= 9) xf = F::zero();
#pragma unroll
      for (int b = 0; b < NB; ++b) acc[a][b] = F::mfma(wf[pr][b], xf, acc[a][b]);
    }
  }
  const int cq = 4 * (lane >> 4);
  float ps[NB][4];
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int t = 0; t < 4; ++t) ps[b][t] = 0.f;
#pragma unroll
  for (int a = 0; a < ROWS; ++a) {
    const size_t m = (size_t)(ti0 + wave * ROWS + a) * W + tj0 + px;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int n = b * 16 + cq;
      if (FWD) {
        float v[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          v[t] = acc[a][b][t] + bias[n + t];
          if (relu) v[t] = v[t] > 0.f ? v[t] : 0.f;
        }
        const uint2 h = uint2{bf16x2_bits(v[0], v[1]), bf16x2_bits(v[2], v[3])};
#pragma unroll
        for (int t = 0; t < 4; ++t) ps[b][t] += bf16x2_half(t < 2 ? h.x : h.y, t & 1);
        *reinterpret_cast<uint2*>(yb.p + m * yb.cs + yb.co + n) = h;
      } else {
        *reinterpret_cast<float4*>(yf.p + m * yf.cs + yf.co + n) =
            float4{acc[a][b][0], acc[a][b][1], acc[a][b][2], acc[a][b][3]};
      }
    }
  }
  if constexpr (FWD && ROWS == 4) {
    if (pout.p) {
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          float v = ps[b][t];
          v += __shfl_xor(v, 1, 64);
          v += __shfl_xor(v, 2, 64);
          ps[b][t] = v;
        }
      if ((px & 3) == 0) {
        const size_t mp = (size_t)(ti0 / 4 + wave) * pout.W + tj0 / 4 + px / 4;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const bf16 q0 = f2bf(ps[b][0] * (1.f / 16.f)), q1 = f2bf(ps[b][1] * (1.f / 16.f));
          const bf16 q2 = f2bf(ps[b][2] * (1.f / 16.f)), q3 = f2bf(ps[b][3] * (1.f / 16.f));
          *reinterpret_cast<uint2*>(pout.p + mp * pout.cs + pout.co + b * 16 + cq) =
              uint2{(uint32_t)q0 | ((uint32_t)q1 << 16), (uint32_t)q2 | ((uint32_t)q3 << 16)};
        }
      }
    }
  }
}

// Weight fragments of a 3x3 convolution with 16 input channels for the
// paired-tap MFMAs of k_conv3_tile16: pair pr holds tap 2pr in lanes 0-31
// and tap 2pr + 1 in lanes 32-63 (the fifth pair's second half is zero).
template <int NO, int NB>
__device__ __forceinline__ void conv16_weights(const bf16* __restrict__ w, int lane,
                                               typename Frag<32>::T (&wf)[5][NB]) {
  typedef Frag<32> F;
  const int px = lane & 15, half = lane >> 5, kc = 8 * ((lane >> 4) & 1);
#pragma unroll
  for (int pr = 0; pr < 5; ++pr) {
    const int tap = 2 * pr + half;
#pragma unroll
    for (int b = 0; b < NB; ++b)
      wf[pr][b] = tap < 9 ? F::load(w + ((size_t)(tap * NO + b * 16 + px) * 16 + kc))
                          : F::zero();
  }
}

// The convolution sums of tile rows 4 wave .. 4 wave + 3 (16 pixels each)
// from an LDS tile sx of 18 x 18 pixels x 16 channels (halo included):
// acc[a][b] = D[channels 16b..16b+15][pixels of row 4 wave + a], taps in
// pair order.
template <int NB, bool FWD>
__device__ __forceinline__ void conv16_mma(const bf16* sx,
                                           const typename Frag<32>::T (&wf)[5][NB],
                                           int lane, int wave, f32x4_t (&acc)[4][NB]) {
  typedef Frag<32> F;
  constexpr int TP = 18, KCH = 16;
  const int px = lane & 15, half = lane >> 5, kc = 8 * ((lane >> 4) & 1);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  // row by row, the next row's five fragments read from LDS before this
  // row's MFMAs (each accumulator still takes the pairs in order 0..4)
  // the pair's two taps are compile-time constants: a lane selects its
  // tap's offset (a runtime tap cost a division by 3 per fragment)
  auto tap_off = [](int tap) {
    const int tp = tap < 9 ? tap : 4;  // the padded tap reads any in-tile value, zeroed
    const int dy = FWD ? tp / 3 - 1 : 1 - tp / 3;
    const int dx = FWD ? tp % 3 - 1 : 1 - tp % 3;
    return (dy * TP + dx) * KCH;
  };
  const bf16* base = sx + ((wave * 4 + 1) * TP + px + 1) * KCH + kc;
  auto frag = [&](int pr, int a) {
    const int off = half ? tap_off(2 * pr + 1) : tap_off(2 * pr);
    typename F::T xf = F::load(base + a * TP * KCH + off);
    return (half && 2 * pr + 1 >= 9) ? F::zero() : xf;
  };
  typename F::T xf[5];
#pragma unroll
  for (int pr = 0; pr < 5; ++pr) xf[pr] = frag(pr, 0);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    typename F::T xn[5];
#pragma unroll
    for (int pr = 0; pr < 5; ++pr) xn[pr] = a < 3 ? frag(pr, a + 1) : xf[pr];
#pragma unroll
    for (int pr = 0; pr < 5; ++pr)
#pragma unroll
      for (int b = 0; b < NB; ++b) acc[a][b] = F::mfma(wf[pr][b], xf[pr], acc[a][b]);
#pragma unroll
    for (int pr = 0; pr < 5; ++pr) xf[pr] = xn[pr];
  }
}

// Constants of the forward epilogue: the lane's four biases per column block
// and, for the fused output layer, its four weights, the output bias and
// the output scale (zero when wout is null).
template <int NB>
struct Conv16Epi {
  float bias[NB][4];
  float wo[4], wo16, ysc;
  __device__ __forceinline__ void load(const float* __restrict__ b, const float* __restrict__ wout,
                                       const float* __restrict__ yscale, int lane) {
#pragma unroll
    for (int k = 0; k < NB; ++k)
#pragma unroll
      for (int u = 0; u < 4; ++u) bias[k][u] = b[k * 16 + 4 * (lane >> 4) + u];
#pragma unroll
    for (int u = 0; u < 4; ++u) wo[u] = 0.f;
    wo16 = 0.f;
    ysc = 0.f;
    if (wout) {
#pragma unroll
      for (int u = 0; u < 4; ++u) wo[u] = wout[4 * (lane >> 4) + u];
      wo16 = wout[16];
      ysc = *yscale;
    }
  }
};

// Forward epilogue of a 16 x 16 tile at (ti0, tj0) of a W-wide level:
// relu(acc + bias) rounded to BF16 and stored into yb (STORE); the 4 x 4
// average pooling into pout (POOL: rows summed in the lane, the four
// columns of a window by lane shuffles, then / 16 and rounded: the sum of 16
// non-negative BF16 values is exact in FP32, as in k_pool_fwd); the 1 x 1
// output layer y = (wo16 + sum_c wo[c] out_c) * ysc into yout (OUT, NB = 1),
// the four channel groups of a pixel summed by lane shuffles.
enum { EPI_STORE = 0, EPI_POOL = 1, EPI_OUT = 2 };
template <int NB, int EPI>
__device__ __forceinline__ void conv16_epilogue(const f32x4_t (&acc)[4][NB],
                                                const Conv16Epi<NB>& ep, int relu, int ti0,
                                                int tj0, int W, View yb, View pout,
                                                float* __restrict__ yout, int lane, int wave) {
  constexpr bool STORE = EPI != EPI_OUT, POOL = EPI == EPI_POOL, OUT = EPI == EPI_OUT;
  static_assert(!OUT || NB == 1, "the output layer follows a 16-channel convolution");
  const int px = lane & 15, cq = 4 * (lane >> 4);
  float ps[NB][4];
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int q = 0; q < 4; ++q) ps[b][q] = 0.f;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int m = (ti0 + wave * 4 + a) * W + tj0 + px;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int n = b * 16 + cq;
      float v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        v[u] = acc[a][b][u] + ep.bias[b][u];
        if (relu) v[u] = v[u] > 0.f ? v[u] : 0.f;
      }
      const uint32_t h0 = bf16x2_bits(v[0], v[1]), h1 = bf16x2_bits(v[2], v[3]);
      float q[4];  // the rounded values
#pragma unroll
      for (int u = 0; u < 4; ++u) q[u] = bf16x2_half(u < 2 ? h0 : h1, u & 1);
      if (POOL) {
#pragma unroll
        for (int u = 0; u < 4; ++u) ps[b][u] += q[u];
      }
      if (STORE) *reinterpret_cast<uint2*>(yb.p + (size_t)m * yb.cs + yb.co + n) = uint2{h0, h1};
      if (OUT) {
        float part = 0.f;
#pragma unroll
        for (int u = 0; u < 4; ++u) part += ep.wo[u] * q[u];
        part += __shfl_xor(part, 16, 64);
        part += __shfl_xor(part, 32, 64);
        if (lane < 16) yout[m] = (ep.wo16 + part) * ep.ysc;
      }
    }
  }
  if (POOL) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        float v = ps[b][u];
        v += __shfl_xor(v, 1, 64);
        v += __shfl_xor(v, 2, 64);
        ps[b][u] = v;
      }
    if ((px & 3) == 0) {
      const int mp = (ti0 / 4 + wave) * pout.W + tj0 / 4 + px / 4;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const bf16 q0 = f2bf(ps[b][0] * (1.f / 16.f)), q1 = f2bf(ps[b][1] * (1.f / 16.f));
        const bf16 q2 = f2bf(ps[b][2] * (1.f / 16.f)), q3 = f2bf(ps[b][3] * (1.f / 16.f));
        *reinterpret_cast<uint2*>(pout.p + (size_t)mp * pout.cs + pout.co + b * 16 + cq) =
            uint2{(uint32_t)q0 | ((uint32_t)q1 << 16), (uint32_t)q2 | ((uint32_t)q3 << 16)};
      }
    }
  }
}

// The tiles of a persistent launch: XCD-aware order.  A launch hands blocks
// to the 8 XCDs round robin, so with the grid a multiple of 8 the blocks of
// XCD x walk the x-th contiguous eighth of the tiles together (neighbouring
// tiles, whose halo lines they share, then sit in one L2); otherwise the
// plain order.  Block b takes first, first + per, ... below end.
struct TileWalk {
  int first, per, end;
  __device__ __forceinline__ TileWalk(int ntiles) {
    const int G = gridDim.x;
    const bool xo = (G & 7) == 0;
    per = xo ? G >> 3 : G;
    const int span = xo ? (ntiles + 7) >> 3 : ntiles;
    const int base = xo ? (int)(blockIdx.x & 7) * span : 0;
    end = min(base + span, ntiles);
    first = base + (xo ? (int)(blockIdx.x >> 3) : (int)blockIdx.x);
  }
};

// Persistent form of k_conv3_tile16 (ROWS = 4): a block walks 16 x 16 tiles
// (numbered along rows, in TileWalk order) and loads its next tile into
// registers while it computes the current one from LDS, so the staging
// latency overlaps the MFMAs; the weights are loaded once per block.
//
// The forward epilogue EPI: the output stored (EPI_STORE), stored and pooled
// into pout (EPI_POOL), or, NO = 16, only the network's 1 x 1 output layer
// applied to it into yout (EPI_OUT: no later kernel reads the layer's own
// output when the network is only applied).
template <int NO, bool FWD, int EPI>
__global__ __launch_bounds__(256, NO == 16 ? 4 : 1) void k_conv3_tile16p(View x, View yb, GView yf,
                                                       const bf16* __restrict__ w,
                                                       const float* __restrict__ bias,
                                                       int relu, View pout, int ntx,
                                                       int ntiles, const float* __restrict__ wout,
                                                       float* __restrict__ yout,
                                                       const float* __restrict__ yscale) {
  constexpr int KCH = 16, NB = NO / 16, TP = 18, CPP = 2;
  constexpr int NE = TP * TP * CPP, NPT = (NE + 255) / 256;
  typedef Frag<32> F;
  __shared__ __attribute__((aligned(16))) bf16 sx[TP * TP * KCH];
  const int H = x.H, W = x.W;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  typename F::T wf[5][NB];
  conv16_weights<NO, NB>(w, lane, wf);
  // Every global load inside the tile loop is the next tile's prefetch,
  // issued unconditionally from a clamped, valid address (masked when
  // staged): the epilogue's constants are loaded before the loop, so the
  // prefetch stays in flight through the MFMAs and the epilogue (a load
  // under a branch, or a later load that must be waited for, makes the
  // compiler drain it early).
  Conv16Epi<NB> ep;  // forward only (the input gradient has no bias)
  if (FWD) ep.load(bias, EPI == EPI_OUT ? wout : nullptr, yscale, lane);
  uint4 pre[NPT];
  bool pok[NPT];
  auto fetch = [&](int t) {
    const int ty = t / ntx;
    const int tj0 = (t - ty * ntx) * 16, ti0 = ty * 16;
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
      const int e = min((int)threadIdx.x + 256 * k, NE - 1);
      const int pix = e / CPP, ch = e - pix * CPP;
      const int r = pix / TP, c = pix - r * TP;
      const int gi = ti0 + r - 1, gj = tj0 + c - 1;
      pok[k] = gi >= 0 && gi < H && gj >= 0 && gj < W;
      const int ci = min(max(gi, 0), H - 1), cj = min(max(gj, 0), W - 1);
      pre[k] = *reinterpret_cast<const uint4*>(x.p + ((size_t)ci * W + cj) * x.cs + x.co + ch * 8);
    }
  };
  const TileWalk tw(ntiles);
  if (tw.first < tw.end) fetch(tw.first);
  for (int t = tw.first; t < tw.end; t += tw.per) {
    const int ty = t / ntx;
    const int tj0 = (t - ty * ntx) * 16, ti0 = ty * 16;
    __syncthreads();  // the previous tile's LDS reads are done
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
      const int e = threadIdx.x + 256 * k;
      if (e < NE)
        *reinterpret_cast<uint4*>(sx + e * 8) = pok[k] ? pre[k] : uint4{0u, 0u, 0u, 0u};
    }
    __syncthreads();
    // the last tile of the block re-reads its own tile (cache hits)
    fetch(t + tw.per < tw.end ? t + tw.per : t);
    f32x4_t acc[4][NB];
    conv16_mma<NB, FWD>(sx, wf, lane, wave, acc);
    if constexpr (FWD) {
      conv16_epilogue<NB, EPI>(acc, ep, relu, ti0, tj0, W, yb, pout, yout, lane, wave);
    } else {
      const int px = lane & 15, cq = 4 * (lane >> 4);
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const size_t m = (size_t)(ti0 + wave * 4 + a) * W + tj0 + px;
#pragma unroll
        for (int b = 0; b < NB; ++b)
          *reinterpret_cast<float4*>(yf.p + m * yf.cs + yf.co + b * 16 + cq) =
              float4{acc[a][b][0], acc[a][b][1], acc[a][b][2], acc[a][b][3]};
      }
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
      for (int b = 0; b < NB; ++b) acc[a][b] = F::mfma(bf[b], af, acc[a][b]);
    }
  }
  // D[(ab, co)][coarse pixel]: four consecutive co of one fine pixel per
  // lane (COUT is a multiple of 16, so the four share ab)
#pragma unroll
  for (int a = 0; a < MP; ++a) {
    const int m = m0 + a * 16 + (lane & 15);
    if (m >= M) continue;
    const int I = m / x.W, J = m - I * x.W;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int n = n0 + b * 16 + 4 * (lane >> 4);
      const int ab = n / COUT, co = n - ab * COUT;
      const int ai = ab >> 2, bj = ab & 3;
      const size_t f = (size_t)(4 * I + ai) * y.W + (4 * J + bj);
      uint32_t h[2];
#pragma unroll
      for (int t = 0; t < 4; t += 2)
        h[t >> 1] = bf16x2_bits(acc[a][b][t] + bias[co + t], acc[a][b][t + 1] + bias[co + t + 1]);
      *reinterpret_cast<uint2*>(y.p + f * y.cs + y.co + co) = uint2{h[0], h[1]};
    }
  }
}

// up-convolution input gradient: dx[m][ci] = sum_{ab,co} dy[fine(m,ab)][co]
// WuT[ab][ci][co]
template <int CIN, int COUT, int NB, int MP>
__global__ __launch_bounds__(256) void k_up_dgrad(View dy, GView dx,
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
          af = F::load(dy.p + f * dy.cs + dy.co + c0 + kq);
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = F::mfma(bf[b], af, acc[a][b]);
      }
    }
  }
  // D[channel][pixel]: four consecutive channels of one pixel per lane
#pragma unroll
  for (int a = 0; a < MP; ++a) {
    const int m = m0 + a * 16 + (lane & 15);
    if (m >= M) continue;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int ci = n0 + b * 16 + 4 * (lane >> 4);
      *reinterpret_cast<float4*>(dx.p + (size_t)m * dx.cs + dx.co + ci) =
          float4{acc[a][b][0], acc[a][b][1], acc[a][b][2], acc[a][b][3]};
    }
  }
}

// ---------------------------------------------------------------------
// Weight gradients on MFMA.  dW[tap][co][ci] = sum_p g(p, tap, co) x(p, tap,
// ci) with K = pixels: each block takes one tap and a chunk of pixels,
// stages 64-pixel slices of both operands in LDS transposed (channel-major,
// pixels contiguous) so every fragment is one 16-byte LDS read, and writes
// FP32 partial sums per chunk (reduced in a fixed order afterwards).
// conv: g = gq[p],           x = X[p + off(tap)], taps 0..8
// up:   g = gq[fine(p, tap)], x = X[p],           taps 0..15
// (gq: the masked BF16 output gradient of k_mask_grad)
template <int CIN, int COUT, bool UP>
__global__ __launch_bounds__(256) void k_wgrad(View dy, View x, int pix_per_chunk,
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
      bf16 v = 0;
      if (p < p1) {
        size_t f = (size_t)p;
        if (UP) {
          const int I = p / x.W, J = p - I * x.W;
          f = (size_t)(4 * I + ai) * dy.W + (4 * J + bj);
        }
        v = dy.p[f * dy.cs + dy.co + c];
      }
      sg[c * PITCH + pp] = v;
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

// Weight gradient of a 3x3 convolution on 16 x 16 pixel tiles through LDS,
// all nine taps per block (levels whose sides are multiples of 16 and whose
// tiles fit: 512 COUT + 1728 CIN bytes <= 64 KiB):
//   dW[tap][co][ci] += sum_p gq[p][co] x[p + off(tap)][ci]
// gT holds the tile's masked BF16 output gradient channel-major; xs holds
// the input tile with its halo rows in three copies shifted by one column
// (dx = -1, 0, +1), channel-major, so that every MFMA operand fragment is an
// aligned 16-byte LDS read.  Block b walks tiles b, b + G, ...; the (tap,
// co-block, ci-block) products are spread over the four waves; partials
// part[((b * 9 + tap) * COUT + co) * CIN + ci] are summed over blocks
// afterwards in a fixed order.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_wgrad_tile(View gq, View x,
                                                    float* __restrict__ part) {
  constexpr int TCO = COUT / 16, TCI = CIN / 16, NCOMB = 9 * TCO * TCI;
  constexpr int CPW = (NCOMB + 3) / 4;  // products per wave
  __shared__ __attribute__((aligned(16))) bf16 gT[COUT * 256];
  __shared__ __attribute__((aligned(16))) bf16 xs[3 * CIN * 18 * 16];
  typedef Frag<32> F;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int H = x.H, W = x.W, ntj = W / 16, ntiles = (H / 16) * ntj;
  f32x4_t acc[CPW];
#pragma unroll
  for (int q = 0; q < CPW; ++q) acc[q] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int ti0 = (t / ntj) * 16, tj0 = (t % ntj) * 16;
    __syncthreads();  // the previous tile's reads are done
    // gT[co][p]: 256 pixels x COUT channels, 8-channel chunks
    for (int e = threadIdx.x; e < 256 * (COUT / 8); e += 256) {
      const int p = e / (COUT / 8), ch = e - p * (COUT / 8);
      const size_t m = (size_t)(ti0 + p / 16) * W + tj0 + (p % 16);
      bf16 v[8];
      __builtin_memcpy(v, gq.p + m * gq.cs + gq.co + ch * 8, 16);
#pragma unroll
      for (int k = 0; k < 8; ++k) gT[(ch * 8 + k) * 256 + p] = v[k];
    }
    // xs[d][ci][row][c] = x[ti0 + row - 1][tj0 + c + d - 1], halo pixel
    // (row, hc) with hc = c + d lands in up to three copies
    for (int e = threadIdx.x; e < 18 * 18 * (CIN / 8); e += 256) {
      const int pix = e / (CIN / 8), ch = e - pix * (CIN / 8);
      const int row = pix / 18, hc = pix - row * 18;
      const int gi = ti0 + row - 1, gj = tj0 + hc - 1;
      bf16 v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (gi >= 0 && gi < H && gj >= 0 && gj < W)
        __builtin_memcpy(v, x.p + ((size_t)gi * W + gj) * x.cs + x.co + ch * 8, 16);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int c = hc - d;
        if (c < 0 || c >= 16) continue;
#pragma unroll
        for (int k = 0; k < 8; ++k)
          xs[((d * CIN + ch * 8 + k) * 18 + row) * 16 + c] = v[k];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
      const int comb = wave + 4 * q;
      if (comb >= NCOMB) break;
      const int tap = comb / (TCO * TCI), rem = comb - tap * (TCO * TCI);
      const int tco = rem / TCI, tci = rem - tco * TCI;
      const int d = tap % 3, rshift = (tap / 3) * 16;  // rows dy + 1
      const bf16* ga = gT + (tco * 16 + (lane & 15)) * 256 + 8 * (lane >> 4);
      const bf16* xb = xs + ((d * CIN + tci * 16 + (lane & 15)) * 18) * 16 + rshift +
                       8 * (lane >> 4);
#pragma unroll
      for (int k0 = 0; k0 < 256; k0 += 32)
        acc[q] = F::mfma(F::load(ga + k0), F::load(xb + k0), acc[q]);
    }
  }
  float* out = part + (size_t)blockIdx.x * 9 * COUT * CIN;
#pragma unroll
  for (int q = 0; q < CPW; ++q) {
    const int comb = wave + 4 * q;
    if (comb >= NCOMB) break;
    const int tap = comb / (TCO * TCI), rem = comb - tap * (TCO * TCI);
    const int tco = rem / TCI, tci = rem - tco * TCI;
    const int ci = tci * 16 + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int co = tco * 16 + 4 * (lane >> 4) + r;
      out[((size_t)tap * COUT + co) * CIN + ci] = acc[q][r];
    }
  }
}

}  // namespace unet
}  // namespace cpgpu
