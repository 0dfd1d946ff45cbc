// U-Net preconditioner on gfx950 (no reference counterpart): parameters,
// forward, backward, Adam, energy-norm training loss, and the C-ABI.
//
// Architecture (4 levels, 16-32-64-128 channels, 3x3 convs + ReLU, 4x4
// average pooling with stride 4 and 4x4 stride-4 transposed convs, skip
// connections by channel concatenation, 1x1 linear output):
//   L0 16 : c0a 1->16, c0b 16->16 ---------------------------- skip --+
//   L1 32 :   pool, c1a 16->32, c1b 32->32 ------------------ skip -+ |
//   L2 64 :     pool, c2a 32->64, c2b 64->64 ------------- skip -+ | |
//   L3 128:       pool, c3a 64->128, c3b 128->128              | | |
//   L2    :     up 128->64, concat, d2a 128->64, d2b 64->64 <--+ | |
//   L1    :   up 64->32, concat, d1a 64->32, d1b 32->32 <--------+ |
//   L0    : up 32->16, concat, d0a 32->16, d0b 16->16 <------------+
//   out   : 1x1 16->1
// Stride-4 pooling gives the 1024^2 grid a 16^2 bottleneck, so one
// application couples the whole domain.  Grid sides must be multiples of 64.
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <memory>

#include "poisson.hpp"
#include "unet_kernels.hpp"

namespace cpgpu {
namespace unet {

enum Kind { K_FIRST, K_CONV, K_UP, K_OUT };

struct LayerDef {
  Kind kind;
  int cin, cout;
};

// fixed layer list (order = parameter order)
static const LayerDef kLayers[] = {
    {K_FIRST, 2, 16},  {K_CONV, 16, 16},  {K_CONV, 16, 32},  {K_CONV, 32, 32},
    {K_CONV, 32, 64},  {K_CONV, 64, 64},  {K_CONV, 64, 128}, {K_CONV, 128, 128},
    {K_UP, 128, 64},   {K_CONV, 128, 64}, {K_CONV, 64, 64},  {K_UP, 64, 32},
    {K_CONV, 64, 32},  {K_CONV, 32, 32},  {K_UP, 32, 16},    {K_CONV, 32, 16},
    {K_CONV, 16, 16},  {K_OUT, 16, 1}};
constexpr int kNL = sizeof(kLayers) / sizeof(kLayers[0]);

inline int ntaps(Kind k) { return k == K_UP ? 16 : (k == K_OUT ? 1 : 9); }

// resolution level (pixels / 16^level) of each layer's output
static const int kLevel[] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 2, 1, 1, 1, 0, 0, 0, 0};
static_assert(sizeof(kLevel) / sizeof(kLevel[0]) == kNL, "one level per layer");
inline int level_of(int l) { return kLevel[l]; }

// chunks of the weight-gradient GEMM over pixels (its K dimension): at least
// 256 pixels each, at most 256 chunks, so that small levels still launch
// many blocks
inline int wg_chunks(int M) {
  int c = M / 256;
  return c < 1 ? 1 : (c < 256 ? c : 256);
}

constexpr int kRedBlocks = 1024;  // blocks of the per-channel reductions

// ---------------------------------------------------------------------
// small kernels

// Wall code of cell (i, j), the network's second input channel: the sum
// over the walls the cell touches of +1 (Dirichlet) or -1 (Neumann); walls
// packs the types of the x-, x+, y-, y+ segments, two bits each (0: no
// wall information).
__device__ __forceinline__ float wall_code(int i, int j, int H, int W, int walls) {
  auto v = [](int c) { return c == CPGPU_BC_DIRICHLET ? 1.f : (c == CPGPU_BC_NEUMANN ? -1.f : 0.f); };
  float s = 0.f;
  if (i == 0) s += v(walls & 3);
  if (i == H - 1) s += v((walls >> 2) & 3);
  if (j == 0) s += v((walls >> 4) & 3);
  if (j == W - 1) s += v((walls >> 6) & 3);
  return s;
}

// first conv 2 -> 16 (FP32 input and wall code, direct):
// a[m][co] = relu(b + sum_tap W[tap][co][0] x + W[tap][co][1] wall)
// (branch-free: variants that skip the wall products away from the walls
// measured slower)
// With scale given, x is the unscaled input and each value is multiplied by
// 1 / *scale on the fly (as k_scale_copy would have stored it).
__global__ __launch_bounds__(256) void k_first_fwd(const float* __restrict__ x,
                                                   int H, int W, int walls, View a,
                                                   const float* __restrict__ w,
                                                   const float* __restrict__ b,
                                                   const float* __restrict__ scale) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= H * W) return;
  const int i = m / W, j = m - i * W;
  float xv[9], wv[9];
  // the nine loads from clamped valid addresses, all in flight together
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int si = min(max(i + t / 3 - 1, 0), H - 1), sj = min(max(j + t % 3 - 1, 0), W - 1);
    xv[t] = x[si * W + sj];
  }
  if (scale) {
    const float sc = 1.f / *scale;
#pragma unroll
    for (int t = 0; t < 9; ++t) xv[t] = xv[t] * sc;
  }
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int si = i + t / 3 - 1, sj = j + t % 3 - 1;
    const bool in = si >= 0 && si < H && sj >= 0 && sj < W;
    xv[t] = in ? xv[t] : 0.f;
    wv[t] = in ? wall_code(si, sj, H, W, walls) : 0.f;
  }
  // explicit fused multiply-adds: the library is built without contraction
  // (for the FP64 solver's bitwise parity), and 288 separate products and
  // sums per pixel made this layer VALU-bound
  bf16 out[16];
#pragma unroll
  for (int co = 0; co < 16; ++co) {
    float s = b[co];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      s = __builtin_fmaf(w[(t * 16 + co) * 2], xv[t], s);
      s = __builtin_fmaf(w[(t * 16 + co) * 2 + 1], wv[t], s);
    }
    out[co] = f2bf(s > 0.f ? s : 0.f);
  }
  __builtin_memcpy(a.p + (size_t)m * a.cs + a.co, out, 32);
}

// ---------------------------------------------------------------------
// Level 0 of an applied network (no backward pass follows) in two passes.
// Training keeps the layer-per-kernel forward, whose activations the
// backward pass reads; an application needs only the skip tensor between
// the encoder and the decoder at full resolution, so the two fused kernels
// below keep every other level-0 activation in LDS.  Each walks 16 x 16
// pixel tiles persistently (TileWalk order), recomputes the halo an inner
// layer needs, and repeats the per-value arithmetic of the layer-per-kernel
// path (same products, same MFMA operand assignment and tap order, same
// BF16 rounding points), so an applied network's output equals the plain
// forward's with the output layer fused.

// Encoder: the first layer (2 -> 16, FP32, as k_first_fwd with the input
// scaled by 1 / *scale as loaded) on the tile's 18 x 18 halo into LDS, then
// the 16 -> 16 convolution with k_conv3_tile16p's pooled epilogue into skip
// (the level-0 skip channels) and pout (the level-1 input).  The first
// layer's input window is 20 x 20.  Where that window avoids the boundary
// rows and columns the wall channel is zero there and its products are
// skipped: fma(w, 0, s) = s up to the sign of a zero, which the ReLU and
// every later product and sum leave without effect.  Item (pixel, channel
// half h) per thread, h wave-uniform so that the weights are scalar loads.
// With sq_part given, the input's rms is formed here from the sq_G partial
// sums of x^2 (as k_rms forms it, in every block; block 0 stores it in
// *scale for the decoder), otherwise read from *scale.
__global__ __launch_bounds__(256, 5) void k_enc0_apply(const float* __restrict__ x,
                                                      float* scale, const double* sq_part,
                                                      int sq_G, int H, int W, int walls,
                                                      const float* __restrict__ w0,
                                                      const float* __restrict__ b0,
                                                      const bf16* __restrict__ w1,
                                                      const float* __restrict__ b1,
                                                      View skip, View pout, int ntx,
                                                      int ntiles) {
  constexpr int XP = 20, TP = 18, NX = XP * XP;
  __shared__ float sxf[NX];
  __shared__ __attribute__((aligned(16))) bf16 sa[TP * TP * 16];
  typedef Frag<32> F;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  F::T wf[5][1];
  conv16_weights<16, 1>(w1, lane, wf);
  Conv16Epi<1> ep;
  ep.load(b1, nullptr, nullptr, lane);
  const int h = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 7);
  const float* w0h = w0 + 16 * h;  // this half's weights: w0h[q * 32 + 2 u + {0, 1}]
  const float* b0h = b0 + 8 * h;
  const int pl = threadIdx.x & 127;
  float pre[2];
  auto fetch = [&](int t) {
    const int ty = t / ntx;
    const int tj0 = (t - ty * ntx) * 16, ti0 = ty * 16;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int e = min((int)threadIdx.x + 256 * k, NX - 1);
      const int r = e / XP, c = e - r * XP;
      const int ci = min(max(ti0 - 2 + r, 0), H - 1), cj = min(max(tj0 - 2 + c, 0), W - 1);
      pre[k] = x[(size_t)ci * W + cj];
    }
  };
  const TileWalk tw(ntiles);
  if (tw.first < tw.end) fetch(tw.first);
  float scv;
  if (sq_part) {  // (the first tile's loads are in flight meanwhile)
    __shared__ double lds[4];
    double tot[1];
    reduce_partials<double, 1>(sq_part, sq_G, tot, lds);
    const double rms = std::sqrt(tot[0] / ((double)H * W));
    scv = rms > 0.0 ? (float)rms : 1.f;
    if (blockIdx.x == 0 && threadIdx.x == 0) *scale = scv;
  } else {
    scv = *scale;
  }
  const float sc = 1.f / scv;
  for (int t = tw.first; t < tw.end; t += tw.per) {
    const int ty = t / ntx;
    const int tj0 = (t - ty * ntx) * 16, ti0 = ty * 16;
    // the thread's indices recomputed per tile (held across the loop, the
    // addresses derived from them outgrow four waves' registers)
    int tv = threadIdx.x;
    asm volatile("" : "+v"(tv));
    const int lane = tv & 63, pl = tv & 127;
    __syncthreads();  // the previous tile's reads of sxf and sa are done
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int e = tv + 256 * k;
      if (e < NX) {
        const int r = e / XP, c = e - r * XP;
        const int gi = ti0 - 2 + r, gj = tj0 - 2 + c;
        const bool in = gi >= 0 && gi < H && gj >= 0 && gj < W;
        sxf[e] = in ? pre[k] * sc : 0.f;
      }
    }
    __syncthreads();
    fetch(t + tw.per < tw.end ? t + tw.per : t);
    const bool edge = walls != 0 && (ti0 == 0 || ti0 + 16 == H || tj0 == 0 || tj0 + 16 == W);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int p = pl + 128 * k;
      if (p < TP * TP) {
        const int r = p / TP, c = p - r * TP;
        const int gi = ti0 - 1 + r, gj = tj0 - 1 + c;
        uint32_t o[4] = {0u, 0u, 0u, 0u};
        if (gi >= 0 && gi < H && gj >= 0 && gj < W) {
          // taps in order, one row of three per iteration (not unrolled,
          // so that only that row's weights occupy scalar registers)
          float s[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) s[u] = b0h[u];
#pragma unroll 1
          for (int dy = 0; dy < 3; ++dy) {
            const float* wr = w0h + dy * 96;  // taps 3 dy .. 3 dy + 2
            float xv[3], wv[3];
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
              xv[dx] = sxf[(r + dy) * XP + c + dx];
              const int si = gi + dy - 1, sj = gj + dx - 1;
              const bool in = si >= 0 && si < H && sj >= 0 && sj < W;
              wv[dx] = edge && in ? wall_code(si, sj, H, W, walls) : 0.f;
            }
            if (edge) {
#pragma unroll
              for (int dx = 0; dx < 3; ++dx)
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                  s[u] = __builtin_fmaf(wr[dx * 32 + 2 * u], xv[dx], s[u]);
                  s[u] = __builtin_fmaf(wr[dx * 32 + 2 * u + 1], wv[dx], s[u]);
                }
            } else {
#pragma unroll
              for (int dx = 0; dx < 3; ++dx)
#pragma unroll
                for (int u = 0; u < 8; ++u)
                  s[u] = __builtin_fmaf(wr[dx * 32 + 2 * u], xv[dx], s[u]);
            }
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
            o[u] = bf16x2_bits(s[2 * u] > 0.f ? s[2 * u] : 0.f,
                               s[2 * u + 1] > 0.f ? s[2 * u + 1] : 0.f);
        }
        *reinterpret_cast<uint4*>(sa + p * 16 + 8 * h) = uint4{o[0], o[1], o[2], o[3]};
      }
    }
    __syncthreads();
    f32x4_t acc[4][1];
    conv16_mma<1, true>(sa, wf, lane, wave, acc);
    conv16_epilogue<1, EPI_POOL>(acc, ep, 1, ti0, tj0, W, skip, pout, nullptr, lane, wave);
  }
}

// Decoder: the 4 x 4 stride-4 up-convolution 32 -> 16 of the level-1 output
// d1 (as k_up_fwd) on the tile's 20 x 20 window, beside the skip channels
// (the concatenation [up | skip], 32 channels, in LDS), then the 32 -> 16
// convolution (as k_conv3_tile, nine K = 32 MFMAs in tap order) on the
// 18 x 18 halo into LDS, then the 16 -> 16 convolution with the output layer
// in k_conv3_tile16p's epilogue into yout, scaled by *yscale.  Zero padding:
// window pixels outside the grid are zero in both LDS tiles.
//   up:  the window's 6 x 6 coarse pixels x 256 columns (ab, co), K = 32;
//        wave w takes the columns of ab = 4w .. 4w + 3
//   mid: 18 rows of 16 pixels (wave w: rows R0(w) .. R0(w) + 4 or + 3, so
//        that a row's fragments serve the three rows beside it) and the
//        two right halo columns as three 16-pixel groups
__global__ __launch_bounds__(256, 3) void k_dec0_apply(View d1, View skip,
                                                      const bf16* __restrict__ wu,
                                                      const float* __restrict__ bu,
                                                      const bf16* __restrict__ wm,
                                                      const float* __restrict__ bm,
                                                      const bf16* __restrict__ wl,
                                                      const float* __restrict__ bl,
                                                      const float* __restrict__ wout,
                                                      const float* __restrict__ yscale,
                                                      float* __restrict__ yout, int ntx,
                                                      int ntiles) {
  constexpr int XP = 20, TP = 18, NC = 36;   // window side, halo side, coarse pixels
  constexpr int SP = 32;                     // window pixel pitch (BF16 values)
  constexpr int NS = XP * XP * 2;            // 16-byte skip chunks of the window
  constexpr int NSP = (NS + 255) / 256;
  // window element (row, column, channel): a pixel's four 16-byte channel
  // chunks are stored in the order chunk ^ ((column >> 1) & 3).  A 128-bit
  // LDS read serves 16 consecutive lanes per cycle; at the plain 64-byte
  // pitch the 16 pixels of a fragment read meet 2-way bank conflicts, with
  // this order none (measured with SQ_LDS_BANK_CONFLICT on gfx950; the key
  // (column >> 2) & 3 leaves them 2-way).  The key depends on the column
  // only: a column's address serves every row at a constant offset.
  auto swz = [](int row, int col, int ch) {
    return (row * XP + col) * SP + (((ch >> 3) ^ ((col >> 1) & 3)) << 3) + (ch & 7);
  };
  __shared__ __attribute__((aligned(16))) bf16 sc[XP * XP * SP];
  __shared__ __attribute__((aligned(16))) bf16 sd[TP * TP * 16];
  __shared__ __attribute__((aligned(16))) bf16 su[NC * 32];
  __shared__ __attribute__((aligned(16))) bf16 swl[9 * 16 * 16];  // last conv [tap][co][ci]
  __shared__ __attribute__((aligned(16))) bf16 swm[9 * 16 * 32];  // middle conv [tap][co][ci]
  typedef Frag<32> F;
  const int H = skip.H, W = skip.W, Hc = d1.H, Wc = d1.W;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int px = lane & 15, kq = 8 * (lane >> 4), cq = 4 * (lane >> 4);
  // the two convolutions' weights in LDS, each taken into registers in its
  // own stage
  for (int e = threadIdx.x; e < 9 * 16 * 2; e += 256)
    *reinterpret_cast<uint4*>(swl + e * 8) = *reinterpret_cast<const uint4*>(wl + e * 8);
  for (int e = threadIdx.x; e < 9 * 16 * 4; e += 256)
    *reinterpret_cast<uint4*>(swm + e * 8) = *reinterpret_cast<const uint4*>(wm + e * 8);
  float bur[4], bmr[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    bur[u] = bu[cq + u];
    bmr[u] = bm[cq + u];
  }
  Conv16Epi<1> ep;
  ep.load(bl, wout, yscale, lane);
  uint4 pre[NSP], prc;
  auto fetch = [&](int t) {
    const int ty = t / ntx;
    const int tj0 = (t - ty * ntx) * 16, ti0 = ty * 16;
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
      const int e = min((int)threadIdx.x + 256 * k, NS - 1);
      const int pix = e >> 1, ch = e & 1;
      const int r = pix / XP, c = pix - r * XP;
      const int ci = min(max(ti0 - 2 + r, 0), H - 1), cj = min(max(tj0 - 2 + c, 0), W - 1);
      pre[k] = *reinterpret_cast<const uint4*>(skip.p + ((size_t)ci * W + cj) * skip.cs +
                                               skip.co + ch * 8);
    }
    const int e = min((int)threadIdx.x, NC * 4 - 1);
    const int cp = e >> 2, ch = e & 3;
    const int I = min(max(ti0 / 4 - 1 + cp / 6, 0), Hc - 1);
    const int J = min(max(tj0 / 4 - 1 + cp % 6, 0), Wc - 1);
    prc = *reinterpret_cast<const uint4*>(d1.p + ((size_t)I * Wc + J) * d1.cs + d1.co + ch * 8);
  };
  // mid-layer rows R0 .. R0 + 4 of this wave (rows past the halo clamp to
  // its last row: a repeated row stores the same values again) and its
  // 16-pixel group of the two right halo columns (none for wave 1)
  const int R0 = wave < 2 ? 5 * wave : 10 + 4 * (wave - 2);
  const int sg = wave == 0 ? 2 : (wave == 1 ? -1 : wave - 2);
  const TileWalk tw(ntiles);
  if (tw.first < tw.end) fetch(tw.first);
  for (int t = tw.first; t < tw.end; t += tw.per) {
    const int ty = t / ntx;
    const int tj0 = (t - ty * ntx) * 16, ti0 = ty * 16;
    // the thread's indices are recomputed per tile: held across the loop,
    // every address derived from them stayed live and the kernel needed 240
    // registers (two waves per SIMD); recomputed, three waves fit
    int tv = threadIdx.x;
    asm volatile("" : "+v"(tv));
    const int px = tv & 15, kq = 8 * ((tv >> 4) & 3), cq = 4 * ((tv >> 4) & 3);
    const int lane = tv & 63;
    // tiles whose window (and coarse window) lies inside the grid, most of
    // them, skip the per-pixel bounds tests
    const bool inner = ti0 > 0 && tj0 > 0 && ti0 + 16 < H && tj0 + 16 < W;
    // A: stage the skip window and the coarse pixels.  No barrier is needed
    // before it: the previous tile's last reads of sc and su precede its
    // barrier before the final convolution.
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
      const int e = tv + 256 * k;
      if (e < NS) {
        const int pix = e >> 1, ch = e & 1;
        const int r = pix / XP, c = pix - r * XP;
        const int gi = ti0 - 2 + r, gj = tj0 - 2 + c;
        const bool in = inner || (gi >= 0 && gi < H && gj >= 0 && gj < W);
        *reinterpret_cast<uint4*>(sc + swz(r, c, 16 + ch * 8)) = in ? pre[k] : uint4{0u, 0u, 0u, 0u};
      }
    }
    if (tv < NC * 4) {
      const int cp = tv >> 2, ch = tv & 3;
      *reinterpret_cast<uint4*>(su + cp * 32 + ch * 8) = prc;
    }
    __syncthreads();
    // up-convolution weights of this wave's columns (ab = 4 wave + j: rows
    // (ab, co)), loaded before the prefetch so that waiting for them does
    // not wait for it
    F::T wuf[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) wuf[j] = F::load(wu + (size_t)((4 * wave + j) * 16 + px) * 32 + kq);
    fetch(t + tw.per < tw.end ? t + tw.per : t);
    // B: up-convolution into channels 0..15 of the window.  Fine pixel
    // (4 I + wave, 4 J + j) of coarse pixel (I, J) sits at window row
    // 4 pi + wave - 2, column 4 pj + j - 2; stores outside the window are
    // masked off.  Lanes past the 36 coarse pixels repeat the last.
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      const int cpl = min(g * 16 + px, NC - 1);
      const int pi = cpl / 6, pj = cpl - 6 * pi;
      const F::T bfr = F::load(su + cpl * 32 + kq);
      const int I = ti0 / 4 - 1 + pi, J = tj0 / 4 - 1 + pj;
      const bool cin = inner || (I >= 0 && I < Hc && J >= 0 && J < Wc);
      const int fr = 4 * pi + wave - 2;
      const bool frok = fr >= 0 && fr < XP;
      // swz(fr, fc, cq) for the window columns fc = 4 pj - 2 + j: the
      // swizzle key (fc >> 1) & 3 is 2 pj - 1 for j = 0, 1 and 2 pj for
      // j = 2, 3 (columns before the window are not stored)
      const int rb = (fr * XP + 4 * pj - 2) * SP + (cq & 7);
      const int k01 = ((cq >> 3) ^ ((2 * pj - 1) & 3)) << 3, k23 = ((cq >> 3) ^ ((2 * pj) & 3)) << 3;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const f32x4_t a = F::mfma(wuf[j], bfr, f32x4_t{0.f, 0.f, 0.f, 0.f});
        const int fc = 4 * pj + j - 2;
        const uint32_t h0 = bf16x2_bits(a[0] + bur[0], a[1] + bur[1]);
        const uint32_t h1 = bf16x2_bits(a[2] + bur[2], a[3] + bur[3]);
        if (frok && fc >= 0 && fc < XP)
          *reinterpret_cast<uint2*>(sc + rb + j * SP + (j < 2 ? k01 : k23)) =
              uint2{cin ? h0 : 0u, cin ? h1 : 0u};
      }
    }
    __syncthreads();
    // C: the 32 -> 16 convolution on the 18 x 18 halo into sd (zero outside
    // the grid)
    auto mid_store = [&](const f32x4_t& a, int R, int c, auto check) {
      const int gi = ti0 - 1 + R, gj = tj0 - 1 + c;
      bool in = true;
      if constexpr (decltype(check)::value) in = gi >= 0 && gi < H && gj >= 0 && gj < W;
      float q[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float s = a[u] + bmr[u];
        q[u] = s > 0.f ? s : 0.f;
      }
      const uint32_t h0 = bf16x2_bits(q[0], q[1]);
      const uint32_t h1 = bf16x2_bits(q[2], q[3]);
      *reinterpret_cast<uint2*>(sd + (R * TP + c) * 16 + cq) = uint2{in ? h0 : 0u, in ? h1 : 0u};
    };
    {
      // five row slots and (waves 0, 2, 3) one group of the right halo
      // columns; wave 1's sixth slot repeats its last row
      const int ii = min(max(sg, 0) * 16 + px, 2 * TP - 1);
      const int Rs = sg >= 0 ? ii >> 1 : min(R0 + 4, TP - 1);
      const int cs = sg >= 0 ? 16 + (ii & 1) : px;
      f32x4_t acc[6];
#pragma unroll
      for (int a = 0; a < 6; ++a) acc[a] = f32x4_t{0.f, 0.f, 0.f, 0.f};
      // software-pipelined: the next tap's weight and six fragments are
      // read from LDS before this tap's six MFMAs, so the reads' latency
      // overlaps them (read just before use, every MFMA waited for its read)
      // fragment addresses split into a row part per slot and a column part
      // per slot kind (the five row slots share column px) and dx, so that a
      // read costs one addition and the tap's row step is the read's
      // immediate offset (swz(row, col, kq) = row XP SP + colpart(col))
      auto colpart = [&](int col) {
        return col * SP + (((kq >> 3) ^ ((col >> 1) & 3)) << 3);
      };
      int rowp[6], colp[2][3];
#pragma unroll
      for (int a = 0; a < 6; ++a) rowp[a] = ((a < 5 ? min(R0 + a, TP - 1) : Rs) + 1) * XP * SP;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        colp[0][d] = colpart(px + d);
        colp[1][d] = colpart(cs + d);
      }
      auto frag = [&](int tap, int a) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        return F::load(sc + (rowp[a] + colp[a < 5 ? 0 : 1][dx + 1]) + dy * XP * SP);
      };
      F::T wt = F::load(swm + px * 32 + kq), af[6];
#pragma unroll
      for (int a = 0; a < 6; ++a) af[a] = frag(0, a);
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) {
        F::T wn = wt, an[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) an[a] = af[a];
        if (tap < 8) {
          wn = F::load(swm + ((tap + 1) * 16 + px) * 32 + kq);
#pragma unroll
          for (int a = 0; a < 6; ++a) an[a] = frag(tap + 1, a);
        }
#pragma unroll
        for (int a = 0; a < 6; ++a) acc[a] = F::mfma(wt, af[a], acc[a]);
        wt = wn;
#pragma unroll
        for (int a = 0; a < 6; ++a) af[a] = an[a];
      }
      if (inner) {
#pragma unroll
        for (int a = 0; a < 5; ++a) mid_store(acc[a], min(R0 + a, TP - 1), px, std::false_type{});
        mid_store(acc[5], Rs, cs, std::false_type{});
      } else {
#pragma unroll
        for (int a = 0; a < 5; ++a) mid_store(acc[a], min(R0 + a, TP - 1), px, std::true_type{});
        mid_store(acc[5], Rs, cs, std::true_type{});
      }
    }
    __syncthreads();
    // D: the 16 -> 16 convolution and the output layer
    F::T wf[5][1];
    conv16_weights<16, 1>(swl, lane, wf);
    f32x4_t acc[4][1];
    conv16_mma<1, true>(sd, wf, lane, wave, acc);
    conv16_epilogue<1, EPI_OUT>(acc, ep, 1, ti0, tj0, W, View{}, View{}, yout, lane, wave);
  }
}

// first conv weight/bias gradient partials: per block kFirstPart values
constexpr int kFirstPart = 9 * 16 * 2 + 16;
__global__ __launch_bounds__(256) void k_first_wgrad(const float* __restrict__ x,
                                                     int H, int W, int walls, GView g,
                                                     View a,
                                                     float* __restrict__ part) {
  // thread t owns output channel co = t % 16 and walks pixels 16 apart
  __shared__ float red[19][256];
  const int co = threadIdx.x & 15;
  float acc[19];
#pragma unroll
  for (int k = 0; k < 19; ++k) acc[k] = 0.f;
  const int M = H * W;
  for (int m = blockIdx.x * 16 + (threadIdx.x >> 4); m < M; m += gridDim.x * 16) {
    const int i = m / W, j = m - i * W;
    float gv = g.p[(size_t)m * g.cs + g.co + co];
    if (!(bf2f(a.p[(size_t)m * a.cs + a.co + co]) > 0.f)) gv = 0.f;
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int si = i + t / 3 - 1, sj = j + t % 3 - 1;
      const bool in = si >= 0 && si < H && sj >= 0 && sj < W;
      const float xv = in ? x[si * W + sj] : 0.f;
      const float wv = in ? wall_code(si, sj, H, W, walls) : 0.f;
      acc[2 * t] += gv * xv;
      acc[2 * t + 1] += gv * wv;
    }
    acc[18] += gv;
  }
#pragma unroll
  for (int k = 0; k < 19; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int s = 128; s >= 16; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int k = 0; k < 19; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    }
    __syncthreads();
  }
  // part[blk*kFirstPart + (tap*16 + co)*2 + ci] (weights [tap][co][ci]),
  // then 16 biases
  if (threadIdx.x < 16) {
    float* q = part + (size_t)blockIdx.x * kFirstPart;
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
      for (int ci = 0; ci < 2; ++ci) q[(t * 16 + threadIdx.x) * 2 + ci] = red[2 * t + ci][threadIdx.x];
    q[288 + threadIdx.x] = red[18][threadIdx.x];
  }
}

// output 1x1 conv 16 -> 1 (FP32 out)
__global__ __launch_bounds__(256) void k_out_fwd(View d, const float* __restrict__ w,
                                                 const float* __restrict__ b,
                                                 float* __restrict__ y, int M) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  bf16 v[16];
  __builtin_memcpy(v, d.p + (size_t)m * d.cs + d.co, 32);
  float s = b[0];
#pragma unroll
  for (int c = 0; c < 16; ++c) s += w[c] * bf2f(v[c]);
  y[m] = s;
}

// output layer backward: g_d[m][c] = dy[m] w[c]; partials of dW (16), db
__global__ __launch_bounds__(256) void k_out_bwd(View d, const float* __restrict__ w,
                                                 const float* __restrict__ dy,
                                                 GView gd, int M,
                                                 float* __restrict__ part) {
  __shared__ float red[256];
  float acc[17];
#pragma unroll
  for (int k = 0; k < 17; ++k) acc[k] = 0.f;
  for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < M;
       m += gridDim.x * blockDim.x) {
    const float g = dy[m];
    bf16 v[16];
    __builtin_memcpy(v, d.p + (size_t)m * d.cs + d.co, 32);
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      gd.p[(size_t)m * gd.cs + gd.co + c] = g * w[c];
      acc[c] += g * bf2f(v[c]);
    }
    acc[16] += g;
  }
  for (int k = 0; k < 17; ++k) {
    red[threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * 17 + k] = red[0];
    __syncthreads();
  }
}

// 4x4 average pooling, C channels: out[I][J][c] = mean of the 4x4 block
__global__ __launch_bounds__(256) void k_pool_fwd(View in, View out, int C) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int Mc = out.H * out.W;
  if (e >= Mc * C) return;
  const int m = e / C, c = e - m * C;
  const int I = m / out.W, J = m - I * out.W;
  float s = 0.f;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
      s += bf2f(in.p[((size_t)(4 * I + a) * in.W + 4 * J + b) * in.cs + in.co + c]);
  out.p[(size_t)m * out.cs + out.co + c] = f2bf(s * (1.f / 16.f));
}

// pooling backward accumulated into the fine gradient: gin += gout/16
__global__ __launch_bounds__(256) void k_pool_bwd(GView gout, GView gin, int C) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int Mf = gin.H * gin.W;
  if (e >= Mf * C) return;
  const int m = e / C, c = e - m * C;
  const int i = m / gin.W, j = m - i * gin.W;
  const size_t mc = (size_t)(i / 4) * gout.W + (j / 4);
  gin.p[(size_t)m * gin.cs + gin.co + c] +=
      gout.p[mc * gout.cs + gout.co + c] * (1.f / 16.f);
}

// Masked output gradient of a layer, once per backward: gq[m][c] =
// bf16(g[m][c] * relu'(y[m][c])) (relu'= 1 when relu == 0), dense with C
// channels, and per-block partial sums of the unrounded masked g per
// channel (the bias gradient).  C divides 256, so thread t always handles
// channel t % C over the (pixel, channel) elements.
__global__ __launch_bounds__(256) void k_mask_grad(GView g, View y, int relu, int C,
                                                   bf16* __restrict__ gq,
                                                   float* __restrict__ part) {
  __shared__ float red[256];
  const size_t n = (size_t)g.H * g.W * C;
  float s = 0.f;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t m = e / C;
    const int c = (int)(e - m * C);
    float v = g.p[m * g.cs + g.co + c];
    if (relu && !(bf2f(y.p[m * y.cs + y.co + c]) > 0.f)) v = 0.f;
    gq[e] = f2bf(v);
    s += v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k >= C; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if ((int)threadIdx.x < C) part[(size_t)blockIdx.x * C + threadIdx.x] = red[threadIdx.x];
}

// out[e] = sum_k part[k*count + e], fixed order
// (16 elements per block, 16 threads per element: thread q sums chunks
// q, q + 16, ...; then a fixed tree over the 16)
__global__ __launch_bounds__(256) void k_sum_chunks(const float* __restrict__ part,
                                                    int nchunks, int count,
                                                    float* __restrict__ out) {
  __shared__ float red[16][17];
  const int el = threadIdx.x & 15, q = threadIdx.x >> 4;
  const int e = blockIdx.x * 16 + el;
  float s = 0.f;
  if (e < count)
    for (int k = q; k < nchunks; k += 16) s += part[(size_t)k * count + e];
  red[q][el] = s;
  __syncthreads();
  for (int h = 8; h > 0; h >>= 1) {
    if (q < h) red[q][el] += red[q + h][el];
    __syncthreads();
  }
  if (q == 0 && e < count) out[e] = red[0][el];
}

inline dim3 sum_grid(int count) { return dim3((count + 15) / 16); }

// Adam with bias correction on the FP32 master parameters
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p,
                                              const float* __restrict__ g,
                                              float* __restrict__ m1,
                                              float* __restrict__ m2, size_t n,
                                              float lr, float b1, float b2,
                                              float eps, float c1, float c2) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const float gv = g[k];
    const float a = b1 * m1[k] + (1.f - b1) * gv;
    const float v = b2 * m2[k] + (1.f - b2) * gv * gv;
    m1[k] = a;
    m2[k] = v;
    p[k] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
  }
}

// BF16 weight copies: w[tap][co][ci] and wt[tap][ci][co]
__global__ __launch_bounds__(256) void k_pack(const float* __restrict__ p,
                                              bf16* __restrict__ w,
                                              bf16* __restrict__ wt, int ntap,
                                              int cout, int cin) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ntap * cout * cin) return;
  const int tap = e / (cout * cin), r = e - tap * cout * cin;
  const int co = r / cin, ci = r - co * cin;
  const bf16 v = f2bf(p[e]);
  w[e] = v;
  wt[((size_t)tap * cin + ci) * cout + co] = v;
}

// FP32 <-> scaled copies for the network input/output
__global__ __launch_bounds__(256) void k_scale_copy(const float* __restrict__ a,
                                                    float* __restrict__ b,
                                                    const float* scale, int inv,
                                                    float extra, size_t n) {
  const float s = inv ? 1.f / *scale : *scale * extra;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    b[k] = a[k] * s;
}

// sum of squares partials (FP32 data, FP64 partial sums)
__global__ __launch_bounds__(256) void k_sumsq(const float* __restrict__ a,
                                               size_t n, double* part) {
  __shared__ double lds[4];
  double v[1] = {0.0};
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    v[0] += (double)a[k] * a[k];
  store_partials<double, 1>(v, part, gridDim.x, blockIdx.x, lds);
}

// rms = sqrt(sum/n) (0 -> 1) into *out
__global__ __launch_bounds__(256) void k_rms(const double* part, int G, size_t n,
                                             float* out) {
  __shared__ double lds[4];
  double t[1];
  reduce_partials<double, 1>(part, G, t, lds);
  if (threadIdx.x == 0) {
    const double r = std::sqrt(t[0] / (double)n);
    out[0] = r > 0.0 ? (float)r : 1.f;
  }
}

// error propagation of one weighted-Jacobi sweep: eout = ein - w D^-1 A ein
__global__ __launch_bounds__(256) void k_smooth_err(Sten<double> s,
                                                    const double* __restrict__ ein,
                                                    double* __restrict__ eout,
                                                    double w, Tiling tl) {
  for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
    int i0, j0;
    tile_origin(t, tl.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < TI / BY; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i < s.nx && j < s.ny) {
        const size_t idx = (size_t)i * s.ny + j;
        const double dg = diag_at(s, i, j);
        eout[idx] = ein[idx] - w * (apply_global(s, ein, i, j, dg) / dg);
      }
    }
  }
}

// training sample: r = A e in FP64, FP32 copies of r and e, partials of e.r
__global__ __launch_bounds__(256) void k_sample(Sten<double> s, const double* __restrict__ e,
                                                float* __restrict__ r32,
                                                float* __restrict__ e32,
                                                double* part, Tiling tl) {
  __shared__ double lds[4];
  double v[1] = {0.0};
  for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
    int i0, j0;
    tile_origin(t, tl.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < TI / BY; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i < s.nx && j < s.ny) {
        const size_t idx = (size_t)i * s.ny + j;
        const double rv = apply_global(s, e, i, j, diag_at(s, i, j));
        r32[idx] = (float)rv;
        e32[idx] = (float)e[idx];
        v[0] += e[idx] * rv;
      }
    }
  }
  store_partials<double, 1>(v, part, gridDim.x, blockIdx.x, lds);
}

// total of G partials into out[0]
__global__ __launch_bounds__(256) void k_total(const double* part, int G,
                                               double* out) {
  __shared__ double lds[4];
  double t[1];
  reduce_partials<double, 1>(part, G, t, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0) out[0] = t[0];
}

// d = s y - e (the error of the scaled network output)
__global__ __launch_bounds__(256) void k_loss_d(const float* __restrict__ y,
                                                const float* __restrict__ e,
                                                const float* scale,
                                                float* __restrict__ d, size_t n) {
  const float s = *scale;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    d[k] = s * y[k] - e[k];
}

// partials of d.q and dL/dy = c s q with c = 2 / (e'Ae * batch)
__global__ __launch_bounds__(256) void k_loss_q(const float* __restrict__ d,
                                                const float* __restrict__ q,
                                                const float* scale, float c,
                                                float* __restrict__ dy,
                                                double* part, size_t n) {
  __shared__ double lds[4];
  const float cs = c * *scale;
  double v[1] = {0.0};
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const float qv = q[k];
    v[0] += (double)d[k] * qv;
    if (dy) dy[k] = cs * qv;
  }
  store_partials<double, 1>(v, part, gridDim.x, blockIdx.x, lds);
}

// out[idx] = total / denom
__global__ __launch_bounds__(256) void k_loss_fin(const double* part, int G,
                                                  double denom, float* out) {
  __shared__ double lds[4];
  double t[1];
  reduce_partials<double, 1>(part, G, t, lds);
  if (threadIdx.x == 0) out[0] = (float)(t[0] / denom);
}

// a = b (first) or a += b
__global__ __launch_bounds__(256) void k_accum(float* __restrict__ a,
                                               const float* __restrict__ b,
                                               size_t n, int first) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    a[k] = first ? b[k] : a[k] + b[k];
}

}  // namespace unet
}  // namespace cpgpu

using namespace cpgpu;
using namespace cpgpu::unet;

struct cpgpu_unet {
  cpgpu_ctx* ctx = nullptr;
  int conv_rows = 4;     // pixel rows per wave of the 16-channel tile convolutions (4 or 8)
  int conv_persist = 4;  // > 0: persistent 16-channel tiles, blocks per compute unit
  int fuse_level0 = 1;   // applied networks: level 0 in two fused kernels (needs conv_persist > 0)
  // bumped whenever the activations move or the launch configuration
  // changes: a captured application is then stale
  long long generation = 0;
  int ncu = 256;         // compute units of the device
  int walls = 0;      // wall types of the current input (2 bits per segment x-, x+, y-, y+)
  size_t woff[kNL], boff[kNL], wbf_off[kNL];
  size_t nparams = 0, nwbf = 0;
  DevBuf<float> params, grads, m1, m2;
  DevBuf<bf16> wbf, wtbf;
  long long adam_t = 0;
  // activation allocation for one grid size
  int H = 0, W = 0;
  DevBuf<bf16> act;
  DevBuf<float> gact;
  // per-level tensors: offsets into act/gact
  struct T {
    size_t off;
    int H, W, C;
  };
  T a0, C0, d0a, d0, p0, a1, C1, d1a, d1, p1, a2, C2, d2a, d2, p2, a3, b3;
  DevBuf<float> x0, y, dy, part, scale;
  DevBuf<double> dpart;
  // training: samples (r = A e, e) in FP32, e' A e per sample, loss scratch
  struct Sample {
    DevBuf<float> r, e;
    double eAe = 0.0;
  };
  std::vector<std::unique_ptr<Sample>> samples;
  int sH = 0, sW = 0;  // grid of the samples
  DevBuf<float> gacc, dbuf, qbuf, lossdev;
  DevBuf<bf16> gq;       // masked BF16 output gradient of the current layer
  DevBuf<float> bpart;   // per-block bias partials
  DevBuf<double> tmp64, tmp64b, lpart;

  View v(const T& t, int co = 0) { return View{act.p + t.off, t.H, t.W, t.C, co}; }
  GView g(const T& t, int co = 0) {
    return GView{gact.p + t.off, t.H, t.W, t.C, co};
  }
};

namespace {

int unet_alloc(cpgpu_unet* U, int H, int W) {
  if (U->H == H && U->W == W) return CPGPU_OK;
  if (H < 64 || W < 64 || (H % 64) || (W % 64))
    return set_error(U->ctx, CPGPU_EUNSUPPORTED,
                     "U-Net grid sides must be multiples of 64 (got %dx%d)", H, W);
  size_t off = 0;
  auto mk = [&](int lvl, int C) {
    cpgpu_unet::T t;
    int s = 1 << (2 * lvl);
    t.H = H / s;
    t.W = W / s;
    t.C = C;
    t.off = off;
    off += (size_t)t.H * t.W * C;
    off = (off + 63) & ~(size_t)63;
    return t;
  };
  U->a0 = mk(0, 16); U->C0 = mk(0, 32); U->d0a = mk(0, 16); U->d0 = mk(0, 16);
  U->p0 = mk(1, 16); U->a1 = mk(1, 32); U->C1 = mk(1, 64); U->d1a = mk(1, 32);
  U->d1 = mk(1, 32);
  U->p1 = mk(2, 32); U->a2 = mk(2, 64); U->C2 = mk(2, 128); U->d2a = mk(2, 64);
  U->d2 = mk(2, 64);
  U->p2 = mk(3, 64); U->a3 = mk(3, 128); U->b3 = mk(3, 128);
  const size_t N = (size_t)H * W;
  // partial sums of the largest weight-gradient launch (chunks x taps x
  // cout x cin), of the bias/first/output reductions otherwise
  size_t npart = (size_t)kRedBlocks * kFirstPart;
  for (int l = 0; l < kNL; ++l) {
    const LayerDef& d = kLayers[l];
    if (d.kind != K_CONV && d.kind != K_UP) continue;
    const int lvl = level_of(l);
    const size_t M = (N >> (4 * lvl)) / (d.kind == K_UP ? 16 : 1);
    npart = std::max(npart, (size_t)wg_chunks((int)M) * ntaps(d.kind) * d.cout * d.cin);
  }
  if (U->act.alloc(off) || U->gact.alloc(off) || U->x0.alloc(N) ||
      U->y.alloc(N) || U->dy.alloc(N) || U->part.alloc(npart) ||
      U->gq.alloc(N * 16) || U->bpart.alloc((size_t)kRedBlocks * 128))
    return set_error(U->ctx, CPGPU_ENOMEM, "U-Net activations");
  U->H = H;
  U->W = W;
  ++U->generation;
  return CPGPU_OK;
}

// Wavefront tiles: 32 pixels x up to 64 channels on the large levels; on
// levels with fewer than kSmallM pixels 16 x 16, so that the few pixels
// still spread over many wavefronts.
constexpr int kSmallM = 16384;

// the LDS-tiled 3x3 convolution: sides multiples of 16, at most 64 input
// channels (18 x 18 x 64 BF16 = 41.5 KB of LDS)
template <int KCH>
inline bool tiled(int H, int W) {
  return KCH <= 64 && H % 16 == 0 && W % 16 == 0;
}

template <int NB, int MP>
inline dim3 tile_grid(int M, int ncols) {
  return dim3((M + 64 * MP - 1) / (64 * MP), ncols / (16 * NB));
}

template <int CIN, int COUT>
// Fused epilogues where the kernel supports them (the persistent 16-channel
// convolution): with pout given, the 4 x 4 pooling of y into pout (bit 1 of
// the result); with wout given, the 1 x 1 output layer into yout scaled by
// *yscale, y itself stored only if store_act (bit 2).
int conv_fwd(cpgpu_unet* U, View x, View y, int l, int relu, View pout = View{},
             const float* wout = nullptr, float* yout = nullptr,
             const float* yscale = nullptr, int store_act = 1) {
  constexpr int NB = COUT / 16 < 4 ? COUT / 16 : 4;
  const int M = x.H * x.W;
  const bf16* w = U->wbf.p + U->wbf_off[l];
  const float* b = U->params.p + U->boff[l];
  // timer families 2 and 3: the 16 -> 16 convolutions of the finest level
  // that store their output (2) and those whose epilogue applies the output
  // layer instead (3): their memory traffic differs
  const bool out_fusable = U->conv_persist > 0 && tiled<CIN>(x.H, x.W) && y.cs % 4 == 0 &&
                           y.co % 4 == 0 && wout && yout && yscale && !store_act;
  const EvPair ev = (CIN == 16 && COUT == 16 && M >= kSmallM)
                        ? timer_events(U->ctx, out_fusable ? 3 : 2) : EvPair();
  if (M < kSmallM) {
    // 4 x 4 pixel groups where the sides allow (the same products per
    // pixel), which lets the epilogue pool
    if (x.H % 4 == 0 && x.W % 4 == 0) {
      const bool fuse = pout.p && pout.cs % 4 == 0 && pout.co % 4 == 0 &&
                        pout.H * 4 == x.H && pout.W * 4 == x.W;
      const View po = fuse ? pout : View{};
      // from 64 x 64 pixels and 64 input channels on, a block takes two
      // pixel groups x two channel groups (its weight fragments then serve
      // two pixel groups, its activations two channel groups): 128 -> 64 at
      // 64^2 9.9 -> 6.9 us, 64 -> 64 5.8 -> 4.8 us; smaller levels and
      // inputs measured fastest with one of each
      if constexpr (CIN >= 64 && COUT % 32 == 0) {
        if (M >= 4096 && (M / 16) % 2 == 0) {
          hipLaunchKernelGGL((k_conv3_fwd_ks<CIN, COUT, true, 2, 2>), dim3(M / 32, COUT / 32),
                             dim3(256), 0, U->ctx->stream, x, y, w, b, relu, po);
          return fuse ? 1 : 0;
        }
      }
      hipLaunchKernelGGL((k_conv3_fwd_ks<CIN, COUT, true>), dim3(M / 16, COUT / 16),
                         dim3(256), 0, U->ctx->stream, x, y, w, b, relu, po);
      return fuse ? 1 : 0;
    }
    hipLaunchKernelGGL((k_conv3_fwd_ks<CIN, COUT, false>), dim3((M + 15) / 16, COUT / 16),
                       dim3(256), 0, U->ctx->stream, x, y, w, b, relu, View{});
    return 0;
  }
  if constexpr (CIN <= 64) {
    if (tiled<CIN>(x.H, x.W) && y.cs % 4 == 0 && y.co % 4 == 0) {  // 8-byte stores
      if constexpr (CIN == 16) {
        const bool fuse = pout.p && pout.cs % 4 == 0 && pout.co % 4 == 0 &&
                          pout.H * 4 == x.H && pout.W * 4 == x.W;
        if (U->conv_persist > 0) {
          const int ntx = x.W / 16, ntiles = ntx * (x.H / 16);
          const int G = std::min(ntiles, U->ncu * U->conv_persist);
          const bool out = COUT == 16 && wout && yout && yscale && !store_act;
          if constexpr (COUT == 16) {
            if (out) {
              CPGPU_LAUNCH(ev, (k_conv3_tile16p<COUT, true, EPI_OUT>), dim3(G), dim3(256), 0,
                           U->ctx->stream, x, y, GView{}, w, b, relu, View{}, ntx, ntiles, wout,
                           yout, yscale);
              return 2;
            }
          }
          if (fuse)
            CPGPU_LAUNCH(ev, (k_conv3_tile16p<COUT, true, EPI_POOL>), dim3(G), dim3(256), 0,
                         U->ctx->stream, x, y, GView{}, w, b, relu, pout, ntx, ntiles, nullptr,
                         nullptr, nullptr);
          else
            CPGPU_LAUNCH(ev, (k_conv3_tile16p<COUT, true, EPI_STORE>), dim3(G), dim3(256), 0,
                         U->ctx->stream, x, y, GView{}, w, b, relu, View{}, ntx, ntiles, nullptr,
                         nullptr, nullptr);
          return fuse ? 1 : 0;
        }
        if (!fuse && x.H % 32 == 0 && U->conv_rows == 8)
          CPGPU_LAUNCH(ev, (k_conv3_tile16<COUT, true, 8>), dim3(x.W / 16, x.H / 32),
                       dim3(256), 0, U->ctx->stream, x, y, GView{}, w, b, relu, View{});
        else
          CPGPU_LAUNCH(ev, (k_conv3_tile16<COUT, true, 4>), dim3(x.W / 16, x.H / 16),
                       dim3(256), 0, U->ctx->stream, x, y, GView{}, w, b, relu,
                       fuse ? pout : View{});
        return fuse ? 1 : 0;
      } else {
        const bool fuse = pout.p && pout.cs % 4 == 0 && pout.co % 4 == 0 &&
                          pout.H * 4 == x.H && pout.W * 4 == x.W;
        CPGPU_LAUNCH(ev, (k_conv3_tile<CIN, COUT, true>), dim3(x.W / 16, x.H / 16),
                     dim3(256), 0, U->ctx->stream, x, y, GView{}, w, b, relu,
                     fuse ? pout : View{});
        return fuse ? 1 : 0;
      }
      return 0;
    }
  }
  CPGPU_LAUNCH(ev, (k_conv3_fwd<CIN, COUT, NB, 2>), (tile_grid<NB, 2>(M, COUT)),
               dim3(256), 0, U->ctx->stream, x, y, w, b, relu);
  return 0;
}

template <int CIN, int COUT>
void up_fwd(cpgpu_unet* U, View x, View y, int l) {
  const int M = x.H * x.W;
  const bf16* w = U->wbf.p + U->wbf_off[l];
  const float* b = U->params.p + U->boff[l];
  // (64 x 64 coarse pixels: two pixel and two column groups per wave,
  // 6.3 -> 4.9 us for 64 -> 32; at 16 x 16 one of each is fastest)
  if (M < kSmallM && M >= 4096)
    hipLaunchKernelGGL((k_up_fwd<CIN, COUT, 2, 2>), (tile_grid<2, 2>(M, 16 * COUT)),
                       dim3(256), 0, U->ctx->stream, x, y, w, b);
  else if (M < kSmallM)
    hipLaunchKernelGGL((k_up_fwd<CIN, COUT, 1, 1>), (tile_grid<1, 1>(M, 16 * COUT)),
                       dim3(256), 0, U->ctx->stream, x, y, w, b);
  else
    hipLaunchKernelGGL((k_up_fwd<CIN, COUT, 4, 2>), (tile_grid<4, 2>(M, 16 * COUT)),
                       dim3(256), 0, U->ctx->stream, x, y, w, b);
}

// final_out (applying only): the scaled output s y is written there directly
// when the last layers fuse (*out_done), U->y and the last activation are
// then not formed.
// raw (applying only): the unscaled input, scaled by 1 / U->scale inside the
// first layer instead of through U->x0.
// sq_part (applying only): sq_G partial sums of raw^2, from which the scale
// is formed (inside the fused encoder where it runs, else by k_rms).
int forward(cpgpu_unet* U, float* final_out = nullptr, bool* out_done = nullptr,
            const float* raw = nullptr, const double* sq_part = nullptr, int sq_G = 0) {
  hipStream_t st = U->ctx->stream;
  const int M = U->H * U->W;
  // an applied network (output and unscaled input given) runs level 0 in
  // the two fused kernels; the first layer's output, d0a, d0 and C0 are
  // then not formed, and the skip channels (C0's second half in the
  // layer-per-kernel path) go densely into a0's buffer: 32 B per pixel,
  // where reading them out of C0's 64-byte pixels fetched whole lines.
  // Only where the layer-per-kernel path takes the tiled convolutions
  // (M >= kSmallM), whose summation order they repeat.
  const bool fused0 = final_out && raw && U->fuse_level0 && U->conv_persist > 0 && M >= kSmallM;
  const int ntx = U->W / 16, ntiles = ntx * (U->H / 16);
  bool pooled0 = false;
  if (fused0) {
    const EvPair ev = timer_events(U->ctx, 8);
    CPGPU_LAUNCH(ev, k_enc0_apply, dim3(std::min(ntiles, U->ncu * 5)), dim3(256), 0,
                 st, raw, U->scale.p, sq_part, sq_G, U->H, U->W, U->walls,
                 U->params.p + U->woff[0],
                 U->params.p + U->boff[0], U->wbf.p + U->wbf_off[1], U->params.p + U->boff[1],
                 U->v(U->a0), U->v(U->p0), ntx, ntiles);
    pooled0 = true;
  } else {
    if (sq_part)
      hipLaunchKernelGGL(k_rms, dim3(1), dim3(256), 0, st, sq_part, sq_G, (size_t)M,
                         U->scale.p);
    hipLaunchKernelGGL(k_first_fwd, dim3((M + 255) / 256), dim3(256), 0, st,
                       raw ? raw : (const float*)U->x0.p, U->H, U->W, U->walls, U->v(U->a0),
                       U->params.p + U->woff[0], U->params.p + U->boff[0],
                       raw ? (const float*)U->scale.p : nullptr);
    pooled0 = conv_fwd<16, 16>(U, U->v(U->a0), U->v(U->C0, 16), 1, 1, U->v(U->p0)) & 1;
  }
  auto pool = [&](View in, View out, int C) {
    const int n = out.H * out.W * C;
    hipLaunchKernelGGL(k_pool_fwd, dim3((n + 255) / 256), dim3(256), 0, st, in,
                       out, C);
  };
  if (!pooled0) pool(U->v(U->C0, 16), U->v(U->p0), 16);
  conv_fwd<16, 32>(U, U->v(U->p0), U->v(U->a1), 2, 1);
  if (!(conv_fwd<32, 32>(U, U->v(U->a1), U->v(U->C1, 32), 3, 1, U->v(U->p1)) & 1))
    pool(U->v(U->C1, 32), U->v(U->p1), 32);
  conv_fwd<32, 64>(U, U->v(U->p1), U->v(U->a2), 4, 1);
  if (!(conv_fwd<64, 64>(U, U->v(U->a2), U->v(U->C2, 64), 5, 1, U->v(U->p2)) & 1))
    pool(U->v(U->C2, 64), U->v(U->p2), 64);
  conv_fwd<64, 128>(U, U->v(U->p2), U->v(U->a3), 6, 1);
  conv_fwd<128, 128>(U, U->v(U->a3), U->v(U->b3), 7, 1);
  up_fwd<128, 64>(U, U->v(U->b3), U->v(U->C2, 0), 8);
  conv_fwd<128, 64>(U, U->v(U->C2), U->v(U->d2a), 9, 1);
  conv_fwd<64, 64>(U, U->v(U->d2a), U->v(U->d2), 10, 1);
  up_fwd<64, 32>(U, U->v(U->d2), U->v(U->C1, 0), 11);
  conv_fwd<64, 32>(U, U->v(U->C1), U->v(U->d1a), 12, 1);
  conv_fwd<32, 32>(U, U->v(U->d1a), U->v(U->d1), 13, 1);
  if (fused0) {
    const EvPair ev = timer_events(U->ctx, 9);
    CPGPU_LAUNCH(ev, k_dec0_apply, dim3(std::min(ntiles, U->ncu * 3)), dim3(256), 0, st,
                 U->v(U->d1), U->v(U->a0), U->wbf.p + U->wbf_off[14],
                 U->params.p + U->boff[14], U->wbf.p + U->wbf_off[15],
                 U->params.p + U->boff[15], U->wbf.p + U->wbf_off[16],
                 U->params.p + U->boff[16], U->params.p + U->woff[17], U->scale.p, final_out,
                 ntx, ntiles);
    *out_done = true;
    CPGPU_HIP(U->ctx, hipGetLastError());
    return CPGPU_OK;
  }
  up_fwd<32, 16>(U, U->v(U->d1), U->v(U->C0, 0), 14);
  conv_fwd<32, 16>(U, U->v(U->C0), U->v(U->d0a), 15, 1);
  // applied only (final_out): the output layer and the output scaling run in
  // the last convolution's epilogue, which then stores nothing else
  const float* wout = U->params.p + U->woff[17];  // 16 weights, then the bias
  const int fused = conv_fwd<16, 16>(U, U->v(U->d0a), U->v(U->d0), 16, 1, View{},
                                     final_out ? wout : nullptr, final_out, U->scale.p,
                                     final_out ? 0 : 1);
  if (final_out && (fused & 2)) {
    *out_done = true;
  } else {
    hipLaunchKernelGGL(k_out_fwd, dim3((M + 255) / 256), dim3(256), 0, st,
                       U->v(U->d0), U->params.p + U->woff[17],
                       U->params.p + U->boff[17], U->y.p, M);
  }
  CPGPU_HIP(U->ctx, hipGetLastError());
  return CPGPU_OK;
}

// gq = masked BF16 output gradient (dense, COUT channels) of `out` at
// channel offset out_co, and the bias gradient of layer l
int mask_grad(cpgpu_unet* U, const cpgpu_unet::T& out, int out_co, int cout,
              int relu, int l) {
  hipStream_t st = U->ctx->stream;
  hipLaunchKernelGGL(k_mask_grad, dim3(kRedBlocks), dim3(256), 0, st,
                     U->g(out, out_co), U->v(out, out_co), relu, cout, U->gq.p,
                     U->bpart.p);
  hipLaunchKernelGGL(k_sum_chunks, sum_grid(cout), dim3(256), 0, st, U->bpart.p,
                     kRedBlocks, cout, U->grads.p + U->boff[l]);
  return CPGPU_OK;
}

template <int CIN, int COUT>
int conv_bwd(cpgpu_unet* U, const cpgpu_unet::T& in, int in_co,
             const cpgpu_unet::T& out, int out_co, int l, bool need_dx) {
  hipStream_t st = U->ctx->stream;
  mask_grad(U, out, out_co, COUT, 1, l);
  const View gq{U->gq.p, out.H, out.W, COUT, 0};
  const int M = in.H * in.W;
  if (need_dx) {
    constexpr int NB = CIN / 16 < 4 ? CIN / 16 : 4;
    const bf16* wt = U->wtbf.p + U->wbf_off[l];
    bool done = false;
    if (M < kSmallM) {
      hipLaunchKernelGGL((k_conv3_dgrad<CIN, COUT, 1, 1>), (tile_grid<1, 1>(M, CIN)),
                         dim3(256), 0, st, gq, U->g(in, in_co), wt);
      done = true;
    }
    if constexpr (COUT <= 64) {
      const GView gx = U->g(in, in_co);
      if (!done && tiled<COUT>(in.H, in.W) && gx.cs % 4 == 0 && gx.co % 4 == 0) {
        if constexpr (COUT == 16) {
          if (U->conv_persist > 0) {
            const int ntx = in.W / 16, ntiles = ntx * (in.H / 16);
            const int G = std::min(ntiles, U->ncu * U->conv_persist);
            hipLaunchKernelGGL((k_conv3_tile16p<CIN, false, EPI_STORE>), dim3(G), dim3(256), 0,
                               st, gq, View{}, gx, wt, (const float*)nullptr, 0, View{}, ntx,
                               ntiles, (const float*)nullptr, (float*)nullptr,
                               (const float*)nullptr);
          } else if (in.H % 32 == 0 && U->conv_rows == 8)
            hipLaunchKernelGGL((k_conv3_tile16<CIN, false, 8>), dim3(in.W / 16, in.H / 32),
                               dim3(256), 0, st, gq, View{}, gx, wt, (const float*)nullptr, 0,
                               View{});
          else
            hipLaunchKernelGGL((k_conv3_tile16<CIN, false, 4>), dim3(in.W / 16, in.H / 16),
                               dim3(256), 0, st, gq, View{}, gx, wt, (const float*)nullptr, 0,
                               View{});
        }
        else
          hipLaunchKernelGGL((k_conv3_tile<COUT, CIN, false>), dim3(in.W / 16, in.H / 16),
                             dim3(256), 0, st, gq, View{}, gx, wt, (const float*)nullptr, 0,
                             View{});
        done = true;
      }
    }
    if (!done)
      hipLaunchKernelGGL((k_conv3_dgrad<CIN, COUT, NB, 2>), (tile_grid<NB, 2>(M, CIN)),
                         dim3(256), 0, st, gq, U->g(in, in_co), wt);
  }
  int nch = wg_chunks(M);
  bool tiled_wg = false;
  if constexpr (512 * COUT + 1728 * CIN <= 65536) {
    if (M >= kSmallM && in.H % 16 == 0 && in.W % 16 == 0) {
      const int ntiles = (in.H / 16) * (in.W / 16);
      nch = ntiles < 256 ? ntiles : 256;
      hipLaunchKernelGGL((k_wgrad_tile<CIN, COUT>), dim3(nch), dim3(256), 0, st, gq,
                         U->v(in, in_co), U->part.p);
      tiled_wg = true;
    }
  }
  if (!tiled_wg) {
    const int ppc = (M + nch - 1) / nch;
    hipLaunchKernelGGL((k_wgrad<CIN, COUT, false>), dim3(nch, 9), dim3(256), 0, st,
                       gq, U->v(in, in_co), ppc, U->part.p);
  }
  const int count = 9 * COUT * CIN;
  hipLaunchKernelGGL(k_sum_chunks, sum_grid(count), dim3(256), 0, st,
                     U->part.p, nch, count, U->grads.p + U->woff[l]);
  return CPGPU_OK;
}

template <int CIN, int COUT>
int up_bwd(cpgpu_unet* U, const cpgpu_unet::T& in, const cpgpu_unet::T& out,
           int l) {
  hipStream_t st = U->ctx->stream;
  // channels [0, COUT) of the concat gradient; the up-conv has no ReLU
  mask_grad(U, out, 0, COUT, 0, l);
  const View gq{U->gq.p, out.H, out.W, COUT, 0};
  const int M = in.H * in.W;
  {
    constexpr int NB = CIN / 16 < 4 ? CIN / 16 : 4;
    const bf16* wt = U->wtbf.p + U->wbf_off[l];
    if (M < kSmallM)
      hipLaunchKernelGGL((k_up_dgrad<CIN, COUT, 1, 1>), (tile_grid<1, 1>(M, CIN)),
                         dim3(256), 0, st, gq, U->g(in), wt);
    else
      hipLaunchKernelGGL((k_up_dgrad<CIN, COUT, NB, 2>), (tile_grid<NB, 2>(M, CIN)),
                         dim3(256), 0, st, gq, U->g(in), wt);
  }
  const int nch = wg_chunks(M);
  const int ppc = (M + nch - 1) / nch;
  hipLaunchKernelGGL((k_wgrad<CIN, COUT, true>), dim3(nch, 16), dim3(256), 0, st,
                     gq, U->v(in), ppc, U->part.p);
  const int count = 16 * COUT * CIN;
  hipLaunchKernelGGL(k_sum_chunks, sum_grid(count), dim3(256), 0, st,
                     U->part.p, nch, count, U->grads.p + U->woff[l]);
  return CPGPU_OK;
}

// gradients of every parameter for dL/dy in U->dy (after forward)
int backward(cpgpu_unet* U) {
  hipStream_t st = U->ctx->stream;
  const int M = U->H * U->W;
  const int nb = kRedBlocks;
  hipLaunchKernelGGL(k_out_bwd, dim3(nb), dim3(256), 0, st, U->v(U->d0),
                     U->params.p + U->woff[17], U->dy.p, U->g(U->d0), M,
                     U->part.p);
  hipLaunchKernelGGL(k_sum_chunks, sum_grid(17), dim3(256), 0, st, U->part.p, nb, 17,
                     U->grads.p + U->woff[17]);  // 16 weights then the bias
  conv_bwd<16, 16>(U, U->d0a, 0, U->d0, 0, 16, true);
  conv_bwd<32, 16>(U, U->C0, 0, U->d0a, 0, 15, true);
  up_bwd<32, 16>(U, U->d1, U->C0, 14);
  conv_bwd<32, 32>(U, U->d1a, 0, U->d1, 0, 13, true);
  conv_bwd<64, 32>(U, U->C1, 0, U->d1a, 0, 12, true);
  up_bwd<64, 32>(U, U->d2, U->C1, 11);
  conv_bwd<64, 64>(U, U->d2a, 0, U->d2, 0, 10, true);
  conv_bwd<128, 64>(U, U->C2, 0, U->d2a, 0, 9, true);
  up_bwd<128, 64>(U, U->b3, U->C2, 8);
  conv_bwd<128, 128>(U, U->a3, 0, U->b3, 0, 7, true);
  conv_bwd<64, 128>(U, U->p2, 0, U->a3, 0, 6, true);
  auto unpool = [&](const cpgpu_unet::T& coarse, const cpgpu_unet::T& fine,
                    int fco, int C) {
    const int n = fine.H * fine.W * C;
    hipLaunchKernelGGL(k_pool_bwd, dim3((n + 255) / 256), dim3(256), 0, st,
                       U->g(coarse), U->g(fine, fco), C);
  };
  unpool(U->p2, U->C2, 64, 64);
  conv_bwd<64, 64>(U, U->a2, 0, U->C2, 64, 5, true);
  conv_bwd<32, 64>(U, U->p1, 0, U->a2, 0, 4, true);
  unpool(U->p1, U->C1, 32, 32);
  conv_bwd<32, 32>(U, U->a1, 0, U->C1, 32, 3, true);
  conv_bwd<16, 32>(U, U->p0, 0, U->a1, 0, 2, true);
  unpool(U->p0, U->C0, 16, 16);
  conv_bwd<16, 16>(U, U->a0, 0, U->C0, 16, 1, true);
  const int nbf = kRedBlocks;
  hipLaunchKernelGGL(k_first_wgrad, dim3(nbf), dim3(256), 0, st, U->x0.p, U->H,
                     U->W, U->walls, U->g(U->a0), U->v(U->a0), U->part.p);
  hipLaunchKernelGGL(k_sum_chunks, sum_grid(kFirstPart), dim3(256), 0, st, U->part.p, nbf,
                     kFirstPart, U->grads.p + U->woff[0]);  // 288 weights then 16 biases
  CPGPU_HIP(U->ctx, hipGetLastError());
  return CPGPU_OK;
}

int pack(cpgpu_unet* U) {
  for (int l = 0; l < kNL; ++l) {
    const LayerDef& d = kLayers[l];
    if (d.kind != K_CONV && d.kind != K_UP) continue;
    const int n = ntaps(d.kind) * d.cout * d.cin;
    hipLaunchKernelGGL(k_pack, dim3((n + 255) / 256), dim3(256), 0,
                       U->ctx->stream, U->params.p + U->woff[l],
                       U->wbf.p + U->wbf_off[l], U->wtbf.p + U->wbf_off[l],
                       ntaps(d.kind), d.cout, d.cin);
  }
  CPGPU_HIP(U->ctx, hipGetLastError());
  return CPGPU_OK;
}

// x0 = in / rms(in); forward; out = rms(in) * y.  Device FP32 arrays.
// sq_part (sq_G partial sums of in^2) given: the rms is reduced from them.
int apply_scaled(cpgpu_unet* U, const float* in, float* out, size_t N,
                 const double* sq_part = nullptr, int sq_G = 0) {
  hipStream_t st = U->ctx->stream;
  if (!sq_part) {
    hipLaunchKernelGGL(k_sumsq, dim3(1024), dim3(256), 0, st, in, N, U->dpart.p);
    sq_part = U->dpart.p;
    sq_G = 1024;
  }
  // an applied network forms its scale in its first kernel; one that a
  // backward pass follows needs it here
  if (!out) {
    hipLaunchKernelGGL(k_rms, dim3(1), dim3(256), 0, st, sq_part, sq_G, N, U->scale.p);
    sq_part = nullptr;
  }
  // a network only applied (out given) scales its input inside the first
  // layer; one that a backward pass follows keeps the scaled input in x0
  if (!out)
    hipLaunchKernelGGL(k_scale_copy, dim3(1024), dim3(256), 0, st, in, U->x0.p,
                       U->scale.p, 1, 1.f, N);
  bool done = false;
  int rc = forward(U, out, &done, out ? in : nullptr, sq_part, sq_G);
  if (rc) return rc;
  if (out && !done)
    hipLaunchKernelGGL(k_scale_copy, dim3(1024), dim3(256), 0, st, U->y.p, out,
                       U->scale.p, 0, 1.f, N);
  CPGPU_HIP(U->ctx, hipGetLastError());
  return CPGPU_OK;
}

int check_problem(cpgpu_unet* U, cpgpu_poisson* P) {
  if (P->ctx != U->ctx)
    return set_error(U->ctx, CPGPU_EINVAL, "U-Net and problem use different contexts");
  if (!U->samples.empty() && (P->nx != U->sH || P->ny != U->sW))
    return set_error(U->ctx, CPGPU_EINVAL,
                     "problem grid %dx%d differs from the samples' %dx%d", P->nx,
                     P->ny, U->sH, U->sW);
  int rc = unet_alloc(U, P->nx, P->ny);
  if (rc) return rc;
  return poisson_prepare_f32(P);
}

// forward on sample slot and its energy-norm loss into *lossout (device);
// with dy_scale > 0 also dL/dy * dy_scale into U->dy
int sample_loss(cpgpu_unet* U, cpgpu_poisson* P, int slot, float dy_scale,
                float* lossout) {
  hipStream_t st = U->ctx->stream;
  cpgpu_unet::Sample& S = *U->samples[slot];
  const size_t N = (size_t)U->H * U->W;
  U->walls = pack_walls(P);
  int rc = apply_scaled(U, S.r.p, nullptr, N);
  if (rc) return rc;
  hipLaunchKernelGGL(k_loss_d, dim3(1024), dim3(256), 0, st, U->y.p, S.e.p,
                     U->scale.p, U->dbuf.p, N);
  hipLaunchKernelGGL(k_apply<float>, dim3(P->tl.G), dim3(TJ, BY), 0, st,
                     P->sten32(), U->dbuf.p, U->qbuf.p, P->tl);
  const float c = (float)(2.0 / S.eAe) * dy_scale;
  hipLaunchKernelGGL(k_loss_q, dim3(1024), dim3(256), 0, st, U->dbuf.p,
                     U->qbuf.p, U->scale.p, c, dy_scale > 0.f ? U->dy.p : nullptr,
                     U->lpart.p, N);
  hipLaunchKernelGGL(k_loss_fin, dim3(1), dim3(256), 0, st, U->lpart.p, 1024,
                     S.eAe, lossout);
  CPGPU_HIP(U->ctx, hipGetLastError());
  return CPGPU_OK;
}

int adam_step(cpgpu_unet* U, float lr, const float* g) {
  U->adam_t++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)U->adam_t);
  const float c2 = 1.f - std::pow(b2, (float)U->adam_t);
  hipLaunchKernelGGL(k_adam, dim3(256), dim3(256), 0, U->ctx->stream,
                     U->params.p, g, U->m1.p, U->m2.p, U->nparams, lr, b1, b2,
                     1e-8f, c1, c2);
  CPGPU_HIP(U->ctx, hipGetLastError());
  return pack(U);
}

// SplitMix64 stream for sample selection
struct SplitMix {
  unsigned long long x;
  unsigned long long next() {
    unsigned long long z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
};

}  // namespace

// ---------------------------------------------------------------------
// internal entry points used by the learned solver
namespace cpgpu {
int pack_walls(const cpgpu_poisson* P) {
  int w = 0;
  for (int k = 0; k < 4; ++k) w |= (P->seg[k].type & 3) << (2 * k);
  return w;
}

int unet_apply_device(cpgpu_unet* U, int H, int W, const float* in, float* out, int walls,
                      const double* sq_part, int sq_G) {
  int rc = unet_alloc(U, H, W);
  if (rc) return rc;
  U->walls = walls;
  return apply_scaled(U, in, out, (size_t)H * W, sq_part, sq_G);
}
cpgpu_ctx* unet_context(const cpgpu_unet* U) { return U->ctx; }
long long unet_generation(const cpgpu_unet* U) { return U->generation; }
}  // namespace cpgpu

extern "C" {

int cpgpu_unet_create(cpgpu_ctx* ctx, unsigned long long seed,
                      cpgpu_unet** out) {
  if (!ctx || !out) return CPGPU_EINVAL;
  *out = nullptr;
  CPGPU_HIP(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<cpgpu_unet> U(new cpgpu_unet());
  U->ctx = ctx;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) ==
          hipSuccess && ncu > 0)
    U->ncu = ncu;
  size_t off = 0, boff = 0;
  for (int l = 0; l < kNL; ++l) {
    const LayerDef& d = kLayers[l];
    const size_t nw = (size_t)ntaps(d.kind) * d.cout * d.cin;
    U->woff[l] = off;
    off += nw;
    U->boff[l] = off;
    off += d.cout;
    U->wbf_off[l] = boff;
    if (d.kind == K_CONV || d.kind == K_UP) boff += nw;
  }
  U->nparams = off;
  U->nwbf = boff;
  if (U->params.alloc(off) || U->grads.alloc(off) || U->m1.alloc(off) ||
      U->m2.alloc(off) || U->wbf.alloc(boff) || U->wtbf.alloc(boff) ||
      U->scale.alloc(4) ||
      U->dpart.alloc(1024))
    return set_error(ctx, CPGPU_ENOMEM, "U-Net parameters");
  // Kaiming-normal weights (fan_in, gain sqrt 2; gain 1 for the linear
  // output layer), zero biases, from xoshiro256** seeded by SplitMix64
  std::vector<float> h(off, 0.f);
  {
    unsigned long long s[4], x = seed;
    for (int k = 0; k < 4; ++k) {
      unsigned long long z = (x += 0x9E3779B97F4A7C15ULL);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
      s[k] = z ^ (z >> 31);
    }
    auto rotl = [](unsigned long long v, int k) { return (v << k) | (v >> (64 - k)); };
    auto next = [&]() {
      const unsigned long long r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
      s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t;
      s[3] = rotl(s[3], 45);
      return r;
    };
    auto unif = [&]() { return (double)(next() >> 11) * 0x1.0p-53; };
    for (int l = 0; l < kNL; ++l) {
      const LayerDef& d = kLayers[l];
      const double fan_in = d.kind == K_UP ? d.cin : (double)ntaps(d.kind) * d.cin;
      const double gain = d.kind == K_OUT ? 1.0 : std::sqrt(2.0);
      const double sd = gain / std::sqrt(fan_in);
      const size_t nw = (size_t)ntaps(d.kind) * d.cout * d.cin;
      for (size_t k = 0; k < nw; ++k) {
        const double u1 = 1.0 - unif(), u2 = unif();
        h[U->woff[l] + k] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) *
                                    std::cos(6.283185307179586 * u2));
      }
    }
  }
  // the context stream is non-blocking: every transfer goes on it
  CPGPU_HIP(ctx, hipMemcpyAsync(U->params.p, h.data(), sizeof(float) * off,
                                hipMemcpyHostToDevice, ctx->stream));
  CPGPU_HIP(ctx, hipMemsetAsync(U->m1.p, 0, sizeof(float) * off, ctx->stream));
  CPGPU_HIP(ctx, hipMemsetAsync(U->m2.p, 0, sizeof(float) * off, ctx->stream));
  CPGPU_HIP(ctx, hipMemsetAsync(U->grads.p, 0, sizeof(float) * off, ctx->stream));
  int rc = pack(U.get());
  if (rc) return rc;
  CPGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = U.release();
  return CPGPU_OK;
}

int cpgpu_unet_set_walls(cpgpu_unet* U, const int* bc_type) {
  if (!U) return CPGPU_EINVAL;
  int w = 0;
  if (bc_type) {
    for (int k = 0; k < 4; ++k) {
      if (bc_type[k] != CPGPU_BC_DIRICHLET && bc_type[k] != CPGPU_BC_NEUMANN)
        return CPGPU_EINVAL;
      w |= bc_type[k] << (2 * k);
    }
  }
  U->walls = w;
  return CPGPU_OK;
}

int cpgpu_unet_set_conv_shape(cpgpu_unet* U, int rows, int persist) {
  if (!U || (rows != 4 && rows != 8) || persist < 0 || persist > 16) return CPGPU_EINVAL;
  U->conv_rows = rows;
  U->conv_persist = persist;
  ++U->generation;
  return CPGPU_OK;
}

int cpgpu_unet_set_fusion(cpgpu_unet* U, int level0) {
  if (!U || level0 < 0 || level0 > 1) return CPGPU_EINVAL;
  U->fuse_level0 = level0;
  ++U->generation;
  return CPGPU_OK;
}

int cpgpu_unet_destroy(cpgpu_unet* U) {
  if (!U) return CPGPU_EINVAL;
  (void)hipStreamSynchronize(U->ctx->stream);
  delete U;
  return CPGPU_OK;
}

long long cpgpu_unet_num_params(const cpgpu_unet* U) {
  return U ? (long long)U->nparams : -1;
}

int cpgpu_unet_get_params(cpgpu_unet* U, float* host, int grads) {
  if (!U || !host) return CPGPU_EINVAL;
  CPGPU_HIP(U->ctx, hipStreamSynchronize(U->ctx->stream));
  CPGPU_HIP(U->ctx, hipMemcpy(host, grads ? U->grads.p : U->params.p,
                              sizeof(float) * U->nparams, hipMemcpyDeviceToHost));
  return CPGPU_OK;
}

int cpgpu_unet_set_params(cpgpu_unet* U, const float* host) {
  if (!U || !host) return CPGPU_EINVAL;
  CPGPU_HIP(U->ctx, hipMemcpyAsync(U->params.p, host, sizeof(float) * U->nparams,
                                   hipMemcpyHostToDevice, U->ctx->stream));
  int rc = pack(U);
  if (rc) return rc;
  CPGPU_HIP(U->ctx, hipStreamSynchronize(U->ctx->stream));
  return CPGPU_OK;
}

// Host in/out forward of the raw network (no input scaling): x [H*W] ->
// y [H*W].  For tests.
int cpgpu_unet_forward(cpgpu_unet* U, int H, int W, const float* x, float* y) {
  if (!U || !x || !y) return CPGPU_EINVAL;
  int rc = unet_alloc(U, H, W);
  if (rc) return rc;
  const size_t N = (size_t)H * W;
  CPGPU_HIP(U->ctx, hipMemcpyAsync(U->x0.p, x, sizeof(float) * N,
                                   hipMemcpyHostToDevice, U->ctx->stream));
  if ((rc = forward(U))) return rc;
  CPGPU_HIP(U->ctx, hipMemcpyAsync(y, U->y.p, sizeof(float) * N,
                                   hipMemcpyDeviceToHost, U->ctx->stream));
  CPGPU_HIP(U->ctx, hipStreamSynchronize(U->ctx->stream));
  return CPGPU_OK;
}

// Internal tensor `id` of the last forward (grad = 0, BF16 widened to
// FP32) or backward (grad != 0, FP32) as H x W x C into host (may be NULL:
// dims only).  ids: 0 input (FP32), 1 a0, 2 C0, 3 d0a, 4 d0, 5 p0, 6 a1,
// 7 C1, 8 d1a, 9 d1, 10 p1, 11 a2, 12 C2, 13 d2a, 14 d2, 15 p2, 16 a3,
// 17 b3, 18 output (FP32).  For layer-by-layer tests.
int cpgpu_unet_tensor(cpgpu_unet* U, int id, int grad, float* host, int* dims) {
  if (!U || !dims || id < 0 || id > 18 || U->H == 0) return CPGPU_EINVAL;
  const size_t N = (size_t)U->H * U->W;
  CPGPU_HIP(U->ctx, hipStreamSynchronize(U->ctx->stream));
  if (id == 0 || id == 18) {
    if (grad && id == 0) return CPGPU_EINVAL;
    dims[0] = U->H;
    dims[1] = U->W;
    dims[2] = 1;
    const float* src = id == 0 ? U->x0.p : (grad ? U->dy.p : U->y.p);
    if (host)
      CPGPU_HIP(U->ctx, hipMemcpy(host, src, sizeof(float) * N, hipMemcpyDeviceToHost));
    return CPGPU_OK;
  }
  const cpgpu_unet::T* tab[] = {nullptr, &U->a0, &U->C0, &U->d0a, &U->d0, &U->p0,
                                &U->a1, &U->C1, &U->d1a, &U->d1, &U->p1, &U->a2,
                                &U->C2, &U->d2a, &U->d2, &U->p2, &U->a3, &U->b3};
  const cpgpu_unet::T& t = *tab[id];
  dims[0] = t.H;
  dims[1] = t.W;
  dims[2] = t.C;
  if (!host) return CPGPU_OK;
  const size_t n = (size_t)t.H * t.W * t.C;
  if (grad) {
    CPGPU_HIP(U->ctx, hipMemcpy(host, U->gact.p + t.off, sizeof(float) * n,
                                hipMemcpyDeviceToHost));
  } else {
    std::vector<bf16> h(n);
    CPGPU_HIP(U->ctx, hipMemcpy(h.data(), U->act.p + t.off, sizeof(bf16) * n,
                                hipMemcpyDeviceToHost));
    for (size_t k = 0; k < n; ++k) host[k] = bf2f(h[k]);
  }
  return CPGPU_OK;
}

// Gradients of sum(dy * y) w.r.t. every parameter after cpgpu_unet_forward
// on the same input; grads (host, num_params) receives them.  For tests.
int cpgpu_unet_backward(cpgpu_unet* U, const float* dy, float* grads) {
  if (!U || !dy || !grads || U->H == 0) return CPGPU_EINVAL;
  const size_t N = (size_t)U->H * U->W;
  CPGPU_HIP(U->ctx, hipMemcpyAsync(U->dy.p, dy, sizeof(float) * N,
                                   hipMemcpyHostToDevice, U->ctx->stream));
  int rc = backward(U);
  if (rc) return rc;
  return cpgpu_unet_get_params(U, grads, 1);
}

// One Adam step on the current gradients.
int cpgpu_unet_adam(cpgpu_unet* U, float lr) {
  if (!U) return CPGPU_EINVAL;
  return adam_step(U, lr, U->grads.p);
}

int cpgpu_unet_set_sample(cpgpu_unet* U, cpgpu_poisson* P, int slot,
                          const double* e, int smooth, double omega) {
  if (!U || !P || !e || slot < 0 || slot > (1 << 20) || smooth < 0 ||
      smooth > 64 || (smooth > 0 && !(omega > 0.0 && omega < 2.0)))
    return CPGPU_EINVAL;
  int rc = check_problem(U, P);
  if (rc) return rc;
  const size_t N = P->N;
  if ((size_t)slot >= U->samples.size()) U->samples.resize(slot + 1);
  if (!U->samples[slot]) U->samples[slot].reset(new cpgpu_unet::Sample());
  cpgpu_unet::Sample& S = *U->samples[slot];
  if (S.r.alloc(N) || S.e.alloc(N) || U->tmp64.alloc(N) ||
      (smooth > 0 && U->tmp64b.alloc(N)) || U->lpart.alloc(kMaxBlocks + 1))
    return set_error(U->ctx, CPGPU_ENOMEM, "U-Net sample");
  U->sH = P->nx;
  U->sW = P->ny;
  hipStream_t st = U->ctx->stream;
  CPGPU_HIP(U->ctx, hipMemcpyAsync(U->tmp64.p, e, sizeof(double) * N,
                                   hipMemcpyHostToDevice, st));
  double* ecur = U->tmp64.p;
  for (int k = 0; k < smooth; ++k) {
    double* enew = ecur == U->tmp64.p ? U->tmp64b.p : U->tmp64.p;
    hipLaunchKernelGGL(k_smooth_err, dim3(P->tl.G), dim3(TJ, BY), 0, st,
                       P->sten64(), (const double*)ecur, enew, omega, P->tl);
    ecur = enew;
  }
  hipLaunchKernelGGL(k_sample, dim3(P->tl.G), dim3(TJ, BY), 0, st, P->sten64(),
                     (const double*)ecur, S.r.p, S.e.p, U->lpart.p, P->tl);
  hipLaunchKernelGGL(k_total, dim3(1), dim3(256), 0, st, U->lpart.p,
                     (int)P->tl.G, U->lpart.p + kMaxBlocks);
  CPGPU_HIP(U->ctx, hipGetLastError());
  double eae = 0.0;
  CPGPU_HIP(U->ctx, hipMemcpyAsync(&eae, U->lpart.p + kMaxBlocks, sizeof(double),
                                   hipMemcpyDeviceToHost, st));
  CPGPU_HIP(U->ctx, hipStreamSynchronize(st));
  if (!(eae > 0.0))
    return set_error(U->ctx, CPGPU_EINVAL, "sample %d has e'Ae = %g (need > 0)",
                     slot, eae);
  S.eAe = eae;
  return CPGPU_OK;
}

int cpgpu_unet_clear_samples(cpgpu_unet* U) {
  if (!U) return CPGPU_EINVAL;
  CPGPU_HIP(U->ctx, hipStreamSynchronize(U->ctx->stream));  // no kernel reads them any more
  U->samples.clear();
  U->sH = U->sW = 0;
  return CPGPU_OK;
}

int cpgpu_unet_num_samples(const cpgpu_unet* U) {
  return U ? (int)U->samples.size() : -1;
}

static int check_samples(cpgpu_unet* U) {
  if (U->samples.empty())
    return set_error(U->ctx, CPGPU_EINVAL, "no training samples");
  for (size_t k = 0; k < U->samples.size(); ++k)
    if (!U->samples[k])
      return set_error(U->ctx, CPGPU_EINVAL, "sample slot %zu is empty", k);
  return CPGPU_OK;
}

static int loss_buffers(cpgpu_unet* U, size_t nloss) {
  const size_t N = (size_t)U->H * U->W;
  if (U->dbuf.alloc(N) || U->qbuf.alloc(N) || U->gacc.alloc(U->nparams) ||
      U->lossdev.alloc(nloss) || U->lpart.alloc(kMaxBlocks + 1))
    return set_error(U->ctx, CPGPU_ENOMEM, "U-Net training buffers");
  return CPGPU_OK;
}

int cpgpu_unet_train(cpgpu_unet* U, cpgpu_poisson* P, int steps, int batch,
                     float lr, unsigned long long seed, float* losses) {
  if (!U || !P || steps < 0 || batch < 1 || !(lr > 0.f)) return CPGPU_EINVAL;
  int rc = check_samples(U);
  if (rc) return rc;
  if ((rc = check_problem(U, P))) return rc;
  if ((rc = loss_buffers(U, (size_t)std::max(steps, 1) * batch))) return rc;
  hipStream_t st = U->ctx->stream;
  SplitMix g{seed};
  const int ns = (int)U->samples.size();
  for (int s = 0; s < steps; ++s) {
    for (int b = 0; b < batch; ++b) {
      const int slot = (int)(g.next() % (unsigned long long)ns);
      if ((rc = sample_loss(U, P, slot, 1.f / batch,
                            U->lossdev.p + (size_t)s * batch + b)))
        return rc;
      if ((rc = backward(U))) return rc;
      hipLaunchKernelGGL(k_accum, dim3(256), dim3(256), 0, st, U->gacc.p,
                         U->grads.p, U->nparams, b == 0 ? 1 : 0);
    }
    if ((rc = adam_step(U, lr, U->gacc.p))) return rc;
  }
  if (losses && steps > 0) {
    std::vector<float> h((size_t)steps * batch);
    CPGPU_HIP(U->ctx, hipMemcpyAsync(h.data(), U->lossdev.p, sizeof(float) * h.size(),
                                     hipMemcpyDeviceToHost, st));
    CPGPU_HIP(U->ctx, hipStreamSynchronize(st));
    for (int s = 0; s < steps; ++s) {
      double m = 0;
      for (int b = 0; b < batch; ++b) m += h[(size_t)s * batch + b];
      losses[s] = (float)(m / batch);
    }
  }
  CPGPU_HIP(U->ctx, hipStreamSynchronize(st));
  return CPGPU_OK;
}

int cpgpu_unet_eval(cpgpu_unet* U, cpgpu_poisson* P, float* losses) {
  if (!U || !P || !losses) return CPGPU_EINVAL;
  int rc = check_samples(U);
  if (rc) return rc;
  if ((rc = check_problem(U, P))) return rc;
  const int ns = (int)U->samples.size();
  if ((rc = loss_buffers(U, ns))) return rc;
  for (int k = 0; k < ns; ++k)
    if ((rc = sample_loss(U, P, k, 0.f, U->lossdev.p + k))) return rc;
  CPGPU_HIP(U->ctx, hipMemcpyAsync(losses, U->lossdev.p, sizeof(float) * ns,
                                   hipMemcpyDeviceToHost, U->ctx->stream));
  CPGPU_HIP(U->ctx, hipStreamSynchronize(U->ctx->stream));
  return CPGPU_OK;
}

}  // extern "C"
