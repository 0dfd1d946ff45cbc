// Flexible conjugate gradients with the U-Net preconditioner (no reference
// counterpart).  FP64 Krylov vectors; the network maps the FP32 copy of the
// residual to the FP32 preconditioned residual z.  One-term truncated FCG
// (Notay's FCG(1)): the direction is A-orthogonalised against the previous
// one only, which keeps the method convergent for a preconditioner that is
// not a fixed symmetric operator:
//   p_k = z_k - (z_k . q_{k-1}) / (p_{k-1} . q_{k-1}) p_{k-1},  q_k = A p_k
//   alpha = (p_k . r_k) / (p_k . q_k),  x += alpha p_k,  r -= alpha q_k
// A breakdown (p.q <= 0) or stagnation hands the current iterate to FP64
// Jacobi-PCG.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "poisson.hpp"

namespace cpgpu {
namespace {

// r = b - A x, r32 = (float) r, partials of r.r and b.b (the reference's
// nrm2 drops the last unknown when norm_ref)
__global__ __launch_bounds__(256) void k_fcg_init(Sten<double> s,
                                                  const double* __restrict__ x,
                                                  const double* __restrict__ b,
                                                  double* __restrict__ r,
                                                  float* __restrict__ r32,
                                                  double* part, Tiling tl,
                                                  int norm_ref) {
  __shared__ double lds[8];
  double v[2] = {0.0, 0.0};
  const size_t nlast = (size_t)s.nx * s.ny - 1;
  for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
    int i0, j0;
    tile_origin(t, tl.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < TI / BY; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i < s.nx && j < s.ny) {
        const size_t idx = (size_t)i * s.ny + j;
        const double bv = b[idx];
        const double rv = bv - apply_global(s, x, i, j, diag_at(s, i, j));
        r[idx] = rv;
        r32[idx] = (float)rv;
        if (!(norm_ref && idx == nlast)) {
          v[0] += rv * rv;
          v[1] += bv * bv;
        }
      }
    }
  }
  store_partials<double, 2>(v, part, gridDim.x, blockIdx.x, lds);
}

// the host's copy of the state (values first, the sequence number last)
__device__ __forceinline__ void post_mail(FMail* mail, double resid, int breakdown, int seq) {
  if (!mail) return;
  mail->resid = resid;
  mail->breakdown = breakdown;
  __hip_atomic_store(&mail->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ __launch_bounds__(256) void k_fcg_start(const double* part, int G,
                                                   FState* st, FMail* mail) {
  __shared__ double lds[8];
  double t[2];
  reduce_partials<double, 2>(part, G, t, lds);
  if (threadIdx.x == 0) {
    const double nb = t[1] > 0.0 ? std::sqrt(t[1]) : 1.0;
    st->normb = nb;
    st->resid = std::sqrt(t[0]) / nb;
    st->pq = 0.0;
    st->iter = 0;
    st->breakdown = 0;
    post_mail(mail, st->resid, 0, 0);
  }
}

// partials of z . q (the previous q; 0 on the first iteration, q null)
// and of sum(z)
__global__ __launch_bounds__(256) void k_fcg_zq(const float* __restrict__ z,
                                                const double* __restrict__ q,
                                                double* part, size_t n) {
  __shared__ double lds[8];
  double v[2] = {0.0, 0.0};
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const double zv = z[k];
    if (q) v[0] += zv * q[k];
    v[1] += zv;
  }
  store_partials<double, 2>(v, part, gridDim.x, blockIdx.x, lds);
}

// p_new = z + beta p_old (beta = -(z.q_prev)/(p.q)_prev, 0 on the first
// iteration), q = A p_new, partials of p.q and p.r.  p_new at the four
// neighbours is recomputed from z and p_old, so p_old and p_new are
// different buffers.  For a singular (all-Neumann) operator z is first
// projected to zero mean (project != 0): the constants are A's null space,
// and a constant part of p would enter p.r through the rounding-level
// mean of r and stall the iteration.
__global__ __launch_bounds__(256) void k_fcg_dir(Sten<double> s,
                                                 const float* __restrict__ z,
                                                 const double* __restrict__ pold,
                                                 double* __restrict__ pnew,
                                                 double* __restrict__ q,
                                                 const double* __restrict__ r,
                                                 const double* partZ, int GZ,
                                                 const FState* st, int first,
                                                 int project, double* part,
                                                 Tiling tl) {
  __shared__ double lds[8];
  double t[2];
  reduce_partials<double, 2>(partZ, GZ, t, lds);
  // (z - m).q_prev = z.q_prev - m sum(q_prev), and sum(A p) = 0
  const double m = project ? t[1] / ((double)s.nx * s.ny) : 0.0;
  const double beta = first ? 0.0 : -t[0] / st->pq;
  double v[2] = {0.0, 0.0};
  const int ny = s.ny;
  constexpr int NR = TI / BY;
  for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
    int i0, j0;
    tile_origin(t, tl.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
    // every load of the thread's rows first (clamped addresses)
    Nbr<float> zn[NR];
    Nbr<double> pn[NR], kc[NR];
    double rr[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      load_nbr(s, z, i, j, zn[k], kc[k]);
      load_nbr(s, pold, i, j, pn[k], kc[k]);
      rr[k] = r[(size_t)min(i, s.nx - 1) * ny + min(j, ny - 1)];
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i < s.nx && j < ny) {
        const size_t idx = (size_t)i * ny + j;
        const double pc = ((double)zn[k].c - m) + beta * pn[k].c;
        double acc = pc * diag_from(s, i, j, kc[k].kx0, kc[k].kx1, kc[k].ky0, kc[k].ky1);
        if (i > 0) acc -= kc[k].kx0 * (((double)zn[k].n - m) + beta * pn[k].n);
        if (j > 0) acc -= kc[k].ky0 * (((double)zn[k].w - m) + beta * pn[k].w);
        if (j < ny - 1) acc -= kc[k].ky1 * (((double)zn[k].e - m) + beta * pn[k].e);
        if (i < s.nx - 1) acc -= kc[k].kx1 * (((double)zn[k].s - m) + beta * pn[k].s);
        pnew[idx] = pc;
        q[idx] = acc;
        v[0] += pc * acc;
        v[1] += pc * rr[k];
      }
    }
  }
  store_partials<double, 2>(v, part, gridDim.x, blockIdx.x, lds);
}

// alpha = (p.r)/(p.q); x += alpha p; r -= alpha q; r32 = (float) r;
// partials of r.r.  Block 0 records p.q for the next direction.
__global__ __launch_bounds__(256) void k_fcg_upd(const double* __restrict__ p,
                                                 const double* __restrict__ q,
                                                 double* __restrict__ x,
                                                 double* __restrict__ r,
                                                 float* __restrict__ r32,
                                                 const double* partD, int GD,
                                                 FState* st, double* part,
                                                 size_t n, int norm_ref) {
  __shared__ double lds[8];
  double t[2];
  reduce_partials<double, 2>(partD, GD, t, lds);
  const double pq = t[0], pr = t[1];
  const bool ok = pq > 0.0 && std::isfinite(pq) && std::isfinite(pr);
  const double alpha = ok ? pr / pq : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->pq = pq;
    if (!ok) st->breakdown = 1;
  }
  double v[1] = {0.0};
  // pairs of elements per thread (16-byte loads and stores); for n odd the
  // last element is block 0's thread 0's
  typedef double d2 __attribute__((ext_vector_type(2)));
  typedef float f2 __attribute__((ext_vector_type(2)));
  const size_t n2 = n / 2;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n2;
       k += (size_t)gridDim.x * blockDim.x) {
    const d2 pv = reinterpret_cast<const d2*>(p)[k], qv = reinterpret_cast<const d2*>(q)[k];
    d2 xv = reinterpret_cast<const d2*>(x)[k], rv = reinterpret_cast<const d2*>(r)[k];
    xv.x = xv.x + alpha * pv.x;
    xv.y = xv.y + alpha * pv.y;
    rv.x = rv.x - alpha * qv.x;
    rv.y = rv.y - alpha * qv.y;
    reinterpret_cast<d2*>(x)[k] = xv;
    reinterpret_cast<d2*>(r)[k] = rv;
    reinterpret_cast<f2*>(r32)[k] = f2{(float)rv.x, (float)rv.y};
    v[0] += rv.x * rv.x;
    if (!(norm_ref && 2 * k + 1 == n - 1)) v[0] += rv.y * rv.y;
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    const size_t k = n - 1;
    {
      x[k] = x[k] + alpha * p[k];
      const double rv = r[k] - alpha * q[k];
      r[k] = rv;
      r32[k] = (float)rv;
      if (!norm_ref) v[0] += rv * rv;
    }
  }
  store_partials<double, 1>(v, part, gridDim.x, blockIdx.x, lds);
}

__global__ __launch_bounds__(256) void k_fcg_check(const double* part, int G,
                                                   FState* st, FMail* mail) {
  __shared__ double lds[4];
  double t[1];
  reduce_partials<double, 1>(part, G, t, lds);
  if (threadIdx.x == 0) {
    st->resid = std::sqrt(t[0]) / st->normb;
    st->iter = st->iter + 1;
    post_mail(mail, st->resid, st->breakdown, st->iter);
  }
}

__global__ __launch_bounds__(256) void k_f64_to_f32(const double* __restrict__ a,
                                                    float* __restrict__ b, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    b[k] = (float)a[k];
}

__global__ __launch_bounds__(256) void k_f32_to_f64(const float* __restrict__ a,
                                                    double* __restrict__ b, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    b[k] = (double)a[k];
}

__global__ __launch_bounds__(256) void k_sum64(const double* __restrict__ a,
                                               double* part, size_t n) {
  __shared__ double lds[4];
  double v[1] = {0.0};
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    v[0] += a[k];
  store_partials<double, 1>(v, part, gridDim.x, blockIdx.x, lds);
}

// x -= mean(x), the mean from G partial sums
__global__ __launch_bounds__(256) void k_sub_mean_part(double* __restrict__ x,
                                                       const double* part, int G,
                                                       size_t n) {
  __shared__ double lds[4];
  double t[1];
  reduce_partials<double, 1>(part, G, t, lds);
  const double mean = t[0] / (double)n;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    x[k] = x[k] - mean;
}

// Weighted Jacobi sweep on A z = r (FP32): zout = zin + w D^-1 (r - A zin);
// zin == nullptr means a zero initial guess (zout = w D^-1 r).  With zadd,
// the sweep starts from zin + zadd (the sum formed on the fly, as a
// separate addition kernel would store it).  ZQ (the cycle's last sweep):
// also the partials of zout . qprev (0 when qprev is null) and of
// sum(zout) into zpart, the values k_fcg_zq would form from the stored z.
template <bool ZQ>
__global__ __launch_bounds__(256) void k_jacobi32(Sten<float> s,
                                                  const float* __restrict__ r,
                                                  const float* __restrict__ zin,
                                                  const float* __restrict__ zadd,
                                                  float* __restrict__ zout,
                                                  float w, Tiling tl,
                                                  const double* __restrict__ qprev,
                                                  double* zpart) {
  double v[2] = {0.0, 0.0};
  for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
    int i0, j0;
    tile_origin(t, tl.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
    constexpr int NR = TI / BY;
    Nbr<float> zn[NR], za[NR], kc[NR];
    float rr[NR];
    double qv[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      // (a zero guess reads r in place of zin: the values are not used)
      load_nbr(s, zin ? zin : r, i, j, zn[k], kc[k]);
      if (zadd) load_nbr(s, zadd, i, j, za[k], kc[k]);
      const size_t ic = (size_t)min(i, s.nx - 1) * s.ny + min(j, s.ny - 1);
      rr[k] = r[ic];
      qv[k] = (ZQ && qprev) ? qprev[ic] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i < s.nx && j < s.ny) {
        const size_t idx = (size_t)i * s.ny + j;
        Nbr<float> z = zn[k];
        if (zadd) {
          z.c = z.c + za[k].c;
          z.n = z.n + za[k].n;
          z.s = z.s + za[k].s;
          z.w = z.w + za[k].w;
          z.e = z.e + za[k].e;
        }
        const float dg = diag_from(s, i, j, kc[k].kx0, kc[k].kx1, kc[k].ky0, kc[k].ky1);
        const float rv = zin ? rr[k] - apply_regs(s, i, j, dg, z, kc[k]) : rr[k];
        const float zo = (zin ? z.c : 0.f) + w * (rv / dg);
        zout[idx] = zo;
        if (ZQ) {
          const double zv = zo;
          v[0] += zv * qv[k];
          v[1] += zv;
        }
      }
    }
  }
  if (ZQ) {
    __shared__ double lds[8];
    store_partials<double, 2>(v, zpart, gridDim.x, blockIdx.x, lds);
  }
}

// Two weighted Jacobi sweeps from a zero guess in one pass: the first
// sweep's values z1 = 0 + w D^-1 r at the five points are formed on the fly
// (the same arithmetic as k_jacobi32 storing them), then
// zout = z1 + w D^-1 (r - A z1).
__global__ __launch_bounds__(256) void k_jacobi32_2z(Sten<float> s,
                                                     const float* __restrict__ r,
                                                     float* __restrict__ zout, float w,
                                                     Tiling tl) {
  const int nx = s.nx, ny = s.ny;
  for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
    int i0, j0;
    tile_origin(t, tl.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
    constexpr int NR = TI / BY;
    Nbr<float> rn[NR], kc[NR];
    float kxm[NR], kxp[NR];
    const int jc = min(j, ny - 1);
    const float kym = s.KY[max(jc - 1, 0)], kyp = s.KY[min(jc + 2, ny)];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      load_nbr(s, r, i, j, rn[k], kc[k]);
      const int ic = min(i, nx - 1);
      kxm[k] = s.KX[max(ic - 1, 0)];
      kxp[k] = s.KX[min(ic + 2, nx)];
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i < nx && j < ny) {
        const float dg = diag_from(s, i, j, kc[k].kx0, kc[k].kx1, kc[k].ky0, kc[k].ky1);
        Nbr<float> z1;
        z1.c = 0.f + w * (rn[k].c / dg);
        z1.n = i > 0 ? 0.f + w * (rn[k].n / diag_from(s, i - 1, j, kxm[k], kc[k].kx0,
                                                      kc[k].ky0, kc[k].ky1))
                     : 0.f;
        z1.s = i < nx - 1 ? 0.f + w * (rn[k].s / diag_from(s, i + 1, j, kc[k].kx1, kxp[k],
                                                           kc[k].ky0, kc[k].ky1))
                          : 0.f;
        z1.w = j > 0 ? 0.f + w * (rn[k].w / diag_from(s, i, j - 1, kc[k].kx0, kc[k].kx1,
                                                      kym, kc[k].ky0))
                     : 0.f;
        z1.e = j < ny - 1 ? 0.f + w * (rn[k].e / diag_from(s, i, j + 1, kc[k].kx0, kc[k].kx1,
                                                           kc[k].ky1, kyp))
                          : 0.f;
        const float rv = rn[k].c - apply_regs(s, i, j, dg, z1, kc[k]);
        zout[(size_t)i * ny + j] = z1.c + w * (rv / dg);
      }
    }
  }
}

// Two weighted Jacobi sweeps from a zero guess and the residual they leave,
// in one pass over an LDS tile: z1 = w D^-1 r on the tile and a halo of two,
// z2 = z1 + w D^-1 (r - A z1) on the tile and a halo of one (stored as zout
// on the tile), rout = r - A z2 on the tile, and the FP64 partial sums of
// rout^2 per block into npart.  Per value the arithmetic of k_jacobi32_2z
// followed by k_resid32 (values outside the grid are never used).
constexpr int kZR = 2;  // halo of the r and z1 tiles
__global__ __launch_bounds__(256) void k_jacobi32_2zr(Sten<float> s,
                                                      const float* __restrict__ r,
                                                      float* __restrict__ zout,
                                                      float* __restrict__ rout, float w,
                                                      Tiling tl, double* npart) {
  constexpr int RH = TI + 2 * kZR, RW = TJ + 2 * kZR;  // r and z1: 20 x 68
  constexpr int ZH = TI + 2, ZW = TJ + 2;              // z2: 18 x 66
  __shared__ float rt[RH][RW], z1t[RH][RW], z2t[ZH][ZW];
  __shared__ float kxs[RH + 1], kys[RW + 1];
  const int nx = s.nx, ny = s.ny;
  const int tid = threadIdx.x + threadIdx.y * TJ;
  double v[1] = {0.0};
  for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
    int i0, j0;
    tile_origin(t, tl.ntj, i0, j0);
    // r and the face coefficients, every load first (clamped addresses)
    constexpr int NE = RH * RW, NPT = (NE + 255) / 256;
    float rv[NPT];
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
      const int e = min(tid + 256 * k, NE - 1);
      const int a = e / RW, b = e - a * RW;
      const int gi = min(max(i0 - kZR + a, 0), nx - 1), gj = min(max(j0 - kZR + b, 0), ny - 1);
      rv[k] = r[(size_t)gi * ny + gj];
    }
    float kv = 0.f;
    if (tid < RH + 1)
      kv = s.KX[min(max(i0 - kZR + tid, 0), nx)];
    else if (tid < RH + 1 + RW + 1)
      kv = s.KY[min(max(j0 - kZR + tid - (RH + 1), 0), ny)];
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
      const int e = tid + 256 * k;
      if (e < NE) {
        const int a = e / RW, b = e - a * RW;
        rt[a][b] = rv[k];
      }
    }
    if (tid < RH + 1)
      kxs[tid] = kv;
    else if (tid < RH + 1 + RW + 1)
      kys[tid - (RH + 1)] = kv;
    __syncthreads();
    // z1 on the tile and a halo of two
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
      const int e = tid + 256 * k;
      if (e < NE) {
        const int a = e / RW, b = e - a * RW;
        const int gi = i0 - kZR + a, gj = j0 - kZR + b;
        float z1 = 0.f;
        if (gi >= 0 && gi < nx && gj >= 0 && gj < ny) {
          const float dg = diag_from(s, gi, gj, kxs[a], kxs[a + 1], kys[b], kys[b + 1]);
          z1 = 0.f + w * (rt[a][b] / dg);
        }
        z1t[a][b] = z1;
      }
    }
    __syncthreads();
    // z2 on the tile and a halo of one; the tile's values stored
    constexpr int NZ = ZH * ZW, NZT = (NZ + 255) / 256;
#pragma unroll
    for (int k = 0; k < NZT; ++k) {
      const int e = tid + 256 * k;
      if (e < NZ) {
        const int a = e / ZW, b = e - a * ZW;  // z2t[a][b] = z1t[a + 1][b + 1]'s cell
        const int gi = i0 - 1 + a, gj = j0 - 1 + b;
        float z2 = 0.f;
        if (gi >= 0 && gi < nx && gj >= 0 && gj < ny) {
          const int ra = a + 1, rb = b + 1;
          const float kx0 = kxs[ra], kx1 = kxs[ra + 1], ky0 = kys[rb], ky1 = kys[rb + 1];
          const float dg = diag_from(s, gi, gj, kx0, kx1, ky0, ky1);
          const float c = z1t[ra][rb];
          const float res = rt[ra][rb] - row_apply(s, gi, gj, dg, c, z1t[ra - 1][rb],
                                                   z1t[ra + 1][rb], z1t[ra][rb - 1],
                                                   z1t[ra][rb + 1], kx0, kx1, ky0, ky1);
          z2 = c + w * (res / dg);
          if (a >= 1 && a <= TI && b >= 1 && b <= TJ) zout[(size_t)gi * ny + gj] = z2;
        }
        z2t[a][b] = z2;
      }
    }
    __syncthreads();
    // rout = r - A z2 on the tile
    const int lj = threadIdx.x + 1, j = j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < TI / BY; ++k) {
      const int li = threadIdx.y + BY * k + 1, i = i0 + li - 1;
      if (i < nx && j < ny) {
        const int ra = li + 1, rb = lj + 1;
        const float kx0 = kxs[ra], kx1 = kxs[ra + 1], ky0 = kys[rb], ky1 = kys[rb + 1];
        const float dg = diag_from(s, i, j, kx0, kx1, ky0, ky1);
        const float ro = rt[ra][rb] - row_apply(s, i, j, dg, z2t[li][lj], z2t[li - 1][lj],
                                                z2t[li + 1][lj], z2t[li][lj - 1],
                                                z2t[li][lj + 1], kx0, kx1, ky0, ky1);
        rout[(size_t)i * ny + j] = ro;
        v[0] += (double)ro * ro;
      }
    }
    __syncthreads();  // the tiles are reused by the block's next tile
  }
  __shared__ double lds[4];
  store_partials<double, 1>(v, npart, gridDim.x, blockIdx.x, lds);
}

// rout = r - A z (FP32).  With npart given, also the FP64 partial sums of
// rout^2 per block (the network input's sum of squares, as k_sumsq forms
// it from the stored rout).
__global__ __launch_bounds__(256) void k_resid32(Sten<float> s,
                                                 const float* __restrict__ r,
                                                 const float* __restrict__ z,
                                                 float* __restrict__ rout, Tiling tl,
                                                 double* npart) {
  double v[1] = {0.0};
  for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
    int i0, j0;
    tile_origin(t, tl.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
    constexpr int NR = TI / BY;
    Nbr<float> zn[NR], kc[NR];
    float rr[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      load_nbr(s, z, i, j, zn[k], kc[k]);
      rr[k] = r[(size_t)min(i, s.nx - 1) * s.ny + min(j, s.ny - 1)];
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i < s.nx && j < s.ny) {
        const float dg = diag_from(s, i, j, kc[k].kx0, kc[k].kx1, kc[k].ky0, kc[k].ky1);
        const float ro = rr[k] - apply_regs(s, i, j, dg, zn[k], kc[k]);
        rout[(size_t)i * s.ny + j] = ro;
        v[0] += (double)ro * ro;
      }
    }
  }
  if (npart) {
    __shared__ double lds[4];
    store_partials<double, 1>(v, npart, gridDim.x, blockIdx.x, lds);
  }
}

// a += b
__global__ __launch_bounds__(256) void k_add32(float* __restrict__ a,
                                               const float* __restrict__ b, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    a[k] = a[k] + b[k];
}

int fcg_alloc(cpgpu_poisson* P) {
  int rc = poisson_prepare_f32(P);
  if (rc) return rc;
  if (P->pf[0].alloc(P->N) || P->pf[1].alloc(P->N) || P->fst.alloc(1) ||
      P->npart.alloc(kMaxBlocks))
    return set_error(P->ctx, CPGPU_ENOMEM, "flexible-CG vectors");
  if (!P->h_fst &&
      hipHostMalloc((void**)&P->h_fst, sizeof(FState), hipHostMallocDefault) !=
          hipSuccess) {
    P->h_fst = nullptr;
    return set_error(P->ctx, CPGPU_ENOMEM, "pinned flexible-CG state");
  }
  if (!P->h_mail) {
    // without a mapped mailbox the state is copied back (read_state)
    if (hipHostMalloc((void**)&P->h_mail, sizeof(FMail),
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void**)&P->d_mail, P->h_mail, 0) != hipSuccess) {
      if (P->h_mail) (void)hipHostFree(P->h_mail);
      P->h_mail = nullptr;
      P->d_mail = nullptr;
      (void)hipGetLastError();
    }
  }
  return CPGPU_OK;
}

// z (P->pa32) = M(r) for r in P->r32: `pre` weighted-Jacobi sweeps from
// zero, the network on the remaining residual, `post` sweeps -- a two-level
// cycle whose coarse correction is the network.  Work buffers: x32, pb32,
// q32, b32.
// z = M(r): pre-smoothing, the network's correction (unless use_net is
// false: the smoothing-only cycle), post-smoothing; output in P->pa32.
// qprev/zpart given: the last post sweep also forms k_fcg_zq's partials
// (tl.G of each) and *zq_done says so.
int apply_precond(cpgpu_poisson* P, bool use_net, const double* qprev = nullptr,
                  double* zpart = nullptr, bool* zq_done = nullptr) {
  if (zq_done) *zq_done = false;
  hipStream_t st = P->ctx->stream;
  const Tiling tl = P->tl;
  const Sten<float> s = P->sten32();
  const float w = (float)P->pc_omega;
  const float* r = P->r32.p;
  float* bufA = P->x32.p;
  float* bufB = P->pb32.p;
  float* z = nullptr;  // null: z = 0
  int k0 = 0;
  // exactly two pre-sweeps before the network: both and the residual the
  // network takes in one pass (its sum-of-squares partials with it)
  const bool zr = use_net && P->pc_pre == 2 && P->npart.p && (int)tl.G <= kMaxBlocks;
  if (zr) {
    hipLaunchKernelGGL(k_jacobi32_2zr, dim3(tl.G), dim3(TJ, BY), 0, st, s, r, bufA,
                       P->q32.p, w, tl, P->npart.p);
    z = bufA;
    k0 = 2;
  } else if (P->pc_pre >= 2) {  // the first two sweeps from zero in one pass
    hipLaunchKernelGGL(k_jacobi32_2z, dim3(tl.G), dim3(TJ, BY), 0, st, s, r, bufA, w, tl);
    z = bufA;
    k0 = 2;
  }
  for (int k = k0; k < P->pc_pre; ++k) {
    float* out = (z == bufA) ? bufB : bufA;
    hipLaunchKernelGGL(k_jacobi32<false>, dim3(tl.G), dim3(TJ, BY), 0, st, s, r,
                       (const float*)z, (const float*)nullptr, out, w, tl,
                       (const double*)nullptr, (double*)nullptr);
    z = out;
  }
  const float* zadd = nullptr;  // a correction the next sweep adds to z
  if (use_net) {
    const float* rn = r;
    const double* sq = nullptr;  // the input's sum-of-squares partials
    if (zr) {
      sq = P->npart.p;
      rn = P->q32.p;
    } else if (z) {
      if (P->npart.p && (int)tl.G <= kMaxBlocks) sq = P->npart.p;
      hipLaunchKernelGGL(k_resid32, dim3(tl.G), dim3(TJ, BY), 0, st, s, r,
                         (const float*)z, P->q32.p, tl, (double*)sq);
      rn = P->q32.p;
    }
    float* c = P->b32.p;
    int rc = unet_apply_device(P->unet, P->nx, P->ny, rn, c, pack_walls(P), sq,
                               sq ? (int)tl.G : 0);
    if (rc) return rc;
    if (!z)
      z = c;
    else if (P->pc_post > 0)
      zadd = c;  // the first post sweep forms z + c
    else
      hipLaunchKernelGGL(k_add32, dim3(1024), dim3(256), 0, st, z, c, P->N);
  }
  // post sweeps alternate between two buffers other than z's, the last one
  // writing P->pa32
  for (int k = 0; k < P->pc_post; ++k) {
    const int left = P->pc_post - 1 - k;
    float* spare = (z == bufA) ? bufB : bufA;
    float* out = (left % 2 == 0) ? P->pa32.p : spare;
    if (out == z) out = (z == bufA) ? bufB : bufA;
    if (left == 0 && zpart && out == P->pa32.p) {
      hipLaunchKernelGGL(k_jacobi32<true>, dim3(tl.G), dim3(TJ, BY), 0, st, s, r,
                         (const float*)z, zadd, out, w, tl, qprev, zpart);
      *zq_done = true;
    } else {
      hipLaunchKernelGGL(k_jacobi32<false>, dim3(tl.G), dim3(TJ, BY), 0, st, s, r,
                         (const float*)z, zadd, out, w, tl, (const double*)nullptr,
                         (double*)nullptr);
    }
    z = out;
    zadd = nullptr;
  }
  if (!z) z = P->r32.p;  // no sweeps and no network: the identity
  if (z != P->pa32.p)
    CPGPU_HIP(P->ctx, hipMemcpyAsync(P->pa32.p, z, sizeof(float) * P->N,
                                     hipMemcpyDeviceToDevice, st));
  CPGPU_HIP(P->ctx, hipGetLastError());
  return CPGPU_OK;
}

// One flexible-CG iteration: the preconditioner cycle, then the Krylov
// kernels.  first: the first iteration (no previous direction); the
// direction buffers alternate with the iteration's parity.
int launch_fcg_iter(cpgpu_poisson* P, bool first, int parity, bool singular,
                    int norm_ref) {
  hipStream_t st = P->ctx->stream;
  const Tiling tl = P->tl;
  const size_t N = P->N;
  const int G1 = 1024;  // blocks of the 1-D kernels
  const double* qprev = first ? nullptr : (const double*)P->q.p;
  bool zq_done = false;
  int rc = apply_precond(P, true, qprev, P->part64.p, &zq_done);
  if (rc) return rc;
  if (!zq_done)
    hipLaunchKernelGGL(k_fcg_zq, dim3(G1), dim3(256), 0, st, P->pa32.p, qprev,
                       P->part64.p, N);  // 2 x G1 partials
  const int GZ = zq_done ? (int)tl.G : G1;
  const double* pold = P->pf[parity].p;
  double* pnew = P->pf[parity ^ 1].p;
  hipLaunchKernelGGL(k_fcg_dir, dim3(tl.G), dim3(TJ, BY), 0, st, P->sten64(), P->pa32.p,
                     pold, pnew, P->q.p, P->r.p, P->part64.p, GZ, P->fst.p, first ? 1 : 0,
                     singular ? 1 : 0, P->partB64.p, tl);
  hipLaunchKernelGGL(k_fcg_upd, dim3(G1), dim3(256), 0, st, pnew, P->q.p, P->x.p, P->r.p,
                     P->r32.p, P->partB64.p, (int)tl.G, P->fst.p, P->partR64.p, N, norm_ref);
  hipLaunchKernelGGL(k_fcg_check, dim3(1), dim3(256), 0, st, P->partR64.p, G1, P->fst.p,
                     P->d_mail);
  CPGPU_HIP(P->ctx, hipGetLastError());
  return CPGPU_OK;
}

// The iteration (not the first) captured once per configuration and
// replayed: one graph launch in place of the cycle's ~30 kernel launches.
// Captured on the context stream itself, which
// apply_precond and the network launch on; the network's activations are
// already allocated (the first iteration ran directly).  Null: launch
// directly.
hipGraphExec_t fcg_graph(cpgpu_poisson* P, int parity, bool singular, int norm_ref) {
  long long om;
  std::memcpy(&om, &P->pc_omega, sizeof(om));
  const std::vector<long long> key{parity, singular ? 1 : 0, norm_ref, (long long)(uintptr_t)P->unet,
                                   unet_generation(P->unet), P->pc_pre, P->pc_post, om,
                                   pack_walls(P), (long long)(uintptr_t)P->pf[0].p};
  for (auto& g : P->fcg_graphs)
    if (g.first == key) return g.second;
  hipStream_t st = P->ctx->stream;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
  const int rc = launch_fcg_iter(P, false, parity, singular, norm_ref);
  const hipError_t ec = hipStreamEndCapture(st, &graph);
  if (rc == CPGPU_OK && ec == hipSuccess && graph &&
      hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
    exec = nullptr;
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipGetLastError();
  if (exec) P->fcg_graphs.push_back({key, exec});
  return exec;
}

// The state after update number `seq` (the start: 0).  With the mailbox the
// host polls the device's own write (a few microseconds after the check
// kernel instead of a copy and a stream synchronisation); a mailbox that
// shows nothing within 5 s of polling (or none at all) falls back to the
// copy, after which the stream is idle and the state final.
int read_state(cpgpu_poisson* P, FState* out, int seq) {
  hipStream_t st = P->ctx->stream;
  if (P->h_mail) {
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned spin = 0;; ++spin) {
      if (__atomic_load_n(&P->h_mail->seq, __ATOMIC_ACQUIRE) == seq) {
        out->resid = P->h_mail->resid;
        out->breakdown = P->h_mail->breakdown;
        out->iter = seq;
        return CPGPU_OK;
      }
      if ((spin & 1023u) == 1023u &&
          std::chrono::steady_clock::now() - t0 > std::chrono::seconds(5))
        break;
    }
  }
  CPGPU_HIP(P->ctx, hipMemcpyAsync(P->h_fst, P->fst.p, sizeof(FState),
                                   hipMemcpyDeviceToHost, st));
  CPGPU_HIP(P->ctx, hipStreamSynchronize(st));
  *out = *P->h_fst;
  return CPGPU_OK;
}

}  // namespace

int unet_warm_start(cpgpu_poisson* P) {
  if (!P->unet) return set_error(P->ctx, CPGPU_EINVAL, "no U-Net attached");
  int rc = poisson_prepare_f32(P);
  if (rc) return rc;
  hipStream_t st = P->ctx->stream;
  hipLaunchKernelGGL(k_f64_to_f32, dim3(1024), dim3(256), 0, st, P->b.p,
                     P->b32.p, P->N);
  if ((rc = unet_apply_device(P->unet, P->nx, P->ny, P->b32.p, P->x32.p, pack_walls(P))))
    return rc;
  hipLaunchKernelGGL(k_f32_to_f64, dim3(1024), dim3(256), 0, st, P->x32.p,
                     P->x.p, P->N);
  CPGPU_HIP(P->ctx, hipGetLastError());
  return CPGPU_OK;
}

int run_fcg_unet(cpgpu_poisson* P, int max_iter, double tol, int norm_ref,
                 int project, double* hist, int hist_cap, int* hist_len,
                 int* iters, double* resid) {
  cpgpu_ctx* ctx = P->ctx;
  if (!P->unet || !(P->unet_flags & CPGPU_UNET_PRECOND))
    return set_error(ctx, CPGPU_EINVAL,
                     "CPGPU_FCG_UNET needs a U-Net attached with CPGPU_UNET_PRECOND");
  int rc = fcg_alloc(P);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  const Tiling tl = P->tl;
  const size_t N = P->N;
  const int G1 = 1024;  // blocks of the 1-D kernels
  *hist_len = 0;
  *iters = 0;
  if (P->unet_flags & CPGPU_UNET_WARM) {
    if ((rc = unet_warm_start(P))) return rc;
  }
  hipLaunchKernelGGL(k_fcg_init, dim3(tl.G), dim3(TJ, BY), 0, st, P->sten64(),
                     P->x.p, P->b.p, P->r.p, P->r32.p, P->partB64.p, tl,
                     norm_ref);
  if (P->h_mail) __atomic_store_n(&P->h_mail->seq, -1, __ATOMIC_RELEASE);
  hipLaunchKernelGGL(k_fcg_start, dim3(1), dim3(256), 0, st, P->partB64.p,
                     (int)tl.G, P->fst.p, P->d_mail);
  CPGPU_HIP(ctx, hipGetLastError());
  FState fs;
  if ((rc = read_state(P, &fs, 0))) return rc;
  double res = fs.resid;
  std::vector<double> trace;
  trace.push_back(res);
  constexpr int kStallWindow = 40;
  constexpr int kProbe = 24;  // first projection of a batched member's progress
  bool singular = true;  // all-Neumann: constants are the null space
  for (int sg = 0; sg < 4; ++sg)
    singular = singular && P->seg[sg].type == CPGPU_BC_NEUMANN;
  double best = res;
  std::vector<double> best_hist(1, res);
  int k = 0;
  bool fallback = false;
  // replayed iteration graphs when asked for, unless kernel timers are on
  // (their events bracket direct launches)
  bool timed = false;
  for (int f = 0; f < cpgpu_ctx::kFamilies; ++f) timed = timed || ctx->timer_on[f];
  const bool use_graphs = P->fcg_graphs_on && !timed;
  while (res > tol) {
    if (k >= max_iter) break;
    hipGraphExec_t ge = (use_graphs && k > 0) ? fcg_graph(P, k & 1, singular, norm_ref) : nullptr;
    if (ge) {
      CPGPU_HIP(ctx, hipGraphLaunch(ge, st));
    } else if ((rc = launch_fcg_iter(P, k == 0, k & 1, singular, norm_ref))) {
      return rc;
    }
    ++k;
    if ((rc = read_state(P, &fs, k))) return rc;
    res = fs.resid;
    trace.push_back(res);
    if (fs.breakdown || !std::isfinite(res)) {
      fallback = true;
      break;
    }
    // stall: the best residual improved by less than 5% over the last
    // kStallWindow iterations (the residual norm of a CG method is not
    // monotone and plateaus for tens of iterations while the error still
    // falls, so short windows and steep thresholds misfire)
    best = std::min(best, res);
    best_hist.push_back(best);
    if (k >= kStallWindow && best > 0.95 * best_hist[k - kStallWindow]) {
      // The cycle no longer reduces the residual: the network maps the
      // residual to corrections that do not reduce it (a
      // deterministic network returns the same correction for a residual
      // that stays put).  Jacobi-PCG finishes from the current iterate: per
      // unit of progress its graph-batched iterations are cheaper than the
      // cycle without the network.
      fallback = true;
      break;
    }
    // Without a fallback of its own (a member of a batched solve, whose
    // batched pass takes over from this iterate), hand over as soon as the
    // mean reduction so far projects past twice the iteration budget.
    if (!P->fcg_fallback && k >= kProbe && k % 8 == 0) {
      const double got = std::log(best / trace[0]), need = std::log(tol / trace[0]);
      if (!(got < 0.0) || (double)k * need / got > 2.0 * max_iter) {
        fallback = true;
        break;
      }
    }
  }
  const int kf = k;
  int code = res <= tol ? CPGPU_OK : CPGPU_MAXITER;
  if (code != CPGPU_OK && P->fcg_fallback && (fallback || k < max_iter)) {
    // Jacobi-PCG from the current iterate with the remaining budget
    int it = std::max(0, max_iter - kf);
    int done = 0;
    double r2 = res;
    if (it > 0) {
      if ((rc = poisson_jpcg64(P, &it, tol, norm_ref, &r2, &done))) return rc;
      k += it;
      res = r2;
      code = done ? CPGPU_OK : CPGPU_MAXITER;
    }
  }
  if (project) {
    hipLaunchKernelGGL(k_sum64, dim3(G1), dim3(256), 0, st, P->x.p,
                       P->partR64.p, N);
    hipLaunchKernelGGL(k_sub_mean_part, dim3(G1), dim3(256), 0, st, P->x.p,
                       P->partR64.p, G1, N);
    CPGPU_HIP(ctx, hipGetLastError());
  }
  CPGPU_HIP(ctx, hipStreamSynchronize(st));
  if (hist && hist_cap > 0) {
    const int nh = std::min<int>((int)trace.size() - 1, hist_cap);
    for (int i = 0; i < nh; ++i) hist[i] = trace[i + 1];
    *hist_len = nh;
  }
  *iters = k;
  *resid = res;
  // the flexible-CG part of the count, for callers that separate them
  P->fcg_iters = kf;
  return code;
}

}  // namespace cpgpu
