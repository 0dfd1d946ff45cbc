// Poisson problems on gfx950: descriptor build (host), right-hand side,
// Jacobi-PCG / CG in FP64 and FP32, FP32-in-FP64 iterative refinement, and
// the C-ABI of include/cpgpu.h for them.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include <mutex>

#include "poisson.hpp"

#pragma clang fp contract(off)

namespace cpgpu {

// ---------------------------------------------------------------------
// kernels other than the Krylov iteration (stencil.hpp)


__global__ __launch_bounds__(256) void k_build_rhs(RhsDesc d,
                                                   const double* __restrict__ src,
                                                   double src_uniform,
                                                   double* __restrict__ b,
                                                   Tiling tl) {
  for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
    int i0, j0;
    tile_origin(t, tl.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
    for (int k = 0; k < TI / BY; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i >= d.nx || j >= d.ny) continue;
      const size_t idx = (size_t)i * d.ny + j;
      double cst = d.vol[i] * (src ? src[idx] : src_uniform);
      if (i == 0) cst = ghost_const(d.seg[0], cst, d.bcval[j]);
      if (j == 0) cst = ghost_const(d.seg[2], cst, d.bcval[2 * d.ny + i]);
      if (j == d.ny - 1)
        cst = ghost_const(d.seg[3], cst, d.bcval[2 * d.ny + d.nx + i]);
      if (i == d.nx - 1) cst = ghost_const(d.seg[1], cst, d.bcval[d.ny + j]);
      b[idx] = cst == 0.0 ? 0.0 : -cst;
    }
  }
}

// Totals of NV partial slots of G blocks into out[0..NV) (one block).
template <int NV>
__global__ __launch_bounds__(256) void k_finalize(const double* __restrict__ partials,
                                                  int G, double* out) {
  __shared__ double lds[4 * NV];
  double tot[NV];
  reduce_partials<double, NV>(partials, G, tot, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = tot[k];
  }
}

// r = b - A x [- A xlo] (FP64, difference form: the true residual of the
// mixed solves and of cpgpu_poisson_residual); partials of b.b, r.r, sum(r)
__global__ __launch_bounds__(256) void k_residual64(Sten<double> s,
                                                    const double* __restrict__ x,
                                                    const double* __restrict__ xlo,
                                                    const double* __restrict__ b,
                                                    double* __restrict__ r,
                                                    double* partials, Tiling tl) {
  __shared__ double lds[12];
  double v[3] = {0.0, 0.0, 0.0};
  for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
    int i0, j0;
    tile_origin(t, tl.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
    for (int k = 0; k < TI / BY; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i >= s.nx || j >= s.ny) continue;
      const size_t idx = (size_t)i * s.ny + j;
      const double bv = b[idx];
      double rv = bv - apply_global_diff(s, x, i, j);
      if (xlo) rv = rv - apply_global_diff(s, xlo, i, j);
      r[idx] = rv;
      v[0] += bv * bv;
      v[1] += rv * rv;
      v[2] += rv;
    }
  }
  store_partials<double, 3>(v, partials, gridDim.x, blockIdx.x, lds);
}

// out32 = (float)(r - mean(r)) (mean only when projecting); e = 0
__global__ __launch_bounds__(256) void k_to_f32(const double* __restrict__ r,
                                                float* __restrict__ out,
                                                float* __restrict__ e,
                                                const RState* st, size_t n,
                                                int project) {
  const double mean = project ? st->rsum / (double)n : 0.0;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    out[k] = (float)(r[k] - mean);
    e[k] = 0.f;
  }
}

// (x, xlo) += (double)e exactly (two-sum); partial sum of the new x
__global__ __launch_bounds__(256) void k_add_f32(double* __restrict__ x,
                                                 double* __restrict__ xlo,
                                                 const float* __restrict__ e,
                                                 double* partials, size_t n) {
  __shared__ double lds[4];
  double v[1] = {0.0};
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    double hi, lo;
    two_sum(x[k], xlo[k] + (double)e[k], hi, lo);
    x[k] = hi;
    xlo[k] = lo;
    v[0] += hi;
  }
  store_partials<double, 1>(v, partials, gridDim.x, blockIdx.x, lds);
}

// x -= total/n, total = *sum
__global__ __launch_bounds__(256) void k_sub_mean(double* __restrict__ x,
                                                  const double* sum, size_t n) {
  const double mean = *sum / (double)n;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    x[k] = x[k] - mean;
}

// (x, xlo) -= total/n exactly (two-sum)
__global__ __launch_bounds__(256) void k_sub_mean2(double* __restrict__ x,
                                                   double* __restrict__ xlo,
                                                   const double* sum, size_t n) {
  const double mean = *sum / (double)n;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    double hi, lo;
    two_sum(x[k], xlo[k] - mean, hi, lo);
    x[k] = hi;
    xlo[k] = lo;
  }
}

// x = x + xlo (the FP64 value nearest the pair), xlo = 0
__global__ __launch_bounds__(256) void k_round_x(double* __restrict__ x,
                                                 double* __restrict__ xlo, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    x[k] = x[k] + xlo[k];
    xlo[k] = 0.0;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T* __restrict__ x, T v, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    x[k] = v;
}

// ---- synthetic right-hand sides (no reference counterpart) ----
// Sum of Gaussian blobs f = sum_k A_k exp(-|x-c_k|^2 / (2 w_k^2)) at cell
// centres, with the partial sums of f for the zero-mean projection.
__global__ __launch_bounds__(256) void k_blobs(int nx, int ny,
                                               const double* __restrict__ cx,
                                               const double* __restrict__ cy,
                                               const double* __restrict__ prm,
                                               int nblobs, double* __restrict__ f,
                                               double* partials, Tiling tl) {
  __shared__ double lds[4];
  double v[1] = {0.0};
  for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
    int i0, j0;
    tile_origin(t, tl.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
    for (int k = 0; k < TI / BY; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i >= nx || j >= ny) continue;
      const double x = cx[i], y = cy[j];
      double s = 0.0;
      for (int b = 0; b < nblobs; ++b) {
        const double dx = x - prm[4 * b], dy = y - prm[4 * b + 1];
        const double w = prm[4 * b + 2];
        s += prm[4 * b + 3] * exp(-(dx * dx + dy * dy) / (2.0 * w * w));
      }
      f[(size_t)i * ny + j] = s;
      v[0] += s;
    }
  }
  store_partials<double, 1>(v, partials, gridDim.x, blockIdx.x, lds);
}

void blob_params(unsigned long long seed, int nblobs, double wmin, double wmax,
                 double* out) {
  Xoshiro g(seed);
  for (int b = 0; b < nblobs; ++b) {
    out[4 * b] = g.uniform();
    out[4 * b + 1] = g.uniform();
    out[4 * b + 2] = wmin + (wmax - wmin) * g.uniform();
    out[4 * b + 3] = g.normal();
  }
}

// ---------------------------------------------------------------------
// host helpers

namespace {

bool all_equal(const double* a, int n) {
  for (int k = 1; k < n; ++k)
    if (std::memcmp(&a[k], &a[0], sizeof(double)) != 0) return false;
  return true;
}

// Face coefficients K = -(-beta*a/d) (s_poisson.hpp:98-135), diagonal
// adjustments and rhs constants of the ghost substitution
// (s_ghost.hpp:140-179), source weights and the boundary values of the
// interior lines, into caller arrays: KX (nx+1), KY (ny+1), adj[4], seg[4],
// vol (nx), bcv (2 ny + 2 nx).
int describe_grid(cpgpu_ctx* ctx, const cpgpu_grid2d* g, double* KX, double* KY,
                  double* adj, RhsSeg* seg, double* vol, double* bcv) {
  const int nx = g->nx, ny = g->ny, gl = g->gl;
  auto cx = [&](int i) { return g->cx[i + gl]; };
  auto sx = [&](int i) { return g->sx[i + gl]; };
  auto cy = [&](int j) { return g->cy[j + gl]; };
  auto sy = [&](int j) { return g->sy[j + gl]; };
  // the 1-D face coefficients are exact only when a = sy[j] (x faces) and
  // sx[i] (y faces) do not vary along the face's own line; other grids take
  // the per-face arrays of general.hip (KX, KY and adj here are then unused)
  const double beta = g->beta;
  for (int f = 0; f <= nx; ++f) {
    const double a = sy(0) * 1.0;
    const double d = std::fabs(cx(f) - cx(f - 1));
    const double coe = -beta * a / d;
    KX[f] = -coe;
  }
  for (int f = 0; f <= ny; ++f) {
    const double a = 1.0 * sx(0);
    const double d = std::fabs(cy(f) - cy(f - 1));
    const double coe = -beta * a / d;
    KY[f] = -coe;
  }
  // segments: 0 x- (ghost i=-1), 1 x+ (i=nx), 2 y- (j=-1), 3 y+ (j=ny)
  for (int s = 0; s < 4; ++s) {
    const int type = g->bc_type[s];
    if (type != CPGPU_BC_DIRICHLET && type != CPGPU_BC_NEUMANN)
      return set_error(ctx, CPGPU_EINVAL, "bad bc type %d on segment %d",
                       type, s);
    double c0, c1, cf, coe;
    if (s == 0) {
      c0 = cx(-1); c1 = cx(0); cf = c1 - sx(0) * 0.5; coe = -KX[0];
    } else if (s == 1) {
      c0 = cx(nx); c1 = cx(nx - 1); cf = c1 + sx(nx - 1) * 0.5; coe = -KX[nx];
    } else if (s == 2) {
      c0 = cy(-1); c1 = cy(0); cf = c1 - sy(0) * 0.5; coe = -KY[0];
    } else {
      c0 = cy(ny); c1 = cy(ny - 1); cf = c1 + sy(ny - 1) * 0.5; coe = -KY[ny];
    }
    const double dl = std::fabs(c1 - c0);
    const double d0 = std::fabs(cf - c0);
    const double d1 = std::fabs(c1 - cf);
    adj[s] = type == CPGPU_BC_DIRICHLET ? (-coe * d0 / d1) : coe;
    seg[s].type = type;
    seg[s].coe = coe;
    seg[s].dl = dl;
    seg[s].d1 = d1;
  }
  for (int i = 0; i < nx; ++i) {
    double v = 1.0;
    for (int d = 0; d < 2; ++d) v *= sx(i);  // s_poisson.hpp:71-77 uses dim 0
    vol[i] = v;
  }
  // boundary values of the interior lines
  const int NX = nx + 2 * gl, NY = ny + 2 * gl;
  for (int k = 0; k < 2 * ny + 2 * nx; ++k) bcv[k] = 0.0;
  if (g->bc_val) {
    for (int j = 0; j < ny; ++j) {
      bcv[j] = g->bc_val[j + gl];
      bcv[ny + j] = g->bc_val[NY + j + gl];
    }
    for (int i = 0; i < nx; ++i) {
      bcv[2 * ny + i] = g->bc_val[2 * NY + i + gl];
      bcv[2 * ny + nx + i] = g->bc_val[2 * NY + NX + i + gl];
    }
  }
  return CPGPU_OK;
}

bool uniform_cells(const cpgpu_grid2d* g) {
  return all_equal(g->sy + g->gl, g->ny) && all_equal(g->sx + g->gl, g->nx);
}

int build_descriptor(cpgpu_poisson* P, const cpgpu_grid2d* g) {
  const int nx = g->nx, ny = g->ny, gl = g->gl;
  P->KX.assign(nx + 1, 0.0);
  P->KY.assign(ny + 1, 0.0);
  P->vol.assign(nx, 0.0);
  P->bcv.assign(2 * ny + 2 * nx, 0.0);
  int rc = describe_grid(P->ctx, g, P->KX.data(), P->KY.data(), P->adj, P->seg,
                         P->vol.data(), P->bcv.data());
  if (rc) return rc;
  if (!uniform_cells(g) && (rc = general_build(P, g))) return rc;
  P->cxi.assign(g->cx + gl, g->cx + gl + nx);
  P->sxi.assign(g->sx + gl, g->sx + gl + nx);
  P->syi.assign(g->sy + gl, g->sy + gl + ny);
  P->cyi.assign(g->cy + gl, g->cy + gl + ny);
  return CPGPU_OK;
}

int upload(cpgpu_ctx* ctx, void* dst, const void* src, size_t bytes) {
  CPGPU_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice,
                                ctx->stream));
  return CPGPU_OK;
}

// interior rows of a host cell field <-> interior-only device vector
int field_to_dev(cpgpu_poisson* P, const double* field, double* dev) {
  const int NY = P->ny + 2 * P->gl;
  CPGPU_HIP(P->ctx, hipMemcpy2DAsync(dev, sizeof(double) * P->ny,
                                     field + (size_t)P->gl * NY + P->gl,
                                     sizeof(double) * NY,
                                     sizeof(double) * P->ny, P->nx,
                                     hipMemcpyHostToDevice, P->ctx->stream));
  CPGPU_HIP(P->ctx, hipStreamSynchronize(P->ctx->stream));
  return CPGPU_OK;
}

int dev_to_field(cpgpu_poisson* P, const double* dev, double* field) {
  const int NY = P->ny + 2 * P->gl;
  CPGPU_HIP(P->ctx, hipMemcpy2DAsync(field + (size_t)P->gl * NY + P->gl,
                                     sizeof(double) * NY, dev,
                                     sizeof(double) * P->ny,
                                     sizeof(double) * P->ny, P->nx,
                                     hipMemcpyDeviceToHost, P->ctx->stream));
  CPGPU_HIP(P->ctx, hipStreamSynchronize(P->ctx->stream));
  return CPGPU_OK;
}

template <typename T>
struct CGVecs {
  T *x, *b, *r, *pa, *pb, *q, *partA, *partB;
};

// Launch of the direction kernel of iteration `it`: variant 0 is the LDS
// tile kernel, variants 1..4 the column march with H = 4, 8, 16, 32.
// Variants 5 and 6 are the aligned march with H = 8, 16, variants 7..10 the
// streaming aligned march over chunks of 32 and 64 rows with the XCD-aware
// unit order, 64 and 128 rows in plain order; variants 11..14 the block-wide
// streaming march over chunks of 16, 32 and 64 rows in XCD order and 32
// rows in plain order, 15 and 16 over 32 rows (XCD, plain order) with
// non-temporal stores there and in the update kernel (16-byte vectors: the
// streaming march of variant 9 when ny is not a multiple of 16 / s); 5..16
// are paired with the vectorised update kernel when ny is even.
constexpr int kDirVariants = 17;
constexpr int kRollRows = 4;  // rows per group of the streaming march

template <typename T, bool PC>
int dir_blocks(int variant, int nx, int ny, const Tiling& tl) {
  if (variant == 0) return tl.G;
  static const int Hs[kDirVariants] = {0, 4, 8, 16, 32, 8, 16, 32, 64, 64, 128,
                                       16, 32, 64, 32, 32, 32};
  constexpr int V = 16 / (int)sizeof(T);
  if (variant >= 11 && ny % V != 0) variant = 9;
  int nunits;
  if (variant >= 11)
    nunits = make_bmarch<T>(nx, ny, Hs[variant], variant != 14 && variant != 16).items;
  else if (variant >= 7)
    nunits = make_roll(nx, ny, Hs[variant]).nunits;
  else
    nunits = (variant >= 5 ? make_amarch(nx, ny, Hs[variant]) : make_march(nx, ny, Hs[variant])).nunits;
  return nunits < kMaxBlocks ? nunits : kMaxBlocks;
}

template <typename T, bool PC>
void launch_dir(int variant, hipStream_t st, const Sten<T>& s, const T* r,
                const T* pold, T* pnew, T* q, KState* kst, const T* partB,
                T* partA, double* hist, int hist_cap, const Tiling& tl,
                int Gprev, int G, const double* partR64 = nullptr,
                int Gref = 0, EvPair ev = EvPair()) {
  const dim3 block(TJ, BY);
  if (variant == 0) {
    CPGPU_LAUNCH(ev, (k_cg_dir<T, PC>), dim3(G), block, 0, st, s, r, pold,
                 pnew, q, kst, partB, partA, hist, hist_cap, tl, Gprev,
                 partR64, Gref);
    return;
  }
#define CPGPU_DIR_MARCH(HV)                                                   \
  CPGPU_LAUNCH(ev, (k_cg_dir_march<T, PC, HV>), dim3(G), block, 0, st, s, r,  \
               pold, pnew, q, kst, partB, partA, hist, hist_cap,              \
               make_march(s.nx, s.ny, HV), Gprev, partR64, Gref)
#define CPGPU_DIR_AMARCH(HV)                                                  \
  CPGPU_LAUNCH(ev, (k_cg_dir_amarch<T, PC, HV>), dim3(G), block, 0, st, s, r, \
               pold, pnew, q, kst, partB, partA, hist, hist_cap,              \
               make_amarch(s.nx, s.ny, HV), Gprev, partR64, Gref)
#define CPGPU_DIR_ROLL(HC, X)                                                       \
  CPGPU_LAUNCH(ev, (k_cg_dir_roll<T, PC, kRollRows>), dim3(G), block, 0, st, s, r, \
               pold, pnew, q, kst, partB, partA, hist, hist_cap,                    \
               make_roll(s.nx, s.ny, HC, X), Gprev, partR64, Gref)
#define CPGPU_DIR_BMARCH(HC, X, NT)                                                       \
  CPGPU_LAUNCH(ev, (k_cg_dir_bmarch<T, PC, kRollRows, NT>), dim3(G), block, 0, st, s, r, \
               pold, pnew, q, kst, partB, partA, hist, hist_cap,                      \
               make_bmarch<T>(s.nx, s.ny, HC, X), Gprev, partR64, Gref)
  if (variant >= 11) {
    // 16-byte vectors need ny % V == 0 and aligned vectors; otherwise the
    // streaming march of variant 9 (any block count G works for it)
    constexpr int V = 16 / (int)sizeof(T);
    const uintptr_t al = (uintptr_t)r | (uintptr_t)pold | (uintptr_t)pnew | (uintptr_t)q;
    if (s.ny % V != 0 || (al & 15)) variant = 9;
  }
  switch (variant) {
    case 11: CPGPU_DIR_BMARCH(16, 1, false); return;
    case 12: CPGPU_DIR_BMARCH(32, 1, false); return;
    case 13: CPGPU_DIR_BMARCH(64, 1, false); return;
    case 14: CPGPU_DIR_BMARCH(32, 0, false); return;
    case 15: CPGPU_DIR_BMARCH(32, 1, true); return;
    case 16: CPGPU_DIR_BMARCH(32, 0, true); return;
    default: break;
  }
#undef CPGPU_DIR_BMARCH
  switch (variant) {
    case 1: CPGPU_DIR_MARCH(4); break;
    case 2: CPGPU_DIR_MARCH(8); break;
    case 3: CPGPU_DIR_MARCH(16); break;
    case 4: CPGPU_DIR_MARCH(32); break;
    case 5: CPGPU_DIR_AMARCH(8); break;
    case 6: CPGPU_DIR_AMARCH(16); break;
    case 7: CPGPU_DIR_ROLL(32, 1); break;
    case 8: CPGPU_DIR_ROLL(64, 1); break;
    case 9: CPGPU_DIR_ROLL(64, 0); break;
    default: CPGPU_DIR_ROLL(128, 0); break;
  }
#undef CPGPU_DIR_MARCH
#undef CPGPU_DIR_AMARCH
#undef CPGPU_DIR_ROLL
}

// The update kernel of a variant: the vectorised one for 5..10 when ny is
// even, else k_cg_upd; both with tl.G blocks (so partial counts match).
template <typename T, bool PC>
void launch_upd(int variant, hipStream_t st, const Sten<T>& s, T* x, const T* p, T* r,
                const T* q, KState* kst, const T* partA, T* partB, const Tiling& tl,
                int norm_ref, int Gprev, EvPair ev = EvPair()) {
  if (variant >= 5 && s.ny % 2 == 0) {
    const int ntj = (s.ny + 127) / 128;
    const int ntiles = ntj * ((s.nx + TI - 1) / TI);
    if (variant >= 15)
      CPGPU_LAUNCH(ev, (k_cg_upd2<T, PC, true>), dim3(tl.G), dim3(TJ, BY), 0, st, s, x, p, r,
                   q, kst, partA, partB, ntiles, ntj, norm_ref, Gprev);
    else
      CPGPU_LAUNCH(ev, (k_cg_upd2<T, PC>), dim3(tl.G), dim3(TJ, BY), 0, st, s, x, p, r, q,
                   kst, partA, partB, ntiles, ntj, norm_ref, Gprev);
  } else {
    CPGPU_LAUNCH(ev, (k_cg_upd<T, PC>), dim3(tl.G), dim3(TJ, BY), 0, st, s, x, p, r, q,
                 kst, partA, partB, tl, norm_ref, Gprev);
  }
}

// One batch of `iters` CG iterations (iters even: the p buffers alternate
// and every batch starts at an odd iteration).  reliable: FP32 iterations
// closed by the FP64 residual replacement (x32 folded into P->x).
template <typename T, bool PC>
void launch_batch(cpgpu_poisson* P, hipStream_t st, const Sten<T>& s,
                  const CGVecs<T>& v, int variant, int Gd, int iters,
                  int norm_ref, double* dhist, int hist_cap, bool timed,
                  bool reliable = false) {
  const Tiling& tl = P->tl;
  const double* pr = reliable ? P->partR64.p : nullptr;
  for (int k = 0; k < iters; ++k) {
    T* pold = (k & 1) ? v.pb : v.pa;
    T* pnew = (k & 1) ? v.pa : v.pb;
    launch_dir<T, PC>(variant, st, s, v.r, pold, pnew, v.q, P->kst.p,
                      v.partB, v.partA, dhist, hist_cap, tl, tl.G, Gd, pr,
                      tl.G, timer_events(timed ? P->ctx : nullptr, 0));
    const EvPair eu = timer_events(timed ? P->ctx : nullptr, 1);
    launch_upd<T, PC>(variant, st, s, v.x, pnew, v.r, v.q, P->kst.p, v.partA, v.partB, tl,
                      norm_ref, Gd, eu);
  }
  if constexpr (sizeof(T) == 4) {
    if (reliable) {
      hipLaunchKernelGGL(k_merge_x, dim3(1024), dim3(256), 0, st, P->x.p, P->xlo.p, v.x,
                         P->N);
      hipLaunchKernelGGL((k_refresh<PC>), dim3(tl.G), dim3(TJ, BY), 0, st,
                         P->sten64(), P->sten32(), P->x.p, P->xlo.p, P->b.p, v.r,
                         v.partB, P->partR64.p, tl);
    }
  }
}

// The instantiated graph of one batch, cached per solve configuration.
template <typename T, bool PC>
hipGraphExec_t batch_graph(cpgpu_poisson* P, const Sten<T>& s,
                           const CGVecs<T>& v, int variant, int Gd, int iters,
                           int norm_ref, double* dhist, int hist_cap,
                           bool reliable) {
  GraphKey key{(const void*)v.x, (const void*)dhist, sizeof(T), PC ? 1 : 0,
               variant, iters, norm_ref, hist_cap, reliable ? 1 : 0};
  for (auto& g : P->graphs)
    if (g.first == key) return g.second;
  hipStream_t cs;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
    return nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) ==
      hipSuccess) {
    launch_batch<T, PC>(P, cs, s, v, variant, Gd, iters, norm_ref, dhist,
                        hist_cap, false, reliable);
    if (hipStreamEndCapture(cs, &graph) == hipSuccess && graph) {
      if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
        exec = nullptr;
      (void)hipGraphDestroy(graph);
    }
  }
  (void)hipStreamDestroy(cs);
  if (exec) P->graphs.push_back({key, exec});
  return exec;
}

// Jacobi-PCG (PC) or CG on device vectors.  Iterations run in batches (a
// replayed hipGraph, or direct launches while kernel timers are on); the
// device state stops them exactly where the reference stops and launches
// after the stop return at once.  The host polls the state of the previous
// batch while the next one runs.
template <typename T, bool PC>
int run_cg(cpgpu_poisson* P, const Sten<T>& s, CGVecs<T> v, int max_iter,
           double tol, int norm_ref, double* dhist, int hist_cap,
           KState* out, bool reliable = false, double normb64 = 1.0) {
  cpgpu_ctx* ctx = P->ctx;
  hipStream_t st = ctx->stream;
  KState init{};
  init.tol = tol;
  init.max_iter = max_iter;
  init.it_dir = 1;
  init.mode = reliable ? 1 : 0;
  init.refresh_every = P->batch;
  init.normb = normb64;
  std::memcpy(&P->h_kst[0], &init, sizeof(KState));
  CPGPU_HIP(ctx, hipMemcpyAsync(P->kst.p, &P->h_kst[0], sizeof(KState),
                                hipMemcpyHostToDevice, st));
  const Tiling tl = P->tl;
  const int variant = sizeof(T) == 8 ? P->dir_variant64 : P->dir_variant32;
  const int Gd = dir_blocks<T, PC>(variant, P->nx, P->ny, tl);
  hipLaunchKernelGGL((k_cg_init<T, PC>), dim3(tl.G), dim3(TJ, BY), 0, st, s,
                     v.x, v.b, v.r, v.partB, tl, norm_ref);
  CPGPU_HIP(ctx, hipGetLastError());
  const bool timed = ctx->timer_on[0] || ctx->timer_on[1];
  // batch size: a handful of iterations for small iteration budgets, then
  // P->batch (even) iterations per replay
  const int B = P->batch;
  hipGraphExec_t ge = nullptr;
  if (P->use_graphs && !timed)
    ge = batch_graph<T, PC>(P, s, v, variant, Gd, B, norm_ref, dhist, hist_cap,
                            reliable);
  // the kernels stop by themselves at max_iter + 1; one batch past it
  // guarantees the final test ran
  const long long total = (long long)max_iter + 2;
  long long launched = 0;
  int nb = 0;
  for (;;) {
    if (ge) {
      CPGPU_HIP(ctx, hipGraphLaunch(ge, st));
    } else {
      launch_batch<T, PC>(P, st, s, v, variant, Gd, B, norm_ref, dhist,
                          hist_cap, timed, reliable);
      CPGPU_HIP(ctx, hipGetLastError());
    }
    launched += B;
    KState* slot = &P->h_kst[1 + (nb & 1)];
    CPGPU_HIP(ctx, hipMemcpyAsync(slot, P->kst.p, sizeof(KState),
                                  hipMemcpyDeviceToHost, st));
    CPGPU_HIP(ctx, hipEventRecord(P->ev_batch[nb & 1], st));
    // check the batch before this one (its copy completed long ago)
    if (nb >= 1) {
      CPGPU_HIP(ctx, hipEventSynchronize(P->ev_batch[(nb - 1) & 1]));
      if (P->h_kst[1 + ((nb - 1) & 1)].done) break;
    }
    ++nb;
    if (launched >= total + B) break;
  }
  CPGPU_HIP(ctx, hipStreamSynchronize(st));
  CPGPU_HIP(ctx, hipMemcpy(&P->h_kst[0], P->kst.p, sizeof(KState),
                           hipMemcpyDeviceToHost));
  *out = P->h_kst[0];
  return CPGPU_OK;
}

int alloc_f32(cpgpu_poisson* P) {
  const size_t N = P->N;
  if (P->b32.alloc(N) || P->x32.alloc(N) || P->r32.alloc(N) ||
      P->pa32.alloc(N) || P->pb32.alloc(N) || P->q32.alloc(N) ||
      P->partA32.alloc(kMaxBlocks) || P->partB32.alloc(3 * kMaxBlocks) ||
      P->partR64.alloc(kMaxBlocks) || P->xlo.alloc(N) ||
      P->fKX.alloc(P->nx + 1) || P->fKY.alloc(P->ny + 1))
    return set_error(P->ctx, CPGPU_ENOMEM, "device allocation failed");
  std::vector<float> kx(P->KX.begin(), P->KX.end()), ky(P->KY.begin(), P->KY.end());
  int rc = upload(P->ctx, P->fKX.p, kx.data(), sizeof(float) * kx.size());
  if (rc) return rc;
  rc = upload(P->ctx, P->fKY.p, ky.data(), sizeof(float) * ky.size());
  if (rc) return rc;
  CPGPU_HIP(P->ctx, hipStreamSynchronize(P->ctx->stream));
  return CPGPU_OK;
}

// r = b - A x [- A xlo] in FP64 with b.b, r.r, sum(r) into P->rst (device)
// and its pinned mirror.
int residual64(cpgpu_poisson* P, bool to_host, bool with_lo = false) {
  cpgpu_ctx* ctx = P->ctx;
  hipStream_t st = ctx->stream;
  hipLaunchKernelGGL(k_residual64, dim3(P->tl.G), dim3(TJ, BY), 0, st,
                     P->sten64(), P->x.p, with_lo ? P->xlo.p : nullptr, P->b.p, P->r.p,
                     P->part64.p, P->tl);
  hipLaunchKernelGGL(k_finalize<3>, dim3(1), dim3(kThreads), 0, st,
                     P->part64.p, (int)P->tl.G, (double*)P->rst.p);
  CPGPU_HIP(ctx, hipGetLastError());
  if (to_host) {
    CPGPU_HIP(ctx, hipMemcpyAsync(P->h_rst, P->rst.p, sizeof(RState),
                                  hipMemcpyDeviceToHost, st));
    CPGPU_HIP(ctx, hipStreamSynchronize(st));
  }
  return CPGPU_OK;
}

// FP32 J-PCG inside FP64 iterative refinement (no reference counterpart).
int run_mixed(cpgpu_poisson* P, int max_inner_total, double tol, int project,
              double* hist, int hist_cap, int* hist_len, int* iters,
              double* resid_out) {
  cpgpu_ctx* ctx = P->ctx;
  hipStream_t st = ctx->stream;
  int rc = alloc_f32(P);
  if (rc) return rc;
  const int nb1 = 1024;
  int total_inner = 0, code = 1, hl = 0;
  double resid = 0;
  hipLaunchKernelGGL(k_fill<double>, dim3(nb1), dim3(256), 0, st, P->xlo.p, 0.0, P->N);
  for (int o = 0; o <= P->max_outer; ++o) {
    if ((rc = residual64(P, true, true))) return rc;
    double normb = std::sqrt(P->h_rst->bb);
    if (normb == 0.0) normb = 1.0;
    double rr = P->h_rst->rr;
    if (project) {  // norm of the projected residual
      const double m = P->h_rst->rsum / (double)P->N;
      rr = std::max(0.0, rr - m * m * (double)P->N);
    }
    resid = std::sqrt(rr) / normb;
    if (hist && hl < hist_cap) hist[hl] = resid;
    hl++;
    if (resid <= tol) {
      code = 0;
      break;
    }
    if (o == P->max_outer || total_inner >= max_inner_total) break;
    hipLaunchKernelGGL(k_to_f32, dim3(nb1), dim3(256), 0, st, P->r.p, P->b32.p,
                       P->x32.p, P->rst.p, P->N, project);
    CGVecs<float> v{P->x32.p, P->b32.p, P->r32.p, P->pa32.p, P->pb32.p,
                    P->q32.p, P->partA32.p, P->partB32.p};
    KState ks;
    rc = run_cg<float, true>(P, P->sten32(), v, max_inner_total - total_inner,
                             P->inner_tol, 0, nullptr, 0, &ks);
    if (rc) return rc;
    total_inner += ks.iter;
    hipLaunchKernelGGL(k_add_f32, dim3(nb1), dim3(256), 0, st, P->x.p, P->xlo.p,
                       P->x32.p, P->part64.p, P->N);
    if (project) {
      hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(kThreads), 0, st,
                         P->part64.p, nb1, &P->rst.p->xsum);
      hipLaunchKernelGGL(k_sub_mean2, dim3(nb1), dim3(256), 0, st, P->x.p, P->xlo.p,
                         &P->rst.p->xsum, P->N);
    }
    CPGPU_HIP(ctx, hipGetLastError());
  }
  hipLaunchKernelGGL(k_round_x, dim3(nb1), dim3(256), 0, st, P->x.p, P->xlo.p, P->N);
  CPGPU_HIP(ctx, hipGetLastError());
  *hist_len = hl;
  *iters = total_inner;
  *resid_out = resid;
  return code;
}

// FP32 Jacobi-PCG with FP64 residual replacement (no reference
// counterpart): one uninterrupted FP32 Krylov iteration whose solution
// increments are folded into the FP64 solution and whose residual is
// replaced by the true FP64 residual every P->batch iterations; the stopping
// test is on the true FP64 relative residual.  hist receives the FP64
// residual at every replacement.
int run_reliable(cpgpu_poisson* P, int max_iter, double tol, int project,
                 double* hist, int hist_cap, int* hist_len, int* iters,
                 double* resid_out) {
  cpgpu_ctx* ctx = P->ctx;
  hipStream_t st = ctx->stream;
  int rc = alloc_f32(P);
  if (rc) return rc;
  if ((rc = residual64(P, true))) return rc;
  double normb = std::sqrt(P->h_rst->bb);
  if (normb == 0.0) normb = 1.0;
  const double resid0 = std::sqrt(P->h_rst->rr) / normb;
  *iters = 0;
  *hist_len = 0;
  *resid_out = resid0;
  if (resid0 <= tol) return CPGPU_OK;
  hipLaunchKernelGGL(k_fill<double>, dim3(1024), dim3(256), 0, st, P->xlo.p, 0.0, P->N);
  // r32 = (float)r64 as the right-hand side of a zero-start FP32 solve
  hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, st, P->r.p, P->b32.p,
                     P->x32.p, P->rst.p, P->N, project);
  if (hist && hist_cap > 0 && P->dhist.alloc(hist_cap))
    return set_error(ctx, CPGPU_ENOMEM, "hist alloc");
  double* dh = hist && hist_cap > 0 ? P->dhist.p : nullptr;
  CGVecs<float> v{P->x32.p, P->b32.p, P->r32.p, P->pa32.p, P->pb32.p,
                  P->q32.p, P->partA32.p, P->partB32.p};
  KState ks;
  rc = run_cg<float, true>(P, P->sten32(), v, max_iter, tol, 0, dh, hist_cap,
                           &ks, true, normb);
  if (rc) return rc;
  // increments since the last replacement (none when it converged there)
  hipLaunchKernelGGL(k_merge_x, dim3(1024), dim3(256), 0, st, P->x.p, P->xlo.p,
                     P->x32.p, P->N);
  if (project) {
    hipLaunchKernelGGL(k_add_f32, dim3(1024), dim3(256), 0, st, P->x.p, P->xlo.p,
                       P->x32.p, P->part64.p, P->N);  // x32 is zero: sums x
    hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(kThreads), 0, st,
                       P->part64.p, 1024, &P->rst.p->xsum);
    hipLaunchKernelGGL(k_sub_mean2, dim3(1024), dim3(256), 0, st, P->x.p, P->xlo.p,
                       &P->rst.p->xsum, P->N);
  }
  hipLaunchKernelGGL(k_round_x, dim3(1024), dim3(256), 0, st, P->x.p, P->xlo.p, P->N);
  CPGPU_HIP(ctx, hipGetLastError());
  const int nh = (ks.iter + P->batch - 1) / P->batch;
  if (dh && nh > 0)
    CPGPU_HIP(ctx, hipMemcpy(hist, dh, sizeof(double) * std::min(nh, hist_cap),
                             hipMemcpyDeviceToHost));
  *hist_len = nh;
  *iters = ks.iter;
  *resid_out = ks.resid;
  return ks.done == 1 ? CPGPU_OK : CPGPU_MAXITER;
}

// Average duration (us) of the direction and update kernels of one variant:
// one real iteration sets up a valid state, then `reps` launches of each
// kernel are captured in a graph and replayed between two events, so the
// figure is GPU time per launch without host launch gaps.  The kernels
// overwrite the Krylov vectors (x included); callers save what they keep.
template <typename T, bool PC>
int time_kernels(cpgpu_poisson* P, const Sten<T>& s, const CGVecs<T>& v,
                 int variant, int reps, float* us_dir, float* us_upd) {
  cpgpu_ctx* ctx = P->ctx;
  hipStream_t st = ctx->stream;
  const Tiling tl = P->tl;
  KState init{};
  // a negative tolerance never converges: a zero right-hand side (resid 0)
  // would otherwise stop the state at once and time empty launches
  init.tol = -1.0;
  init.max_iter = 1 << 30;
  init.it_dir = 1;
  std::memcpy(&P->h_kst[0], &init, sizeof(KState));
  CPGPU_HIP(ctx, hipMemcpyAsync(P->kst.p, &P->h_kst[0], sizeof(KState),
                                hipMemcpyHostToDevice, st));
  const int Gd = dir_blocks<T, PC>(variant, P->nx, P->ny, tl);
  hipLaunchKernelGGL((k_cg_init<T, PC>), dim3(tl.G), dim3(TJ, BY), 0, st, s,
                     v.x, v.b, v.r, v.partB, tl, 0);
  launch_dir<T, PC>(variant, st, s, v.r, v.pa, v.pb, v.q, P->kst.p, v.partB,
                    v.partA, nullptr, 0, tl, tl.G, Gd);
  launch_upd<T, PC>(variant, st, s, v.x, v.pb, v.r, v.q, P->kst.p, v.partA, v.partB, tl, 0,
                    Gd);
  CPGPU_HIP(ctx, hipGetLastError());
  CPGPU_HIP(ctx, hipStreamSynchronize(st));
  float out[2] = {0.f, 0.f};
  for (int which = 0; which < 2; ++which) {
    hipStream_t cs;
    CPGPU_HIP(ctx, hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    CPGPU_HIP(ctx, hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < reps; ++k) {
      if (which == 0)
        launch_dir<T, PC>(variant, cs, s, v.r, v.pb, v.pa, v.q, P->kst.p,
                          v.partB, v.partA, nullptr, 0, tl, tl.G, Gd);
      else
        launch_upd<T, PC>(variant, cs, s, v.x, v.pa, v.r, v.q, P->kst.p, v.partA, v.partB,
                          tl, 0, Gd);
    }
    CPGPU_HIP(ctx, hipStreamEndCapture(cs, &graph));
    CPGPU_HIP(ctx, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    CPGPU_HIP(ctx, hipGraphLaunch(exec, st));  // warm
    hipEvent_t e0, e1;
    CPGPU_HIP(ctx, hipEventCreate(&e0));
    CPGPU_HIP(ctx, hipEventCreate(&e1));
    CPGPU_HIP(ctx, hipEventRecord(e0, st));
    CPGPU_HIP(ctx, hipGraphLaunch(exec, st));
    CPGPU_HIP(ctx, hipEventRecord(e1, st));
    CPGPU_HIP(ctx, hipEventSynchronize(e1));
    float ms = 0.f;
    CPGPU_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
    out[which] = ms * 1e3f / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipGraphExecDestroy(exec);
    (void)hipStreamDestroy(cs);
  }
  *us_dir = out[0];
  *us_upd = out[1];
  return CPGPU_OK;
}

// Times every direction-kernel variant and returns the fastest.
template <typename T, bool PC>
int tune_dir(cpgpu_poisson* P, const Sten<T>& s, CGVecs<T> v, int reps,
             int* best, float* best_ms) {
  cpgpu_ctx* ctx = P->ctx;
  hipStream_t st = ctx->stream;
  const Tiling tl = P->tl;
  float bt = 1e30f, up = 0.f;
  int bv = 3;
  // the fastest (direction + update) pair: variants 5 and 6 change both
  for (int variant = 0; variant < kDirVariants; ++variant) {
    float ud = 0.f, uu = 0.f;
    int rc = time_kernels<T, PC>(P, s, v, variant, reps, &ud, &uu);
    if (rc) return rc;
    if (ud + uu < bt) {
      bt = ud + uu;
      bv = variant;
      up = uu;
    }
  }
  (void)up;
  (void)st;
  (void)tl;
  *best = bv;
  *best_ms = bt / 1e3f;
  return CPGPU_OK;
}

// ---------------------------------------------------------------------
// Scratch-free solve (cpgpu_poisson2d_solve_static; no reference
// counterpart).  Every work vector, coefficient array, reduction partial and
// the Krylov state live in one file-scope device array per GPU, `scratch`;
// the caller's device arrays are the only other memory touched.

constexpr size_t kStaticMaxCells = 512 * 512;  // cells of the largest grid
constexpr int kStaticMaxSide = 4096;           // cells per side
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }
// byte offsets: b, r, p0, p1, q | KX, KY | vol | bcv | partA | partB | state
constexpr size_t kOffVec = 0;
constexpr size_t kVecBytes = align_up(kStaticMaxCells * sizeof(double));
constexpr size_t kOffKX = kOffVec + 5 * kVecBytes;
constexpr size_t kOffKY = kOffKX + align_up((kStaticMaxSide + 1) * sizeof(double));
constexpr size_t kOffVol = kOffKY + align_up((kStaticMaxSide + 1) * sizeof(double));
constexpr size_t kOffBcv = kOffVol + align_up(kStaticMaxSide * sizeof(double));
constexpr size_t kOffPartA = kOffBcv + align_up(4 * kStaticMaxSide * sizeof(double));
constexpr size_t kOffPartB = kOffPartA + align_up(kMaxBlocks * sizeof(double));
constexpr size_t kOffState = kOffPartB + align_up(3 * kMaxBlocks * sizeof(double));
constexpr size_t kScratchBytes = kOffState + align_up(sizeof(KState));

// Host staging of the descriptor arrays and the state, guarded by g_stage_mu
// (the copies complete before the lock is released).
struct StaticStage {
  double KX[kStaticMaxSide + 1], KY[kStaticMaxSide + 1], vol[kStaticMaxSide],
      bcv[4 * kStaticMaxSide];
  KState st;
};
StaticStage g_stage;
std::mutex g_stage_mu;

}  // namespace

void drop_graphs(cpgpu_poisson* P) {
  for (auto& g : P->graphs) (void)hipGraphExecDestroy(g.second);
  P->graphs.clear();
  for (auto& g : P->fcg_graphs) (void)hipGraphExecDestroy(g.second);
  P->fcg_graphs.clear();
}

// The file-scope device array (one copy per GPU, held until the process
// ends).  In host code its name is a host-side stand-in: every use goes
// through the per-device address from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;
}  // namespace

int static_scratch_setup(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return CPGPU_EINVAL;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    // the device is current (cpgpu_create set it)
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
      return CPGPU_EHIP;
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return CPGPU_OK;
}

namespace {

int solve_static(cpgpu_ctx* ctx, const cpgpu_grid2d* g, const double* d_src,
                 double src_uniform, double* d_phi, int method, int* max_iter,
                 double* tol, int norm_mode) {
  const int nx = g->nx, ny = g->ny;
  const size_t N = (size_t)nx * ny;
  unsigned char* base = (unsigned char*)ctx->scratch;
  if (!base) return set_error(ctx, CPGPU_EHIP, "static scratch not set up on this device");
  hipStream_t st = ctx->stream;
  double* b = (double*)(base + kOffVec);
  double* r = (double*)(base + kOffVec + kVecBytes);
  double* pa = (double*)(base + kOffVec + 2 * kVecBytes);
  double* pb = (double*)(base + kOffVec + 3 * kVecBytes);
  double* q = (double*)(base + kOffVec + 4 * kVecBytes);
  double* dKX = (double*)(base + kOffKX);
  double* dKY = (double*)(base + kOffKY);
  double* dvol = (double*)(base + kOffVol);
  double* dbcv = (double*)(base + kOffBcv);
  double* partA = (double*)(base + kOffPartA);
  double* partB = (double*)(base + kOffPartB);
  KState* kst = (KState*)(base + kOffState);
  Sten<double> s;
  s.nx = nx;
  s.ny = ny;
  s.KX = dKX;
  s.KY = dKY;
  RhsDesc d;
  d.nx = nx;
  d.ny = ny;
  d.vol = dvol;
  d.bcval = dbcv;
  const Tiling tl = make_tiling(nx, ny);
  const int PCm = method == CPGPU_JPCG;
  const int nref = norm_mode == CPGPU_NORM_REF && N > 1;
  // a new problem's direction kernel: the dot products are partitioned as
  // in cpgpu_poisson2d_solve, so the results are bitwise its own
  const int variant = kDefaultDirVariant;
  const int Gd = PCm ? dir_blocks<double, true>(variant, nx, ny, tl)
                     : dir_blocks<double, false>(variant, nx, ny, tl);
  {
    std::lock_guard<std::mutex> lock(g_stage_mu);
    StaticStage& h = g_stage;
    if (!uniform_cells(g))
      return set_error(ctx, CPGPU_EUNSUPPORTED, "static solve: uniform cell sizes only");
    int rc = describe_grid(ctx, g, h.KX, h.KY, s.adj, d.seg, h.vol, h.bcv);
    if (rc) return rc;
    std::memset(&h.st, 0, sizeof(KState));
    h.st.tol = *tol;
    h.st.max_iter = *max_iter;
    h.st.it_dir = 1;
    h.st.refresh_every = 16;
    h.st.normb = 1.0;
    CPGPU_HIP(ctx, hipMemcpyAsync(dKX, h.KX, sizeof(double) * (nx + 1),
                                  hipMemcpyHostToDevice, st));
    CPGPU_HIP(ctx, hipMemcpyAsync(dKY, h.KY, sizeof(double) * (ny + 1),
                                  hipMemcpyHostToDevice, st));
    CPGPU_HIP(ctx, hipMemcpyAsync(dvol, h.vol, sizeof(double) * nx,
                                  hipMemcpyHostToDevice, st));
    CPGPU_HIP(ctx, hipMemcpyAsync(dbcv, h.bcv, sizeof(double) * (2 * nx + 2 * ny),
                                  hipMemcpyHostToDevice, st));
    CPGPU_HIP(ctx, hipMemcpyAsync(kst, &h.st, sizeof(KState), hipMemcpyHostToDevice, st));
    CPGPU_HIP(ctx, hipStreamSynchronize(st));
  }
  s.shift = 0.0;
  hipLaunchKernelGGL(k_build_rhs, dim3(tl.G), dim3(TJ, BY), 0, st, d, d_src, src_uniform,
                     b, tl);
  if (PCm)
    hipLaunchKernelGGL((k_cg_init<double, true>), dim3(tl.G), dim3(TJ, BY), 0, st, s,
                       d_phi, b, r, partB, tl, nref);
  else
    hipLaunchKernelGGL((k_cg_init<double, false>), dim3(tl.G), dim3(TJ, BY), 0, st, s,
                       d_phi, b, r, partB, tl, nref);
  CPGPU_HIP(ctx, hipGetLastError());
  // batches of B iterations; the kernels stop by themselves at the test,
  // later launches return at entry; the host reads the state after each
  // batch (the direction kernel of iteration max_iter + 1 makes the final
  // test, so the loop ends at most one batch past it)
  const int B = 16;
  const long long total = (long long)*max_iter + 2;
  KState hs{};
  for (long long launched = 0;; launched += B) {
    for (int k = 0; k < B; ++k) {
      double* pold = (k & 1) ? pb : pa;
      double* pnew = (k & 1) ? pa : pb;
      if (PCm) {
        launch_dir<double, true>(variant, st, s, r, pold, pnew, q, kst, partB, partA,
                                 nullptr, 0, tl, tl.G, Gd, nullptr, tl.G);
        launch_upd<double, true>(variant, st, s, d_phi, pnew, r, q, kst, partA, partB, tl,
                                 nref, Gd);
      } else {
        launch_dir<double, false>(variant, st, s, r, pold, pnew, q, kst, partB, partA,
                                  nullptr, 0, tl, tl.G, Gd, nullptr, tl.G);
        launch_upd<double, false>(variant, st, s, d_phi, pnew, r, q, kst, partA, partB,
                                  tl, nref, Gd);
      }
    }
    CPGPU_HIP(ctx, hipGetLastError());
    CPGPU_HIP(ctx, hipMemcpyAsync(&hs, kst, sizeof(KState), hipMemcpyDeviceToHost, st));
    CPGPU_HIP(ctx, hipStreamSynchronize(st));
    if (hs.done || launched + B >= total + B) break;
  }
  *tol = hs.resid;
  if (hs.done == 1) {
    *max_iter = hs.iter;
    return CPGPU_OK;
  }
  return CPGPU_MAXITER;
}

}  // namespace

int poisson_rhs_from_device(cpgpu_poisson* P, const double* dsrc,
                            double src_uniform) {
  if (P->general) return general_rhs(P, dsrc, src_uniform);
  RhsDesc d;
  d.nx = P->nx;
  d.ny = P->ny;
  d.vol = P->dvol.p;
  d.bcval = P->dbcv.p;
  for (int s = 0; s < 4; ++s) d.seg[s] = P->seg[s];
  hipLaunchKernelGGL(k_build_rhs, dim3(P->tl.G), dim3(TJ, BY), 0,
                     P->ctx->stream, d, dsrc, src_uniform, P->b.p, P->tl);
  CPGPU_HIP(P->ctx, hipGetLastError());
  return CPGPU_OK;
}

int poisson_prepare_f32(cpgpu_poisson* P) { return alloc_f32(P); }

int poisson_jpcg64(cpgpu_poisson* P, int* iters, double tol, int norm_ref,
                   double* resid, int* done) {
  CGVecs<double> v{P->x.p, P->b.p, P->r.p, P->p0.p, P->p1.p, P->q.p,
                   P->partA64.p, P->partB64.p};
  KState ks;
  int rc = run_cg<double, true>(P, P->sten64(), v, *iters, tol, norm_ref,
                                nullptr, 0, &ks);
  if (rc) return rc;
  *iters = ks.iter;
  *resid = ks.resid;
  *done = ks.done == 1;
  return CPGPU_OK;
}

int poisson_solve_internal(cpgpu_poisson* P, int method, int* max_iter,
                           double* tol, int norm_mode, int opts, double* hist,
                           int hist_cap, int* hist_len) {
  cpgpu_ctx* ctx = P->ctx;
  int hl = 0;
  if (P->general)
    return general_solve(P, method, max_iter, tol, norm_mode == CPGPU_NORM_REF && P->N > 1,
                         hist, hist_cap, hist_len);
  if (method == CPGPU_BICGSTAB || method == CPGPU_JACOBI) {
    const int nref = norm_mode == CPGPU_NORM_REF && P->N > 1;
    int code = method == CPGPU_BICGSTAB
                   ? run_bicgstab(P, max_iter, tol, nref, hist, hist_cap, &hl)
                   : run_jacobi(P, max_iter, tol, nref, hist, hist_cap, &hl);
    if (code < 0) return code;
    if (hist_len) *hist_len = hl;
    return code;
  }
  if (method == CPGPU_FCG_UNET) {
    const int nref = norm_mode == CPGPU_NORM_REF && P->N > 1;
    int iters = 0;
    double resid = 0;
    int code = run_fcg_unet(P, *max_iter, *tol, nref,
                            opts & CPGPU_OPT_PROJECT_MEAN, hist, hist_cap, &hl,
                            &iters, &resid);
    if (code < 0) return code;
    if (hist_len) *hist_len = hl;
    *max_iter = iters;
    *tol = resid;
    return code;
  }
  if (P->unet && (P->unet_flags & CPGPU_UNET_WARM)) {
    int rc = unet_warm_start(P);
    if (rc) return rc;
  }
  if (method == CPGPU_CG || method == CPGPU_JPCG) {
    if (hist && hist_cap > 0 && P->dhist.alloc(hist_cap))
      return set_error(ctx, CPGPU_ENOMEM, "hist alloc");
    CGVecs<double> v{P->x.p, P->b.p, P->r.p, P->p0.p, P->p1.p, P->q.p,
                     P->partA64.p, P->partB64.p};
    KState ks;
    // carpio::nrm2 drops the last element except for n == 1, where it
    // returns |x0| (BLAS_Level1.hpp:565-567)
    const int nref = norm_mode == CPGPU_NORM_REF && P->N > 1;
    double* dh = hist && hist_cap > 0 ? P->dhist.p : nullptr;
    int rc = method == CPGPU_JPCG
                 ? run_cg<double, true>(P, P->sten64(), v, *max_iter, *tol,
                                        nref, dh, hist_cap, &ks)
                 : run_cg<double, false>(P, P->sten64(), v, *max_iter, *tol,
                                         nref, dh, hist_cap, &ks);
    if (rc) return rc;
    hl = ks.iter;
    if (dh && hl > 0)
      CPGPU_HIP(ctx, hipMemcpy(hist, dh, sizeof(double) * std::min(hl, hist_cap),
                               hipMemcpyDeviceToHost));
    if (hist_len) *hist_len = hl;
    *tol = ks.resid;
    if (ks.done == 1) {
      *max_iter = ks.iter;
      return CPGPU_OK;
    }
    return CPGPU_MAXITER;
  }
  if (method == CPGPU_MIXED || method == CPGPU_MIXED_RESTART) {
    int iters = 0;
    double resid = 0;
    const int proj = opts & CPGPU_OPT_PROJECT_MEAN;
    int code = method == CPGPU_MIXED
                   ? run_reliable(P, *max_iter, *tol, proj, hist, hist_cap,
                                  &hl, &iters, &resid)
                   : run_mixed(P, *max_iter, *tol, proj, hist, hist_cap, &hl,
                               &iters, &resid);
    if (code < 0) return code;
    if (hist_len) *hist_len = hl;
    *max_iter = iters;
    *tol = resid;
    return code;
  }
  return set_error(ctx, CPGPU_EUNSUPPORTED, "method %d not available", method);
}

}  // namespace cpgpu

using namespace cpgpu;

extern "C" {

int cpgpu_grid_uniform_axis(int n, int gl, double vmin, double vmax, double* c,
                            double* s) {
  if (n < 1 || gl < 0 || !c || !s || !(vmax > vmin)) return CPGPU_EINVAL;
  const int ng = n + 2 * gl;
  const double ds = (vmax - vmin) / n;
  for (int k = 0; k < ng; ++k) s[k] = ds;
  double csum = 0.0;
  for (int k = 0; k < ng; ++k) {
    c[k] = csum + 0.5 * s[k];
    csum += s[k];
  }
  double llg = 0.0;
  for (int k = 0; k < gl; ++k) llg += s[k];
  const double sh = vmin - llg;
  for (int k = 0; k < ng; ++k) c[k] += sh;
  return CPGPU_OK;
}

int cpgpu_grid_axis_sizes(int n, int gl, double vmin, const double* sizes, double* c,
                          double* s) {
  if (n < 1 || gl < 0 || !sizes || !c || !s) return CPGPU_EINVAL;
  for (int k = 0; k < n; ++k)
    if (!(sizes[k] > 0.0)) return CPGPU_EINVAL;
  const int ng = n + 2 * gl;
  for (int k = 0; k < ng; ++k) s[k] = k < gl ? sizes[0] : (k - gl >= n ? sizes[n - 1] : sizes[k - gl]);
  double csum = 0.0;
  for (int k = 0; k < ng; ++k) {
    c[k] = csum + 0.5 * s[k];
    csum += s[k];
  }
  double llg = 0.0;
  for (int k = 0; k < gl; ++k) llg += s[k];
  const double sh = vmin - llg;
  for (int k = 0; k < ng; ++k) c[k] += sh;
  return CPGPU_OK;
}

int cpgpu_poisson_create(cpgpu_ctx* ctx, const cpgpu_grid2d* g,
                         cpgpu_poisson** out) {
  if (!ctx || !g || !out || g->nx < 1 || g->ny < 1 || g->gl < 1 || !g->cx ||
      !g->sx || !g->cy || !g->sy)
    return set_error(ctx, CPGPU_EINVAL, "invalid poisson descriptor");
  *out = nullptr;
  CPGPU_HIP(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<cpgpu_poisson> P(new cpgpu_poisson());
  P->ctx = ctx;
  P->nx = g->nx;
  P->ny = g->ny;
  P->gl = g->gl;
  P->N = (size_t)g->nx * g->ny;
  P->beta = g->beta;
  P->tl = make_tiling(g->nx, g->ny);
  int rc = build_descriptor(P.get(), g);
  if (rc) return rc;
  const size_t N = P->N;
  if (P->x.alloc(N) || P->b.alloc(N) || P->r.alloc(N) || P->p0.alloc(N) ||
      P->p1.alloc(N) || P->q.alloc(N) || P->partA64.alloc(kMaxBlocks) ||
      P->partB64.alloc(3 * kMaxBlocks) || P->part64.alloc(3 * kMaxBlocks) ||
      P->kst.alloc(1) || P->rst.alloc(1) || P->dKX.alloc(P->nx + 1) ||
      P->dKY.alloc(P->ny + 1) || P->dvol.alloc(P->nx) ||
      P->dbcv.alloc(P->bcv.size()))
    return set_error(ctx, CPGPU_ENOMEM, "device allocation failed");
  if ((rc = upload(ctx, P->dKX.p, P->KX.data(), sizeof(double) * P->KX.size())) ||
      (rc = upload(ctx, P->dKY.p, P->KY.data(), sizeof(double) * P->KY.size())) ||
      (rc = upload(ctx, P->dvol.p, P->vol.data(), sizeof(double) * P->vol.size())) ||
      (rc = upload(ctx, P->dbcv.p, P->bcv.data(), sizeof(double) * P->bcv.size())))
    return rc;
  hipLaunchKernelGGL(k_fill<double>, dim3(1024), dim3(256), 0, ctx->stream,
                     P->x.p, 0.0, N);
  hipLaunchKernelGGL(k_fill<double>, dim3(1024), dim3(256), 0, ctx->stream,
                     P->b.p, 0.0, N);
  CPGPU_HIP(ctx, hipGetLastError());
  CPGPU_HIP(ctx, hipHostMalloc((void**)&P->h_kst, 3 * sizeof(KState), 0));
  for (auto& e : P->ev_batch)
    CPGPU_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  CPGPU_HIP(ctx, hipHostMalloc((void**)&P->h_rst, sizeof(RState), 0));
  CPGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = P.release();
  return CPGPU_OK;
}

int cpgpu_poisson_destroy(cpgpu_poisson* P) {
  if (!P) return CPGPU_EINVAL;
  (void)hipStreamSynchronize(P->ctx->stream);
  delete P;
  return CPGPU_OK;
}

int cpgpu_poisson_set_source(cpgpu_poisson* P, const double* src_field,
                             double src_uniform) {
  if (!P) return CPGPU_EINVAL;
  if (src_field) {
    if (P->src.alloc(P->N)) return set_error(P->ctx, CPGPU_ENOMEM, "src");
    int rc = field_to_dev(P, src_field, P->src.p);
    if (rc) return rc;
    P->src_kind = 1;
    return poisson_rhs_from_device(P, P->src.p, 0.0);
  }
  P->src_kind = 2;
  P->src_uni = src_uniform;
  return poisson_rhs_from_device(P, nullptr, src_uniform);
}

int cpgpu_poisson_set_source_device(cpgpu_poisson* P, const double* d_src) {
  if (!P || !d_src) return CPGPU_EINVAL;
  P->src_kind = 3;
  return poisson_rhs_from_device(P, d_src, 0.0);
}

int cpgpu_poisson_fill_phi(cpgpu_poisson* P, double value) {
  if (!P) return CPGPU_EINVAL;
  hipLaunchKernelGGL(k_fill<double>, dim3(1024), dim3(256), 0, P->ctx->stream,
                     P->x.p, value, P->N);
  CPGPU_HIP(P->ctx, hipGetLastError());
  return CPGPU_OK;
}

int cpgpu_blob_params(unsigned long long seed, int nblobs, double wmin,
                      double wmax, double* out) {
  if (nblobs < 1 || !out || !(wmax >= wmin) || !(wmin > 0.0))
    return CPGPU_EINVAL;
  blob_params(seed, nblobs, wmin, wmax, out);
  return CPGPU_OK;
}

int cpgpu_poisson_set_source_blobs(cpgpu_poisson* P, unsigned long long seed,
                                   int nblobs, double wmin, double wmax) {
  if (!P || nblobs < 1 || nblobs > 4096 || !(wmax >= wmin) || !(wmin > 0.0))
    return CPGPU_EINVAL;
  cpgpu_ctx* ctx = P->ctx;
  std::vector<double> prm(4 * (size_t)nblobs);
  blob_params(seed, nblobs, wmin, wmax, prm.data());
  DevBuf<double> dprm, dcx, dcy;
  if (P->src.alloc(P->N) || dprm.alloc(prm.size()) || dcx.alloc(P->nx) ||
      dcy.alloc(P->ny))
    return set_error(ctx, CPGPU_ENOMEM, "blob buffers");
  int rc;
  if ((rc = upload(ctx, dprm.p, prm.data(), sizeof(double) * prm.size())) ||
      (rc = upload(ctx, dcx.p, P->cxi.data(), sizeof(double) * P->nx)) ||
      (rc = upload(ctx, dcy.p, P->cyi.data(), sizeof(double) * P->ny)))
    return rc;
  hipLaunchKernelGGL(k_blobs, dim3(P->tl.G), dim3(TJ, BY), 0, ctx->stream,
                     P->nx, P->ny, dcx.p, dcy.p, dprm.p, nblobs, P->src.p,
                     P->part64.p, P->tl);
  hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(kThreads), 0, ctx->stream,
                     P->part64.p, (int)P->tl.G, &P->rst.p->rsum);
  hipLaunchKernelGGL(k_sub_mean, dim3(1024), dim3(256), 0, ctx->stream,
                     P->src.p, &P->rst.p->rsum, P->N);
  CPGPU_HIP(ctx, hipGetLastError());
  P->src_kind = 1;
  rc = poisson_rhs_from_device(P, P->src.p, 0.0);
  if (rc) return rc;
  // the parameter buffers must outlive the kernels
  CPGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CPGPU_OK;
}

int cpgpu_poisson_set_phi(cpgpu_poisson* P, const double* phi_field) {
  if (!P || !phi_field) return CPGPU_EINVAL;
  return field_to_dev(P, phi_field, P->x.p);
}

int cpgpu_poisson_set_alpha(cpgpu_poisson* P, double alpha) {
  if (!P) return CPGPU_EINVAL;
  // b of a caller's device source cannot be rebuilt with the new weight
  if (P->src_kind == 3)
    return set_error(P->ctx, CPGPU_EINVAL,
                     "alpha after a device source: set alpha first, then the source");
  // uniform cells only: one shift for every row
  for (double v : P->sxi)
    if (v != P->sxi[0]) return set_error(P->ctx, CPGPU_EUNSUPPORTED, "alpha: non-uniform x sizes");
  for (double v : P->syi)
    if (v != P->syi[0]) return set_error(P->ctx, CPGPU_EUNSUPPORTED, "alpha: non-uniform y sizes");
  // v = product of the sizes over d (StencilHelmholtz_::coe_row,
  // s_poisson.hpp:248-258): the shift v*alpha and the source weight v
  double v = 1.0;
  v *= P->sxi[0];
  v *= P->syi[0];
  for (int i = 0; i < P->nx; ++i) P->vol[i] = v;
  P->shift = v * alpha;
  CPGPU_HIP(P->ctx, hipSetDevice(P->ctx->device));
  CPGPU_HIP(P->ctx, hipStreamSynchronize(P->ctx->stream));
  // the captured iteration graphs hold the old diagonal in their kernel
  // arguments: drop them (the next solve captures new ones)
  drop_graphs(P);
  int rc = upload(P->ctx, P->dvol.p, P->vol.data(), sizeof(double) * P->vol.size());
  if (rc) return rc;
  // b carries the source weight v: rebuild it from the kept source
  if (P->src_kind == 1) rc = poisson_rhs_from_device(P, P->src.p, 0.0);
  else if (P->src_kind == 2) rc = poisson_rhs_from_device(P, nullptr, P->src_uni);
  if (rc) return rc;
  CPGPU_HIP(P->ctx, hipStreamSynchronize(P->ctx->stream));
  return CPGPU_OK;
}

int cpgpu_poisson_set_bc_values(cpgpu_poisson* P, const double* bc_val) {
  if (!P || !bc_val) return CPGPU_EINVAL;
  const int nx = P->nx, ny = P->ny, gl = P->gl;
  const int NX = nx + 2 * gl, NY = ny + 2 * gl;
  for (int j = 0; j < ny; ++j) {
    P->bcv[j] = bc_val[j + gl];
    P->bcv[ny + j] = bc_val[NY + j + gl];
  }
  for (int i = 0; i < nx; ++i) {
    P->bcv[2 * ny + i] = bc_val[2 * NY + i + gl];
    P->bcv[2 * ny + nx + i] = bc_val[2 * NY + NX + i + gl];
  }
  CPGPU_HIP(P->ctx, hipSetDevice(P->ctx->device));
  int rc = upload(P->ctx, P->dbcv.p, P->bcv.data(), sizeof(double) * P->bcv.size());
  if (rc) return rc;
  CPGPU_HIP(P->ctx, hipStreamSynchronize(P->ctx->stream));
  return CPGPU_OK;
}

int cpgpu_poisson_get_phi(cpgpu_poisson* P, double* phi_field) {
  if (!P || !phi_field) return CPGPU_EINVAL;
  return dev_to_field(P, P->x.p, phi_field);
}

double* cpgpu_poisson_device_vec(cpgpu_poisson* P, int which) {
  if (!P) return nullptr;
  return which == 0 ? P->x.p : which == 1 ? P->b.p : which == 2 ? P->r.p
                                                                 : nullptr;
}

int cpgpu_poisson_get_vec(cpgpu_poisson* P, int which, double* field) {
  double* d = cpgpu_poisson_device_vec(P, which);
  if (!d || !field) return CPGPU_EINVAL;
  return dev_to_field(P, d, field);
}

int cpgpu_poisson_set_mixed(cpgpu_poisson* P, double inner_tol, int max_outer) {
  if (!P || !(inner_tol > 0.0) || inner_tol >= 1.0 || max_outer < 1)
    return CPGPU_EINVAL;
  P->inner_tol = inner_tol;
  P->max_outer = max_outer;
  return CPGPU_OK;
}

int cpgpu_poisson_set_variant(cpgpu_poisson* P, int v64, int v32) {
  if (!P || v64 < 0 || v64 >= kDirVariants || v32 < 0 || v32 >= kDirVariants)
    return CPGPU_EINVAL;
  P->dir_variant64 = v64;
  P->dir_variant32 = v32;
  return CPGPU_OK;
}

int cpgpu_poisson_autotune(cpgpu_poisson* P, int reps, int* v64, int* v32,
                           float* us64, float* us32) {
  if (!P || reps < 1) return CPGPU_EINVAL;
  cpgpu_ctx* ctx = P->ctx;
  if (P->general) {  // one kernel family: nothing to choose
    if (v64) *v64 = -1;
    if (v32) *v32 = -1;
    if (us64) *us64 = 0.f;
    if (us32) *us32 = 0.f;
    return CPGPU_OK;
  }
  CPGPU_HIP(ctx, hipSetDevice(ctx->device));
  // keep phi: the tuning runs Krylov kernels on the problem's vectors
  DevBuf<double> save;
  if (save.alloc(P->N)) return set_error(ctx, CPGPU_ENOMEM, "tune buffer");
  CPGPU_HIP(ctx, hipMemcpyAsync(save.p, P->x.p, sizeof(double) * P->N,
                                hipMemcpyDeviceToDevice, ctx->stream));
  int b64 = 3, b32 = 3;
  float t64 = 0.f, t32 = 0.f;
  CGVecs<double> v{P->x.p, P->b.p, P->r.p, P->p0.p, P->p1.p, P->q.p,
                   P->partA64.p, P->partB64.p};
  int rc = tune_dir<double, true>(P, P->sten64(), v, reps, &b64, &t64);
  if (rc) return rc;
  if ((rc = alloc_f32(P))) return rc;
  hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, ctx->stream, P->b.p,
                     P->b32.p, P->x32.p, P->rst.p, P->N, 0);
  CGVecs<float> w{P->x32.p, P->b32.p, P->r32.p, P->pa32.p, P->pb32.p,
                  P->q32.p, P->partA32.p, P->partB32.p};
  rc = tune_dir<float, true>(P, P->sten32(), w, reps, &b32, &t32);
  if (rc) return rc;
  CPGPU_HIP(ctx, hipMemcpyAsync(P->x.p, save.p, sizeof(double) * P->N,
                                hipMemcpyDeviceToDevice, ctx->stream));
  CPGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  P->dir_variant64 = b64;
  P->dir_variant32 = b32;
  if (v64) *v64 = b64;
  if (v32) *v32 = b32;
  if (us64) *us64 = t64 * 1e3f;
  if (us32) *us32 = t32 * 1e3f;
  return CPGPU_OK;
}

int cpgpu_poisson_kernel_times(cpgpu_poisson* P, int precision, int reps,
                               double* us_dir, double* us_upd) {
  if (!P || reps < 1 || !us_dir || !us_upd ||
      (precision != 32 && precision != 64))
    return CPGPU_EINVAL;
  cpgpu_ctx* ctx = P->ctx;
  if (P->general)
    return set_error(ctx, CPGPU_EUNSUPPORTED, "kernel times: uniform cell sizes only");
  CPGPU_HIP(ctx, hipSetDevice(ctx->device));
  DevBuf<double> save;
  if (save.alloc(P->N)) return set_error(ctx, CPGPU_ENOMEM, "timing buffer");
  CPGPU_HIP(ctx, hipMemcpyAsync(save.p, P->x.p, sizeof(double) * P->N,
                                hipMemcpyDeviceToDevice, ctx->stream));
  float d = 0.f, u = 0.f;
  int rc;
  if (precision == 64) {
    CGVecs<double> v{P->x.p, P->b.p, P->r.p, P->p0.p, P->p1.p, P->q.p,
                     P->partA64.p, P->partB64.p};
    rc = time_kernels<double, true>(P, P->sten64(), v, P->dir_variant64, reps,
                                    &d, &u);
  } else {
    if ((rc = alloc_f32(P))) return rc;
    hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, ctx->stream, P->b.p,
                       P->b32.p, P->x32.p, P->rst.p, P->N, 0);
    CGVecs<float> w{P->x32.p, P->b32.p, P->r32.p, P->pa32.p, P->pb32.p,
                    P->q32.p, P->partA32.p, P->partB32.p};
    rc = time_kernels<float, true>(P, P->sten32(), w, P->dir_variant32, reps,
                                   &d, &u);
  }
  if (rc) return rc;
  CPGPU_HIP(ctx, hipMemcpyAsync(P->x.p, save.p, sizeof(double) * P->N,
                                hipMemcpyDeviceToDevice, ctx->stream));
  CPGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *us_dir = d;
  *us_upd = u;
  return CPGPU_OK;
}

int cpgpu_poisson_set_batch(cpgpu_poisson* P, int iters, int use_graphs) {
  if (!P || iters < 2 || (iters & 1)) return CPGPU_EINVAL;
  P->batch = iters;
  if (use_graphs < 0 || use_graphs > 2) return CPGPU_EINVAL;
  P->use_graphs = use_graphs != 0;
  P->fcg_graphs_on = use_graphs == 2;
  return CPGPU_OK;
}

int cpgpu_poisson_set_fcg(cpgpu_poisson* P, int pre, int post, double omega) {
  if (!P || pre < 0 || post < 0 || pre > 64 || post > 64 || !(omega > 0.0) ||
      !(omega < 2.0))
    return CPGPU_EINVAL;
  P->pc_pre = pre;
  P->pc_post = post;
  P->pc_omega = omega;
  return CPGPU_OK;
}

int cpgpu_poisson_set_unet(cpgpu_poisson* P, cpgpu_unet* U, int flags) {
  if (!P || (flags & ~(CPGPU_UNET_WARM | CPGPU_UNET_PRECOND))) return CPGPU_EINVAL;
  if (U && unet_context(U) != P->ctx)
    return set_error(P->ctx, CPGPU_EINVAL, "U-Net and problem use different contexts");
  if (U && (P->nx % 64 || P->ny % 64))
    return set_error(P->ctx, CPGPU_EUNSUPPORTED,
                     "the U-Net needs grid sides that are multiples of 64 (got %dx%d)",
                     P->nx, P->ny);
  P->unet = U;
  P->unet_flags = U ? flags : 0;
  return CPGPU_OK;
}

int cpgpu_poisson_apply(cpgpu_poisson* P, const double* x_field,
                        double* y_field) {
  if (!P || !x_field || !y_field) return CPGPU_EINVAL;
  int rc = field_to_dev(P, x_field, P->p0.p);
  if (rc) return rc;
  if (P->general) {
    if ((rc = general_apply(P, P->p0.p, P->q.p))) return rc;
  } else {
    hipLaunchKernelGGL(k_apply<double>, dim3(P->tl.G), dim3(TJ, BY), 0,
                       P->ctx->stream, P->sten64(), P->p0.p, P->q.p, P->tl);
    CPGPU_HIP(P->ctx, hipGetLastError());
  }
  return dev_to_field(P, P->q.p, y_field);
}

int cpgpu_poisson_residual(cpgpu_poisson* P, double* rel) {
  if (!P || !rel) return CPGPU_EINVAL;
  if (P->general) return general_residual(P, rel);
  int rc = residual64(P, true);
  if (rc) return rc;
  double nb = std::sqrt(P->h_rst->bb);
  *rel = std::sqrt(P->h_rst->rr) / (nb == 0.0 ? 1.0 : nb);
  return CPGPU_OK;
}

int cpgpu_poisson_solve(cpgpu_poisson* P, int method, int* max_iter,
                        double* tol, int norm_mode, int opts, double* hist,
                        int hist_cap, int* hist_len) {
  if (!P || !max_iter || !tol || *max_iter < 0 || hist_cap < 0)
    return CPGPU_EINVAL;
  if (norm_mode != CPGPU_NORM_REF && norm_mode != CPGPU_NORM_TRUE)
    return set_error(P->ctx, CPGPU_EINVAL, "bad norm mode");
  CPGPU_HIP(P->ctx, hipSetDevice(P->ctx->device));
  return poisson_solve_internal(P, method, max_iter, tol, norm_mode, opts,
                                hist, hist_cap, hist_len);
}

int cpgpu_poisson2d_solve(cpgpu_ctx* ctx, const cpgpu_grid2d* g,
                          const double* src_field, double src_uniform,
                          double* phi_field, int method, int* max_iter,
                          double* tol, int norm_mode, double* hist,
                          int hist_cap, int* hist_len) {
  if (!ctx || !g || !phi_field) return CPGPU_EINVAL;
  cpgpu_poisson* P = nullptr;
  int rc = cpgpu_poisson_create(ctx, g, &P);
  if (rc) return rc;
  if ((rc = cpgpu_poisson_set_source(P, src_field, src_uniform)) ||
      (rc = cpgpu_poisson_set_phi(P, phi_field))) {
    cpgpu_poisson_destroy(P);
    return rc;
  }
  int code = cpgpu_poisson_solve(P, method, max_iter, tol, norm_mode, 0, hist,
                                 hist_cap, hist_len);
  if (code >= 0) {
    rc = cpgpu_poisson_get_phi(P, phi_field);
    if (rc) code = rc;
  }
  cpgpu_poisson_destroy(P);
  return code;
}

int cpgpu_poisson2d_solve_static(cpgpu_ctx* ctx, const cpgpu_grid2d* g,
                                 const double* d_src, double src_uniform, double* d_phi,
                                 int method, int* max_iter, double* tol, int norm_mode) {
  if (!ctx || !g || !d_phi || !max_iter || !tol || !g->cx || !g->sx || !g->cy || !g->sy)
    return set_error(ctx, CPGPU_EINVAL, "invalid arguments");
  if (method != CPGPU_JPCG && method != CPGPU_CG)
    return set_error(ctx, CPGPU_EUNSUPPORTED, "static solve: CPGPU_JPCG or CPGPU_CG only");
  if (g->nx < 1 || g->ny < 1 || g->gl < 1 || *max_iter < 0)
    return set_error(ctx, CPGPU_EINVAL, "invalid grid");
  if ((size_t)g->nx * g->ny > kStaticMaxCells || g->nx > kStaticMaxSide ||
      g->ny > kStaticMaxSide)
    return set_error(ctx, CPGPU_EINVAL,
                     "static solve: %d x %d cells exceed the static scratch (at most %zu "
                     "cells and %d per side)",
                     g->nx, g->ny, kStaticMaxCells, kStaticMaxSide);
  CPGPU_HIP(ctx, hipSetDevice(ctx->device));
  return solve_static(ctx, g, d_src, src_uniform, d_phi, method, max_iter, tol, norm_mode);
}

}  // extern "C"
