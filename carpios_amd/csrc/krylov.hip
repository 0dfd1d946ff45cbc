// The reference's other two structured solvers on the device, matrix-free:
//   carpio::Dia_BiCGSTAB    (lib/algebra/solver_matrix.hpp:599-672)
//   carpio::Solver_Jacobi_  (lib/algebra/solver_matrix.hpp:820-910)
// with their quirks: BiCGSTAB's half-step exit on ||s|| leaves max_iter
// unchanged (:642-647) and codes 2 (rho = 0) / 3 (omega = 0); Jacobi returns
// the previous iterate on convergence (:898-903).  FP64, no FMA contraction
// (the Makefile builds with -ffp-contract=off): the elementwise arithmetic
// and the operator rows are the reference's; dot products are associated
// per block (deterministic), so iterates agree with the reference to
// rounding.
#include <algorithm>
#include <cmath>
#include <vector>

#include "bicg.hpp"
#include "poisson.hpp"

namespace cpgpu {
namespace {

constexpr int kG1 = 1024;  // blocks of the 1-D kernels

// partials of x.y (NV = 1) or of x.y and u.w (NV = 2); with norm_ref the
// last element is left out of the sums (carpio::nrm2, BLAS_Level1.hpp:569)
template <int NV>
__global__ __launch_bounds__(256) void k_dots(const double* __restrict__ x,
                                              const double* __restrict__ y,
                                              const double* __restrict__ u,
                                              const double* __restrict__ w,
                                              size_t n, int skip_last,
                                              double* part) {
  __shared__ double lds[4 * NV];
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  const size_t m = skip_last ? n - 1 : n;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < m;
       k += (size_t)gridDim.x * blockDim.x) {
    v[0] += x[k] * y[k];
    if constexpr (NV == 2) v[1] += u[k] * w[k];
  }
  store_partials<double, NV>(v, part, gridDim.x, blockIdx.x, lds);
}

template <int NV>
__global__ __launch_bounds__(256) void k_total_n(const double* part, int G,
                                                 double* out) {
  __shared__ double lds[4 * NV];
  double t[NV];
  reduce_partials<double, NV>(part, G, t, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = t[k];
  }
}


// y = A x
__global__ __launch_bounds__(256) void k_apply64(Sten<double> s, const double* __restrict__ x,
                                                 double* __restrict__ y, Tiling tl) {
  for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
    int i0, j0;
    tile_origin(t, tl.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < TI / BY; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i < s.nx && j < s.ny)
        y[(size_t)i * s.ny + j] = apply_global(s, x, i, j, diag_at(s, i, j));
    }
  }
}





// Dia_BiCGSTAB's operator kernels on the structured operator (the scalar
// kernels, the update kernel and the batch loop are shared with the CSR
// matrix).  Both stage a (TI+2) x (TJ+2) tile with its halo in LDS, as
// k_cg_dir does: every global load of the tile issued first from clamped
// valid addresses, ph (or sh) formed once per staged cell, then the rows of
// A x from LDS in the reference's CSR term order (row_apply).

// dinv = 1 / a_ii (the factor of the reference's ph = p * (1 / a_ii))
__global__ __launch_bounds__(256) void k_dinv64(Sten<double> s, double* __restrict__ dinv,
                                                Tiling tl) {
  for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
    int i0, j0;
    tile_origin(t, tl.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < TI / BY; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i < s.nx && j < s.ny) dinv[(size_t)i * s.ny + j] = 1. / diag_at(s, i, j);
    }
  }
}

// staged element k of this thread: its tile coordinates and clamped global index
__device__ __forceinline__ size_t stage_index(const Sten<double>& s, int i0, int j0, int k) {
  const int tid = threadIdx.x + TJ * threadIdx.y;
  const int e = min(tid + TJ * BY * k, kStageE - 1);
  const int ii = e / (TJ + 2), jj = e - ii * (TJ + 2);
  const int gi = min(max(i0 - 1 + ii, 0), s.nx - 1), gj = min(max(j0 - 1 + jj, 0), s.ny - 1);
  return (size_t)gi * s.ny + gj;
}

// the face coefficients of the tile's rows and columns into LDS
__device__ __forceinline__ void stage_coeffs(const Sten<double>& s, int i0, int j0,
                                             double ckx, double cky, TileLds<double>& L) {
  const int tid = threadIdx.x + TJ * threadIdx.y;
  if (tid < TI + 3) L.skx[tid] = ckx;
  if (tid < TJ + 3) L.sky[tid] = cky;
}

// row (i, j) of A x from the staged tile
__device__ __forceinline__ double tile_row(const Sten<double>& s, const TileLds<double>& L,
                                           int li, int lj, int i, int j) {
  const double kx0 = L.skx[li], kx1 = L.skx[li + 1], ky0 = L.sky[lj], ky1 = L.sky[lj + 1];
  const double dg = diag_from(s, i, j, kx0, kx1, ky0, ky1);
  return row_apply(s, i, j, dg, L.tile[li][lj], L.tile[li - 1][lj], L.tile[li + 1][lj],
                   L.tile[li][lj - 1], L.tile[li][lj + 1], kx0, kx1, ky0, ky1);
}

// p = r (first) or r + beta (p_old - omega v_old), v = A (p D^-1);
// partials of rt.v
__global__ __launch_bounds__(256) void k_sbicg_dir(
    Sten<double> s, const double* __restrict__ d, const double* __restrict__ r,
    const double* __restrict__ rt, const double* __restrict__ pold,
    const double* __restrict__ vold, double* __restrict__ pnew, double* __restrict__ vnew,
    const BState* st, double* part, Tiling tl) {
  __shared__ double lds[4];
  __shared__ TileLds<double> L;
  if (st->stop) return;
  const bool first = st->it == 0;
  const double beta = st->beta, omega = st->omega;
  const int nx = s.nx, ny = s.ny;
  const int tid = threadIdx.x + TJ * threadIdx.y;
  double v[1] = {0.0};
  for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
    int i0, j0;
    tile_origin(t, tl.ntj, i0, j0);
    double br[kStageK], bd[kStageK], bp[kStageK], bv[kStageK];
#pragma unroll
    for (int k = 0; k < kStageK; ++k) {
      const size_t m = stage_index(s, i0, j0, k);
      br[k] = r[m];
      bd[k] = d[m];
      bp[k] = first ? 0.0 : pold[m];
      bv[k] = first ? 0.0 : vold[m];
    }
    const double ckx = s.KX[min(max(i0 - 1 + tid, 0), nx)];
    const double cky = s.KY[min(max(j0 - 1 + tid, 0), ny)];
    __syncthreads();  // the previous tile's reads are done
    stage_coeffs(s, i0, j0, ckx, cky, L);
#pragma unroll
    for (int k = 0; k < kStageK; ++k) {
      const int e = tid + TJ * BY * k;
      if (e >= kStageE) continue;
      const int ii = e / (TJ + 2), jj = e - ii * (TJ + 2);
      const int gi = i0 - 1 + ii, gj = j0 - 1 + jj;
      double ph = 0.0;
      if (gi >= 0 && gi < nx && gj >= 0 && gj < ny) {
        const double pv = first ? br[k] : br[k] + beta * (bp[k] - omega * bv[k]);
        ph = pv * bd[k];
        if (ii >= 1 && ii <= TI && jj >= 1 && jj <= TJ) pnew[(size_t)gi * ny + gj] = pv;
      }
      L.tile[ii][jj] = ph;
    }
    __syncthreads();
    const int lj = threadIdx.x + 1, j = j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < TI / BY; ++k) {
      const int li = threadIdx.y + BY * k + 1, i = i0 + li - 1;
      if (i >= nx || j >= ny) continue;
      const double a = tile_row(s, L, li, lj, i, j);
      const size_t idx = (size_t)i * ny + j;
      vnew[idx] = a;
      v[0] += rt[idx] * a;
    }
  }
  store_partials<double, 1>(v, part, gridDim.x, blockIdx.x, lds);
}

// s = r - alpha v, t = A (s D^-1); partials of s.s (norm_ref: without the
// last unknown), t.s, t.t
__global__ __launch_bounds__(256) void k_sbicg_st(
    Sten<double> s, const double* __restrict__ d, const double* __restrict__ r,
    const double* __restrict__ vv, double* __restrict__ sv, double* __restrict__ tv,
    const BState* st, int norm_ref, double* part, Tiling tl) {
  __shared__ double lds[12];
  __shared__ TileLds<double> L;
  __shared__ double S[TI][TJ];
  if (st->stop) return;
  const double alpha = st->alpha;
  const int nx = s.nx, ny = s.ny;
  const int tid = threadIdx.x + TJ * threadIdx.y;
  const size_t nlast = (size_t)nx * ny - 1;
  double v[3] = {0.0, 0.0, 0.0};
  for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
    int i0, j0;
    tile_origin(t, tl.ntj, i0, j0);
    double br[kStageK], bd[kStageK], bv[kStageK];
#pragma unroll
    for (int k = 0; k < kStageK; ++k) {
      const size_t m = stage_index(s, i0, j0, k);
      br[k] = r[m];
      bd[k] = d[m];
      bv[k] = vv[m];
    }
    const double ckx = s.KX[min(max(i0 - 1 + tid, 0), nx)];
    const double cky = s.KY[min(max(j0 - 1 + tid, 0), ny)];
    __syncthreads();
    stage_coeffs(s, i0, j0, ckx, cky, L);
#pragma unroll
    for (int k = 0; k < kStageK; ++k) {
      const int e = tid + TJ * BY * k;
      if (e >= kStageE) continue;
      const int ii = e / (TJ + 2), jj = e - ii * (TJ + 2);
      const int gi = i0 - 1 + ii, gj = j0 - 1 + jj;
      double sh = 0.0;
      if (gi >= 0 && gi < nx && gj >= 0 && gj < ny) {
        const double si = br[k] - alpha * bv[k];
        sh = si * bd[k];
        if (ii >= 1 && ii <= TI && jj >= 1 && jj <= TJ) {
          sv[(size_t)gi * ny + gj] = si;
          S[ii - 1][jj - 1] = si;
        }
      }
      L.tile[ii][jj] = sh;
    }
    __syncthreads();
    const int lj = threadIdx.x + 1, j = j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < TI / BY; ++k) {
      const int li = threadIdx.y + BY * k + 1, i = i0 + li - 1;
      if (i >= nx || j >= ny) continue;
      const double ti = tile_row(s, L, li, lj, i, j);
      const double si = S[li - 1][lj - 1];
      const size_t idx = (size_t)i * ny + j;
      tv[idx] = ti;
      if (!(norm_ref && idx == nlast)) v[0] += si * si;
      v[1] += ti * si;
      v[2] += ti * ti;
    }
  }
  store_partials<double, 3>(v, part, gridDim.x, blockIdx.x, lds);
}

// r = b - A x
__global__ __launch_bounds__(256) void k_resid64(Sten<double> s, const double* __restrict__ x,
                                                 const double* __restrict__ b,
                                                 double* __restrict__ r, Tiling tl) {
  for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
    int i0, j0;
    tile_origin(t, tl.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < TI / BY; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i < s.nx && j < s.ny) {
        const size_t idx = (size_t)i * s.ny + j;
        r[idx] = b[idx] - apply_global(s, x, i, j, diag_at(s, i, j));
      }
    }
  }
}

// Jacobi sweep of Solver_Jacobi_: newx = T x + C with T = -D^-1 (L + U)
// stored row by row in CSR order (the diagonal entry zeroed, off-diagonals
// -a_ik / a_ii) and C = b / a_ii; the product accumulates in the CSR order
// (i-1,j), (i,j-1), (i,j), (i,j+1), (i+1,j) from 0 like MatrixSCR_::operator*.
// Runs only while the device state says "not done".
__global__ __launch_bounds__(256) void k_jacobi_sweep(Sten<double> s,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ b,
                                                      double* __restrict__ newx,
                                                      const JState* st, Tiling tl) {
  if (st->done) return;
  const int ny = s.ny;
  for (int t = blockIdx.x; t < tl.ntiles; t += gridDim.x) {
    int i0, j0;
    tile_origin(t, tl.ntj, i0, j0);
    const int j = j0 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < TI / BY; ++k) {
      const int i = i0 + threadIdx.y + BY * k;
      if (i < s.nx && j < ny) {
        const size_t idx = (size_t)i * ny + j;
        const double dv = diag_at(s, i, j);
        double acc = 0.0;
        if (i > 0) acc = acc + x[idx - ny] * (-(-s.KX[i]) / dv);
        if (j > 0) acc = acc + x[idx - 1] * (-(-s.KY[j]) / dv);
        acc = acc + x[idx] * 0.0;
        if (j < ny - 1) acc = acc + x[idx + 1] * (-(-s.KY[j + 1]) / dv);
        if (i < s.nx - 1) acc = acc + x[idx + ny] * (-(-s.KX[i + 1]) / dv);
        newx[idx] = acc + b[idx] / dv;
      }
    }
  }
}

// r = b - A newx; partials of r.r (norm_ref: without the last unknown)
__global__ __launch_bounds__(256) void k_jacobi_resid(Sten<double> s,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ b,
                                                      double* part, const JState* st,
                                                      int norm_ref, Tiling tl) {
  __shared__ double lds[4];
  if (st->done) return;
  double v[1] = {0.0};
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
        const double rv = b[idx] - apply_global(s, x, i, j, diag_at(s, i, j));
        if (!(norm_ref && idx == nlast)) v[0] += rv * rv;
      }
    }
  }
  store_partials<double, 1>(v, part, gridDim.x, blockIdx.x, lds);
}

// resid = ||r|| / ||b||; stop at resid <= tol (the previous iterate stays
// in x: the sweep that converged wrote newx only); history entry
__global__ __launch_bounds__(256) void k_jacobi_check(const double* part, int G,
                                                      JState* st, double tol,
                                                      int max_iter, double* hist,
                                                      int hist_cap) {
  __shared__ double lds[4];
  if (st->done) return;
  double t[1];
  reduce_partials<double, 1>(part, G, t, lds);
  if (threadIdx.x == 0) {
    const int it = st->iter + 1;
    const double res = std::sqrt(t[0]) / st->normb;
    st->resid = res;
    st->iter = it;
    if (hist && it <= hist_cap) hist[it - 1] = res;
    if (res <= tol) st->done = 1;
    else if (it >= max_iter) st->done = 2;
  }
}

int total(cpgpu_poisson* P, int nv, double* out) {
  hipStream_t st = P->ctx->stream;
  if (nv == 1)
    hipLaunchKernelGGL(k_total_n<1>, dim3(1), dim3(256), 0, st, P->part64.p, kG1,
                       P->partB64.p + 2 * kMaxBlocks);
  else
    hipLaunchKernelGGL(k_total_n<2>, dim3(1), dim3(256), 0, st, P->part64.p, kG1,
                       P->partB64.p + 2 * kMaxBlocks);
  CPGPU_HIP(P->ctx, hipMemcpyAsync(out, P->partB64.p + 2 * kMaxBlocks,
                                   sizeof(double) * nv, hipMemcpyDeviceToHost, st));
  CPGPU_HIP(P->ctx, hipStreamSynchronize(st));
  return CPGPU_OK;
}

// host-side value of dot(x, y) (and dot(u, w))
int dots(cpgpu_poisson* P, const double* x, const double* y, const double* u,
         const double* w, int skip_last, double* out) {
  hipStream_t st = P->ctx->stream;
  if (u)
    hipLaunchKernelGGL(k_dots<2>, dim3(kG1), dim3(256), 0, st, x, y, u, w, P->N,
                       skip_last, P->part64.p);
  else
    hipLaunchKernelGGL(k_dots<1>, dim3(kG1), dim3(256), 0, st, x, y, u, w, P->N,
                       skip_last, P->part64.p);
  CPGPU_HIP(P->ctx, hipGetLastError());
  return total(P, u ? 2 : 1, out);
}

}  // namespace

namespace {
struct StenBicgOps {
  hipStream_t s;
  Sten<double> sten;
  Tiling tl;
  size_t N;
  double *d, *x, *r, *rt, *sv, *tv;
  double* P[2];
  double* V[2];
  const BState* state;
  double* part;
  int G, norm_ref;
  void rho0() {
    hipLaunchKernelGGL(k_dots<1>, dim3(G), dim3(256), 0, s, (const double*)rt, (const double*)r,
                       nullptr, nullptr, N, 0, part);
  }
  void dir(int prev, int cur) {
    hipLaunchKernelGGL(k_sbicg_dir, dim3(G), dim3(TJ, BY), 0, s, sten, (const double*)d,
                       (const double*)r, (const double*)rt, (const double*)P[prev],
                       (const double*)V[prev], P[cur], V[cur], state, part, tl);
  }
  void st(int cur) {
    hipLaunchKernelGGL(k_sbicg_st, dim3(G), dim3(TJ, BY), 0, s, sten, (const double*)d,
                       (const double*)r, (const double*)V[cur], sv, tv, state, norm_ref, part,
                       tl);
  }
  void xr(int cur) {
    hipLaunchKernelGGL(k_bicg_xr, dim3(G), dim3(256), 0, s, N, (const double*)d, x, r,
                       (const double*)rt, (const double*)P[cur], (const double*)sv,
                       (const double*)tv, state, norm_ref, part);
  }
};
}  // namespace

// Dia_BiCGSTAB (solver_matrix.hpp:599-672): the initial residual test on
// the host, then the device-scalar iterations of bicg.hpp
int run_bicgstab(cpgpu_poisson* P, int* max_iter, double* tol, int norm_ref,
                 double* hist, int hist_cap, int* hist_len) {
  cpgpu_ctx* ctx = P->ctx;
  hipStream_t st = ctx->stream;
  const size_t N = P->N;
  const Tiling tl = P->tl;
  const Sten<double> s = P->sten64();
  for (auto& v : P->bv)
    if (v.alloc(N)) return set_error(ctx, CPGPU_ENOMEM, "BiCGSTAB vectors");
  if (P->bst.alloc(1)) return set_error(ctx, CPGPU_ENOMEM, "BiCGSTAB state");
  if (hist && hist_cap > 0 && P->dhist.alloc(hist_cap))
    return set_error(ctx, CPGPU_ENOMEM, "hist alloc");
  double* x = P->x.p;
  double* r = P->r.p;
  double* rt = P->bv[0].p;
  double* dinv = P->bv[7].p;
  const dim3 g2(tl.G), b2(TJ, BY);
  double normb = 0, resid = 0;
  int rc;
  if ((rc = dots(P, P->b.p, P->b.p, nullptr, nullptr, norm_ref, &normb))) return rc;
  normb = std::sqrt(normb);
  hipLaunchKernelGGL(k_resid64, g2, b2, 0, st, s, x, P->b.p, r, tl);
  hipLaunchKernelGGL(k_dinv64, g2, b2, 0, st, s, dinv, tl);
  CPGPU_HIP(ctx, hipMemcpyAsync(rt, r, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
  if (normb == 0.0) normb = 1;
  if ((rc = dots(P, r, r, nullptr, nullptr, norm_ref, &resid))) return rc;
  resid = std::sqrt(resid) / normb;
  *hist_len = 0;
  if (resid <= *tol) {
    *tol = resid;
    *max_iter = 0;
    if (hist && hist_cap > 0) hist[0] = resid;
    *hist_len = 1;
    return CPGPU_OK;
  }
  if (*max_iter < 1) {
    *tol = resid;
    return CPGPU_MAXITER;
  }
  StenBicgOps ops{st, s, tl, N, dinv, x, r, rt, P->bv[5].p, P->bv[6].p,
                  {P->bv[1].p, P->bv[2].p}, {P->bv[3].p, P->bv[4].p}, P->bst.p,
                  P->part64.p, (int)tl.G, norm_ref};
  return bicg_run(ctx, st, ops, P->bst.p, normb, resid, max_iter, tol,
                  hist && hist_cap > 0 ? P->dhist.p : nullptr, hist_cap, hist, hist_len);
}

int run_jacobi(cpgpu_poisson* P, int* max_iter, double* tol, int norm_ref,
               double* hist, int hist_cap, int* hist_len) {
  cpgpu_ctx* ctx = P->ctx;
  hipStream_t st = ctx->stream;
  const size_t N = P->N;
  const Tiling tl = P->tl;
  const Sten<double> s = P->sten64();
  if (P->bv[0].alloc(N) || P->jst.alloc(1))
    return set_error(ctx, CPGPU_ENOMEM, "Jacobi vectors");
  if (hist && hist_cap > 0 && P->dhist.alloc(hist_cap))
    return set_error(ctx, CPGPU_ENOMEM, "hist alloc");
  double* dh = hist && hist_cap > 0 ? P->dhist.p : nullptr;
  const dim3 g2(tl.G), b2(TJ, BY);
  double normb = 0, resid = 0;
  int rc;
  if ((rc = dots(P, P->b.p, P->b.p, nullptr, nullptr, norm_ref, &normb))) return rc;
  normb = std::sqrt(normb);
  hipLaunchKernelGGL(k_resid64, g2, b2, 0, st, s, P->x.p, P->b.p, P->r.p, tl);
  if ((rc = dots(P, P->r.p, P->r.p, nullptr, nullptr, norm_ref, &resid))) return rc;
  if (normb == 0.0) normb = 1;
  resid = std::sqrt(resid) / normb;
  *hist_len = 0;
  if (resid <= *tol) {
    *tol = resid;
    *max_iter = 0;
    return CPGPU_OK;
  }
  JState h0{normb, resid, 0, 0};
  CPGPU_HIP(ctx, hipMemcpyAsync(P->jst.p, &h0, sizeof(JState), hipMemcpyHostToDevice, st));
  // iterate k lives in buf[k & 1] (buf[0] = x); sweeps after the stop return
  double* buf[2] = {P->x.p, P->bv[0].p};
  const int B = 16;
  JState hs{};
  for (int k0 = 0; k0 < *max_iter; k0 += B) {
    for (int k = k0; k < std::min(k0 + B, *max_iter); ++k) {
      hipLaunchKernelGGL(k_jacobi_sweep, g2, b2, 0, st, s, (const double*)buf[k & 1],
                         (const double*)P->b.p, buf[(k + 1) & 1], (const JState*)P->jst.p, tl);
      hipLaunchKernelGGL(k_jacobi_resid, g2, b2, 0, st, s, (const double*)buf[(k + 1) & 1],
                         (const double*)P->b.p, P->part64.p, (const JState*)P->jst.p,
                         norm_ref, tl);
      hipLaunchKernelGGL(k_jacobi_check, dim3(1), dim3(256), 0, st,
                         (const double*)P->part64.p, (int)tl.G, P->jst.p, *tol,
                         *max_iter, dh, hist_cap);
    }
    CPGPU_HIP(ctx, hipGetLastError());
    CPGPU_HIP(ctx, hipMemcpyAsync(&hs, P->jst.p, sizeof(JState), hipMemcpyDeviceToHost, st));
    CPGPU_HIP(ctx, hipStreamSynchronize(st));
    if (hs.done) break;
  }
  const int it = hs.iter;
  // converged at iteration it: x is iterate it - 1 (in buf[(it - 1) & 1]);
  // stopped at max_iter: x is iterate it (the reference copied newx)
  const int keep = hs.done == 1 ? it - 1 : it;
  if (keep & 1)
    CPGPU_HIP(ctx, hipMemcpyAsync(P->x.p, buf[1], sizeof(double) * N,
                                  hipMemcpyDeviceToDevice, st));
  if (dh && it > 0)
    CPGPU_HIP(ctx, hipMemcpyAsync(hist, dh, sizeof(double) * std::min(it, hist_cap),
                                  hipMemcpyDeviceToHost, st));
  CPGPU_HIP(ctx, hipStreamSynchronize(st));
  *hist_len = it;
  *tol = it > 0 ? hs.resid : resid;  // no sweep ran (max_iter <= 0)
  if (hs.done == 1) {
    *max_iter = it;
    return CPGPU_OK;
  }
  return CPGPU_MAXITER;
}

}  // namespace cpgpu
