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
#include <cmath>
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

__global__ __launch_bounds__(256) void k_fcg_start(const double* part, int G,
                                                   FState* st) {
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
  }
}

// partials of z . q (the previous q)
__global__ __launch_bounds__(256) void k_fcg_zq(const float* __restrict__ z,
                                                const double* __restrict__ q,
                                                double* part, size_t n) {
  __shared__ double lds[4];
  double v[1] = {0.0};
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    v[0] += (double)z[k] * q[k];
  store_partials<double, 1>(v, part, gridDim.x, blockIdx.x, lds);
}

// p_new = z + beta p_old (beta = -(z.q_prev)/(p.q)_prev, 0 on the first
// iteration), q = A p_new, partials of p.q and p.r.  p_new at the four
// neighbours is recomputed from z and p_old, so p_old and p_new are
// different buffers.
__global__ __launch_bounds__(256) void k_fcg_dir(Sten<double> s,
                                                 const float* __restrict__ z,
                                                 const double* __restrict__ pold,
                                                 double* __restrict__ pnew,
                                                 double* __restrict__ q,
                                                 const double* __restrict__ r,
                                                 const double* partZ, int GZ,
                                                 const FState* st, int first,
                                                 double* part, Tiling tl) {
  __shared__ double lds[8];
  double beta = 0.0;
  if (!first) {
    double t[1];
    reduce_partials<double, 1>(partZ, GZ, t, lds);
    beta = -t[0] / st->pq;
  }
  double v[2] = {0.0, 0.0};
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
        const double pc = (double)z[idx] + beta * pold[idx];
        double acc = pc * diag_at(s, i, j);
        if (i > 0) acc -= s.KX[i] * ((double)z[idx - ny] + beta * pold[idx - ny]);
        if (j > 0) acc -= s.KY[j] * ((double)z[idx - 1] + beta * pold[idx - 1]);
        if (j < ny - 1) acc -= s.KY[j + 1] * ((double)z[idx + 1] + beta * pold[idx + 1]);
        if (i < s.nx - 1)
          acc -= s.KX[i + 1] * ((double)z[idx + ny] + beta * pold[idx + ny]);
        pnew[idx] = pc;
        q[idx] = acc;
        v[0] += pc * acc;
        v[1] += pc * r[idx];
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
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    x[k] = x[k] + alpha * p[k];
    const double rv = r[k] - alpha * q[k];
    r[k] = rv;
    r32[k] = (float)rv;
    if (!(norm_ref && k == n - 1)) v[0] += rv * rv;
  }
  store_partials<double, 1>(v, part, gridDim.x, blockIdx.x, lds);
}

__global__ __launch_bounds__(256) void k_fcg_check(const double* part, int G,
                                                   FState* st) {
  __shared__ double lds[4];
  double t[1];
  reduce_partials<double, 1>(part, G, t, lds);
  if (threadIdx.x == 0) {
    st->resid = std::sqrt(t[0]) / st->normb;
    st->iter = st->iter + 1;
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

int fcg_alloc(cpgpu_poisson* P) {
  int rc = poisson_prepare_f32(P);
  if (rc) return rc;
  if (P->pf[0].alloc(P->N) || P->pf[1].alloc(P->N) || P->fst.alloc(1))
    return set_error(P->ctx, CPGPU_ENOMEM, "flexible-CG vectors");
  if (!P->h_fst &&
      hipHostMalloc((void**)&P->h_fst, sizeof(FState), hipHostMallocDefault) !=
          hipSuccess) {
    P->h_fst = nullptr;
    return set_error(P->ctx, CPGPU_ENOMEM, "pinned flexible-CG state");
  }
  return CPGPU_OK;
}

int read_state(cpgpu_poisson* P, FState* out) {
  hipStream_t st = P->ctx->stream;
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
  if ((rc = unet_apply_device(P->unet, P->nx, P->ny, P->b32.p, P->x32.p)))
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
  hipLaunchKernelGGL(k_fcg_start, dim3(1), dim3(256), 0, st, P->partB64.p,
                     (int)tl.G, P->fst.p);
  CPGPU_HIP(ctx, hipGetLastError());
  FState fs;
  if ((rc = read_state(P, &fs))) return rc;
  double res = fs.resid;
  std::vector<double> trace;
  trace.push_back(res);
  int k = 0;
  bool fallback = false;
  while (res > tol) {
    if (k >= max_iter) break;
    // z = preconditioned residual
    if ((rc = unet_apply_device(P->unet, P->nx, P->ny, P->r32.p, P->pa32.p)))
      return rc;
    if (k > 0)
      hipLaunchKernelGGL(k_fcg_zq, dim3(G1), dim3(256), 0, st, P->pa32.p,
                         P->q.p, P->partA64.p, N);
    const double* pold = P->pf[k & 1].p;
    double* pnew = P->pf[(k + 1) & 1].p;
    hipLaunchKernelGGL(k_fcg_dir, dim3(tl.G), dim3(TJ, BY), 0, st, P->sten64(),
                       P->pa32.p, pold, pnew, P->q.p, P->r.p, P->partA64.p, G1,
                       P->fst.p, k == 0 ? 1 : 0, P->partB64.p, tl);
    hipLaunchKernelGGL(k_fcg_upd, dim3(G1), dim3(256), 0, st, pnew, P->q.p,
                       P->x.p, P->r.p, P->r32.p, P->partB64.p, (int)tl.G,
                       P->fst.p, P->partR64.p, N, norm_ref);
    hipLaunchKernelGGL(k_fcg_check, dim3(1), dim3(256), 0, st, P->partR64.p, G1,
                       P->fst.p);
    CPGPU_HIP(ctx, hipGetLastError());
    ++k;
    if ((rc = read_state(P, &fs))) return rc;
    res = fs.resid;
    trace.push_back(res);
    if (fs.breakdown || !std::isfinite(res)) {
      fallback = true;
      break;
    }
    // stagnation: less than a factor 2 over the last 16 iterations
    if (k >= 16 && res > 0.5 * trace[k - 16]) {
      fallback = true;
      break;
    }
  }
  const int kf = k;
  int code = res <= tol ? CPGPU_OK : CPGPU_MAXITER;
  if (code != CPGPU_OK && (fallback || k < max_iter)) {
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
