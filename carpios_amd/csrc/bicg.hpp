// Dia_BiCGSTAB (lib/algebra/solver_matrix.hpp:599-672) with its scalars on
// the device, shared by the matrix-free structured operator (krylov.hip)
// and the CSR matrix (csr.hip).  An iteration of the reference
//
//   rho_1 = rt.r                      (rho_1 == 0: code 2)
//   beta = (rho_1 / rho_2)(alpha / omega)     (from the second iteration)
//   p = r (first) or r + beta (p - omega v);  ph = p D^-1;  v = A ph
//   alpha = rho_1 / rt.v
//   s = r - alpha v;  ||s|| / ||b|| < tol: x += alpha ph, exit (max_iter kept)
//   sh = s D^-1;  t = A sh;  omega = t.s / t.t
//   x += alpha ph + omega sh;  r = s - omega t;  rho_2 = rho_1
//   resid = ||r|| / ||b|| < tol: exit;  omega == 0: code 3
//
// runs as three operator steps and three one-block scalar kernels:
//   dir:   p and v = A ph, partials of rt.v
//   alpha: alpha
//   st:    s and t = A sh, partials of s.s (the reference nrm2: without the
//          last unknown when norm_ref), t.s and t.t
//   omega: the half-step test, else omega
//   xr:    x and r (the half-step exit: x += alpha ph only), partials of
//          r.r and of rt.r, the next iteration's rho_1
//   test:  the stopping tests, rho_2, the next rho_1 and beta
// Every kernel returns at entry once the state says the run has stopped, so
// the host launches batches of iterations and reads the state once per
// batch.  The structured operator forms ph (sh) on an LDS tile with its
// halo inside the product kernel, from the previous p and v, which
// therefore alternate between two buffers (a tile's halo reads the
// neighbouring tiles' previous p and v while they write the new ones); the
// CSR matrix forms ph (sh) in an elementwise pass before its product.
#pragma once

#include <cmath>

#include "poisson.hpp"

namespace cpgpu {
namespace {

constexpr int kBicgBatch = 16;  // iterations per host read; even (p, v alternate)

__global__ __launch_bounds__(256) void k_bicg_rho0(const double* part, int G, BState* st) {
  __shared__ double lds[4];
  double t[1];
  reduce_partials<double, 1>(part, G, t, lds);
  if (threadIdx.x == 0) {
    st->rho_1 = t[0];
    if (t[0] == 0) {
      st->stop = 1;
      st->code = CPGPU_BREAKDOWN_RHO;
    }
  }
}

__global__ __launch_bounds__(256) void k_bicg_alpha(const double* part, int G, BState* st) {
  __shared__ double lds[4];
  if (st->stop) return;
  double t[1];
  reduce_partials<double, 1>(part, G, t, lds);
  if (threadIdx.x == 0) st->alpha = st->rho_1 / t[0];
}

__global__ __launch_bounds__(256) void k_bicg_omega(const double* part, int G, BState* st,
                                                    double* hist) {
  __shared__ double lds[12];
  if (st->stop) return;
  double t[3];
  reduce_partials<double, 3>(part, G, t, lds);
  if (threadIdx.x == 0) {
    const double ns = std::sqrt(t[0]) / st->normb;
    if (ns < st->tol) {
      st->half = 1;
      st->resid = ns;
      if (hist && st->hl < st->hist_cap) hist[st->hl] = ns;
      st->hl += 1;
      st->code = CPGPU_OK;
    } else {
      st->omega = t[1] / t[2];
    }
  }
}

__global__ __launch_bounds__(256) void k_bicg_test(const double* part, int G, BState* st,
                                                   double* hist) {
  __shared__ double lds[8];
  if (st->stop) return;
  if (st->half) {  // the update kernel has taken the half step
    if (threadIdx.x == 0) st->stop = 1;
    return;
  }
  double t[2];
  reduce_partials<double, 2>(part, G, t, lds);
  if (threadIdx.x == 0) {
    const double resid = std::sqrt(t[0]) / st->normb;
    st->resid = resid;
    if (hist && st->hl < st->hist_cap) hist[st->hl] = resid;
    st->hl += 1;
    st->it += 1;
    st->rho_2 = st->rho_1;
    if (resid < st->tol) {
      st->stop = 1;
      st->code = CPGPU_OK;
    } else if (st->omega == 0) {
      st->stop = 1;
      st->code = CPGPU_BREAKDOWN_OMEGA;
    } else if (st->it >= st->max_iter) {
      st->stop = 1;
      st->code = CPGPU_MAXITER;
    } else {
      st->rho_1 = t[1];
      if (t[1] == 0) {
        st->stop = 1;
        st->code = CPGPU_BREAKDOWN_RHO;
      } else {
        st->beta = (st->rho_1 / st->rho_2) * (st->alpha / st->omega);
      }
    }
  }
}

// x += alpha ph + omega sh, r = s - omega t (ph = p D^-1, sh = s D^-1 with
// dinv = 1 / a_ii, the bits of the reference's ph and sh); partials of r.r
// (without the last unknown when skip_last) and rt.r.  After a half-step
// exit only x += alpha ph.
__global__ __launch_bounds__(256) void k_bicg_xr(size_t n, const double* __restrict__ d,
                                                 double* __restrict__ x, double* __restrict__ r,
                                                 const double* __restrict__ rt,
                                                 const double* __restrict__ p,
                                                 const double* __restrict__ sv,
                                                 const double* __restrict__ tv,
                                                 const BState* st, int skip_last,
                                                 double* part) {
  __shared__ double lds[8];
  if (st->stop) return;
  const double alpha = st->alpha;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t k0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (st->half) {
    for (size_t k = k0; k < n; k += stride) x[k] += alpha * (p[k] * d[k]);
    return;
  }
  const double omega = st->omega;
  double v[2] = {0.0, 0.0};
  const size_t m = skip_last ? n - 1 : n;
  for (size_t k = k0; k < n; k += stride) {
    x[k] += alpha * (p[k] * d[k]) + omega * (sv[k] * d[k]);
    const double rk = sv[k] - omega * tv[k];
    r[k] = rk;
    if (k < m) v[0] += rk * rk;
    v[1] += rt[k] * rk;
  }
  store_partials<double, 2>(v, part, gridDim.x, blockIdx.x, lds);
}

// The iterations after the initial residual test (which the caller has
// done: resid = ||r0|| / ||b|| > tol, max_iter >= 1, rt = r0).  Ops supplies
// G (blocks of its kernels, the partials count), part, and the launches
// rho0() (partials of rt.r), dir(prev, cur), st(cur), xr(cur), each on `s`
// with the state pointer it holds.  Returns the reference's status; tol,
// max_iter (set on a full-step exit only) and the history as the reference.
template <class Ops>
int bicg_run(cpgpu_ctx* ctx, hipStream_t s, Ops& ops, BState* dst, double normb, double resid,
             int* max_iter, double* tol, double* dh, int cap, double* hist, int* hist_len) {
  BState h0{};
  h0.normb = normb;
  h0.tol = *tol;
  h0.resid = resid;
  h0.max_iter = *max_iter;
  h0.hist_cap = dh ? cap : 0;
  h0.code = CPGPU_MAXITER;
  CPGPU_HIP(ctx, hipMemcpyAsync(dst, &h0, sizeof(BState), hipMemcpyHostToDevice, s));
  ops.rho0();
  hipLaunchKernelGGL(k_bicg_rho0, dim3(1), dim3(256), 0, s, (const double*)ops.part, ops.G,
                     dst);
  BState h{};
  for (;;) {
    for (int k = 0; k < kBicgBatch; ++k) {
      const int cur = (k + 1) & 1;
      ops.dir(k & 1, cur);
      hipLaunchKernelGGL(k_bicg_alpha, dim3(1), dim3(256), 0, s, (const double*)ops.part,
                         ops.G, dst);
      ops.st(cur);
      hipLaunchKernelGGL(k_bicg_omega, dim3(1), dim3(256), 0, s, (const double*)ops.part,
                         ops.G, dst, dh);
      ops.xr(cur);
      hipLaunchKernelGGL(k_bicg_test, dim3(1), dim3(256), 0, s, (const double*)ops.part,
                         ops.G, dst, dh);
    }
    CPGPU_HIP(ctx, hipGetLastError());
    CPGPU_HIP(ctx, hipMemcpyAsync(&h, dst, sizeof(BState), hipMemcpyDeviceToHost, s));
    CPGPU_HIP(ctx, hipStreamSynchronize(s));
    if (h.stop) break;
  }
  if (hist && dh && cap > 0 && h.hl > 0) {
    CPGPU_HIP(ctx, hipMemcpyAsync(hist, dh, sizeof(double) * (h.hl < cap ? h.hl : cap),
                                  hipMemcpyDeviceToHost, s));
    CPGPU_HIP(ctx, hipStreamSynchronize(s));
  }
  *hist_len = h.hl;
  *tol = h.resid;
  if (h.code == CPGPU_OK && !h.half) *max_iter = h.it;
  return h.code;
}

}  // namespace
}  // namespace cpgpu
