// General CSR Krylov solvers on the device: the reference's free-function
// solvers of lib/algebra/solver_matrix.hpp over a MatrixSCR_ (CSR) matrix,
//   carpio::CG              (solver_matrix.hpp:264-316)
//   IC_CG's loop with DiaPre(A,1) (:318-374, preconditioner.hpp:182-262)
//   carpio::Dia_BiCGSTAB    (:599-672)
//   carpio::Solver_Jacobi_  (:820-910)
// with the same status codes, max_iter/tol in/out meaning, residual history
// and quirks as the structured path (half-step BiCGSTAB exit, Jacobi's
// previous iterate, nrm2 without the last element in reference-norm mode).
// The product y = A x sums each row in CSR order from 0, like
// MatrixSCR_::operator* (matrix_SparCompRow.hpp:236-249), one row per lane:
// the operator is the reference's to the bit.  Dot products are associated
// per block (deterministic), so iterates agree with the reference to
// rounding.  IC_CG / IC_CGS (incomplete-Cholesky triangular solves) are
// sequential by construction and stay on the host.
#include <algorithm>
#include <cmath>
#include <vector>

#include "poisson.hpp"

struct cpgpu_csr {
  cpgpu_ctx* ctx = nullptr;
  int n = 0, nnz = 0;
  cpgpu::DevBuf<int> rp, ci;
  cpgpu::DevBuf<double> val;
  cpgpu::DevBuf<double> dinv;   // DiaPre: 1 / a_ii
  cpgpu::DevBuf<double> tval;   // Solver_Jacobi_'s T = -D^-1 (L + U)
  cpgpu::DevBuf<double> vec[10];
  cpgpu::DevBuf<double> part;   // block partials + totals
  cpgpu::DevBuf<int> flag;
};

namespace cpgpu {
namespace {

constexpr int kCsrG = 1024;    // blocks of the 1-D kernels
constexpr double kSmall = 1e-10;  // SMALL, lib/type_define.hpp:16

inline int csr_grid(int n) { return std::max(1, std::min(kCsrG, (n + 255) / 256)); }

// y = A x: row i summed in CSR order from 0 (matrix_SparCompRow.hpp:236-249)
__global__ __launch_bounds__(256) void k_csr_mult(int n, const int* __restrict__ rp,
                                                  const int* __restrict__ ci,
                                                  const double* __restrict__ val,
                                                  const double* __restrict__ x,
                                                  double* __restrict__ y) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) acc += x[ci[k]] * val[k];
    y[i] = acc;
  }
}

// r = b - A x, the product formed first (array_list.hpp:820-825)
__global__ __launch_bounds__(256) void k_csr_resid(int n, const int* __restrict__ rp,
                                                   const int* __restrict__ ci,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ b,
                                                   double* __restrict__ r) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) acc += x[ci[k]] * val[k];
    r[i] = b[i] - acc;
  }
}

// DiaPre(A, 1): 1 / a_ii from the first diagonal entry of each row; a zero
// or missing diagonal raises the flag (the reference exit()s)
__global__ __launch_bounds__(256) void k_csr_diapre(int n, const int* __restrict__ rp,
                                                    const int* __restrict__ ci,
                                                    const double* __restrict__ val,
                                                    double* __restrict__ d, int* flag) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double dv = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (ci[k] == i) {
        dv = val[k];
        break;
      }
    d[i] = dv == 0.0 ? 0.0 : 1. / dv;
    if (dv == 0.0) *flag = 1;
  }
}

// Solver_Jacobi_'s T and C (solver_matrix.hpp:866-885): the first diagonal
// entry of a row becomes 0 and gives C_i = b_i / a_ii; the other entries
// become -a_ik / a_ii; a row without a diagonal keeps C_i = 1 / SMALL
__global__ __launch_bounds__(256) void k_csr_jacobi_setup(
    int n, const int* __restrict__ rp, const int* __restrict__ ci,
    const double* __restrict__ val, const double* __restrict__ b, double* __restrict__ tv,
    double* __restrict__ c) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double dv = 0.0;
    int fl = -1;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (ci[k] == i) {
        fl = k;
        dv = val[k];
        break;
      }
    c[i] = 1.0 / kSmall;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      if (k == fl) {
        tv[k] = 0.0;
        c[i] = b[ci[k]] / dv;
      } else {
        tv[k] = -val[k] / dv;
      }
    }
  }
}

// newx = T x + C (the product in CSR order, then + C)
__global__ __launch_bounds__(256) void k_csr_jacobi_sweep(int n, const int* __restrict__ rp,
                                                          const int* __restrict__ ci,
                                                          const double* __restrict__ tv,
                                                          const double* __restrict__ c,
                                                          const double* __restrict__ x,
                                                          double* __restrict__ newx) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) acc += x[ci[k]] * tv[k];
    newx[i] = acc + c[i];
  }
}

// partials of x.y (and u.w); skip_last leaves the last element out
template <int NV>
__global__ __launch_bounds__(256) void k_csr_dots(const double* __restrict__ x,
                                                  const double* __restrict__ y,
                                                  const double* __restrict__ u,
                                                  const double* __restrict__ w, int n,
                                                  int skip_last, double* part) {
  __shared__ double lds[4 * NV];
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  const int m = skip_last ? n - 1 : n;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
    v[0] += x[k] * y[k];
    if constexpr (NV == 2) v[1] += u[k] * w[k];
  }
  store_partials<double, NV>(v, part, gridDim.x, blockIdx.x, lds);
}

template <int NV>
__global__ __launch_bounds__(256) void k_csr_total(const double* part, int G, double* out) {
  __shared__ double lds[4 * NV];
  double t[NV];
  reduce_partials<double, NV>(part, G, t, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = t[k];
  }
}

// z = r (no preconditioner) or r * d
__global__ __launch_bounds__(256) void k_csr_precond(int n, const double* __restrict__ r,
                                                     const double* __restrict__ d,
                                                     double* __restrict__ z) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    z[k] = d ? r[k] * d[k] : r[k];
}

// p = z (first) or z + beta p
__global__ __launch_bounds__(256) void k_csr_dir(int n, const double* __restrict__ z,
                                                 double* __restrict__ p, double beta, int first) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    p[k] = first ? z[k] : z[k] + beta * p[k];
}

// x += alpha p; r -= alpha q
__global__ __launch_bounds__(256) void k_csr_cg_upd(int n, double* __restrict__ x,
                                                    double* __restrict__ r,
                                                    const double* __restrict__ p,
                                                    const double* __restrict__ q, double alpha) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    x[k] += alpha * p[k];
    r[k] -= alpha * q[k];
  }
}

// p = r (first) or r + beta (p - omega v); ph = p * d
__global__ __launch_bounds__(256) void k_csr_bicg_p(int n, const double* __restrict__ r,
                                                    double* __restrict__ p,
                                                    const double* __restrict__ v,
                                                    const double* __restrict__ d,
                                                    double* __restrict__ ph, double beta,
                                                    double omega, int first) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const double pv = first ? r[k] : r[k] + beta * (p[k] - omega * v[k]);
    p[k] = pv;
    ph[k] = pv * d[k];
  }
}

// s = r - alpha v
__global__ __launch_bounds__(256) void k_csr_bicg_s(int n, const double* __restrict__ r,
                                                    const double* __restrict__ v,
                                                    double* __restrict__ s, double alpha) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    s[k] = r[k] - alpha * v[k];
}

// x += a y
__global__ __launch_bounds__(256) void k_csr_axpy(int n, double* __restrict__ x,
                                                  const double* __restrict__ y, double a) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    x[k] += a * y[k];
}

// x += alpha ph + omega sh; r = s - omega t
__global__ __launch_bounds__(256) void k_csr_bicg_x(int n, double* __restrict__ x,
                                                    double* __restrict__ r,
                                                    const double* __restrict__ ph,
                                                    const double* __restrict__ sh,
                                                    const double* __restrict__ s,
                                                    const double* __restrict__ t, double alpha,
                                                    double omega) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    x[k] += alpha * ph[k] + omega * sh[k];
    r[k] = s[k] - omega * t[k];
  }
}

struct CsrRun {
  cpgpu_csr* A;
  hipStream_t st;
  int G;
  int norm_ref;
  double* hist;
  int hist_cap;
  int hl = 0;

  void push(double v) {
    if (hist && hl < hist_cap) hist[hl] = v;
    ++hl;
  }
  // host value of x.y (and u.w)
  int dots(const double* x, const double* y, const double* u, const double* w, int skip_last,
           double* out) {
    const int n = A->n;
    if (u)
      hipLaunchKernelGGL(k_csr_dots<2>, dim3(G), dim3(256), 0, st, x, y, u, w, n, skip_last,
                         A->part.p);
    else
      hipLaunchKernelGGL(k_csr_dots<1>, dim3(G), dim3(256), 0, st, x, y, u, w, n, skip_last,
                         A->part.p);
    double* tot = A->part.p + 2 * kCsrG;
    if (u)
      hipLaunchKernelGGL(k_csr_total<2>, dim3(1), dim3(256), 0, st, A->part.p, G, tot);
    else
      hipLaunchKernelGGL(k_csr_total<1>, dim3(1), dim3(256), 0, st, A->part.p, G, tot);
    CPGPU_HIP(A->ctx, hipGetLastError());
    CPGPU_HIP(A->ctx, hipMemcpyAsync(out, tot, sizeof(double) * (u ? 2 : 1),
                                     hipMemcpyDeviceToHost, st));
    CPGPU_HIP(A->ctx, hipStreamSynchronize(st));
    return CPGPU_OK;
  }
  // carpio::nrm2 (reference mode: without the last element; n == 1: |x0|)
  // or the Euclidean norm
  int nrm(const double* x, double* out) {
    int rc = dots(x, x, nullptr, nullptr, norm_ref && A->n > 1, out);
    *out = std::sqrt(*out);
    return rc;
  }
  void mult(const double* x, double* y) {
    hipLaunchKernelGGL(k_csr_mult, dim3(G), dim3(256), 0, st, A->n, (const int*)A->rp.p,
                       (const int*)A->ci.p, (const double*)A->val.p, x, y);
  }
  void resid(const double* x, const double* b, double* r) {
    hipLaunchKernelGGL(k_csr_resid, dim3(G), dim3(256), 0, st, A->n, (const int*)A->rp.p,
                       (const int*)A->ci.p, (const double*)A->val.p, x, b, r);
  }
  int diapre() {
    CPGPU_HIP(A->ctx, hipMemsetAsync(A->flag.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_csr_diapre, dim3(G), dim3(256), 0, st, A->n, (const int*)A->rp.p,
                       (const int*)A->ci.p, (const double*)A->val.p, A->dinv.p, A->flag.p);
    int f = 0;
    CPGPU_HIP(A->ctx, hipMemcpyAsync(&f, A->flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    CPGPU_HIP(A->ctx, hipStreamSynchronize(st));
    return f ? set_error(A->ctx, CPGPU_EZERODIAG, "DiaPre: zero or missing diagonal entry")
             : CPGPU_OK;
  }
};

// CG / Jacobi-PCG (solver_matrix.hpp:264-374 with DiaPre(A,1))
int csr_cg(CsrRun& R, double* x, const double* b, int* max_iter, double* tol, bool precond) {
  cpgpu_csr* A = R.A;
  const int n = A->n, G = R.G;
  hipStream_t st = R.st;
  double *r = A->vec[0].p, *z = A->vec[1].p, *p = A->vec[2].p, *q = A->vec[3].p;
  int rc;
  if (precond && (rc = R.diapre())) return rc;
  double normb = 0, resid = 0;
  if ((rc = R.nrm(b, &normb))) return rc;
  R.resid(x, b, r);
  if (normb == 0.0) normb = 1;
  if ((rc = R.nrm(r, &resid))) return rc;
  if ((resid /= normb) <= *tol) {
    *tol = resid;
    *max_iter = 0;
    return CPGPU_OK;
  }
  double rho = 0, rho_1 = 0;
  for (int i = 1; i <= *max_iter; ++i) {
    hipLaunchKernelGGL(k_csr_precond, dim3(G), dim3(256), 0, st, n, (const double*)r,
                       precond ? (const double*)A->dinv.p : nullptr, z);
    if ((rc = R.dots(r, z, nullptr, nullptr, 0, &rho))) return rc;
    const double beta = i == 1 ? 0.0 : rho / rho_1;
    hipLaunchKernelGGL(k_csr_dir, dim3(G), dim3(256), 0, st, n, (const double*)z, p, beta,
                       i == 1 ? 1 : 0);
    R.mult(p, q);
    double pq = 0;
    if ((rc = R.dots(p, q, nullptr, nullptr, 0, &pq))) return rc;
    const double alpha = rho / pq;
    hipLaunchKernelGGL(k_csr_cg_upd, dim3(G), dim3(256), 0, st, n, x, r, (const double*)p,
                       (const double*)q, alpha);
    if ((rc = R.nrm(r, &resid))) return rc;
    resid /= normb;
    R.push(resid);
    if (resid <= *tol) {
      *tol = resid;
      *max_iter = i;
      return CPGPU_OK;
    }
    rho_1 = rho;
  }
  *tol = resid;
  return CPGPU_MAXITER;
}

// Dia_BiCGSTAB (solver_matrix.hpp:599-672): strict '<' tests, the half-step
// exit leaves max_iter unchanged, codes 2 (rho = 0) and 3 (omega = 0)
int csr_bicgstab(CsrRun& R, double* x, const double* b, int* max_iter, double* tol) {
  cpgpu_csr* A = R.A;
  const int n = A->n, G = R.G;
  hipStream_t st = R.st;
  double *r = A->vec[0].p, *rt = A->vec[1].p, *p = A->vec[2].p, *ph = A->vec[3].p;
  double *v = A->vec[4].p, *s = A->vec[5].p, *sh = A->vec[6].p, *t = A->vec[7].p;
  int rc;
  if ((rc = R.diapre())) return rc;
  CPGPU_HIP(A->ctx, hipMemsetAsync(p, 0, sizeof(double) * n, st));
  CPGPU_HIP(A->ctx, hipMemsetAsync(v, 0, sizeof(double) * n, st));
  double normb = 0, resid = 0;
  if ((rc = R.nrm(b, &normb))) return rc;
  R.resid(x, b, r);
  CPGPU_HIP(A->ctx, hipMemcpyAsync(rt, r, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  if (normb == 0.0) normb = 1;
  if ((rc = R.nrm(r, &resid))) return rc;
  if ((resid /= normb) <= *tol) {
    *tol = resid;
    *max_iter = 0;
    R.push(resid);
    return CPGPU_OK;
  }
  double rho_1 = 0, rho_2 = 0, alpha = 0, beta = 0, omega = 0;
  for (int i = 1; i <= *max_iter; ++i) {
    if ((rc = R.dots(rt, r, nullptr, nullptr, 0, &rho_1))) return rc;
    if (rho_1 == 0) {
      if ((rc = R.nrm(r, &resid))) return rc;
      *tol = resid / normb;
      return CPGPU_BREAKDOWN_RHO;
    }
    if (i > 1) beta = (rho_1 / rho_2) * (alpha / omega);
    hipLaunchKernelGGL(k_csr_bicg_p, dim3(G), dim3(256), 0, st, n, (const double*)r, p,
                       (const double*)v, (const double*)A->dinv.p, ph, beta, omega,
                       i == 1 ? 1 : 0);
    R.mult(ph, v);
    double rtv = 0;
    if ((rc = R.dots(rt, v, nullptr, nullptr, 0, &rtv))) return rc;
    alpha = rho_1 / rtv;
    hipLaunchKernelGGL(k_csr_bicg_s, dim3(G), dim3(256), 0, st, n, (const double*)r,
                       (const double*)v, s, alpha);
    double ns = 0;
    if ((rc = R.nrm(s, &ns))) return rc;
    if ((resid = ns / normb) < *tol) {
      hipLaunchKernelGGL(k_csr_axpy, dim3(G), dim3(256), 0, st, n, x, (const double*)ph,
                         alpha);
      *tol = resid;
      R.push(resid);
      return CPGPU_OK;  // max_iter unchanged, as the reference
    }
    hipLaunchKernelGGL(k_csr_precond, dim3(G), dim3(256), 0, st, n, (const double*)s,
                       (const double*)A->dinv.p, sh);
    R.mult(sh, t);
    double ts_tt[2];
    if ((rc = R.dots(t, s, t, t, 0, ts_tt))) return rc;
    omega = ts_tt[0] / ts_tt[1];
    hipLaunchKernelGGL(k_csr_bicg_x, dim3(G), dim3(256), 0, st, n, x, r, (const double*)ph,
                       (const double*)sh, (const double*)s, (const double*)t, alpha, omega);
    rho_2 = rho_1;
    if ((rc = R.nrm(r, &resid))) return rc;
    resid /= normb;
    R.push(resid);
    if (resid < *tol) {
      *tol = resid;
      *max_iter = i;
      return CPGPU_OK;
    }
    if (omega == 0) {
      *tol = resid;
      return CPGPU_BREAKDOWN_OMEGA;
    }
  }
  *tol = resid;
  return CPGPU_MAXITER;
}

// Solver_Jacobi_ (solver_matrix.hpp:842-909): on convergence x keeps the
// previous iterate (the converged newx is not copied back)
int csr_jacobi(CsrRun& R, double* x, const double* b, int* max_iter, double* tol) {
  cpgpu_csr* A = R.A;
  const int n = A->n, G = R.G;
  hipStream_t st = R.st;
  double *r = A->vec[0].p, *c = A->vec[1].p, *newx = A->vec[2].p;
  if (A->tval.alloc(A->nnz)) return set_error(A->ctx, CPGPU_ENOMEM, "Jacobi matrix");
  int rc;
  double normb = 0, resid = 0;
  if ((rc = R.nrm(b, &normb))) return rc;
  R.resid(x, b, r);
  if (normb == 0.0) normb = 1;
  if ((rc = R.nrm(r, &resid))) return rc;
  if ((resid /= normb) <= *tol) {
    *tol = resid;
    *max_iter = 0;
    return CPGPU_OK;
  }
  hipLaunchKernelGGL(k_csr_jacobi_setup, dim3(G), dim3(256), 0, st, n, (const int*)A->rp.p,
                     (const int*)A->ci.p, (const double*)A->val.p, b, A->tval.p, c);
  for (int i = 1; i <= *max_iter; ++i) {
    hipLaunchKernelGGL(k_csr_jacobi_sweep, dim3(G), dim3(256), 0, st, n, (const int*)A->rp.p,
                       (const int*)A->ci.p, (const double*)A->tval.p, (const double*)c,
                       (const double*)x, newx);
    R.resid(newx, b, r);
    if ((rc = R.nrm(r, &resid))) return rc;
    resid /= normb;
    R.push(resid);
    if (resid <= *tol) {
      *tol = resid;
      *max_iter = i;
      return CPGPU_OK;
    }
    CPGPU_HIP(A->ctx, hipMemcpyAsync(x, newx, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  }
  *tol = resid;
  return CPGPU_MAXITER;
}

}  // namespace
}  // namespace cpgpu

using namespace cpgpu;

int cpgpu_csr_create(cpgpu_ctx* ctx, int n, const int* row_ptr, const int* col_ind,
                     const double* val, cpgpu_csr** out) {
  if (!ctx || !out || n <= 0 || !row_ptr || !col_ind || !val) return CPGPU_EINVAL;
  *out = nullptr;
  if (row_ptr[0] != 0) return set_error(ctx, CPGPU_EINVAL, "CSR row_ptr[0] must be 0");
  for (int i = 0; i < n; ++i)
    if (row_ptr[i + 1] < row_ptr[i])
      return set_error(ctx, CPGPU_EINVAL, "CSR row_ptr must be non-decreasing");
  const int nnz = row_ptr[n];
  for (int k = 0; k < nnz; ++k)
    if (col_ind[k] < 0 || col_ind[k] >= n)
      return set_error(ctx, CPGPU_EINVAL, "CSR column index %d out of range at entry %d",
                       col_ind[k], k);
  CPGPU_HIP(ctx, hipSetDevice(ctx->device));
  cpgpu_csr* A = new cpgpu_csr;
  A->ctx = ctx;
  A->n = n;
  A->nnz = nnz;
  bool bad = A->rp.alloc(n + 1) || A->ci.alloc(std::max(nnz, 1)) ||
             A->val.alloc(std::max(nnz, 1)) || A->dinv.alloc(n) ||
             A->part.alloc(2 * kCsrG + 2) || A->flag.alloc(1);
  for (auto& v : A->vec) bad = bad || v.alloc(n);
  if (bad) {
    delete A;
    return set_error(ctx, CPGPU_ENOMEM, "CSR matrix of %d rows, %d entries", n, nnz);
  }
  hipStream_t st = ctx->stream;
  if (hipMemcpyAsync(A->rp.p, row_ptr, sizeof(int) * (n + 1), hipMemcpyHostToDevice, st) !=
          hipSuccess ||
      (nnz > 0 &&
       (hipMemcpyAsync(A->ci.p, col_ind, sizeof(int) * nnz, hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(A->val.p, val, sizeof(double) * nnz, hipMemcpyHostToDevice, st) !=
            hipSuccess)) ||
      hipStreamSynchronize(st) != hipSuccess) {
    delete A;
    return set_error(ctx, CPGPU_EHIP, "CSR upload");
  }
  *out = A;
  return CPGPU_OK;
}

int cpgpu_csr_destroy(cpgpu_csr* A) {
  if (!A) return CPGPU_EINVAL;
  (void)hipSetDevice(A->ctx->device);
  (void)hipStreamSynchronize(A->ctx->stream);
  delete A;
  return CPGPU_OK;
}

int cpgpu_csr_mult(cpgpu_csr* A, const double* x, double* y) {
  if (!A || !x || !y) return CPGPU_EINVAL;
  cpgpu_ctx* ctx = A->ctx;
  CPGPU_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int n = A->n;
  CPGPU_HIP(ctx, hipMemcpyAsync(A->vec[8].p, x, sizeof(double) * n, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_csr_mult, dim3(csr_grid(n)), dim3(256), 0, st, n, (const int*)A->rp.p,
                     (const int*)A->ci.p, (const double*)A->val.p, (const double*)A->vec[8].p,
                     A->vec[9].p);
  CPGPU_HIP(ctx, hipGetLastError());
  CPGPU_HIP(ctx, hipMemcpyAsync(y, A->vec[9].p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  CPGPU_HIP(ctx, hipStreamSynchronize(st));
  return CPGPU_OK;
}

int cpgpu_csr_solve(cpgpu_csr* A, int method, double* x, const double* b, int* max_iter,
                    double* tol, int norm_mode, double* hist, int hist_cap, int* hist_len) {
  if (!A || !x || !b || !max_iter || !tol || hist_cap < 0) return CPGPU_EINVAL;
  if (norm_mode != CPGPU_NORM_REF && norm_mode != CPGPU_NORM_TRUE) return CPGPU_EINVAL;
  if (method != CPGPU_CG && method != CPGPU_JPCG && method != CPGPU_BICGSTAB &&
      method != CPGPU_JACOBI)
    return set_error(A->ctx, CPGPU_EUNSUPPORTED,
                     "CSR solve: method %d (CG, JPCG, BICGSTAB and JACOBI run on the device)",
                     method);
  cpgpu_ctx* ctx = A->ctx;
  CPGPU_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int n = A->n;
  double* dx = A->vec[8].p;
  double* db = A->vec[9].p;
  CPGPU_HIP(ctx, hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, st));
  CPGPU_HIP(ctx, hipMemcpyAsync(db, b, sizeof(double) * n, hipMemcpyHostToDevice, st));
  CsrRun R{A, st, csr_grid(n), norm_mode == CPGPU_NORM_REF ? 1 : 0, hist, hist_cap};
  int code;
  switch (method) {
    case CPGPU_CG: code = csr_cg(R, dx, db, max_iter, tol, false); break;
    case CPGPU_JPCG: code = csr_cg(R, dx, db, max_iter, tol, true); break;
    case CPGPU_BICGSTAB: code = csr_bicgstab(R, dx, db, max_iter, tol); break;
    default: code = csr_jacobi(R, dx, db, max_iter, tol); break;
  }
  if (hist_len) *hist_len = R.hl;
  if (code < 0) return code;
  CPGPU_HIP(ctx, hipGetLastError());
  CPGPU_HIP(ctx, hipMemcpyAsync(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  CPGPU_HIP(ctx, hipStreamSynchronize(st));
  return code;
}
