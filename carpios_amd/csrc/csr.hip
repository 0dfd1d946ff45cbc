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

#include "bicg.hpp"
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
  cpgpu::DevBuf<double> dhist;  // device residual history of the device-scalar CG
  void* hst = nullptr;          // pinned host copy of its state
  ~cpgpu_csr() {
    if (hst) (void)hipHostFree(hst);
  }
};

namespace cpgpu {
namespace {

constexpr int kCsrG = 1024;    // blocks of the 1-D kernels
constexpr double kSmall = 1e-10;  // SMALL, lib/type_define.hpp:16

inline int csr_grid(int n) { return std::max(1, std::min(kCsrG, (n + 255) / 256)); }

// sum over k in [k, e) of f(ci[k]) * w[k], left to right as the reference
// adds a row: the entries are taken four at a time so that the four
// gathers are in flight together
template <typename F>
__device__ __forceinline__ double row_sum(int k, int e, const int* __restrict__ ci,
                                          const double* __restrict__ w, F f) {
  double acc = 0.0;
  for (; k + 4 <= e; k += 4) {
    const int c0 = ci[k], c1 = ci[k + 1], c2 = ci[k + 2], c3 = ci[k + 3];
    const double w0 = w[k], w1 = w[k + 1], w2 = w[k + 2], w3 = w[k + 3];
    const double x0 = f(c0), x1 = f(c1), x2 = f(c2), x3 = f(c3);
    acc += x0 * w0;
    acc += x1 * w1;
    acc += x2 * w2;
    acc += x3 * w3;
  }
  for (; k < e; ++k) acc += f(ci[k]) * w[k];
  return acc;
}

// y = A x: row i summed in CSR order from 0 (matrix_SparCompRow.hpp:236-249)
__global__ __launch_bounds__(256) void k_csr_mult(int n, const int* __restrict__ rp,
                                                  const int* __restrict__ ci,
                                                  const double* __restrict__ val,
                                                  const double* __restrict__ x,
                                                  double* __restrict__ y) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    y[i] = row_sum(rp[i], rp[i + 1], ci, val, [&](int c) { return x[c]; });
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
    r[i] = b[i] - row_sum(rp[i], rp[i + 1], ci, val, [&](int c) { return x[c]; });
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
    newx[i] = row_sum(rp[i], rp[i + 1], ci, tv, [&](int j) { return x[j]; }) + c[i];
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

// ---------------------------------------------------------------------
// CG / Jacobi-PCG with its scalars on the device.  An iteration of
// solver_matrix.hpp:264-374 (z = M r; rho = r.z; p = z + (rho/rho_1) p;
// q = A p; alpha = rho/p.q; x += alpha p; r -= alpha q; resid = nrm(r)/|b|)
// runs as five launches that read and write the state below (direction,
// product with p.q, alpha, update with the preconditioner, stopping test);
// iterations after the stop return at entry, so the host launches batches
// and reads the state once per batch.  z and r.z of the next iteration are
// formed in the update kernel from the updated r, the values the next
// iteration's first two steps compute.  The direction is a pass of its own:
// forming p at every column of the product would gather z and p_old per
// matrix entry instead of p alone.
struct CState {
  double normb, tol, resid, rho, rho_1, alpha;
  int it, max_iter, done, hist_cap;
};
static_assert(sizeof(CState) <= 8 * sizeof(double), "the CSR state slot holds 8 doubles");

// p = z (first) or z + beta p
__global__ __launch_bounds__(256) void k_csr_cg_p(int n, const double* __restrict__ z,
                                                  double* __restrict__ p, const CState* st) {
  if (st->done) return;
  const bool first = st->it == 0;
  const double beta = first ? 0.0 : st->rho / st->rho_1;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    p[k] = first ? z[k] : z[k] + beta * p[k];
}

// q = A p; partials of p.q
__global__ __launch_bounds__(256) void k_csr_cg_q(int n, const int* __restrict__ rp,
                                                  const int* __restrict__ ci,
                                                  const double* __restrict__ val,
                                                  const double* __restrict__ p,
                                                  double* __restrict__ q, const CState* st,
                                                  double* part) {
  __shared__ double lds[4];
  if (st->done) return;
  double v[1] = {0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double qi = row_sum(rp[i], rp[i + 1], ci, val, [&](int c) { return p[c]; });
    q[i] = qi;
    v[0] += p[i] * qi;
  }
  store_partials<double, 1>(v, part, gridDim.x, blockIdx.x, lds);
}

__global__ __launch_bounds__(256) void k_csr_alpha(const double* part, int G, CState* st) {
  __shared__ double lds[4];
  if (st->done) return;
  double t[1];
  reduce_partials<double, 1>(part, G, t, lds);
  if (threadIdx.x == 0) st->alpha = st->rho / t[0];
}

// x += alpha p; r -= alpha q; z = M r; partials of r.r (the reference nrm2
// leaves the last element out when skip_last) and r.z
__global__ __launch_bounds__(256) void k_csr_updz(int n, double* __restrict__ x,
                                                  double* __restrict__ r,
                                                  const double* __restrict__ p,
                                                  const double* __restrict__ q,
                                                  const double* __restrict__ d,
                                                  double* __restrict__ z, const CState* st,
                                                  int skip_last, double* part) {
  __shared__ double lds[8];
  if (st->done) return;
  const double alpha = st->alpha;
  double v[2] = {0.0, 0.0};
  const int m = skip_last ? n - 1 : n;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    x[k] += alpha * p[k];
    const double rk = r[k] - alpha * q[k];
    r[k] = rk;
    const double zk = d ? rk * d[k] : rk;
    z[k] = zk;
    if (k < m) v[0] += rk * rk;
    v[1] += rk * zk;
  }
  store_partials<double, 2>(v, part, gridDim.x, blockIdx.x, lds);
}

// the stopping test: resid = sqrt(r.r)/|b| into the history; converged,
// or max_iter reached; rho_1 <- rho, rho <- r.z of the next iteration
__global__ __launch_bounds__(256) void k_csr_test(const double* part, int G, CState* st,
                                                  double* hist) {
  __shared__ double lds[8];
  if (st->done) return;
  double t[2];
  reduce_partials<double, 2>(part, G, t, lds);
  if (threadIdx.x == 0) {
    const double resid = std::sqrt(t[0]) / st->normb;
    if (hist && st->it < st->hist_cap) hist[st->it] = resid;
    st->it += 1;
    st->resid = resid;
    if (resid <= st->tol) st->done = 1;
    else if (st->it >= st->max_iter) st->done = 2;
    st->rho_1 = st->rho;
    st->rho = t[1];
  }
}

// the first z = M r and rho = r.z
__global__ __launch_bounds__(256) void k_csr_pz(int n, const double* __restrict__ r,
                                                const double* __restrict__ d,
                                                double* __restrict__ z, double* part) {
  __shared__ double lds[4];
  double v[1] = {0.0};
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const double zk = d ? r[k] * d[k] : r[k];
    z[k] = zk;
    v[0] += r[k] * zk;
  }
  store_partials<double, 1>(v, part, gridDim.x, blockIdx.x, lds);
}

__global__ __launch_bounds__(256) void k_csr_rho0(const double* part, int G, CState* st) {
  __shared__ double lds[4];
  double t[1];
  reduce_partials<double, 1>(part, G, t, lds);
  if (threadIdx.x == 0) st->rho = t[0];
}

// Dia_BiCGSTAB's operator kernels on the CSR matrix (the scalar kernels,
// the update kernel and the batch loop are shared with the structured
// operator).
// p = r (first) or r + beta (p - omega v); ph = p D^-1
__global__ __launch_bounds__(256) void k_csr_bicg_p(int n, const double* __restrict__ r,
                                                    double* __restrict__ p,
                                                    const double* __restrict__ v,
                                                    const double* __restrict__ d,
                                                    double* __restrict__ ph, const BState* st) {
  if (st->stop) return;
  const bool first = st->it == 0;
  const double beta = st->beta, omega = st->omega;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const double pv = first ? r[k] : r[k] + beta * (p[k] - omega * v[k]);
    p[k] = pv;
    ph[k] = pv * d[k];
  }
}

// v = A ph; partials of rt.v
__global__ __launch_bounds__(256) void k_csr_bicg_v(int n, const int* __restrict__ rp,
                                                    const int* __restrict__ ci,
                                                    const double* __restrict__ val,
                                                    const double* __restrict__ ph,
                                                    const double* __restrict__ rt,
                                                    double* __restrict__ v, const BState* st,
                                                    double* part) {
  __shared__ double lds[4];
  if (st->stop) return;
  double acc[1] = {0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double vi = row_sum(rp[i], rp[i + 1], ci, val, [&](int c) { return ph[c]; });
    v[i] = vi;
    acc[0] += rt[i] * vi;
  }
  store_partials<double, 1>(acc, part, gridDim.x, blockIdx.x, lds);
}

// s = r - alpha v; sh = s D^-1
__global__ __launch_bounds__(256) void k_csr_bicg_s(int n, const double* __restrict__ r,
                                                    const double* __restrict__ v,
                                                    const double* __restrict__ d,
                                                    double* __restrict__ sv,
                                                    double* __restrict__ sh, const BState* st) {
  if (st->stop) return;
  const double alpha = st->alpha;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const double si = r[k] - alpha * v[k];
    sv[k] = si;
    sh[k] = si * d[k];
  }
}

// t = A sh; partials of s.s (without the last unknown when skip_last), t.s,
// t.t
__global__ __launch_bounds__(256) void k_csr_bicg_t(int n, const int* __restrict__ rp,
                                                    const int* __restrict__ ci,
                                                    const double* __restrict__ val,
                                                    const double* __restrict__ sh,
                                                    const double* __restrict__ sv,
                                                    double* __restrict__ tv, const BState* st,
                                                    int skip_last, double* part) {
  __shared__ double lds[12];
  if (st->stop) return;
  double v[3] = {0.0, 0.0, 0.0};
  const int m = skip_last ? n - 1 : n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double ti = row_sum(rp[i], rp[i + 1], ci, val, [&](int c) { return sh[c]; });
    const double si = sv[i];
    tv[i] = ti;
    if (i < m) v[0] += si * si;
    v[1] += ti * si;
    v[2] += ti * ti;
  }
  store_partials<double, 3>(v, part, gridDim.x, blockIdx.x, lds);
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

// CG / Jacobi-PCG (solver_matrix.hpp:264-374 with DiaPre(A,1)): the initial
// residual test on the host, then batches of kBatch device-scalar
// iterations, the state read back once per batch
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
  if (*max_iter < 1) {
    *tol = resid;
    return CPGPU_MAXITER;
  }
  if (!A->hst && hipHostMalloc(&A->hst, sizeof(CState), hipHostMallocDefault) != hipSuccess) {
    A->hst = nullptr;
    return set_error(A->ctx, CPGPU_ENOMEM, "pinned CSR solver state");
  }
  const int cap = R.hist ? R.hist_cap : 0;
  if (cap > 0 && A->dhist.alloc(cap)) return set_error(A->ctx, CPGPU_ENOMEM, "CSR history");
  // the state sits after the partials (2 kCsrG) and the totals (2)
  CState* dst = reinterpret_cast<CState*>(A->part.p + 2 * kCsrG + 8);
  CState* hs = static_cast<CState*>(A->hst);
  CState init{};
  init.normb = normb;
  init.tol = *tol;
  init.resid = resid;
  init.max_iter = *max_iter;
  init.hist_cap = cap;
  *hs = init;
  CPGPU_HIP(A->ctx, hipMemcpyAsync(dst, hs, sizeof(CState), hipMemcpyHostToDevice, st));
  const double* dd = precond ? (const double*)A->dinv.p : nullptr;
  double* part = A->part.p;
  double* dh = cap > 0 ? A->dhist.p : nullptr;
  hipLaunchKernelGGL(k_csr_pz, dim3(G), dim3(256), 0, st, n, (const double*)r, dd, z, part);
  hipLaunchKernelGGL(k_csr_rho0, dim3(1), dim3(256), 0, st, (const double*)part, G, dst);
  constexpr int kBatch = 16;
  const int skip = R.norm_ref && n > 1 ? 1 : 0;
  for (;;) {
    for (int k = 0; k < kBatch; ++k) {
      hipLaunchKernelGGL(k_csr_cg_p, dim3(G), dim3(256), 0, st, n, (const double*)z, p,
                         (const CState*)dst);
      hipLaunchKernelGGL(k_csr_cg_q, dim3(G), dim3(256), 0, st, n, (const int*)A->rp.p,
                         (const int*)A->ci.p, (const double*)A->val.p, (const double*)p, q,
                         (const CState*)dst, part);
      hipLaunchKernelGGL(k_csr_alpha, dim3(1), dim3(256), 0, st, (const double*)part, G, dst);
      hipLaunchKernelGGL(k_csr_updz, dim3(G), dim3(256), 0, st, n, x, r, (const double*)p,
                         (const double*)q, dd, z, (const CState*)dst, skip, part);
      hipLaunchKernelGGL(k_csr_test, dim3(1), dim3(256), 0, st, (const double*)part, G, dst,
                         dh);
    }
    CPGPU_HIP(A->ctx, hipGetLastError());
    CPGPU_HIP(A->ctx, hipMemcpyAsync(hs, dst, sizeof(CState), hipMemcpyDeviceToHost, st));
    CPGPU_HIP(A->ctx, hipStreamSynchronize(st));
    if (hs->done) break;
  }
  const CState f = *hs;
  if (cap > 0) {
    const int nh = std::min(f.it, cap);
    CPGPU_HIP(A->ctx, hipMemcpyAsync(R.hist, dh, sizeof(double) * nh, hipMemcpyDeviceToHost, st));
    CPGPU_HIP(A->ctx, hipStreamSynchronize(st));
  }
  R.hl = f.it;
  *tol = f.resid;
  if (f.done == 1) {
    *max_iter = f.it;
    return CPGPU_OK;
  }
  return CPGPU_MAXITER;
}

// Dia_BiCGSTAB (solver_matrix.hpp:599-672): strict '<' tests, the half-step
// exit leaves max_iter unchanged, codes 2 (rho = 0) and 3 (omega = 0); the
// initial residual test on the host, then the device-scalar iterations
struct CsrBicgOps {
  CsrRun* R;
  double* x;
  double *r, *rt, *p, *v, *s, *t, *ph, *sh;
  const BState* state;
  double* part;
  int G, skip;
  void rho0() {
    hipLaunchKernelGGL(k_csr_dots<1>, dim3(G), dim3(256), 0, R->st, (const double*)rt,
                       (const double*)r, nullptr, nullptr, R->A->n, 0, part);
  }
  void dir(int, int) {
    cpgpu_csr* A = R->A;
    hipLaunchKernelGGL(k_csr_bicg_p, dim3(G), dim3(256), 0, R->st, A->n, (const double*)r, p,
                       (const double*)v, (const double*)A->dinv.p, ph, state);
    hipLaunchKernelGGL(k_csr_bicg_v, dim3(G), dim3(256), 0, R->st, A->n, (const int*)A->rp.p,
                       (const int*)A->ci.p, (const double*)A->val.p, (const double*)ph,
                       (const double*)rt, v, state, part);
  }
  void st(int) {
    cpgpu_csr* A = R->A;
    hipLaunchKernelGGL(k_csr_bicg_s, dim3(G), dim3(256), 0, R->st, A->n, (const double*)r,
                       (const double*)v, (const double*)A->dinv.p, s, sh, state);
    hipLaunchKernelGGL(k_csr_bicg_t, dim3(G), dim3(256), 0, R->st, A->n, (const int*)A->rp.p,
                       (const int*)A->ci.p, (const double*)A->val.p, (const double*)sh,
                       (const double*)s, t, state, skip, part);
  }
  void xr(int) {
    cpgpu_csr* A = R->A;
    hipLaunchKernelGGL(k_bicg_xr, dim3(G), dim3(256), 0, R->st, (size_t)A->n,
                       (const double*)A->dinv.p, x, r, (const double*)rt, (const double*)p,
                       (const double*)s, (const double*)t, state, skip, part);
  }
};

int csr_bicgstab(CsrRun& R, double* x, const double* b, int* max_iter, double* tol) {
  cpgpu_csr* A = R.A;
  const int n = A->n;
  hipStream_t st = R.st;
  double *r = A->vec[0].p, *rt = A->vec[1].p;
  int rc;
  if ((rc = R.diapre())) return rc;
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
  if (*max_iter < 1) {
    *tol = resid;
    return CPGPU_MAXITER;
  }
  const int cap = R.hist ? R.hist_cap : 0;
  if (cap > 0 && A->dhist.alloc(cap)) return set_error(A->ctx, CPGPU_ENOMEM, "CSR history");
  BState* dst = reinterpret_cast<BState*>(A->part.p + 3 * kCsrG + 8);
  CsrBicgOps ops{&R, x, r, rt, A->vec[2].p, A->vec[3].p, A->vec[4].p, A->vec[5].p,
                 A->vec[6].p, A->vec[7].p, dst, A->part.p, R.G, R.norm_ref && n > 1 ? 1 : 0};
  int hl = 0;
  const int code = bicg_run(A->ctx, st, ops, dst, normb, resid, max_iter, tol,
                            cap > 0 ? A->dhist.p : nullptr, cap, R.hist, &hl);
  R.hl = hl;
  return code;
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
             A->part.alloc(3 * kCsrG + 24) || A->flag.alloc(1);
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
