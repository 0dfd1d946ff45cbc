// Matrix-free 5-point Poisson operator and Krylov solvers for gfx950.
//
// The operator is structure::StencilPoisson_<2> with GhostRegular_<2>
// substitution (reference lib/structure/s_poisson.hpp:60-181,
// s_ghost.hpp:140-357) applied without assembling CSR: per-face
// coefficients K = beta*a/d are 1-D arrays (the grid is uniform along the
// other axis), the diagonal is rebuilt per cell in the reference's order of
// accumulation, and boundary rows fold the ghost terms exactly as the
// reference's substitution does.  Vectors are interior-only nx*ny arrays,
// j contiguous (the interior of a structure::Scalar_<2> buffer).
//
// FP64 arithmetic is not contracted into FMAs (fp contract off) and the
// SpMV adds the five terms in the reference's CSR order
// (matrix_SparCompRow.hpp:236-249), so A*x is bitwise the reference's.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include "core.hpp"

#pragma clang fp contract(off)

namespace cpgpu {

// ---------------------------------------------------------------------
// operator description
template <typename T>
struct Sten {
  int nx, ny;
  const T* KX;  // nx+1 face coefficients along x: face f between cells f-1, f
  const T* KY;  // ny+1
  T adj[4];     // diagonal adjustment per boundary segment x-, x+, y-, y+
};

// Centre coefficient: coe(index,0,_C_) (s_poisson.hpp:64-77) then the ghost
// substitutions in Index_compare_ order x-, y-, y+, x+ (s_ghost.hpp:334-357).
template <typename T>
__device__ __forceinline__ T diag_at(const Sten<T>& s, int i, int j) {
  T d = s.KX[i] + s.KX[i + 1];
  d = d + s.KY[j];
  d = d + s.KY[j + 1];
  if (i == 0) d = d + s.adj[0];
  if (j == 0) d = d + s.adj[2];
  if (j == s.ny - 1) d = d + s.adj[3];
  if (i == s.nx - 1) d = d + s.adj[1];
  return d;
}

// Device state of one Krylov solve.  Scalars are kept in double; for an FP32
// solve they hold FP32 values exactly.
struct KState {
  double normb, resid, tol;
  double rho, rho_old, pq, alpha, omega;
  double rr;
  int iter, max_iter, done, code;
  unsigned int ticket;
  int pad;
};

constexpr int TJ = 64;  // tile width (j, contiguous) = one wavefront
constexpr int TI = 16;  // tile height (i)
constexpr int BY = 4;   // thread rows per block; each thread does TI/BY rows

inline dim3 tile_grid(int nx, int ny) {
  return dim3((ny + TJ - 1) / TJ, (nx + TI - 1) / TI);
}

template <typename T>
__device__ __forceinline__ T apply_global(const Sten<T>& s, const T* __restrict__ x,
                                          int i, int j, T dg) {
  const int ny = s.ny;
  const size_t k = (size_t)i * ny + j;
  T acc = T(0);
  if (i > 0) acc = acc + x[k - ny] * (-s.KX[i]);
  if (j > 0) acc = acc + x[k - 1] * (-s.KY[j]);
  acc = acc + x[k] * dg;
  if (j < ny - 1) acc = acc + x[k + 1] * (-s.KY[j + 1]);
  if (i < s.nx - 1) acc = acc + x[k + ny] * (-s.KX[i + 1]);
  return acc;
}

// ---------------------------------------------------------------------
// kernels

// y = A x
template <typename T>
__global__ __launch_bounds__(256) void k_apply(Sten<T> s, const T* __restrict__ x,
                                               T* __restrict__ y) {
  const int j = blockIdx.x * TJ + threadIdx.x;
  for (int k = 0; k < TI / BY; ++k) {
    const int i = blockIdx.y * TI + threadIdx.y + BY * k;
    if (i < s.nx && j < s.ny)
      y[(size_t)i * s.ny + j] = apply_global(s, x, i, j, diag_at(s, i, j));
  }
}

// b = -(vol*src + ghost constants)   (s_poisson.hpp:79-97, s_ghost.hpp:161-175,
// s_solver.hpp:122)
struct RhsSeg {
  int type;       // 1 Dirichlet, 2 Neumann
  double coe;     // coefficient of the ghost term (negative)
  double dl, d1;  // distances of the substitution
};
struct RhsDesc {
  int nx, ny;
  const double* vol;    // nx: (1.0*sx)*sx
  const double* bcval;  // packed interior lines: x- ny | x+ ny | y- nx | y+ nx
  RhsSeg seg[4];
};

__device__ __forceinline__ double ghost_const(const RhsSeg& g, double cst,
                                              double val) {
  if (g.type == CPGPU_BC_DIRICHLET) return cst + (g.coe * g.dl * val / g.d1);
  return cst + (-val * g.dl * g.coe);
}

__global__ __launch_bounds__(256) void k_build_rhs(RhsDesc d,
                                                   const double* __restrict__ src,
                                                   double src_uniform,
                                                   double* __restrict__ b) {
  const int j = blockIdx.x * TJ + threadIdx.x;
  for (int k = 0; k < TI / BY; ++k) {
    const int i = blockIdx.y * TI + threadIdx.y + BY * k;
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

// r = b - A x; z = M^-1 r; partial sums of b.b, r.r (nrm2 modes) and r.z.
// Finalises normb, resid, rho and the convergence flag (solver_matrix.hpp:
// 277-286 / 329-338).
template <typename T, bool PC>
__global__ __launch_bounds__(256) void k_cg_init(Sten<T> s, const T* __restrict__ x,
                                                 const T* __restrict__ b,
                                                 T* __restrict__ r, KState* st,
                                                 T* partials, int norm_ref) {
  const int j = blockIdx.x * TJ + threadIdx.x;
  const size_t nlast = (size_t)s.nx * s.ny - 1;
  T v[3] = {T(0), T(0), T(0)};
  for (int k = 0; k < TI / BY; ++k) {
    const int i = blockIdx.y * TI + threadIdx.y + BY * k;
    if (i >= s.nx || j >= s.ny) continue;
    const size_t idx = (size_t)i * s.ny + j;
    const T dg = diag_at(s, i, j);
    const T bv = b[idx];
    const T rv = bv - apply_global(s, x, i, j, dg);
    r[idx] = rv;
    const T z = PC ? rv * (T(1) / dg) : rv;
    if (!(norm_ref && idx == nlast)) {
      v[0] += bv * bv;
      v[1] += rv * rv;
    }
    v[2] += rv * z;
  }
  T tot[3];
  const int nb = gridDim.x * gridDim.y;
  if (grid_sum<T, 3>(v, partials, &st->ticket, nb, tot) && threadIdx.x == 0 &&
      threadIdx.y == 0) {
    double normb = std::sqrt((double)tot[0]);
    if (normb == 0.0) normb = 1;
    st->normb = normb;
    st->resid = std::sqrt((double)tot[1]) / normb;
    st->rho = (double)tot[2];
    st->rho_old = 0.0;
    st->iter = 0;
    st->done = st->resid <= st->tol ? 1 : 0;
    st->code = st->done ? 0 : 1;
  }
}

// Search direction and operator in one pass: p = z + beta p_old over a
// (TI+2)x(TJ+2) LDS tile (halo recomputed from r and p_old), then q = A p
// from LDS and the partial p.q.  p is double-buffered because neighbouring
// tiles read p_old in their halos.
template <typename T, bool PC>
__global__ __launch_bounds__(256) void k_cg_dir(Sten<T> s, const T* __restrict__ r,
                                                const T* __restrict__ pold,
                                                T* __restrict__ pnew,
                                                T* __restrict__ q, KState* st,
                                                T* partials) {
  if (st->done || st->iter >= st->max_iter) return;
  __shared__ T tile[TI + 2][TJ + 2];
  const bool first = st->iter == 0;
  const T beta = first ? T(0) : (T)st->rho / (T)st->rho_old;
  const int i0 = blockIdx.y * TI, j0 = blockIdx.x * TJ;
  const int tid = threadIdx.x + TJ * threadIdx.y;
  const int nx = s.nx, ny = s.ny;
  for (int t = tid; t < (TI + 2) * (TJ + 2); t += TJ * BY) {
    const int ii = t / (TJ + 2), jj = t - ii * (TJ + 2);
    const int gi = i0 - 1 + ii, gj = j0 - 1 + jj;
    T pv = T(0);
    if (gi >= 0 && gi < nx && gj >= 0 && gj < ny) {
      const size_t k = (size_t)gi * ny + gj;
      const T rv = r[k];
      const T z = PC ? rv * (T(1) / diag_at(s, gi, gj)) : rv;
      pv = first ? z : z + beta * pold[k];
      if (ii >= 1 && ii <= TI && jj >= 1 && jj <= TJ) pnew[k] = pv;
    }
    tile[ii][jj] = pv;
  }
  __syncthreads();
  T v[1] = {T(0)};
  const int lj = threadIdx.x + 1, j = j0 + threadIdx.x;
  for (int k = 0; k < TI / BY; ++k) {
    const int li = threadIdx.y + BY * k + 1, i = i0 + li - 1;
    if (i >= nx || j >= ny) continue;
    const T dg = diag_at(s, i, j);
    T acc = T(0);
    if (i > 0) acc = acc + tile[li - 1][lj] * (-s.KX[i]);
    if (j > 0) acc = acc + tile[li][lj - 1] * (-s.KY[j]);
    acc = acc + tile[li][lj] * dg;
    if (j < ny - 1) acc = acc + tile[li][lj + 1] * (-s.KY[j + 1]);
    if (i < nx - 1) acc = acc + tile[li + 1][lj] * (-s.KX[i + 1]);
    q[(size_t)i * ny + j] = acc;
    v[0] += tile[li][lj] * acc;
  }
  T tot[1];
  const int nb = gridDim.x * gridDim.y;
  if (grid_sum<T, 1>(v, partials, &st->ticket, nb, tot) && threadIdx.x == 0 &&
      threadIdx.y == 0)
    st->pq = (double)tot[0];
}

// x += alpha p; r -= alpha q; partial r.z and r.r; finalises the iteration:
// residual history, rho shift and the stopping test (solver_matrix.hpp:
// 356-372).
template <typename T, bool PC>
__global__ __launch_bounds__(256) void k_cg_upd(Sten<T> s, T* __restrict__ x,
                                                const T* __restrict__ p,
                                                T* __restrict__ r,
                                                const T* __restrict__ q,
                                                KState* st, T* partials,
                                                double* hist, int hist_cap,
                                                int norm_ref) {
  if (st->done || st->iter >= st->max_iter) return;
  const T alpha = (T)st->rho / (T)st->pq;
  const int j = blockIdx.x * TJ + threadIdx.x;
  const size_t nlast = (size_t)s.nx * s.ny - 1;
  T v[2] = {T(0), T(0)};
  for (int k = 0; k < TI / BY; ++k) {
    const int i = blockIdx.y * TI + threadIdx.y + BY * k;
    if (i >= s.nx || j >= s.ny) continue;
    const size_t idx = (size_t)i * s.ny + j;
    x[idx] = x[idx] + alpha * p[idx];
    const T rv = r[idx] - alpha * q[idx];
    r[idx] = rv;
    const T z = PC ? rv * (T(1) / diag_at(s, i, j)) : rv;
    v[0] += rv * z;
    if (!(norm_ref && idx == nlast)) v[1] += rv * rv;
  }
  T tot[2];
  const int nb = gridDim.x * gridDim.y;
  if (grid_sum<T, 2>(v, partials, &st->ticket, nb, tot) && threadIdx.x == 0 &&
      threadIdx.y == 0) {
    const int it = st->iter + 1;
    const double resid = std::sqrt((double)tot[1]) / st->normb;
    st->rho_old = st->rho;
    st->rho = (double)tot[0];
    st->resid = resid;
    st->iter = it;
    if (hist && it - 1 < hist_cap) hist[it - 1] = resid;
    if (resid <= st->tol) {
      st->done = 1;
      st->code = 0;
    }
  }
}

// ---- helpers for refinement: FP64 residual with true norms and sums ----
struct RState {
  double bb, rr, rsum, xsum;
  unsigned int ticket;
  int pad;
};

// r = b - A x (FP64, reference order); partial b.b, r.r, sum(r)
__global__ __launch_bounds__(256) void k_residual64(Sten<double> s,
                                                    const double* __restrict__ x,
                                                    const double* __restrict__ b,
                                                    double* __restrict__ r,
                                                    RState* st, double* partials) {
  const int j = blockIdx.x * TJ + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < TI / BY; ++k) {
    const int i = blockIdx.y * TI + threadIdx.y + BY * k;
    if (i >= s.nx || j >= s.ny) continue;
    const size_t idx = (size_t)i * s.ny + j;
    const double bv = b[idx];
    const double rv = bv - apply_global(s, x, i, j, diag_at(s, i, j));
    r[idx] = rv;
    v[0] += bv * bv;
    v[1] += rv * rv;
    v[2] += rv;
  }
  double tot[3];
  const int nb = gridDim.x * gridDim.y;
  if (grid_sum<double, 3>(v, partials, &st->ticket, nb, tot) &&
      threadIdx.x == 0 && threadIdx.y == 0) {
    st->bb = tot[0];
    st->rr = tot[1];
    st->rsum = tot[2];
  }
}

// out32 = (float)(r - shift*mean(r)); mean from the residual sums
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

// x += (double)e, and the sum of the new x for the mean projection
__global__ __launch_bounds__(256) void k_add_f32(double* __restrict__ x,
                                                 const float* __restrict__ e,
                                                 RState* st, double* partials,
                                                 size_t n) {
  double v[1] = {0.0};
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const double nv = x[k] + (double)e[k];
    x[k] = nv;
    v[0] += nv;
  }
  double tot[1];
  if (grid_sum<double, 1>(v, partials, &st->ticket, gridDim.x, tot) &&
      threadIdx.x == 0)
    st->xsum = tot[0];
}

__global__ __launch_bounds__(256) void k_sub_mean(double* __restrict__ x,
                                                  const RState* st, size_t n) {
  const double mean = st->xsum / (double)n;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    x[k] = x[k] - mean;
}

// ---- synthetic right-hand sides (no reference counterpart) ----
// Sum of Gaussian blobs f = sum_k A_k exp(-|x-c_k|^2 / (2 w_k^2)) at cell
// centres, with the partial sum of f for the zero-mean projection.
__global__ __launch_bounds__(256) void k_blobs(int nx, int ny,
                                               const double* __restrict__ cx,
                                               const double* __restrict__ cy,
                                               const double* __restrict__ prm,
                                               int nblobs, double* __restrict__ f,
                                               RState* st, double* partials) {
  const int j = blockIdx.x * TJ + threadIdx.x;
  double v[1] = {0.0};
  for (int k = 0; k < TI / BY; ++k) {
    const int i = blockIdx.y * TI + threadIdx.y + BY * k;
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
  double tot[1];
  const int nb = gridDim.x * gridDim.y;
  if (grid_sum<double, 1>(v, partials, &st->ticket, nb, tot) &&
      threadIdx.x == 0 && threadIdx.y == 0)
    st->rsum = tot[0];
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T* __restrict__ x, T v, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    x[k] = v;
}

__global__ __launch_bounds__(256) void k_sub_rsum(double* __restrict__ f,
                                                  const RState* st, size_t n) {
  const double mean = st->rsum / (double)n;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    f[k] = f[k] - mean;
}

// SplitMix64-seeded xoshiro256** (Blackman & Vigna); doubles from the top
// 53 bits; normals by Box-Muller.  Per blob: cx, cy ~ U[0,1), w ~ U[wmin,
// wmax), A ~ N(0,1).
struct Xoshiro {
  unsigned long long s[4];
  static unsigned long long splitmix(unsigned long long& x) {
    unsigned long long z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  explicit Xoshiro(unsigned long long seed) {
    for (int k = 0; k < 4; ++k) s[k] = splitmix(seed);
  }
  static unsigned long long rotl(unsigned long long x, int k) {
    return (x << k) | (x >> (64 - k));
  }
  unsigned long long next() {
    const unsigned long long r = rotl(s[1] * 5, 7) * 9;
    const unsigned long long t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  double normal() {
    const double u1 = 1.0 - uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

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

}  // namespace cpgpu

using namespace cpgpu;

// ---------------------------------------------------------------------
// problem object

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  int alloc(size_t count) {
    if (n >= count && p) return 0;
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
    if (hipMalloc(&p, sizeof(T) * (count ? count : 1)) != hipSuccess) return -1;
    n = count;
    return 0;
  }
  ~DevBuf() {
    if (p) hipFree(p);
  }
};

struct cpgpu_poisson {
  cpgpu_ctx* ctx = nullptr;
  int nx = 0, ny = 0, gl = 0;
  size_t N = 0;
  double beta = 1.0;
  // host copies of the descriptor
  std::vector<double> KX, KY, vol, bcv;
  std::vector<double> cxi, cyi;  // interior cell centres
  double adj[4] = {0, 0, 0, 0};
  RhsSeg seg[4];
  // device operator
  DevBuf<double> dKX, dKY, dvol, dbcv;
  DevBuf<float> fKX, fKY;
  // FP64 vectors
  DevBuf<double> x, b, r, p0, p1, q, tmp, src;
  // FP32 vectors of the inner solve
  DevBuf<float> b32, x32, r32, pa32, pb32, q32;
  // reductions and states
  DevBuf<double> part64;
  DevBuf<float> part32;
  DevBuf<KState> kst;
  DevBuf<RState> rst;
  DevBuf<double> dhist;
  KState* h_kst = nullptr;  // pinned mirrors
  RState* h_rst = nullptr;
  double inner_tol = 1e-4;
  int max_outer = 40;

  Sten<double> sten64() const {
    Sten<double> s;
    s.nx = nx;
    s.ny = ny;
    s.KX = dKX.p;
    s.KY = dKY.p;
    for (int k = 0; k < 4; ++k) s.adj[k] = adj[k];
    return s;
  }
  Sten<float> sten32() const {
    Sten<float> s;
    s.nx = nx;
    s.ny = ny;
    s.KX = fKX.p;
    s.KY = fKY.p;
    for (int k = 0; k < 4; ++k) s.adj[k] = (float)adj[k];
    return s;
  }
  ~cpgpu_poisson() {
    if (h_kst) hipHostFree(h_kst);
    if (h_rst) hipHostFree(h_rst);
  }
};

namespace {

bool all_equal(const double* a, int n) {
  for (int k = 1; k < n; ++k)
    if (std::memcmp(&a[k], &a[0], sizeof(double)) != 0) return false;
  return true;
}

// Host side of the descriptor: face coefficients K = -(-beta*a/d)
// (s_poisson.hpp:98-135), diagonal adjustments and rhs constants of the
// ghost substitution (s_ghost.hpp:140-179).
int build_descriptor(cpgpu_poisson* P, const cpgpu_grid2d* g) {
  const int nx = g->nx, ny = g->ny, gl = g->gl;
  auto cx = [&](int i) { return g->cx[i + gl]; };
  auto sx = [&](int i) { return g->sx[i + gl]; };
  auto cy = [&](int j) { return g->cy[j + gl]; };
  auto sy = [&](int j) { return g->sy[j + gl]; };
  // the kernels use 1-D face coefficients: a = sy[j] (x faces) and sx[i]
  // (y faces) must not vary along the face's own line
  if (!all_equal(g->sy + gl, ny) || !all_equal(g->sx + gl, nx))
    return set_error(P->ctx, CPGPU_EUNSUPPORTED,
                     "non-uniform cell sizes are not supported");
  const double beta = g->beta;
  P->KX.assign(nx + 1, 0.0);
  P->KY.assign(ny + 1, 0.0);
  for (int f = 0; f <= nx; ++f) {
    const double a = sy(0) * 1.0;
    const double d = std::fabs(cx(f) - cx(f - 1));
    const double coe = -beta * a / d;
    P->KX[f] = -coe;
  }
  for (int f = 0; f <= ny; ++f) {
    const double a = 1.0 * sx(0);
    const double d = std::fabs(cy(f) - cy(f - 1));
    const double coe = -beta * a / d;
    P->KY[f] = -coe;
  }
  // segments: 0 x- (ghost i=-1), 1 x+ (i=nx), 2 y- (j=-1), 3 y+ (j=ny)
  for (int s = 0; s < 4; ++s) {
    const int type = g->bc_type[s];
    if (type != CPGPU_BC_DIRICHLET && type != CPGPU_BC_NEUMANN)
      return set_error(P->ctx, CPGPU_EINVAL, "bad bc type %d on segment %d",
                       type, s);
    double c0, c1, cf, coe;
    if (s == 0) {
      c0 = cx(-1); c1 = cx(0); cf = c1 - sx(0) * 0.5; coe = -P->KX[0];
    } else if (s == 1) {
      c0 = cx(nx); c1 = cx(nx - 1); cf = c1 + sx(nx - 1) * 0.5; coe = -P->KX[nx];
    } else if (s == 2) {
      c0 = cy(-1); c1 = cy(0); cf = c1 - sy(0) * 0.5; coe = -P->KY[0];
    } else {
      c0 = cy(ny); c1 = cy(ny - 1); cf = c1 + sy(ny - 1) * 0.5; coe = -P->KY[ny];
    }
    const double dl = std::fabs(c1 - c0);
    const double d0 = std::fabs(cf - c0);
    const double d1 = std::fabs(c1 - cf);
    P->adj[s] = type == CPGPU_BC_DIRICHLET ? (-coe * d0 / d1) : coe;
    P->seg[s].type = type;
    P->seg[s].coe = coe;
    P->seg[s].dl = dl;
    P->seg[s].d1 = d1;
  }
  P->cxi.assign(g->cx + gl, g->cx + gl + nx);
  P->cyi.assign(g->cy + gl, g->cy + gl + ny);
  P->vol.assign(nx, 0.0);
  for (int i = 0; i < nx; ++i) {
    double v = 1.0;
    for (int d = 0; d < 2; ++d) v *= sx(i);  // s_poisson.hpp:71-77 uses dim 0
    P->vol[i] = v;
  }
  // boundary values of the interior lines
  const int NX = nx + 2 * gl, NY = ny + 2 * gl;
  P->bcv.assign(2 * ny + 2 * nx, 0.0);
  if (g->bc_val) {
    for (int j = 0; j < ny; ++j) {
      P->bcv[j] = g->bc_val[j + gl];
      P->bcv[ny + j] = g->bc_val[NY + j + gl];
    }
    for (int i = 0; i < nx; ++i) {
      P->bcv[2 * ny + i] = g->bc_val[2 * NY + i + gl];
      P->bcv[2 * ny + nx + i] = g->bc_val[2 * NY + NX + i + gl];
    }
  }
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

int ensure_hist(cpgpu_poisson* P, int cap) {
  if (cap > 0 && P->dhist.alloc(cap))
    return set_error(P->ctx, CPGPU_ENOMEM, "hist alloc");
  return CPGPU_OK;
}

// Vectors of one CG run in precision T.
template <typename T>
struct CGVecs {
  T *x, *b, *r, *pa, *pb, *q, *part;
};

// Jacobi-PCG (PC) or CG on device vectors.  Kernels launched in batches;
// the device flag stops the iteration exactly where the reference stops.
template <typename T, bool PC>
int run_cg(cpgpu_poisson* P, const Sten<T>& s, CGVecs<T> v, int max_iter,
           double tol, int norm_ref, double* dhist, int hist_cap,
           KState* out) {
  cpgpu_ctx* ctx = P->ctx;
  hipStream_t st = ctx->stream;
  KState init{};
  init.tol = tol;
  init.max_iter = max_iter;
  init.code = 1;
  std::memcpy(P->h_kst, &init, sizeof(KState));
  CPGPU_HIP(ctx, hipMemcpyAsync(P->kst.p, P->h_kst, sizeof(KState),
                                hipMemcpyHostToDevice, st));
  const dim3 grid = tile_grid(P->nx, P->ny), block(TJ, BY);
  hipLaunchKernelGGL((k_cg_init<T, PC>), grid, block, 0, st, s, v.x, v.b, v.r,
                     P->kst.p, v.part, norm_ref);
  CPGPU_HIP(ctx, hipGetLastError());
  int launched = 0;
  int batch = 8;
  while (launched < max_iter) {
    const int nb = std::min(batch, max_iter - launched);
    for (int k = 0; k < nb; ++k) {
      const int it = launched + k;
      T* pold = (it & 1) ? v.pb : v.pa;
      T* pnew = (it & 1) ? v.pa : v.pb;
      {
        TimerScope ts(ctx, 0);
        hipLaunchKernelGGL((k_cg_dir<T, PC>), grid, block, 0, st, s, v.r, pold,
                           pnew, v.q, P->kst.p, v.part);
      }
      {
        TimerScope ts(ctx, 1);
        hipLaunchKernelGGL((k_cg_upd<T, PC>), grid, block, 0, st, s, v.x, pnew,
                           v.r, v.q, P->kst.p, v.part, dhist, hist_cap,
                           norm_ref);
      }
    }
    CPGPU_HIP(ctx, hipGetLastError());
    launched += nb;
    CPGPU_HIP(ctx, hipMemcpyAsync(P->h_kst, P->kst.p, sizeof(KState),
                                  hipMemcpyDeviceToHost, st));
    CPGPU_HIP(ctx, hipStreamSynchronize(st));
    if (P->h_kst->done || P->h_kst->iter >= max_iter) break;
    batch = std::min(batch * 2, 64);
  }
  CPGPU_HIP(ctx, hipMemcpyAsync(P->h_kst, P->kst.p, sizeof(KState),
                                hipMemcpyDeviceToHost, st));
  CPGPU_HIP(ctx, hipStreamSynchronize(st));
  *out = *P->h_kst;
  return CPGPU_OK;
}

int alloc_f64(cpgpu_poisson* P) {
  const size_t N = P->N;
  const size_t nb = (size_t)tile_grid(P->nx, P->ny).x * tile_grid(P->nx, P->ny).y;
  if (P->x.alloc(N) || P->b.alloc(N) || P->r.alloc(N) || P->p0.alloc(N) ||
      P->p1.alloc(N) || P->q.alloc(N) || P->part64.alloc(3 * std::max<size_t>(nb, 1024)) ||
      P->kst.alloc(1) || P->rst.alloc(1))
    return set_error(P->ctx, CPGPU_ENOMEM, "device allocation failed");
  return CPGPU_OK;
}

int alloc_f32(cpgpu_poisson* P) {
  const size_t N = P->N;
  const size_t nb = (size_t)tile_grid(P->nx, P->ny).x * tile_grid(P->nx, P->ny).y;
  if (P->b32.alloc(N) || P->x32.alloc(N) || P->r32.alloc(N) ||
      P->pa32.alloc(N) || P->pb32.alloc(N) || P->q32.alloc(N) ||
      P->part32.alloc(3 * std::max<size_t>(nb, 1024)) || P->fKX.alloc(P->nx + 1) ||
      P->fKY.alloc(P->ny + 1))
    return set_error(P->ctx, CPGPU_ENOMEM, "device allocation failed");
  std::vector<float> kx(P->KX.begin(), P->KX.end()), ky(P->KY.begin(), P->KY.end());
  int rc = upload(P->ctx, P->fKX.p, kx.data(), sizeof(float) * kx.size());
  if (rc) return rc;
  rc = upload(P->ctx, P->fKY.p, ky.data(), sizeof(float) * ky.size());
  if (rc) return rc;
  CPGPU_HIP(P->ctx, hipStreamSynchronize(P->ctx->stream));
  return CPGPU_OK;
}

// FP32 J-PCG inside FP64 iterative refinement.
int run_mixed(cpgpu_poisson* P, int max_inner_total, double tol, int project,
              double* hist, int hist_cap, int* hist_len, int* iters,
              double* resid_out) {
  cpgpu_ctx* ctx = P->ctx;
  hipStream_t st = ctx->stream;
  int rc = alloc_f32(P);
  if (rc) return rc;
  const dim3 grid = tile_grid(P->nx, P->ny), block(TJ, BY);
  const int nb1 = 1024;
  int total_inner = 0, code = 1, hl = 0;
  double resid = 0;
  for (int o = 0; o <= P->max_outer; ++o) {
    CPGPU_HIP(ctx, hipMemsetAsync(P->rst.p, 0, sizeof(RState), st));
    hipLaunchKernelGGL(k_residual64, grid, block, 0, st, P->sten64(), P->x.p,
                       P->b.p, P->r.p, P->rst.p, P->part64.p);
    CPGPU_HIP(ctx, hipMemcpyAsync(P->h_rst, P->rst.p, sizeof(RState),
                                  hipMemcpyDeviceToHost, st));
    CPGPU_HIP(ctx, hipStreamSynchronize(st));
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
                    P->q32.p, P->part32.p};
    KState ks;
    rc = run_cg<float, true>(P, P->sten32(), v, max_inner_total - total_inner,
                             P->inner_tol, 0, nullptr, 0, &ks);
    if (rc) return rc;
    total_inner += ks.iter;
    hipLaunchKernelGGL(k_add_f32, dim3(nb1), dim3(256), 0, st, P->x.p,
                       P->x32.p, P->rst.p, P->part64.p, P->N);
    if (project)
      hipLaunchKernelGGL(k_sub_mean, dim3(nb1), dim3(256), 0, st, P->x.p,
                         P->rst.p, P->N);
    CPGPU_HIP(ctx, hipGetLastError());
  }
  *hist_len = hl;
  *iters = total_inner;
  *resid_out = resid;
  return code;
}

}  // namespace

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
  int rc = build_descriptor(P.get(), g);
  if (rc) return rc;
  rc = alloc_f64(P.get());
  if (rc) return rc;
  if (P->dKX.alloc(P->nx + 1) || P->dKY.alloc(P->ny + 1) ||
      P->dvol.alloc(P->nx) || P->dbcv.alloc(P->bcv.size()))
    return set_error(ctx, CPGPU_ENOMEM, "device allocation failed");
  if ((rc = upload(ctx, P->dKX.p, P->KX.data(), sizeof(double) * P->KX.size())) ||
      (rc = upload(ctx, P->dKY.p, P->KY.data(), sizeof(double) * P->KY.size())) ||
      (rc = upload(ctx, P->dvol.p, P->vol.data(), sizeof(double) * P->vol.size())) ||
      (rc = upload(ctx, P->dbcv.p, P->bcv.data(), sizeof(double) * P->bcv.size())))
    return rc;
  CPGPU_HIP(ctx, hipMemsetAsync(P->x.p, 0, sizeof(double) * P->N, ctx->stream));
  CPGPU_HIP(ctx, hipMemsetAsync(P->b.p, 0, sizeof(double) * P->N, ctx->stream));
  CPGPU_HIP(ctx, hipHostMalloc((void**)&P->h_kst, sizeof(KState), 0));
  CPGPU_HIP(ctx, hipHostMalloc((void**)&P->h_rst, sizeof(RState), 0));
  CPGPU_HIP(ctx, hipMemsetAsync(P->kst.p, 0, sizeof(KState), ctx->stream));
  CPGPU_HIP(ctx, hipMemsetAsync(P->rst.p, 0, sizeof(RState), ctx->stream));
  CPGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = P.release();
  return CPGPU_OK;
}

int cpgpu_poisson_destroy(cpgpu_poisson* P) {
  if (!P) return CPGPU_EINVAL;
  hipStreamSynchronize(P->ctx->stream);
  delete P;
  return CPGPU_OK;
}

static int rhs_from_device(cpgpu_poisson* P, const double* dsrc,
                           double src_uniform) {
  RhsDesc d;
  d.nx = P->nx;
  d.ny = P->ny;
  d.vol = P->dvol.p;
  d.bcval = P->dbcv.p;
  for (int s = 0; s < 4; ++s) d.seg[s] = P->seg[s];
  hipLaunchKernelGGL(k_build_rhs, tile_grid(P->nx, P->ny), dim3(TJ, BY), 0,
                     P->ctx->stream, d, dsrc, src_uniform, P->b.p);
  CPGPU_HIP(P->ctx, hipGetLastError());
  return CPGPU_OK;
}

int cpgpu_poisson_set_source(cpgpu_poisson* P, const double* src_field,
                             double src_uniform) {
  if (!P) return CPGPU_EINVAL;
  if (src_field) {
    if (P->src.alloc(P->N)) return set_error(P->ctx, CPGPU_ENOMEM, "src");
    int rc = field_to_dev(P, src_field, P->src.p);
    if (rc) return rc;
    return rhs_from_device(P, P->src.p, 0.0);
  }
  return rhs_from_device(P, nullptr, src_uniform);
}

int cpgpu_poisson_set_source_device(cpgpu_poisson* P, const double* d_src) {
  if (!P || !d_src) return CPGPU_EINVAL;
  return rhs_from_device(P, d_src, 0.0);
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
  CPGPU_HIP(ctx, hipMemsetAsync(P->rst.p, 0, sizeof(RState), ctx->stream));
  hipLaunchKernelGGL(k_blobs, tile_grid(P->nx, P->ny), dim3(TJ, BY), 0,
                     ctx->stream, P->nx, P->ny, dcx.p, dcy.p, dprm.p, nblobs,
                     P->src.p, P->rst.p, P->part64.p);
  hipLaunchKernelGGL(k_sub_rsum, dim3(1024), dim3(256), 0, ctx->stream,
                     P->src.p, P->rst.p, P->N);
  CPGPU_HIP(ctx, hipGetLastError());
  rc = rhs_from_device(P, P->src.p, 0.0);
  if (rc) return rc;
  // keep the parameter buffers alive until the kernels ran
  CPGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CPGPU_OK;
}

int cpgpu_poisson_set_phi(cpgpu_poisson* P, const double* phi_field) {
  if (!P || !phi_field) return CPGPU_EINVAL;
  return field_to_dev(P, phi_field, P->x.p);
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

int cpgpu_poisson_apply(cpgpu_poisson* P, const double* x_field,
                        double* y_field) {
  if (!P || !x_field || !y_field) return CPGPU_EINVAL;
  int rc = field_to_dev(P, x_field, P->p0.p);
  if (rc) return rc;
  hipLaunchKernelGGL(k_apply<double>, tile_grid(P->nx, P->ny), dim3(TJ, BY), 0,
                     P->ctx->stream, P->sten64(), P->p0.p, P->q.p);
  CPGPU_HIP(P->ctx, hipGetLastError());
  return dev_to_field(P, P->q.p, y_field);
}

int cpgpu_poisson_residual(cpgpu_poisson* P, double* rel) {
  if (!P || !rel) return CPGPU_EINVAL;
  cpgpu_ctx* ctx = P->ctx;
  CPGPU_HIP(ctx, hipMemsetAsync(P->rst.p, 0, sizeof(RState), ctx->stream));
  hipLaunchKernelGGL(k_residual64, tile_grid(P->nx, P->ny), dim3(TJ, BY), 0,
                     ctx->stream, P->sten64(), P->x.p, P->b.p, P->r.p, P->rst.p,
                     P->part64.p);
  CPGPU_HIP(ctx, hipMemcpyAsync(P->h_rst, P->rst.p, sizeof(RState),
                                hipMemcpyDeviceToHost, ctx->stream));
  CPGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
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
  cpgpu_ctx* ctx = P->ctx;
  CPGPU_HIP(ctx, hipSetDevice(ctx->device));
  int hl = 0;
  if (method == CPGPU_CG || method == CPGPU_JPCG) {
    int rc = ensure_hist(P, hist ? hist_cap : 0);
    if (rc) return rc;
    CGVecs<double> v{P->x.p, P->b.p, P->r.p, P->p0.p, P->p1.p, P->q.p,
                     P->part64.p};
    KState ks;
    const int nref = norm_mode == CPGPU_NORM_REF;
    double* dh = hist && hist_cap > 0 ? P->dhist.p : nullptr;
    rc = method == CPGPU_JPCG
             ? run_cg<double, true>(P, P->sten64(), v, *max_iter, *tol, nref,
                                    dh, hist_cap, &ks)
             : run_cg<double, false>(P, P->sten64(), v, *max_iter, *tol, nref,
                                     dh, hist_cap, &ks);
    if (rc) return rc;
    hl = ks.iter;
    if (dh && hl > 0) {
      CPGPU_HIP(ctx, hipMemcpy(hist, dh, sizeof(double) * std::min(hl, hist_cap),
                               hipMemcpyDeviceToHost));
    }
    if (hist_len) *hist_len = hl;
    *tol = ks.resid;
    if (ks.done) {
      *max_iter = ks.iter;
      return CPGPU_OK;
    }
    return CPGPU_MAXITER;
  }
  if (method == CPGPU_MIXED) {
    int iters = 0;
    double resid = 0;
    int code = run_mixed(P, *max_iter, *tol, opts & CPGPU_OPT_PROJECT_MEAN,
                         hist, hist_cap, &hl, &iters, &resid);
    if (code < 0) return code;
    if (hist_len) *hist_len = hl;
    *max_iter = iters;
    *tol = resid;
    return code;
  }
  return set_error(ctx, CPGPU_EUNSUPPORTED, "method %d not available", method);
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

}  // extern "C"
