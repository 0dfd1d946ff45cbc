// The Poisson operator of grids whose cell sizes vary along both axes
// (structure::Grid_ built from size arrays, Grid_::_set_non_uniform_grid,
// s_grid.hpp:397-434), matrix-free with per-face coefficient arrays.
//
// The uniform path keeps 1-D face coefficients; here the x-face coefficient
// beta*a/d has a = sy(j) and d = |cx(i) - cx(i-1)|, so it varies along both
// axes, and every boundary row carries its own ghost substitution.  The host
// computes each coefficient exactly as StencilPoisson_ and Ghost_ do
// (s_poisson.hpp:79-135, s_ghost.hpp:140-179): FX[f][j] and FY[i][f] the
// face coefficients -coe (f the face between cells f-1 and f), D the centre
// coefficient after the substitutions, and per boundary line the ghost
// term's constant data.  A x adds the five terms in the reference's CSR
// order without FMAs, so it is bitwise the reference's SpMV, and the
// Jacobi-PCG / CG iterations are those of the uniform path (the same device
// state and deterministic reductions), one cell per lane.
//
// Algorithmic bytes of one Jacobi-PCG iteration: 10 N s of vectors as on a
// uniform grid, plus FX, FY and D (3 N s) read by the direction kernel and D
// by the update kernel.
#include <cmath>
#include <cstring>
#include <vector>

#include "poisson.hpp"

#pragma clang fp contract(off)

namespace cpgpu {
namespace {

struct GenOp {
  int nx, ny;
  const double* FX;  // (nx+1) x ny: face f of row j at FX[f * ny + j]
  const double* FY;  // nx x (ny+1): face f of column i at FY[i * (ny+1) + f]
  const double* D;   // nx x ny
};

__device__ __forceinline__ double gen_row(const GenOp& o, const double* __restrict__ x, int i,
                                          int j) {
  const int ny = o.ny;
  const size_t k = (size_t)i * ny + j;
  double acc = 0.0;
  if (i > 0) acc = acc + x[k - ny] * (-o.FX[(size_t)i * ny + j]);
  if (j > 0) acc = acc + x[k - 1] * (-o.FY[(size_t)i * (ny + 1) + j]);
  acc = acc + x[k] * o.D[k];
  if (j < ny - 1) acc = acc + x[k + 1] * (-o.FY[(size_t)i * (ny + 1) + j + 1]);
  if (i < o.nx - 1) acc = acc + x[k + ny] * (-o.FX[(size_t)(i + 1) * ny + j]);
  return acc;
}

constexpr int kGenThreads = 256;

inline int gen_blocks(size_t n) {
  const size_t b = (n + kGenThreads - 1) / kGenThreads;
  return (int)(b < (size_t)kMaxBlocks ? b : (size_t)kMaxBlocks);
}

__global__ __launch_bounds__(256) void k_gen_apply(GenOp o, const double* __restrict__ x,
                                                   double* __restrict__ y) {
  const size_t n = (size_t)o.nx * o.ny;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    y[k] = gen_row(o, x, (int)(k / o.ny), (int)(k % o.ny));
}

// b = -(vol*src + the ghost constants in Index_compare_ order x-, y-, y+,
// x+); seg holds the ghost data per boundary line: x- (ny), x+ (ny),
// y- (nx), y+ (nx).
__global__ __launch_bounds__(256) void k_gen_rhs(int nx, int ny, const double* __restrict__ vol,
                                                 const RhsSeg* __restrict__ seg,
                                                 const double* __restrict__ bcval,
                                                 const double* __restrict__ src, double src_uniform,
                                                 double* __restrict__ b) {
  const size_t n = (size_t)nx * ny;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(k / ny), j = (int)(k % ny);
    double cst = vol[i] * (src ? src[k] : src_uniform);
    if (i == 0) cst = ghost_const(seg[j], cst, bcval[j]);
    if (j == 0) cst = ghost_const(seg[2 * ny + i], cst, bcval[2 * ny + i]);
    if (j == ny - 1) cst = ghost_const(seg[2 * ny + nx + i], cst, bcval[2 * ny + nx + i]);
    if (i == nx - 1) cst = ghost_const(seg[ny + j], cst, bcval[ny + j]);
    b[k] = cst == 0.0 ? 0.0 : -cst;
  }
}

// r = b - A x; partials of r.z, r.r, b.b (the reference's nrm2 drops the
// last unknown when norm_ref)
template <bool PC>
__global__ __launch_bounds__(256) void k_gen_init(GenOp o, const double* __restrict__ x,
                                                  const double* __restrict__ b,
                                                  double* __restrict__ r,
                                                  double* __restrict__ partB, int norm_ref) {
  __shared__ double lds[16];
  const size_t n = (size_t)o.nx * o.ny;
  double v[3] = {0.0, 0.0, 0.0};
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const double bv = b[k];
    const double rv = bv - gen_row(o, x, (int)(k / o.ny), (int)(k % o.ny));
    r[k] = rv;
    const double z = PC ? rv * (1.0 / o.D[k]) : rv;
    v[SLOT_RZ] += rv * z;
    if (!(norm_ref && k == n - 1)) {
      v[SLOT_RR] += rv * rv;
      v[SLOT_BB] += bv * bv;
    }
  }
  store_partials<double, 3>(v, partB, gridDim.x, blockIdx.x, lds);
}

// p of iteration it at (i, j): z + beta p_old (z alone on the first)
template <bool PC>
__device__ __forceinline__ double gen_p(const GenOp& o, const double* __restrict__ r,
                                        const double* __restrict__ pold, int i, int j,
                                        double beta, bool first) {
  if (i < 0 || i >= o.nx || j < 0 || j >= o.ny) return 0.0;
  const size_t k = (size_t)i * o.ny + j;
  const double rv = r[k];
  const double z = PC ? rv * (1.0 / o.D[k]) : rv;
  return first ? z : z + beta * pold[k];
}

// Iteration `it`, first kernel: the prologue's stopping test and beta
// (dir_prologue, as the uniform path), p = z + beta p_old, q = A p with
// the neighbours' p formed in place, partial p.q.
template <bool PC>
__global__ __launch_bounds__(256) void k_gen_dir(GenOp o, const double* __restrict__ r,
                                                 const double* __restrict__ pold,
                                                 double* __restrict__ pnew, double* __restrict__ q,
                                                 KState* st, const double* __restrict__ partB,
                                                 double* __restrict__ partA, double* hist,
                                                 int hist_cap, int Gprev) {
  if (st->done) return;
  __shared__ double lds[16];
  double beta;
  bool first;
  if (!dir_prologue<double>(st, partB, Gprev, nullptr, 0, hist, hist_cap, lds, &beta, &first))
    return;
  const int nx = o.nx, ny = o.ny;
  const size_t n = (size_t)nx * ny;
  double v[1] = {0.0};
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(k / ny), j = (int)(k % ny);
    const double pc = gen_p<PC>(o, r, pold, i, j, beta, first);
    const double pn = gen_p<PC>(o, r, pold, i - 1, j, beta, first);
    const double pw = gen_p<PC>(o, r, pold, i, j - 1, beta, first);
    const double pe = gen_p<PC>(o, r, pold, i, j + 1, beta, first);
    const double ps = gen_p<PC>(o, r, pold, i + 1, j, beta, first);
    double acc = 0.0;
    if (i > 0) acc = acc + pn * (-o.FX[(size_t)i * ny + j]);
    if (j > 0) acc = acc + pw * (-o.FY[(size_t)i * (ny + 1) + j]);
    acc = acc + pc * o.D[k];
    if (j < ny - 1) acc = acc + pe * (-o.FY[(size_t)i * (ny + 1) + j + 1]);
    if (i < nx - 1) acc = acc + ps * (-o.FX[(size_t)(i + 1) * ny + j]);
    pnew[k] = pc;
    q[k] = acc;
    v[0] += pc * acc;
  }
  store_partials<double, 1>(v, partA, gridDim.x, blockIdx.x, lds);
}

// Iteration `it`, second kernel: alpha, x += alpha p, r -= alpha q,
// partials of r.z and r.r; publishes the next iteration index.
template <bool PC>
__global__ __launch_bounds__(256) void k_gen_upd(GenOp o, double* __restrict__ x,
                                                 const double* __restrict__ p,
                                                 double* __restrict__ r,
                                                 const double* __restrict__ q, KState* st,
                                                 const double* __restrict__ partA,
                                                 double* __restrict__ partB, int norm_ref,
                                                 int Gprev) {
  if (st->done) return;
  __shared__ double lds[16];
  const int it = st->it_upd;
  double pq[1];
  reduce_partials<double, 1>(partA, Gprev, pq, lds);
  const double alpha = st->rho[it & 1] / pq[0];
  const size_t n = (size_t)o.nx * o.ny;
  double v[3] = {0.0, 0.0, 0.0};
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    x[k] = x[k] + alpha * p[k];
    const double rv = r[k] - alpha * q[k];
    r[k] = rv;
    const double z = PC ? rv * (1.0 / o.D[k]) : rv;
    v[SLOT_RZ] += rv * z;
    if (!(norm_ref && k == n - 1)) v[SLOT_RR] += rv * rv;
  }
  store_partials<double, 3>(v, partB, gridDim.x, blockIdx.x, lds);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->it_dir = it + 1;
}

// partials of b.b and r.r of r = b - A x (the true residual)
__global__ __launch_bounds__(256) void k_gen_resid(GenOp o, const double* __restrict__ x,
                                                   const double* __restrict__ b,
                                                   double* __restrict__ partials) {
  __shared__ double lds[8];
  const size_t n = (size_t)o.nx * o.ny;
  double v[2] = {0.0, 0.0};
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const double bv = b[k];
    const double rv = bv - gen_row(o, x, (int)(k / o.ny), (int)(k % o.ny));
    v[0] += bv * bv;
    v[1] += rv * rv;
  }
  store_partials<double, 2>(v, partials, gridDim.x, blockIdx.x, lds);
}

GenOp gen_op(const cpgpu_poisson* P) {
  GenOp o;
  o.nx = P->nx;
  o.ny = P->ny;
  o.FX = P->gFX.p;
  o.FY = P->gFY.p;
  o.D = P->gD.p;
  return o;
}

}  // namespace

// The rows of StencilPoisson_ with GhostRegular_ substitution
// (s_poisson.hpp:79-135, s_ghost.hpp:140-179, :334-357) per cell.
int general_build(cpgpu_poisson* P, const cpgpu_grid2d* g) {
  const int nx = g->nx, ny = g->ny, gl = g->gl;
  const double beta = g->beta;
  auto cax = [&](int d, int k) { return d == 0 ? g->cx[k + gl] : g->cy[k + gl]; };
  auto sax = [&](int d, int k) { return d == 0 ? g->sx[k + gl] : g->sy[k + gl]; };
  // _coe_other: -beta*a/d, a the product of the other sizes (z size 1),
  // d the distance between the centres (s_poisson.hpp:98-135)
  auto coe_other = [&](int i, int j, int d, int ori) {
    double a, c, co;
    if (d == 0) {
      a = sax(1, j) * 1.0;
      c = cax(0, i);
      co = cax(0, i + ori);
    } else {
      a = 1.0 * sax(0, i);
      c = cax(1, j);
      co = cax(1, j + ori);
    }
    return -beta * a / std::fabs(c - co);
  };
  std::vector<double> FX((size_t)(nx + 1) * ny), FY((size_t)nx * (ny + 1)), D((size_t)nx * ny);
  std::vector<RhsSeg> seg(2 * (size_t)ny + 2 * (size_t)nx);
  for (int f = 0; f <= nx; ++f)
    for (int j = 0; j < ny; ++j)
      FX[(size_t)f * ny + j] = -(f < nx ? coe_other(f, j, 0, -1) : coe_other(nx - 1, j, 0, 1));
  for (int i = 0; i < nx; ++i)
    for (int f = 0; f <= ny; ++f)
      FY[(size_t)i * (ny + 1) + f] =
          -(f < ny ? coe_other(i, f, 1, -1) : coe_other(i, ny - 1, 1, 1));
  // one ghost term: centre adjustment and the constant's data (a the axis,
  // ori -1 or +1, (i, j) the boundary cell)
  auto subst = [&](int i, int j, int a, int ori, double coe, double& cen, RhsSeg& sg, int type) {
    const int ci = a == 0 ? i : j;
    const double c0 = cax(a, ci + ori), c1 = cax(a, ci);
    const double cf = ori > 0 ? c1 + sax(a, ci) * 0.5 : c1 - sax(a, ci) * 0.5;
    const double dl = std::fabs(c1 - c0), d0 = std::fabs(cf - c0), d1 = std::fabs(c1 - cf);
    cen = type == CPGPU_BC_DIRICHLET ? cen + (-coe * d0 / d1) : cen + coe;
    sg.type = type;
    sg.coe = coe;
    sg.dl = dl;
    sg.d1 = d1;
  };
  for (int s = 0; s < 4; ++s)
    if (g->bc_type[s] != CPGPU_BC_DIRICHLET && g->bc_type[s] != CPGPU_BC_NEUMANN)
      return set_error(P->ctx, CPGPU_EINVAL, "bad bc type %d on segment %d", g->bc_type[s], s);
  for (int i = 0; i < nx; ++i) {
    for (int j = 0; j < ny; ++j) {
      const double cm0 = coe_other(i, j, 0, -1), cp0 = coe_other(i, j, 0, 1);
      const double cm1 = coe_other(i, j, 1, -1), cp1 = coe_other(i, j, 1, 1);
      double cen = 0.0;
      cen += -cm0;
      cen += -cp0;
      cen += -cm1;
      cen += -cp1;
      // ghost terms in Index_compare_ order: (i-1, j), (i, j-1), (i, j+1), (i+1, j)
      RhsSeg dummy;
      if (i == 0) subst(i, j, 0, -1, cm0, cen, seg[j], g->bc_type[0]);
      if (j == 0) subst(i, j, 1, -1, cm1, cen, seg[2 * ny + i], g->bc_type[2]);
      if (j == ny - 1) subst(i, j, 1, 1, cp1, cen, seg[2 * ny + nx + i], g->bc_type[3]);
      if (i == nx - 1) subst(i, j, 0, 1, cp0, cen, seg[ny + j], g->bc_type[1]);
      (void)dummy;
      D[(size_t)i * ny + j] = cen;
    }
  }
  cpgpu_ctx* ctx = P->ctx;
  if (P->gFX.alloc(FX.size()) || P->gFY.alloc(FY.size()) || P->gD.alloc(D.size()) ||
      P->gseg.alloc(seg.size()))
    return set_error(ctx, CPGPU_ENOMEM, "general operator arrays");
  hipStream_t st = ctx->stream;
  CPGPU_HIP(ctx, hipMemcpyAsync(P->gFX.p, FX.data(), sizeof(double) * FX.size(),
                                hipMemcpyHostToDevice, st));
  CPGPU_HIP(ctx, hipMemcpyAsync(P->gFY.p, FY.data(), sizeof(double) * FY.size(),
                                hipMemcpyHostToDevice, st));
  CPGPU_HIP(ctx, hipMemcpyAsync(P->gD.p, D.data(), sizeof(double) * D.size(),
                                hipMemcpyHostToDevice, st));
  CPGPU_HIP(ctx, hipMemcpyAsync(P->gseg.p, seg.data(), sizeof(RhsSeg) * seg.size(),
                                hipMemcpyHostToDevice, st));
  CPGPU_HIP(ctx, hipStreamSynchronize(st));
  P->general = true;
  return CPGPU_OK;
}

int general_rhs(cpgpu_poisson* P, const double* dsrc, double src_uniform) {
  const size_t n = P->N;
  hipLaunchKernelGGL(k_gen_rhs, dim3(gen_blocks(n)), dim3(kGenThreads), 0, P->ctx->stream,
                     P->nx, P->ny, P->dvol.p, P->gseg.p, P->dbcv.p, dsrc, src_uniform, P->b.p);
  CPGPU_HIP(P->ctx, hipGetLastError());
  return CPGPU_OK;
}

int general_apply(cpgpu_poisson* P, const double* dx, double* dy) {
  hipLaunchKernelGGL(k_gen_apply, dim3(gen_blocks(P->N)), dim3(kGenThreads), 0, P->ctx->stream,
                     gen_op(P), dx, dy);
  CPGPU_HIP(P->ctx, hipGetLastError());
  return CPGPU_OK;
}

int general_residual(cpgpu_poisson* P, double* rel) {
  cpgpu_ctx* ctx = P->ctx;
  hipStream_t st = ctx->stream;
  const int G = gen_blocks(P->N);
  hipLaunchKernelGGL(k_gen_resid, dim3(G), dim3(kGenThreads), 0, st, gen_op(P), P->x.p, P->b.p,
                     P->part64.p);
  CPGPU_HIP(ctx, hipGetLastError());
  std::vector<double> part(2 * (size_t)G);
  CPGPU_HIP(ctx, hipMemcpyAsync(part.data(), P->part64.p, sizeof(double) * part.size(),
                                hipMemcpyDeviceToHost, st));
  CPGPU_HIP(ctx, hipStreamSynchronize(st));
  double bb = 0.0, rr = 0.0;
  for (int k = 0; k < G; ++k) {
    bb += part[k];
    rr += part[G + k];
  }
  *rel = std::sqrt(rr) / (bb > 0.0 ? std::sqrt(bb) : 1.0);
  return CPGPU_OK;
}

// carpio::CG / IC_CG's loop with DiaPre(A,1) (solver_matrix.hpp:264-374):
// the same device state, stopping test and direct launches polled every 16
// iterations as the uniform path's unbatched mode.
int general_solve(cpgpu_poisson* P, int method, int* max_iter, double* tol, int norm_ref,
                  double* hist, int hist_cap, int* hist_len) {
  cpgpu_ctx* ctx = P->ctx;
  if (method != CPGPU_JPCG && method != CPGPU_CG)
    return set_error(ctx, CPGPU_EUNSUPPORTED,
                     "cell sizes that vary: Jacobi-PCG and CG only (method %d)", method);
  hipStream_t st = ctx->stream;
  const GenOp o = gen_op(P);
  const int G = gen_blocks(P->N);
  const bool pc = method == CPGPU_JPCG;
  if (hist && hist_cap > 0 && P->dhist.alloc(hist_cap))
    return set_error(ctx, CPGPU_ENOMEM, "hist alloc");
  double* dh = hist && hist_cap > 0 ? P->dhist.p : nullptr;
  KState init{};
  init.tol = *tol;
  init.max_iter = *max_iter;
  init.it_dir = 1;
  init.refresh_every = 16;
  init.normb = 1.0;
  KState* hs = &P->h_kst[0];
  *hs = init;
  CPGPU_HIP(ctx, hipMemcpyAsync(P->kst.p, hs, sizeof(KState), hipMemcpyHostToDevice, st));
  if (pc)
    hipLaunchKernelGGL(k_gen_init<true>, dim3(G), dim3(kGenThreads), 0, st, o, P->x.p, P->b.p,
                       P->r.p, P->partB64.p, norm_ref);
  else
    hipLaunchKernelGGL(k_gen_init<false>, dim3(G), dim3(kGenThreads), 0, st, o, P->x.p, P->b.p,
                       P->r.p, P->partB64.p, norm_ref);
  CPGPU_HIP(ctx, hipGetLastError());
  const int B = 16;
  const long long total = (long long)*max_iter + 2;
  for (long long launched = 0;; launched += B) {
    for (int k = 0; k < B; ++k) {
      double* pold = (k & 1) ? P->p1.p : P->p0.p;
      double* pnew = (k & 1) ? P->p0.p : P->p1.p;
      if (pc) {
        hipLaunchKernelGGL(k_gen_dir<true>, dim3(G), dim3(kGenThreads), 0, st, o, P->r.p, pold,
                           pnew, P->q.p, P->kst.p, P->partB64.p, P->partA64.p, dh, hist_cap, G);
        hipLaunchKernelGGL(k_gen_upd<true>, dim3(G), dim3(kGenThreads), 0, st, o, P->x.p, pnew,
                           P->r.p, P->q.p, P->kst.p, P->partA64.p, P->partB64.p, norm_ref, G);
      } else {
        hipLaunchKernelGGL(k_gen_dir<false>, dim3(G), dim3(kGenThreads), 0, st, o, P->r.p, pold,
                           pnew, P->q.p, P->kst.p, P->partB64.p, P->partA64.p, dh, hist_cap, G);
        hipLaunchKernelGGL(k_gen_upd<false>, dim3(G), dim3(kGenThreads), 0, st, o, P->x.p, pnew,
                           P->r.p, P->q.p, P->kst.p, P->partA64.p, P->partB64.p, norm_ref, G);
      }
    }
    CPGPU_HIP(ctx, hipGetLastError());
    CPGPU_HIP(ctx, hipMemcpyAsync(hs, P->kst.p, sizeof(KState), hipMemcpyDeviceToHost, st));
    CPGPU_HIP(ctx, hipStreamSynchronize(st));
    if (hs->done || launched + B >= total + B) break;
  }
  const int hl = hs->iter;
  if (dh && hl > 0)
    CPGPU_HIP(ctx, hipMemcpy(hist, dh, sizeof(double) * std::min(hl, hist_cap),
                             hipMemcpyDeviceToHost));
  if (hist_len) *hist_len = hl;
  *tol = hs->resid;
  if (hs->done == 1) {
    *max_iter = hs->iter;
    return CPGPU_OK;
  }
  return CPGPU_MAXITER;
}

}  // namespace cpgpu
