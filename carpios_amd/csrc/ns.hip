// Device-resident NS_explicit_ projection step on gfx950.
//
// One step of structure::NS_explicit_<2>::run_one_step (reference
// lib/structure/s_ns_explicit.hpp:118-155) followed by Time_::advance
// (s_time.hpp:61-64), with every field resident in HBM:
//   k_maxmag + k_dt ........ NS_::restrict_time (s_ns.hpp:146-169)
//   k_ns_predict ........... in one pass over LDS tiles: Equation_::apply_bc
//                            of u, v (s_equation.hpp:340-378), TVD advection
//                            + explicit diffusion for both components
//                            (s_ns_explicit.hpp:173-229, s_operation.hpp:
//                            549-620, :664-688), apply_bc of u*, v*,
//                            InterpolateC2F + Div (s_operation.hpp:100-120,
//                            :166-187)
//   Poisson solve .......... the projection, all-Neumann, warm-started from
//                            the previous phi (s_poisson.hpp:320-420)
//   k_grad_correct ......... Grad, p = phi/dt, u*f -= grad phi,
//                            u* -= F2C(grad phi), u <- u*, uf <- u*f
//                            (s_operation.hpp:46-72, :141-164, :732-845)
// Fields are ghost-padded (nx+2gl) x (ny+2gl) buffers, j contiguous, exactly
// the reference's Scalar_ layout.  Arithmetic follows the reference's
// evaluation order without FMA contraction; the face values that a
// neighbouring cell would write are recomputed with the same operands, which
// gives the same bits because IEEE addition and multiplication commute.
//
// Corner ghost cells (outside the domain on both axes) are filled from the
// freshly filled edge ghosts; the reference's sweep leaves them depending on
// the visiting order, and no stencil of the step reads them.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include "poisson.hpp"

namespace cpgpu {

constexpr double kSmall = 1e-10;  // lib/type_define.hpp:16

struct Axes {
  int nx, ny, gl, NX, NY;
  const double *cx, *sx, *cy, *sy;  // ghost-inclusive, indexed by I = i+gl
};

__device__ __forceinline__ size_t fidx(const Axes& a, int i, int j) {
  return (size_t)(i + a.gl) * a.NY + (j + a.gl);
}
__device__ __forceinline__ double cax(const Axes& a, int d, int k) {
  return d == 0 ? a.cx[k + a.gl] : a.cy[k + a.gl];
}
__device__ __forceinline__ double sax(const Axes& a, int d, int k) {
  return d == 0 ? a.sx[k + a.gl] : a.sy[k + a.gl];
}

struct BCDev {
  int type[4];
  const double* val;  // packed per ghost line: x- NY | x+ NY | y- NX | y+ NX
};

struct NSDev {
  double dt, t;
  double maxmag;
};

// Limiters, s_operation.hpp:501-546 (std::min/std::max semantics).
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }

template <int LIM>
__device__ __forceinline__ double limiter(double r, double R) {
  if (LIM == 0) return 0.5 * R * (r + fabs(r)) / (R - 1 + r);  // VanLeer
  if (LIM == 1) return dmax(dmax(0.0, dmin(R * r, 1.0)), dmin(r, R));  // superbee
  if (LIM == 2) {  // WAHYD
    if (r <= 1.0) return 0.5 * R * (r + fabs(r)) / (R - 1 + r);
    return dmin((r + R * r * fabs(r)) / (R + r * r), R);
  }
  if (LIM == 3) return 0.75 * r + 0.25;                 // QUICK
  if (LIM == 4) return 2.0 / 3.0 * r + 1.0 / 3.0;       // CUI
  return r;                                             // upwind2
}

// ---------------------------------------------------------------------
// kernels

// partial max of sqrt(u^2 + v^2) over the interior (s_vector.hpp:89-103)
__global__ __launch_bounds__(256) void k_maxmag(Axes a, const double* __restrict__ u,
                                                const double* __restrict__ v,
                                                double* partials) {
  __shared__ double lds[4];
  double m = 0.0;
  const size_t n = (size_t)a.nx * a.ny;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(k / a.ny), j = (int)(k - (size_t)i * a.ny);
    const size_t f = fidx(a, i, j);
    double sum = 0.0;
    sum += u[f] * u[f];
    sum += v[f] * v[f];
    m = dmax(m, sqrt(sum));
  }
  // max is order independent: wave then block then store
  for (int off = 32; off > 0; off >>= 1) m = dmax(m, __shfl_xor(m, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    partials[blockIdx.x] = dmax(dmax(dmax(lds[0], lds[1]), lds[2]), lds[3]);
}

// dt <- min(dt, CFL*min_size/max|u|, DN/kappa/sum(1/h^2)) when smaller
// (s_ns.hpp:146-169, s_operation.hpp:847-867)
__global__ __launch_bounds__(256) void k_dt(const double* partials, int G,
                                            NSDev* st, double cfl,
                                            double min_size, double dnum,
                                            double ddenom, double dcoef) {
  __shared__ double lds[4];
  double m = 0.0;
  for (int b = threadIdx.x; b < G; b += blockDim.x) m = dmax(m, partials[b]);
  for (int off = 32; off > 0; off >>= 1) m = dmax(m, __shfl_xor(m, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double mx = dmax(dmax(dmax(lds[0], lds[1]), lds[2]), lds[3]);
    st->maxmag = mx;
    const double rdt = cfl * min_size / mx;
    double rdt2 = st->dt;
    if (dnum < 100) rdt2 = dnum / dcoef / ddenom;
    const double mn = dmin(rdt, rdt2);
    if (st->dt > mn) st->dt = mn;
  }
}

// Ghost fill of one field (s_equation.hpp:340-378).  A ghost cell (i, j)
// takes its boundary segment and its source cell (oi, oj) from the first
// axis on which it lies outside the domain, so a corner's source is an edge
// ghost; its value is the line through the segment's value at the face and
// the source value (Interp_::Linear, arithmetic.hpp:428-431) for a
// Dirichlet segment, source - value * distance for a Neumann one.
__device__ __forceinline__ void ghost_src(const Axes& a, int i, int j, int& seg, int& oi,
                                          int& oj) {
  if (i < 0) { seg = 0; oi = 0; oj = j; }
  else if (i >= a.nx) { seg = 1; oi = a.nx - 1; oj = j; }
  else if (j < 0) { seg = 2; oi = i; oj = 0; }
  else { seg = 3; oi = i; oj = a.ny - 1; }
}

__device__ __forceinline__ double ghost_formula(const Axes& a, const BCDev& bc, int i, int j,
                                                int seg, int oi, int oj, double fo) {
  const int ax = seg < 2 ? 0 : 1, ori = (seg & 1) ? 1 : -1;
  const int line = (ax == 0 ? j : i) + a.gl;
  const int off = seg == 0 ? 0 : seg == 1 ? a.NY : seg == 2 ? 2 * a.NY : 2 * a.NY + a.NX;
  const double val = bc.val ? bc.val[off + line] : 0.0;
  const double xg = cax(a, ax, ax == 0 ? i : j);
  const double xo = cax(a, ax, ax == 0 ? oi : oj);
  const double so = sax(a, ax, ax == 0 ? oi : oj);
  const double xf = ori > 0 ? xo + so * 0.5 : xo - so * 0.5;
  if (bc.type[seg] == CPGPU_BC_DIRICHLET) {
    const double u = (xg - xf) / (xo - xf);
    return (val * (1.0 - u)) + (fo * u);
  }
  const double dl = fabs(xg - xo);
  return fo - val * dl;
}

// One thread per ghost cell; edge ghosts first (k_ghost_edges), corners
// second (k_ghost_corners), which read the edge ghosts.
__device__ __forceinline__ void ghost_one(const Axes& a, double* f,
                                          const BCDev& bc, int i, int j) {
  int seg, oi, oj;
  ghost_src(a, i, j, seg, oi, oj);
  f[fidx(a, i, j)] = ghost_formula(a, bc, i, j, seg, oi, oj, f[fidx(a, oi, oj)]);
}

// The value ghost_one leaves in ghost cell (i, j) (edge or corner) of a
// field whose interior values f(i, j) gives: an edge ghost from its source
// cell, a corner from the edge ghost that is its source.
template <class F>
__device__ __forceinline__ double ghost_of(const Axes& a, const BCDev& bc, int i, int j,
                                           const F& f) {
  int seg, oi, oj;
  ghost_src(a, i, j, seg, oi, oj);
  double fo;
  if (oi >= 0 && oi < a.nx && oj >= 0 && oj < a.ny) {
    fo = f(oi, oj);
  } else {
    int s2, oi2, oj2;
    ghost_src(a, oi, oj, s2, oi2, oj2);
    fo = ghost_formula(a, bc, oi, oj, s2, oi2, oj2, f(oi2, oj2));
  }
  return ghost_formula(a, bc, i, j, seg, oi, oj, fo);
}

// edge ghosts of two fields (blockIdx.y selects the field)
__global__ __launch_bounds__(256) void k_ghost_edges(Axes a, double* f0, BCDev b0,
                                                     double* f1, BCDev b1) {
  double* f = blockIdx.y == 0 ? f0 : f1;
  const BCDev& bc = blockIdx.y == 0 ? b0 : b1;
  // x edges: gl*ny each side, y edges: gl*nx each side
  const int nxe = 2 * a.gl * a.ny, nye = 2 * a.gl * a.nx;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nxe + nye;
       k += gridDim.x * blockDim.x) {
    int i, j;
    if (k < nxe) {
      const int side = k / (a.gl * a.ny), r = k - side * a.gl * a.ny;
      const int g = r / a.ny;
      j = r - g * a.ny;
      i = side == 0 ? -1 - g : a.nx + g;
    } else {
      const int kk = k - nxe;
      const int side = kk / (a.gl * a.nx), r = kk - side * a.gl * a.nx;
      const int g = r / a.nx;
      i = r - g * a.nx;
      j = side == 0 ? -1 - g : a.ny + g;
    }
    ghost_one(a, f, bc, i, j);
  }
}

__global__ __launch_bounds__(256) void k_ghost_corners(Axes a, double* f0, BCDev b0,
                                                       double* f1, BCDev b1) {
  double* f = blockIdx.y == 0 ? f0 : f1;
  const BCDev& bc = blockIdx.y == 0 ? b0 : b1;
  const int per = a.gl * a.gl;
  for (int k = threadIdx.x; k < 4 * per; k += blockDim.x) {
    const int c = k / per, r = k - c * per;
    const int gi = r / a.gl, gj = r - gi * a.gl;
    const int i = (c & 1) ? a.nx + gi : -1 - gi;
    const int j = (c & 2) ? a.ny + gj : -1 - gj;
    ghost_one(a, f, bc, i, j);
  }
}

// A scalar field read at cell (i, j): a ghost-padded device field, or a
// tile of it staged in LDS (the in-place advection sweep).
struct FieldAt {
  const double* __restrict__ p;
  Axes a;
  __device__ __forceinline__ double operator()(int i, int j) const { return p[fidx(a, i, j)]; }
};
struct TileAt {
  const double* t;  // LDS window of pitch TP: cell (i0 - H + r, j0 - H + c) at t[r * TP + c]
  int i0, j0;
  static constexpr int H = 2, TP = 20;
  __device__ __forceinline__ double operator()(int i, int j) const {
    return t[(i - i0 + H) * TP + (j - j0 + H)];
  }
};

// VdotNabla_TVD for one cell and component (s_operation.hpp:549-596); the
// face velocities through accessors (device fields or LDS tiles)
template <int LIM, class FU, class FV, class PHI>
__device__ __forceinline__ double tvd_cell_at(const Axes& a, const FU& uf, const FV& vf,
                                              const PHI& phi, int i, int j) {
  double vdd[2];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    const int ei = d == 0 ? 1 : 0, ej = d == 0 ? 0 : 1;
    const int c = d == 0 ? i : j;
    const double vfp = d == 0 ? uf(i, j) : vf(i, j);
    const double vfm = d == 0 ? uf(i - ei, j - ej) : vf(i - ei, j - ej);
    const double veo = (vfp + vfm) * 0.5;
    int Cm, Um, Dm, Cp, Up, Dp;
    if (veo >= 0.0) {
      Dm = 0; Cm = -1; Um = -2;
      Dp = 1; Cp = 0; Up = -1;
    } else {
      Dm = -1; Cm = 0; Um = 1;
      Dp = 0; Cp = 1; Up = 2;
    }
#define VAL(o) phi(i + ei * (o), j + ej * (o))
#define SZ(o) sax(a, d, c + (o))
    double sC = SZ(Cp), sD = SZ(Dp), sU = SZ(Up);
    double vC = VAL(Cp), vD = VAL(Dp), vU = VAL(Up);
    double R = (sD + sC) / sC;
    double r = (vC - vU) * (sD + sC) / (vD - vC + kSmall) / (sC + sU);
    const double fp = vC + limiter<LIM>(r, R) / R * (vD - vC);
    sC = SZ(Cm); sD = SZ(Dm); sU = SZ(Um);
    vC = VAL(Cm); vD = VAL(Dm); vU = VAL(Um);
    R = (sD + sC) / sC;
    r = (vC - vU) * (sD + sC) / (vD - vC + kSmall) / (sC + sU);
    const double fm = vC + limiter<LIM>(r, R) / R * (vD - vC);
#undef VAL
#undef SZ
    vdd[d] = (vfp * fp - vfm * fm) / sax(a, d, c);
  }
  double res = 0.0;
  res += vdd[0];
  res += vdd[1];
  return res;
}

template <int LIM, class PHI>
__device__ __forceinline__ double tvd_cell(const Axes& a,
                                           const double* __restrict__ uf,
                                           const double* __restrict__ vf,
                                           const PHI& phi, int i, int j) {
  return tvd_cell_at<LIM>(a, FieldAt{uf, a}, FieldAt{vf, a}, phi, i, j);
}

// NablaMuNabla with uniform mu (s_operation.hpp:664-688)
template <class PHI>
__device__ __forceinline__ double diffusion_cell_at(const Axes& a, double mu, const PHI& phi,
                                                    int i, int j) {
  double td[2];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    const int ei = d == 0 ? 1 : 0, ej = d == 0 ? 0 : 1, c = d == 0 ? i : j;
    const double f = phi(i, j);
    const double fp = phi(i + ei, j + ej);
    const double fm = phi(i - ei, j - ej);
    const double dm = mu * (f - fm) / (cax(a, d, c) - cax(a, d, c - 1));
    const double dp = mu * (fp - f) / (cax(a, d, c + 1) - cax(a, d, c));
    td[d] = (dp - dm) / sax(a, d, c);
  }
  double res = 0.0;
  res += td[0];
  res += td[1];
  return res;
}

__device__ __forceinline__ double diffusion_cell(const Axes& a, double mu,
                                                 const double* __restrict__ phi,
                                                 int i, int j) {
  return diffusion_cell_at(a, mu, FieldAt{phi, a}, i, j);
}

// face value of InterpolateC2F between cells k and k+1 along d
__device__ __forceinline__ double c2f(const Axes& a, int d, int k, double ck,
                                      double ckp) {
  const double hs = sax(a, d, k) * 0.5;
  const double hsp = sax(a, d, k + 1) * 0.5;
  return (ckp * hs + ck * hsp) / (hs + hsp);
}

// u*f, v*f and div (s_operation.hpp:100-120, :166-187); div also into the
// interior-only source of the projection.
__global__ __launch_bounds__(256) void k_c2f_div(Axes a, const double* __restrict__ us,
                                                 const double* __restrict__ vs,
                                                 double* __restrict__ usf,
                                                 double* __restrict__ vsf,
                                                 double* __restrict__ div,
                                                 double* __restrict__ div_int) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= a.nx || j >= a.ny) return;
  const double uc = us[fidx(a, i, j)], vc = vs[fidx(a, i, j)];
  const double fxp = c2f(a, 0, i, uc, us[fidx(a, i + 1, j)]);
  const double fxm = c2f(a, 0, i - 1, us[fidx(a, i - 1, j)], uc);
  const double fyp = c2f(a, 1, j, vc, vs[fidx(a, i, j + 1)]);
  const double fym = c2f(a, 1, j - 1, vs[fidx(a, i, j - 1)], vc);
  usf[fidx(a, i, j)] = fxp;
  vsf[fidx(a, i, j)] = fyp;
  if (i == 0) usf[fidx(a, -1, j)] = fxm;
  if (j == 0) vsf[fidx(a, i, -1)] = fym;
  double dv = 0.0;
  dv += (fxp - fxm) / sax(a, 0, i);
  dv += (fyp - fym) / sax(a, 1, j);
  div[fidx(a, i, j)] = dv;
  div_int[(size_t)i * a.ny + j] = dv;
}

// ---------------------------------------------------------------------
// The explicit predictor in one pass over LDS tiles: apply_bc(u, v), u*, v*
// (TVD advection + diffusion), apply_bc(u*, v*), InterpolateC2F and Div
// (s_ns_explicit.hpp:173-229, s_equation.hpp:340-378, s_operation.hpp:
// 100-120, :166-187).  A block owns a PI x PJ tile of cells.  It stages u and
// v with a three-cell halo and the face velocities with a two-cell halo,
// fills the ghost cells of the staged window from the staged interior (the
// values apply_bc writes), forms u*, v* on the tile and a one-cell ring,
// fills the ring's ghost cells of u*, v*, and then forms the tile's faces
// and divergence.  It writes u*, v*, u*f, v*f, div, the projection's
// interior-only source and the ghost cells of u, v, u*, v* that border its
// tile (corners included).  Every value is computed with the operands and
// operation order of the ghost kernels (k_ghost_*), of the cell functions
// tvd_cell / diffusion_cell, and of k_c2f_div, so the results are the bits a
// layer-by-layer evaluation gives.
struct WinAt {
  const double* t;  // cell (i, j) at t[(i - ib) * pitch + (j - jb)]
  int ib, jb, pitch;
  __device__ __forceinline__ double operator()(int i, int j) const {
    return t[(i - ib) * pitch + (j - jb)];
  }
};

// Per-position geometry of one axis over a tile's window, in LDS: for cell
// k the size s[k]; for face k (between cells k and k+1) the sum
// ss[k] = s[k] + s[k+1], the TVD ratios rc[k] = ss[k] / s[k] (upwind cell
// below the face) and rd[k] = ss[k] / s[k+1] (above), the centre spacing
// dc[k] = c[k+1] - c[k] and the C2F weight sum hh[k] = s[k]/2 + s[k+1]/2.
// Each is formed once per tile from the same operands as the reference's
// expressions (addition commutes bit for bit), so it is the value those
// compute per cell; the TVD ratio R then costs no division per face.
template <int N>
struct GeoAxis {
  double s[N], ss[N], rc[N], rd[N], dc[N], hh[N];
};

template <int N>
__device__ __forceinline__ void geo_fill(GeoAxis<N>& g, const double* __restrict__ c,
                                         const double* __restrict__ sz, int kb, int n, int gl,
                                         int tid) {
  for (int t = tid; t < N; t += 256) {
    const int k = kb + t;
    const bool ok0 = k >= -gl && k < n + gl, ok1 = k + 1 >= -gl && k + 1 < n + gl;
    const double s0 = ok0 ? sz[k + gl] : 1.0, s1 = ok1 ? sz[k + 1 + gl] : 1.0;
    const double c0 = ok0 ? c[k + gl] : 0.0, c1 = ok1 ? c[k + 1 + gl] : 1.0;
    const double sum = s0 + s1;
    g.s[t] = s0;
    g.ss[t] = sum;
    g.rc[t] = sum / s0;
    g.rd[t] = sum / s1;
    g.dc[t] = c1 - c0;
    g.hh[t] = s0 * 0.5 + s1 * 0.5;
  }
}

// tvd_cell_at's arithmetic with the geometry from LDS (GeoAxis of the x and
// y windows starting at cells xb, yb); the per-cell upwind choice and every
// rounding are the reference's (s_operation.hpp:549-596, limiters :501-546).
template <int LIM, int NX_, int NY_, class FU, class FV, class PHI>
__device__ __forceinline__ double tvd_cell_geo(const GeoAxis<NX_>& gx, const GeoAxis<NY_>& gy,
                                               int xb, int yb, const FU& uf, const FV& vf,
                                               const PHI& phi, int i, int j) {
  double vdd[2];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    const int ei = d == 0 ? 1 : 0, ej = d == 0 ? 0 : 1;
    const int c = d == 0 ? i - xb : j - yb;  // position in the axis window
    const double vfp = d == 0 ? uf(i, j) : vf(i, j);
    const double vfm = d == 0 ? uf(i - ei, j - ej) : vf(i - ei, j - ej);
    const double veo = (vfp + vfm) * 0.5;
#define VAL(o) phi(i + ei * (o), j + ej * (o))
#define GEO(f, k) (d == 0 ? gx.f[k] : gy.f[k])
    double fp, fm;
    if (veo >= 0.0) {
      // face p: C = c, D = c+1, U = c-1; face m: C = c-1, D = c, U = c-2
      double vC = VAL(0), vD = VAL(1), vU = VAL(-1);
      double R = GEO(rc, c);
      double r = (vC - vU) * GEO(ss, c) / (vD - vC + kSmall) / GEO(ss, c - 1);
      fp = vC + limiter<LIM>(r, R) / R * (vD - vC);
      vC = VAL(-1); vD = VAL(0); vU = VAL(-2);
      R = GEO(rc, c - 1);
      r = (vC - vU) * GEO(ss, c - 1) / (vD - vC + kSmall) / GEO(ss, c - 2);
      fm = vC + limiter<LIM>(r, R) / R * (vD - vC);
    } else {
      // face p: C = c+1, D = c, U = c+2; face m: C = c, D = c-1, U = c+1
      double vC = VAL(1), vD = VAL(0), vU = VAL(2);
      double R = GEO(rd, c);
      double r = (vC - vU) * GEO(ss, c) / (vD - vC + kSmall) / GEO(ss, c + 1);
      fp = vC + limiter<LIM>(r, R) / R * (vD - vC);
      vC = VAL(0); vD = VAL(-1); vU = VAL(1);
      R = GEO(rd, c - 1);
      r = (vC - vU) * GEO(ss, c - 1) / (vD - vC + kSmall) / GEO(ss, c);
      fm = vC + limiter<LIM>(r, R) / R * (vD - vC);
    }
    vdd[d] = (vfp * fp - vfm * fm) / GEO(s, c);
#undef VAL
#undef GEO
  }
  double res = 0.0;
  res += vdd[0];
  res += vdd[1];
  return res;
}

// diffusion_cell_at's arithmetic with the geometry from LDS
template <int NX_, int NY_, class PHI>
__device__ __forceinline__ double diffusion_cell_geo(const GeoAxis<NX_>& gx,
                                                     const GeoAxis<NY_>& gy, int xb, int yb,
                                                     double mu, const PHI& phi, int i, int j) {
  double td[2];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    const int ei = d == 0 ? 1 : 0, ej = d == 0 ? 0 : 1;
    const int c = d == 0 ? i - xb : j - yb;
    const double f = phi(i, j);
    const double fp = phi(i + ei, j + ej);
    const double fm = phi(i - ei, j - ej);
    const double dm = mu * (f - fm) / (d == 0 ? gx.dc[c - 1] : gy.dc[c - 1]);
    const double dp = mu * (fp - f) / (d == 0 ? gx.dc[c] : gy.dc[c]);
    td[d] = (dp - dm) / (d == 0 ? gx.s[c] : gy.s[c]);
  }
  double res = 0.0;
  res += td[0];
  res += td[1];
  return res;
}

// c2f's arithmetic (face between cells k and k+1 of a window) from LDS
template <int N>
__device__ __forceinline__ double c2f_geo(const GeoAxis<N>& g, int k, double ck, double ckp) {
  const double hs = g.s[k] * 0.5;
  const double hsp = g.s[k + 1] * 0.5;
  return (ckp * hs + ck * hsp) / g.hh[k];
}

// The tile of k_ns_predict: PI x PJ cells, u and v staged with a three-cell
// halo, the faces from two cells below the tile, u*, v* on a one-cell ring.
template <int PI, int PJ>
struct PredTile {
  static constexpr int UWI = PI + 6, UWJ = PJ + 6;  // u, v: tile +- 3
  static constexpr int FWI = PI + 3, FWJ = PJ + 3;  // faces: from tile - 2
  static constexpr int RI = PI + 2, RJ = PJ + 2;    // ring: tile +- 1
  static constexpr int RPER = (RI * RJ + 255) / 256;
  static constexpr int GX = PI + 6, GY = PJ + 6;    // axis windows: tile +- 3
};

template <int LIM, int PI, int PJ>
__global__ __launch_bounds__(256) void k_ns_predict(
    Axes a, double* __restrict__ u, double* __restrict__ v, const double* __restrict__ uf,
    const double* __restrict__ vf, double* __restrict__ us, double* __restrict__ vs,
    double* __restrict__ usf, double* __restrict__ vsf, double* __restrict__ div,
    double* __restrict__ div_int, BCDev bu, BCDev bv, const NSDev* st, double rho,
    double mu, int ntj) {
  typedef PredTile<PI, PJ> TL;
  __shared__ double su[TL::UWI * TL::UWJ], sv[TL::UWI * TL::UWJ];
  __shared__ double sf[2][TL::FWI * TL::FWJ];  // u f, v f; later u*, v* on the ring
  __shared__ GeoAxis<TL::GX> gx;
  __shared__ GeoAxis<TL::GY> gy;
  const int tid = threadIdx.x + 64 * threadIdx.y;
  const int ti = blockIdx.x / ntj, tj = blockIdx.x - ti * ntj;
  const int i0 = ti * PI, j0 = tj * PJ;
  const int nx = a.nx, ny = a.ny, gl = a.gl;
  const int xb = i0 - 3, yb = j0 - 3;  // first cell of the axis windows
  const double dt = st->dt;
  auto inbuf = [&](int i, int j) { return i >= -gl && i < nx + gl && j >= -gl && j < ny + gl; };
  auto inner = [&](int i, int j) { return i >= 0 && i < nx && j >= 0 && j < ny; };
  // ghost cells bordering the tile: rows [ib, ie) x columns [jb, je) minus
  // the interior
  const int ib = ti == 0 ? -gl : i0, ie = i0 + PI >= nx ? nx + gl : i0 + PI;
  const int jb = tj == 0 ? -gl : j0, je = j0 + PJ >= ny ? ny + gl : j0 + PJ;
  const bool edge_tile = ti == 0 || tj == 0 || i0 + PI >= nx || j0 + PJ >= ny;
  // the ring's stencils read rows i0-3 .. i0+PI+2 and columns alike
  const bool win_edge = i0 < 3 || j0 < 3 || i0 + PI + 2 >= nx || j0 + PJ + 2 >= ny;

  for (int e = tid; e < TL::UWI * TL::UWJ; e += 256) {
    const int r = e / TL::UWJ, c = e - r * TL::UWJ;
    const int i = i0 - 3 + r, j = j0 - 3 + c;
    const bool ok = inbuf(i, j);
    const size_t f = ok ? fidx(a, i, j) : 0;
    su[e] = ok ? u[f] : 0.0;
    sv[e] = ok ? v[f] : 0.0;
  }
  for (int e = tid; e < TL::FWI * TL::FWJ; e += 256) {
    const int r = e / TL::FWJ, c = e - r * TL::FWJ;
    const int i = i0 - 2 + r, j = j0 - 2 + c;
    const bool ok = inbuf(i, j);
    const size_t f = ok ? fidx(a, i, j) : 0;
    sf[0][e] = ok ? uf[f] : 0.0;
    sf[1][e] = ok ? vf[f] : 0.0;
  }
  geo_fill(gx, a.cx, a.sx, xb, nx, gl, tid);
  geo_fill(gy, a.cy, a.sy, yb, ny, gl, tid);
  __syncthreads();
  const WinAt U{su, i0 - 3, j0 - 3, TL::UWJ}, V{sv, i0 - 3, j0 - 3, TL::UWJ};
  // apply_bc(u, v): the window's ghost cells (reads interior cells only,
  // writes ghost cells only) and the tile's own ghost cells in memory
  if (win_edge) {
    for (int e = tid; e < TL::UWI * TL::UWJ; e += 256) {
      const int r = e / TL::UWJ, c = e - r * TL::UWJ;
      const int i = i0 - 3 + r, j = j0 - 3 + c;
      if (inbuf(i, j) && !inner(i, j)) {
        su[e] = ghost_of(a, bu, i, j, U);
        sv[e] = ghost_of(a, bv, i, j, V);
      }
    }
  }
  if (edge_tile) {
    const int w = je - jb;
    for (int e = tid; e < (ie - ib) * w; e += 256) {
      const int i = ib + e / w, j = jb + e % w;
      if (!inner(i, j)) {
        u[fidx(a, i, j)] = ghost_of(a, bu, i, j, U);
        v[fidx(a, i, j)] = ghost_of(a, bv, i, j, V);
      }
    }
  }
  __syncthreads();
  // u*, v* on the ring's interior cells (s_ns_explicit.hpp:196-211)
  const WinAt UF{sf[0], i0 - 2, j0 - 2, TL::FWJ}, VF{sf[1], i0 - 2, j0 - 2, TL::FWJ};
  double ru[TL::RPER], rv[TL::RPER];
#pragma unroll
  for (int k = 0; k < TL::RPER; ++k) {
    const int e = tid + 256 * k;
    const int r = e / TL::RJ, c = e - r * TL::RJ;
    const int i = i0 - 1 + r, j = j0 - 1 + c;
    ru[k] = rv[k] = 0.0;
    if (e < TL::RI * TL::RJ && inner(i, j)) {
      const double src = 0.0;
      double adv = tvd_cell_geo<LIM>(gx, gy, xb, yb, UF, VF, U, i, j);
      double dif = diffusion_cell_geo(gx, gy, xb, yb, mu, U, i, j);
      ru[k] = (-adv + (dif + src) / rho) * dt + U(i, j);
      adv = tvd_cell_geo<LIM>(gx, gy, xb, yb, UF, VF, V, i, j);
      dif = diffusion_cell_geo(gx, gy, xb, yb, mu, V, i, j);
      rv[k] = (-adv + (dif + src) / rho) * dt + V(i, j);
    }
  }
  __syncthreads();  // every read of the face window done
  double* rus = sf[0];
  double* rvs = sf[1];
#pragma unroll
  for (int k = 0; k < TL::RPER; ++k) {
    const int e = tid + 256 * k;
    if (e < TL::RI * TL::RJ) {
      rus[e] = ru[k];
      rvs[e] = rv[k];
    }
  }
  __syncthreads();
  const WinAt US{rus, i0 - 1, j0 - 1, TL::RJ}, VS{rvs, i0 - 1, j0 - 1, TL::RJ};
  if (edge_tile) {
    // apply_bc(u*, v*): the ring's edge ghosts (the faces read them) and the
    // tile's own ghost cells in memory
    for (int e = tid; e < TL::RI * TL::RJ; e += 256) {
      const int r = e / TL::RJ, c = e - r * TL::RJ;
      const int i = i0 - 1 + r, j = j0 - 1 + c;
      const bool iin = i >= 0 && i < nx, jin = j >= 0 && j < ny;
      if (inbuf(i, j) && iin != jin) {
        rus[e] = ghost_of(a, bu, i, j, US);
        rvs[e] = ghost_of(a, bv, i, j, VS);
      }
    }
    const int w = je - jb;
    for (int e = tid; e < (ie - ib) * w; e += 256) {
      const int i = ib + e / w, j = jb + e % w;
      if (!inner(i, j)) {
        us[fidx(a, i, j)] = ghost_of(a, bu, i, j, US);
        vs[fidx(a, i, j)] = ghost_of(a, bv, i, j, VS);
      }
    }
  }
  __syncthreads();
  // faces and divergence of the tile (k_c2f_div's arithmetic)
  for (int e = tid; e < PI * PJ; e += 256) {
    const int i = i0 + e / PJ, j = j0 + e % PJ;
    if (i >= nx || j >= ny) continue;
    const int cx = i - xb, cy = j - yb;
    const double uc = US(i, j), vc = VS(i, j);
    const double fxp = c2f_geo(gx, cx, uc, US(i + 1, j));
    const double fxm = c2f_geo(gx, cx - 1, US(i - 1, j), uc);
    const double fyp = c2f_geo(gy, cy, vc, VS(i, j + 1));
    const double fym = c2f_geo(gy, cy - 1, VS(i, j - 1), vc);
    const size_t f = fidx(a, i, j);
    us[f] = uc;
    vs[f] = vc;
    usf[f] = fxp;
    vsf[f] = fyp;
    if (i == 0) usf[fidx(a, -1, j)] = fxm;
    if (j == 0) vsf[fidx(a, i, -1)] = fym;
    double dv = 0.0;
    dv += (fxp - fxm) / gx.s[cx];
    dv += (fyp - fym) / gy.s[cy];
    div[f] = dv;
    div_int[(size_t)i * ny + j] = dv;
  }
}

// Grad of phi (Neumann value 0 ghosts equal the boundary cell), p = phi/dt,
// face and centre corrections (s_ns_explicit.hpp:140-153).  phi is the
// projection's interior-only vector.
__device__ __forceinline__ double phi_at(const Axes& a, const double* __restrict__ x,
                                         int i, int j) {
  i = i < 0 ? 0 : (i >= a.nx ? a.nx - 1 : i);
  j = j < 0 ? 0 : (j >= a.ny ? a.ny - 1 : j);
  return x[(size_t)i * a.ny + j];
}

// KIM: p also loses (0.5/rho) * NablaMuNabla(mu, phi) (s_ns_kim.hpp:180-192).
template <bool KIM>
__global__ __launch_bounds__(256) void k_grad_correct(Axes a, const double* __restrict__ x,
                                                      double* __restrict__ usf,
                                                      double* __restrict__ vsf,
                                                      double* __restrict__ us,
                                                      double* __restrict__ vs,
                                                      double* __restrict__ p,
                                                      double* __restrict__ phi,
                                                      const NSDev* st, double mu,
                                                      double rho) {
  const int J = blockIdx.x * blockDim.x + threadIdx.x;
  const int I = blockIdx.y * blockDim.y + threadIdx.y;
  if (I >= a.NX || J >= a.NY) return;
  const int i = I - a.gl, j = J - a.gl;
  const bool inx = i >= 0 && i < a.nx, iny = j >= 0 && j < a.ny;
  const size_t f = (size_t)I * a.NY + J;
  if (!inx && !iny) {  // corner: p and phi carry the ghost value
    const double ph = phi_at(a, x, i, j);
    phi[f] = ph;
    p[f] = ph;
    return;
  }
  if (!inx || !iny) {  // edge ghost
    const double ph = phi_at(a, x, i, j);
    phi[f] = ph;
    p[f] = ph;
    // the m-face slot of the boundary row/column carries grad phi
    if (i == -1 && iny) {
      const double g = (ph - phi_at(a, x, 0, j)) / (cax(a, 0, -1) - cax(a, 0, 0));
      usf[f] = usf[f] - g;
    }
    if (j == -1 && inx) {
      const double g = (ph - phi_at(a, x, i, 0)) / (cax(a, 1, -1) - cax(a, 1, 0));
      vsf[f] = vsf[f] - g;
    }
    return;
  }
  const double ph = x[(size_t)i * a.ny + j];
  const double pxp = phi_at(a, x, i + 1, j), pxm = phi_at(a, x, i - 1, j);
  const double pyp = phi_at(a, x, i, j + 1), pym = phi_at(a, x, i, j - 1);
  const double gxp = (pxp - ph) / (cax(a, 0, i + 1) - cax(a, 0, i));
  const double gyp = (pyp - ph) / (cax(a, 1, j + 1) - cax(a, 1, j));
  const double gxm = (ph - pxm) / (cax(a, 0, i) - cax(a, 0, i - 1));
  const double gym = (ph - pym) / (cax(a, 1, j) - cax(a, 1, j - 1));
  phi[f] = ph;
  double pv = ph * (1.0 / st->dt);
  if (KIM) {
    // NablaMuNabla(mu, phi) over phi's Neumann ghosts (s_operation.hpp:664-688)
    const double dmx = mu * (ph - pxm) / (cax(a, 0, i) - cax(a, 0, i - 1));
    const double dpx = mu * (pxp - ph) / (cax(a, 0, i + 1) - cax(a, 0, i));
    const double dmy = mu * (ph - pym) / (cax(a, 1, j) - cax(a, 1, j - 1));
    const double dpy = mu * (pyp - ph) / (cax(a, 1, j + 1) - cax(a, 1, j));
    double lap = 0.0;
    lap += (dpx - dmx) / sax(a, 0, i);
    lap += (dpy - dmy) / sax(a, 1, j);
    pv = pv - lap * (0.5 / rho);
  }
  p[f] = pv;
  usf[f] = usf[f] - gxp;
  vsf[f] = vsf[f] - gyp;
  us[f] = us[f] - (gxp + gxm) * 0.5;
  vs[f] = vs[f] - (gyp + gym) * 0.5;
}

// VdotNabla_upwind1 at one cell (s_operation.hpp:400-422), the NS_ default
// advection (s_ns.hpp:225); faces as in k_c2f_div (uf at (i, j) is the
// x+ face of cell (i, j)).
template <class PHI>
__device__ __forceinline__ double upwind1_cell(const Axes& a, const double* __restrict__ uf,
                                               const double* __restrict__ vf,
                                               const PHI& s, int i, int j) {
  double t[2] = {0.0, 0.0};
  const double c = s(i, j);
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    const int ei = d == 0 ? 1 : 0, ej = d == 0 ? 0 : 1, k = d == 0 ? i : j;
    const double* f = d == 0 ? uf : vf;
    const double veo = (f[fidx(a, i, j)] + f[fidx(a, i - ei, j - ej)]) * 0.5;
    if (veo > 0.0)
      t[d] = veo * (c - s(i - ei, j - ej)) / (cax(a, d, k) - cax(a, d, k - 1));
    else if (veo < 0.0)
      t[d] = veo * (s(i + ei, j + ej) - c) / (cax(a, d, k + 1) - cax(a, d, k));
  }
  double r = 0.0;
  r += t[0];
  r += t[1];
  return r;
}

// NS_kim_ predictor source of both components, interior-only (i*ny + j)
// (s_ns_kim.hpp:238-265): (adv - (0.5*dif + 0)/rho)*dt - u with adv the
// upwind term at the first step and 1.5*a - 0.5*a' after.  a' is the upwind
// term of um on umf: run_one_step copies u (ghosts included) into um and uf
// into umf before apply_bc(u) refreshes u's ghosts (:132-133, :225-227), so
// um differs from u in its ghost cells only (they keep the walls applied to
// the previous step's u* before its correction) and umf equals uf.
__global__ __launch_bounds__(256) void k_kim_src(Axes a, const double* __restrict__ u,
                                                 const double* __restrict__ v,
                                                 const double* __restrict__ um,
                                                 const double* __restrict__ vm,
                                                 const double* __restrict__ uf,
                                                 const double* __restrict__ vf,
                                                 double* __restrict__ su,
                                                 double* __restrict__ sv,
                                                 const NSDev* st, double rho, double mu,
                                                 int first) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= a.nx || j >= a.ny) return;
  const double dt = st->dt;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const double* s = c == 0 ? u : v;
    const double advn = upwind1_cell(a, uf, vf, FieldAt{s, a}, i, j);
    const double advnm1 = upwind1_cell(a, uf, vf, FieldAt{c == 0 ? um : vm, a}, i, j);
    const double adv = 1.5 * advn - 0.5 * advnm1;
    const double dif = diffusion_cell(a, mu, s, i, j);
    const double src = 0.0;
    const double cur = s[fidx(a, i, j)];
    const double r = first ? (advn - (0.5 * dif + src) / rho) * dt - cur
                           : (adv - (0.5 * dif + src) / rho) * dt - cur;
    (c == 0 ? su : sv)[(size_t)i * a.ny + j] = r;
  }
}

// interior of two ghost-padded fields <-> two interior-only vectors
__global__ __launch_bounds__(256) void k_field_to_x2(Axes a, const double* __restrict__ f0,
                                                     const double* __restrict__ f1,
                                                     double* __restrict__ x0,
                                                     double* __restrict__ x1) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= a.nx || j >= a.ny) return;
  x0[(size_t)i * a.ny + j] = f0[fidx(a, i, j)];
  x1[(size_t)i * a.ny + j] = f1[fidx(a, i, j)];
}

__global__ __launch_bounds__(256) void k_x_to_field2(Axes a, const double* __restrict__ x0,
                                                     const double* __restrict__ x1,
                                                     double* __restrict__ f0,
                                                     double* __restrict__ f1) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= a.nx || j >= a.ny) return;
  f0[fidx(a, i, j)] = x0[(size_t)i * a.ny + j];
  f1[fidx(a, i, j)] = x1[(size_t)i * a.ny + j];
}

__global__ void k_advance(NSDev* st) {
  if (threadIdx.x == 0) st->t += st->dt;
}

// ---------------------------------------------------------------------
// Advection_ (s_advection.hpp:20-217): explicit Euler steps of
// d phi/dt + v . grad phi = 0 with a fixed face velocity.

// VdotNabla_center (s_operation.hpp:442-459)
template <class PHI>
__device__ __forceinline__ double center_cell(const Axes& a, const double* __restrict__ uf,
                                              const double* __restrict__ vf, const PHI& s,
                                              int i, int j) {
  double t[2];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    const int ei = d == 0 ? 1 : 0, ej = d == 0 ? 0 : 1, k = d == 0 ? i : j;
    const double* f = d == 0 ? uf : vf;
    const double veo = (f[fidx(a, i, j)] + f[fidx(a, i - ei, j - ej)]) * 0.5;
    t[d] = veo * (s(i + ei, j + ej) - s(i - ei, j - ej)) / (cax(a, d, k + 1) - cax(a, d, k - 1));
  }
  double r = 0.0;
  r += t[0];
  r += t[1];
  return r;
}

// VdotNabla_center4 (s_operation.hpp:461-480).  The reference forms the
// averaged velocity and leaves it out of the result: the term is the
// fourth-order gradient alone, reproduced as such.
template <class PHI>
__device__ __forceinline__ double center4_cell(const Axes& a, const PHI& s, int i, int j) {
  double t[2];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    const int ei = d == 0 ? 1 : 0, ej = d == 0 ? 0 : 1, k = d == 0 ? i : j;
    const double inn = (s(i + ei, j + ej) - s(i - ei, j - ej)) /
                       (cax(a, d, k + 1) - cax(a, d, k - 1));
    const double out = (s(i + 2 * ei, j + 2 * ej) - s(i - 2 * ei, j - 2 * ej)) /
                       (cax(a, d, k + 2) - cax(a, d, k - 2));
    t[d] = (4.0 * inn - out) / 3.0;
  }
  double r = 0.0;
  r += t[0];
  r += t[1];
  return r;
}

// the advection term of scheme SCH (CPGPU_ADV_*: the limiter ids select the
// TVD family)
template <int SCH, class PHI>
__device__ __forceinline__ double adv_cell(const Axes& a, const double* __restrict__ uf,
                                           const double* __restrict__ vf, const PHI& s, int i,
                                           int j) {
  if constexpr (SCH == CPGPU_ADV_UPWIND1) return upwind1_cell(a, uf, vf, s, i, j);
  else if constexpr (SCH == CPGPU_ADV_CENTER) return center_cell(a, uf, vf, s, i, j);
  else if constexpr (SCH == CPGPU_ADV_CENTER4) return center4_cell(a, s, i, j);
  else return tvd_cell<SCH>(a, uf, vf, s, i, j);
}

// One tile diagonal of the in-place sweep of Advection_::run_one_step
// (s_advection.hpp:200-210).  The reference visits cells i-fastest and
// overwrites phi as it goes, so a cell sees the new values of the cells
// before it in that order and the old values of those after it.  Its
// stencil is a cross of radius two, whose cells before it are (i - 1, j),
// (i - 2, j), (i, j - 1), (i, j - 2): along anti-diagonals i + j those lie on
// the two previous diagonals and the cells after it on the next two.
// Sweeping anti-diagonals in order therefore reproduces the reference's
// values exactly; at the tile level, tiles (ti, tj) of one diagonal
// ti + tj = D are independent (no cross stencil spans two of them), and the
// tiles of diagonals D - 1 (updated) and D + 1 (untouched) are what the
// cross reads.  One wavefront per 16 x 16 tile stages its cells and a
// two-cell halo in LDS, then walks the tile's 31 local anti-diagonals.
template <int SCH>
__global__ __launch_bounds__(64) void k_adv_sweep(Axes a, const double* __restrict__ uf,
                                                  const double* __restrict__ vf,
                                                  double* __restrict__ phi, double dt, int D,
                                                  int nti, int ntj) {
  constexpr int T = 16, H = TileAt::H, TP = TileAt::TP;
  __shared__ double w[TP * TP];
  const int ti = max(0, D - (ntj - 1)) + (int)blockIdx.x;
  const int tj = D - ti;
  if (ti >= nti || tj < 0) return;
  const int i0 = ti * T, j0 = tj * T;
  // stage the window (cells outside the field's ghost frame are never read)
  for (int e = threadIdx.x; e < TP * TP; e += 64) {
    const int r = e / TP, c = e - r * TP;
    const int i = i0 - H + r, j = j0 - H + c;
    double v = 0.0;
    if (i >= -a.gl && i < a.nx + a.gl && j >= -a.gl && j < a.ny + a.gl) v = phi[fidx(a, i, j)];
    w[e] = v;
  }
  __syncthreads();
  const TileAt at{w, i0, j0};
  for (int d = 0; d < 2 * T - 1; ++d) {
    const int li = threadIdx.x, lj = d - li;
    const int i = i0 + li, j = j0 + lj;
    if (li < T && lj >= 0 && lj < T && i < a.nx && j < a.ny) {
      const double adv = adv_cell<SCH>(a, uf, vf, at, i, j);
      const double phin = at(i, j);
      w[(li + H) * TP + (lj + H)] = phin - dt * adv;
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < T * T; e += 64) {
    const int li = e / T, lj = e - li * T;
    const int i = i0 + li, j = j0 + lj;
    if (i < a.nx && j < a.ny) phi[fidx(a, i, j)] = w[(li + H) * TP + (lj + H)];
  }
}

// face velocities from the centre velocities (InterpolateC2F,
// s_operation.hpp:100-120): every face of the field, the m-face of the first
// cell included
__global__ __launch_bounds__(256) void k_adv_faces(Axes a, const double* __restrict__ u,
                                                   const double* __restrict__ v,
                                                   double* __restrict__ uf,
                                                   double* __restrict__ vf) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= a.nx || j >= a.ny) return;
  const double uc = u[fidx(a, i, j)], vc = v[fidx(a, i, j)];
  uf[fidx(a, i, j)] = c2f(a, 0, i, uc, u[fidx(a, i + 1, j)]);
  vf[fidx(a, i, j)] = c2f(a, 1, j, vc, v[fidx(a, i, j + 1)]);
  if (i == 0) uf[fidx(a, -1, j)] = c2f(a, 0, -1, u[fidx(a, -1, j)], uc);
  if (j == 0) vf[fidx(a, i, -1)] = c2f(a, 1, -1, v[fidx(a, i, -1)], vc);
}

}  // namespace cpgpu

using namespace cpgpu;

struct cpgpu_ns {
  cpgpu_ctx* ctx = nullptr;
  int nx = 0, ny = 0, gl = 0, NX = 0, NY = 0;
  size_t NT = 0;
  cpgpu_poisson* P = nullptr;  // all-Neumann projection problem
  DevBuf<double> cx, sx, cy, sy;
  DevBuf<double> bcu_val, bcv_val;
  int bcu_type[4], bcv_type[4];
  // fields: u, v, uf, vf, us, vs, usf, vsf, p, phi, div
  DevBuf<double> F[11];
  int slot[11];  // field id -> buffer (u/us and uf/usf swap every step)
  DevBuf<double> div_int, part;
  DevBuf<double> err_ref;  // exact field of cpgpu_ns_error_norms
  DevBuf<double> prev[11];  // snapshots of cpgpu_ns_field_change, per field id
  bool has_prev[11] = {false, false, false, false, false, false,
                       false, false, false, false, false};
  DevBuf<NSDev> st;
  NSDev* h_st = nullptr;
  double rho, mu, cfl, dnum, dcoef, min_size, ddenom;
  int limiter, solver, max_iter, norm_mode, opts;
  double tol;
  int last_iters = 0;
  double last_resid = 0.0;
  long long total_iters = 0;
  // host copies of the axes and velocity boundary values (Helmholtz set-up)
  std::vector<double> hcx, hsx, hcy, hsy, hbc[2];
  // NS_kim_ mode: Helmholtz predictor problems (alpha 1, beta dt*0.5/rho),
  // rebuilt when dt changes
  bool kim = false;
  int kim_solver = CPGPU_JACOBI, kim_max_iter = 1000;
  double kim_tol = 1e-4, kim_dt = -1.0;
  long long kim_step = 0;
  cpgpu_poisson* H[2] = {nullptr, nullptr};
  DevBuf<double> ksrc[2];
  DevBuf<double> kum[2];  // u, v as the step found them (ghosts included)
  int helm_iters[2] = {0, 0};

  double* fld(int id) { return F[slot[id]].p; }
  Axes axes() const {
    Axes a;
    a.nx = nx; a.ny = ny; a.gl = gl; a.NX = NX; a.NY = NY;
    a.cx = cx.p; a.sx = sx.p; a.cy = cy.p; a.sy = sy.p;
    return a;
  }
  ~cpgpu_ns() {
    if (P) cpgpu_poisson_destroy(P);
    for (int c = 0; c < 2; ++c)
      if (H[c]) cpgpu_poisson_destroy(H[c]);
    if (h_st) (void)hipHostFree(h_st);
  }
};

namespace {

enum { F_U, F_V, F_UF, F_VF, F_US, F_VS, F_USF, F_VSF, F_P, F_PHI, F_DIV };

int ns_upload(cpgpu_ctx* ctx, void* dst, const void* src, size_t bytes) {
  CPGPU_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice,
                                ctx->stream));
  return CPGPU_OK;
}

template <int LIM, int PI, int PJ>
void launch_predict_t(cpgpu_ns* S, const BCDev& bu, const BCDev& bv) {
  const int ntj = (S->ny + PJ - 1) / PJ, nti = (S->nx + PI - 1) / PI;
  const EvPair ev = timer_events(S->ctx, 4);
  CPGPU_LAUNCH(ev, (k_ns_predict<LIM, PI, PJ>), dim3(nti * ntj), dim3(64, 4), 0,
               S->ctx->stream, S->axes(), S->fld(F_U), S->fld(F_V), S->fld(F_UF),
               S->fld(F_VF), S->fld(F_US), S->fld(F_VS), S->fld(F_USF), S->fld(F_VSF),
               S->fld(F_DIV), S->div_int.p, bu, bv, S->st.p, S->rho, S->mu, ntj);
}

// 8 x 64 tiles from 1024^2 cells on, 8 x 32 below (more blocks); measured
// at 4096^2 against 16 x 64 and 8 x 32 tiles: 1.09, 1.20, 1.15 ms
template <int LIM>
void launch_predict(cpgpu_ns* S, const BCDev& bu, const BCDev& bv) {
  if ((size_t)S->nx * S->ny >= (size_t)1024 * 1024)
    launch_predict_t<LIM, 8, 64>(S, bu, bv);
  else
    launch_predict_t<LIM, 8, 32>(S, bu, bv);
}

int ns_project(cpgpu_ns* S, bool kim, bool faces_done = false);

int ns_one_step(cpgpu_ns* S) {
  cpgpu_ctx* ctx = S->ctx;
  hipStream_t st = ctx->stream;
  const Axes a = S->axes();
  BCDev bu{{S->bcu_type[0], S->bcu_type[1], S->bcu_type[2], S->bcu_type[3]},
           S->bcu_val.p};
  BCDev bv{{S->bcv_type[0], S->bcv_type[1], S->bcv_type[2], S->bcv_type[3]},
           S->bcv_val.p};
  const int Gm = 512;
  // restrict_time
  hipLaunchKernelGGL(k_maxmag, dim3(Gm), dim3(256), 0, st, a, S->fld(F_U),
                     S->fld(F_V), S->part.p);
  hipLaunchKernelGGL(k_dt, dim3(1), dim3(256), 0, st, S->part.p, Gm, S->st.p,
                     S->cfl, S->min_size, S->dnum, S->ddenom, S->dcoef);
  // apply_bc(u, v), u*, v*, apply_bc(u*, v*), faces and divergence
  switch (S->limiter) {
    case 0: launch_predict<0>(S, bu, bv); break;
    case 1: launch_predict<1>(S, bu, bv); break;
    case 2: launch_predict<2>(S, bu, bv); break;
    case 3: launch_predict<3>(S, bu, bv); break;
    case 4: launch_predict<4>(S, bu, bv); break;
    default: launch_predict<5>(S, bu, bv); break;
  }
  CPGPU_HIP(ctx, hipGetLastError());
  return ns_project(S, false, true);
}

// C2F + Div of u* (unless the predictor formed them), the projection solve,
// grad phi, p, the corrections and the time advance; u <- u*, uf <- u*f.
int ns_project(cpgpu_ns* S, bool kim, bool faces_done) {
  cpgpu_ctx* ctx = S->ctx;
  hipStream_t st = ctx->stream;
  const Axes a = S->axes();
  const dim3 b2(64, 4), g2((S->ny + 63) / 64, (S->nx + 3) / 4);
  if (!faces_done) {
    const EvPair ev5 = timer_events(ctx, 5);
    CPGPU_LAUNCH(ev5, k_c2f_div, g2, b2, 0, st, a, S->fld(F_US), S->fld(F_VS),
                 S->fld(F_USF), S->fld(F_VSF), S->fld(F_DIV), S->div_int.p);
    CPGPU_HIP(ctx, hipGetLastError());
  }
  // projection
  int rc = poisson_rhs_from_device(S->P, S->div_int.p, 0.0);
  if (rc) return rc;
  int mi = S->max_iter;
  double tol = S->tol;
  int hl = 0;
  int code = poisson_solve_internal(S->P, S->solver, &mi, &tol, S->norm_mode,
                                    S->opts, nullptr, 0, &hl);
  if (code < 0) return code;
  S->last_iters = mi;
  S->last_resid = tol;
  S->total_iters += mi;
  // grad, corrections, p
  const dim3 g3((S->NY + 63) / 64, (S->NX + 3) / 4);
  const EvPair ev6 = timer_events(ctx, 6);
  if (kim)
    CPGPU_LAUNCH(ev6, k_grad_correct<true>, g3, b2, 0, st, a, S->P->x.p,
                 S->fld(F_USF), S->fld(F_VSF), S->fld(F_US), S->fld(F_VS),
                 S->fld(F_P), S->fld(F_PHI), S->st.p, S->mu, S->rho);
  else
    CPGPU_LAUNCH(ev6, k_grad_correct<false>, g3, b2, 0, st, a, S->P->x.p,
                 S->fld(F_USF), S->fld(F_VSF), S->fld(F_US), S->fld(F_VS),
                 S->fld(F_P), S->fld(F_PHI), S->st.p, S->mu, S->rho);
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, S->st.p);
  CPGPU_HIP(ctx, hipGetLastError());
  // u <- u*, uf <- u*f by exchanging the buffers
  std::swap(S->slot[F_U], S->slot[F_US]);
  std::swap(S->slot[F_V], S->slot[F_VS]);
  std::swap(S->slot[F_UF], S->slot[F_USF]);
  std::swap(S->slot[F_VF], S->slot[F_VSF]);
  return code;
}

// The two Helmholtz predictor problems of the current dt (Poisson_ with
// set_uniform_alpha(1) and uniform beta dt*0.5/rho, s_ns_kim.hpp:267-283).
// The predictor is a Poisson_ of its own (:105) whose BoundaryIndex stays
// empty (s_equation.hpp:104), so every wall of its rows takes the default
// condition, BC2 with value 0 (s_boundary.hpp:29-36, :195-207); the
// velocity walls reach u* only through apply_bc(us) after the solves.
int ns_kim_build(cpgpu_ns* S) {
  const double dt = S->h_st->dt;
  if (S->H[0] && S->H[1] && S->kim_dt == dt) return CPGPU_OK;
  for (int c = 0; c < 2; ++c)
    if (S->H[c]) {
      cpgpu_poisson_destroy(S->H[c]);
      S->H[c] = nullptr;
    }
  cpgpu_grid2d gd;
  gd.nx = S->nx; gd.ny = S->ny; gd.gl = S->gl;
  gd.cx = S->hcx.data(); gd.sx = S->hsx.data();
  gd.cy = S->hcy.data(); gd.sy = S->hsy.data();
  gd.beta = dt * 0.5 / S->rho;
  for (int k = 0; k < 4; ++k) gd.bc_type[k] = CPGPU_BC_NEUMANN;
  gd.bc_val = nullptr;
  for (int c = 0; c < 2; ++c) {
    int rc = cpgpu_poisson_create(S->ctx, &gd, &S->H[c]);
    if (rc) return rc;
    rc = cpgpu_poisson_set_alpha(S->H[c], 1.0);
    if (rc) return rc;
  }
  S->kim_dt = dt;
  return CPGPU_OK;
}

// NS_kim_::run_one_step + Time_::advance (s_ns_kim.hpp:130-205): no
// restrict_time; the predictor solves start from the current u*.
int ns_kim_step(cpgpu_ns* S) {
  cpgpu_ctx* ctx = S->ctx;
  hipStream_t st = ctx->stream;
  const Axes a = S->axes();
  int rc = ns_kim_build(S);
  if (rc) return rc;
  const size_t n = (size_t)S->nx * S->ny;
  if (S->ksrc[0].alloc(n) || S->ksrc[1].alloc(n) || S->kum[0].alloc(S->NT) ||
      S->kum[1].alloc(S->NT))
    return set_error(ctx, CPGPU_ENOMEM, "kim source");
  // um <- u, vm <- v before the ghost refresh (Operation_::Copy, IJK sweep)
  for (int c = 0; c < 2; ++c)
    CPGPU_HIP(ctx, hipMemcpyAsync(S->kum[c].p, S->fld(F_U + c), sizeof(double) * S->NT,
                                  hipMemcpyDeviceToDevice, st));
  BCDev bu{{S->bcu_type[0], S->bcu_type[1], S->bcu_type[2], S->bcu_type[3]},
           S->bcu_val.p};
  BCDev bv{{S->bcv_type[0], S->bcv_type[1], S->bcv_type[2], S->bcv_type[3]},
           S->bcv_val.p};
  const dim3 gedge(64, 2), gcorner(1, 2);
  const dim3 b2(64, 4), g2((S->ny + 63) / 64, (S->nx + 3) / 4);
  hipLaunchKernelGGL(k_ghost_edges, gedge, dim3(256), 0, st, a, S->fld(F_U), bu,
                     S->fld(F_V), bv);
  hipLaunchKernelGGL(k_ghost_corners, gcorner, dim3(256), 0, st, a, S->fld(F_U),
                     bu, S->fld(F_V), bv);
  hipLaunchKernelGGL(k_kim_src, g2, b2, 0, st, a, S->fld(F_U), S->fld(F_V),
                     S->kum[0].p, S->kum[1].p,
                     S->fld(F_UF), S->fld(F_VF), S->ksrc[0].p, S->ksrc[1].p,
                     S->st.p, S->rho, S->mu, S->kim_step == 0 ? 1 : 0);
  hipLaunchKernelGGL(k_field_to_x2, g2, b2, 0, st, a, S->fld(F_US), S->fld(F_VS),
                     S->H[0]->x.p, S->H[1]->x.p);
  CPGPU_HIP(ctx, hipGetLastError());
  for (int c = 0; c < 2; ++c) {
    rc = poisson_rhs_from_device(S->H[c], S->ksrc[c].p, 0.0);
    if (rc) return rc;
    int mi = S->kim_max_iter;
    double tol = S->kim_tol;
    int hl = 0;
    const int code = poisson_solve_internal(S->H[c], S->kim_solver, &mi, &tol,
                                            S->norm_mode, 0, nullptr, 0, &hl);
    if (code < 0) return code;
    S->helm_iters[c] = mi;
  }
  hipLaunchKernelGGL(k_x_to_field2, g2, b2, 0, st, a, S->H[0]->x.p, S->H[1]->x.p,
                     S->fld(F_US), S->fld(F_VS));
  hipLaunchKernelGGL(k_ghost_edges, gedge, dim3(256), 0, st, a, S->fld(F_US),
                     bu, S->fld(F_VS), bv);
  hipLaunchKernelGGL(k_ghost_corners, gcorner, dim3(256), 0, st, a,
                     S->fld(F_US), bu, S->fld(F_VS), bv);
  CPGPU_HIP(ctx, hipGetLastError());
  const int code = ns_project(S, true);
  if (code < 0) return code;
  // the reference copies u* into u (whole arrays); u* keeps the same values
  const size_t bytes = sizeof(double) * S->NT;
  for (int k = 0; k < 4; ++k)
    CPGPU_HIP(ctx, hipMemcpyAsync(S->fld(F_US + k), S->fld(F_U + k), bytes,
                                  hipMemcpyDeviceToDevice, st));
  S->kim_step++;
  return code;
}

}  // namespace

namespace {
bool ns_solver_ok(int m) {
  return m == CPGPU_JPCG || m == CPGPU_MIXED || m == CPGPU_CG ||
         m == CPGPU_FCG_UNET;
}
// the methods that are not the reference's report a zero-mean phi
int ns_solver_opts(int m) {
  return (m == CPGPU_MIXED || m == CPGPU_FCG_UNET) ? CPGPU_OPT_PROJECT_MEAN : 0;
}
}  // namespace

extern "C" {

int cpgpu_ns_create(cpgpu_ctx* ctx, const cpgpu_ns_desc* d, cpgpu_ns** out) {
  if (!ctx || !d || !out) return CPGPU_EINVAL;
  *out = nullptr;
  const cpgpu_grid2d& g = d->grid;
  if (g.nx < 1 || g.ny < 1 || g.gl < 2 || !g.cx || !g.sx || !g.cy || !g.sy)
    return set_error(ctx, CPGPU_EINVAL, "NS needs gl >= 2 and axis arrays");
  if (d->limiter < 0 || d->limiter > 5 || !(d->rho > 0.0) ||
      !(d->dt > 0.0) || !ns_solver_ok(d->solver))
    return set_error(ctx, CPGPU_EINVAL, "bad NS parameters");
  for (int s = 0; s < 4; ++s)
    if ((d->bcu_type[s] != 1 && d->bcu_type[s] != 2) ||
        (d->bcv_type[s] != 1 && d->bcv_type[s] != 2))
      return set_error(ctx, CPGPU_EINVAL, "bad velocity bc type");
  CPGPU_HIP(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<cpgpu_ns> S(new cpgpu_ns());
  S->ctx = ctx;
  S->nx = g.nx; S->ny = g.ny; S->gl = g.gl;
  S->NX = g.nx + 2 * g.gl; S->NY = g.ny + 2 * g.gl;
  S->NT = (size_t)S->NX * S->NY;
  // projection: all-Neumann, value 0 (an empty BoundaryIndex, s_ns.hpp:222)
  cpgpu_grid2d pg = g;
  pg.beta = 1.0;
  for (int s = 0; s < 4; ++s) pg.bc_type[s] = CPGPU_BC_NEUMANN;
  pg.bc_val = nullptr;
  int rc = cpgpu_poisson_create(ctx, &pg, &S->P);
  if (rc) return rc;
  for (int k = 0; k < 11; ++k) {
    if (S->F[k].alloc(S->NT)) return set_error(ctx, CPGPU_ENOMEM, "ns fields");
    CPGPU_HIP(ctx, hipMemsetAsync(S->F[k].p, 0, sizeof(double) * S->NT,
                                  ctx->stream));
    S->slot[k] = k;
  }
  const size_t nbc = 2 * (size_t)S->NY + 2 * (size_t)S->NX;
  if (S->cx.alloc(S->NX) || S->sx.alloc(S->NX) || S->cy.alloc(S->NY) ||
      S->sy.alloc(S->NY) || S->bcu_val.alloc(nbc) || S->bcv_val.alloc(nbc) ||
      S->div_int.alloc((size_t)g.nx * g.ny) || S->part.alloc(1024) ||
      S->st.alloc(1))
    return set_error(ctx, CPGPU_ENOMEM, "ns buffers");
  std::vector<double> zero(nbc, 0.0);
  if ((rc = ns_upload(ctx, S->cx.p, g.cx, sizeof(double) * S->NX)) ||
      (rc = ns_upload(ctx, S->sx.p, g.sx, sizeof(double) * S->NX)) ||
      (rc = ns_upload(ctx, S->cy.p, g.cy, sizeof(double) * S->NY)) ||
      (rc = ns_upload(ctx, S->sy.p, g.sy, sizeof(double) * S->NY)) ||
      (rc = ns_upload(ctx, S->bcu_val.p, d->bcu_val ? d->bcu_val : zero.data(),
                      sizeof(double) * nbc)) ||
      (rc = ns_upload(ctx, S->bcv_val.p, d->bcv_val ? d->bcv_val : zero.data(),
                      sizeof(double) * nbc)))
    return rc;
  for (int s = 0; s < 4; ++s) {
    S->bcu_type[s] = d->bcu_type[s];
    S->bcv_type[s] = d->bcv_type[s];
  }
  S->hcx.assign(g.cx, g.cx + S->NX);
  S->hsx.assign(g.sx, g.sx + S->NX);
  S->hcy.assign(g.cy, g.cy + S->NY);
  S->hsy.assign(g.sy, g.sy + S->NY);
  S->hbc[0].assign(d->bcu_val ? d->bcu_val : zero.data(),
                   (d->bcu_val ? d->bcu_val : zero.data()) + nbc);
  S->hbc[1].assign(d->bcv_val ? d->bcv_val : zero.data(),
                   (d->bcv_val ? d->bcv_val : zero.data()) + nbc);
  S->rho = d->rho; S->mu = d->mu; S->cfl = d->cfl; S->dnum = d->dnum;
  S->dcoef = d->dcoef; S->limiter = d->limiter; S->solver = d->solver;
  S->max_iter = d->max_iter; S->tol = d->tol; S->norm_mode = d->norm_mode;
  S->opts = ns_solver_opts(d->solver);
  // Grid_::min_size over the ghost-inclusive size arrays; the diffusion
  // denominator sum over d of 1/(min_size(d)^2) (s_operation.hpp:856-867)
  double msx = g.sx[0], msy = g.sy[0];
  for (int k = 1; k < S->NX; ++k) msx = g.sx[k] < msx ? g.sx[k] : msx;
  for (int k = 1; k < S->NY; ++k) msy = g.sy[k] < msy ? g.sy[k] : msy;
  S->min_size = msy < msx ? msy : msx;
  double den = 0.0;
  den += 1.0 / (msx * msx);
  den += 1.0 / (msy * msy);
  S->ddenom = den;
  CPGPU_HIP(ctx, hipHostMalloc((void**)&S->h_st, sizeof(NSDev), 0));
  S->h_st->dt = d->dt;
  S->h_st->t = 0.0;
  S->h_st->maxmag = 0.0;
  CPGPU_HIP(ctx, hipMemcpyAsync(S->st.p, S->h_st, sizeof(NSDev),
                                hipMemcpyHostToDevice, ctx->stream));
  CPGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = S.release();
  return CPGPU_OK;
}

int cpgpu_ns_destroy(cpgpu_ns* S) {
  if (!S) return CPGPU_EINVAL;
  (void)hipStreamSynchronize(S->ctx->stream);
  delete S;
  return CPGPU_OK;
}

int cpgpu_ns_set_field(cpgpu_ns* S, int which, const double* field) {
  if (!S || !field || which < 0 || which > 10) return CPGPU_EINVAL;
  if (which == F_PHI) {  // the warm start lives in the projection's vector
    int rc = cpgpu_poisson_set_phi(S->P, field);
    if (rc) return rc;
  }
  CPGPU_HIP(S->ctx, hipMemcpyAsync(S->fld(which), field, sizeof(double) * S->NT,
                                   hipMemcpyHostToDevice, S->ctx->stream));
  CPGPU_HIP(S->ctx, hipStreamSynchronize(S->ctx->stream));
  return CPGPU_OK;
}

int cpgpu_ns_get_field(cpgpu_ns* S, int which, double* field) {
  if (!S || !field || which < 0 || which > 10) return CPGPU_EINVAL;
  CPGPU_HIP(S->ctx, hipMemcpyAsync(field, S->fld(which), sizeof(double) * S->NT,
                                   hipMemcpyDeviceToHost, S->ctx->stream));
  CPGPU_HIP(S->ctx, hipStreamSynchronize(S->ctx->stream));
  return CPGPU_OK;
}

int cpgpu_ns_error_norms(cpgpu_ns* S, int which, const double* exact_field,
                         double* out) {
  if (!S || !exact_field || !out || which < 0 || which > 10) return CPGPU_EINVAL;
  CPGPU_HIP(S->ctx, hipSetDevice(S->ctx->device));
  if (S->err_ref.alloc(S->NT)) return set_error(S->ctx, CPGPU_ENOMEM, "error norm buffer");
  CPGPU_HIP(S->ctx, hipMemcpyAsync(S->err_ref.p, exact_field, sizeof(double) * S->NT,
                                   hipMemcpyHostToDevice, S->ctx->stream));
  const size_t off = (size_t)S->gl * S->NY + S->gl;
  return error_norms_device(S->ctx, S->nx, S->ny, S->err_ref.p + off, S->NY,
                            S->fld(which) + off, S->NY, S->sx.p + S->gl, S->sy.p + S->gl,
                            S->part.p, S->part.n, out);
}

int cpgpu_ns_field_change(cpgpu_ns* S, int which, double* out) {
  if (!S || !out || which < 0 || which > 10) return CPGPU_EINVAL;
  cpgpu_ctx* ctx = S->ctx;
  CPGPU_HIP(ctx, hipSetDevice(ctx->device));
  if (S->prev[which].alloc(S->NT)) return set_error(ctx, CPGPU_ENOMEM, "field snapshot");
  if (S->has_prev[which]) {
    const size_t off = (size_t)S->gl * S->NY + S->gl;
    int rc = error_norms_device(ctx, S->nx, S->ny, S->prev[which].p + off, S->NY,
                                S->fld(which) + off, S->NY, S->sx.p + S->gl,
                                S->sy.p + S->gl, S->part.p, S->part.n, out);
    if (rc) return rc;
  } else {
    out[0] = out[1] = out[2] = -1.0;
  }
  CPGPU_HIP(ctx, hipMemcpyAsync(S->prev[which].p, S->fld(which), sizeof(double) * S->NT,
                                hipMemcpyDeviceToDevice, ctx->stream));
  S->has_prev[which] = true;
  return CPGPU_OK;
}

int cpgpu_ns_step(cpgpu_ns* S, int nsteps, int* iters, double* resid) {
  if (!S || nsteps < 0) return CPGPU_EINVAL;
  CPGPU_HIP(S->ctx, hipSetDevice(S->ctx->device));
  int code = CPGPU_OK;
  for (int k = 0; k < nsteps; ++k) {
    code = S->kim ? ns_kim_step(S) : ns_one_step(S);
    if (code < 0) return code;
    if (iters) iters[k] = S->last_iters;
    if (resid) resid[k] = S->last_resid;
  }
  return code;
}

int cpgpu_ns_time(cpgpu_ns* S, double* t, double* dt) {
  if (!S) return CPGPU_EINVAL;
  CPGPU_HIP(S->ctx, hipMemcpyAsync(S->h_st, S->st.p, sizeof(NSDev),
                                   hipMemcpyDeviceToHost, S->ctx->stream));
  CPGPU_HIP(S->ctx, hipStreamSynchronize(S->ctx->stream));
  if (t) *t = S->h_st->t;
  if (dt) *dt = S->h_st->dt;
  return CPGPU_OK;
}

cpgpu_poisson* cpgpu_ns_projection(cpgpu_ns* S) { return S ? S->P : nullptr; }

int cpgpu_ns_set_time(cpgpu_ns* S, double t, double dt) {
  if (!S || !(dt > 0.0)) return CPGPU_EINVAL;
  CPGPU_HIP(S->ctx, hipStreamSynchronize(S->ctx->stream));
  S->h_st->t = t;
  S->h_st->dt = dt;
  S->h_st->maxmag = 0.0;
  CPGPU_HIP(S->ctx, hipMemcpyAsync(S->st.p, S->h_st, sizeof(NSDev),
                                   hipMemcpyHostToDevice, S->ctx->stream));
  CPGPU_HIP(S->ctx, hipStreamSynchronize(S->ctx->stream));
  return CPGPU_OK;
}

int cpgpu_ns_set_kim(cpgpu_ns* S, int enable, int helm_solver,
                     int helm_max_iter, double helm_tol) {
  if (!S || helm_max_iter < 1 || !(helm_tol > 0.0)) return CPGPU_EINVAL;
  if (helm_solver != CPGPU_CG && helm_solver != CPGPU_JPCG &&
      helm_solver != CPGPU_BICGSTAB && helm_solver != CPGPU_JACOBI)
    return CPGPU_EINVAL;
  // the explicit steps move dt on the device only: take the device's
  CPGPU_HIP(S->ctx, hipSetDevice(S->ctx->device));
  CPGPU_HIP(S->ctx, hipMemcpyAsync(S->h_st, S->st.p, sizeof(NSDev),
                                   hipMemcpyDeviceToHost, S->ctx->stream));
  CPGPU_HIP(S->ctx, hipStreamSynchronize(S->ctx->stream));
  S->kim = enable != 0;
  S->kim_solver = helm_solver;
  S->kim_max_iter = helm_max_iter;
  S->kim_tol = helm_tol;
  S->kim_step = 0;
  return CPGPU_OK;
}

int cpgpu_ns_kim_iters(cpgpu_ns* S, int* iters) {
  if (!S || !iters) return CPGPU_EINVAL;
  iters[0] = S->helm_iters[0];
  iters[1] = S->helm_iters[1];
  return CPGPU_OK;
}

int cpgpu_ns_set_bc_values(cpgpu_ns* S, int comp, const double* bc_val) {
  if (!S || !bc_val || comp < 0 || comp > 1) return CPGPU_EINVAL;
  const size_t nbc = 2 * (size_t)S->NY + 2 * (size_t)S->NX;
  S->hbc[comp].assign(bc_val, bc_val + nbc);
  CPGPU_HIP(S->ctx, hipSetDevice(S->ctx->device));
  CPGPU_HIP(S->ctx, hipMemcpyAsync(comp == 0 ? S->bcu_val.p : S->bcv_val.p,
                                   S->hbc[comp].data(), sizeof(double) * nbc,
                                   hipMemcpyHostToDevice, S->ctx->stream));
  CPGPU_HIP(S->ctx, hipStreamSynchronize(S->ctx->stream));
  // the NS_kim_ predictor rows keep their own walls (BC2 value 0)
  return CPGPU_OK;
}

int cpgpu_ns_set_solver(cpgpu_ns* S, int solver, int max_iter, double tol,
                        int norm_mode) {
  if (!S || !ns_solver_ok(solver) || max_iter < 1 || !(tol > 0.0))
    return CPGPU_EINVAL;
  S->solver = solver;
  S->max_iter = max_iter;
  S->tol = tol;
  S->norm_mode = norm_mode;
  S->opts = ns_solver_opts(solver);
  return CPGPU_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------
// Advection_ driver: fields u, v, uf, vf, phi resident on the device.
struct cpgpu_advection {
  cpgpu_ctx* ctx = nullptr;
  int nx = 0, ny = 0, gl = 0, NX = 0, NY = 0;
  size_t NT = 0;
  DevBuf<double> cx, sx, cy, sy;
  DevBuf<double> bc_val[3];  // phi, u, v: packed per ghost line
  int bc_type[3][4];
  DevBuf<double> F[5];       // u, v, uf, vf, phi
  int scheme = CPGPU_ADV_UPWIND1;
  double dt = 0.0, t = 0.0;
  long long steps = 0;
  Axes axes() const {
    Axes a;
    a.nx = nx; a.ny = ny; a.gl = gl; a.NX = NX; a.NY = NY;
    a.cx = cx.p; a.sx = sx.p; a.cy = cy.p; a.sy = sy.p;
    return a;
  }
  BCDev bcdev(int k) const {
    return BCDev{{bc_type[k][0], bc_type[k][1], bc_type[k][2], bc_type[k][3]}, bc_val[k].p};
  }
};

namespace {

bool adv_scheme_ok(int s) {
  return (s >= CPGPU_LIM_VANLEER && s <= CPGPU_LIM_UPWIND2) || s == CPGPU_ADV_UPWIND1 ||
         s == CPGPU_ADV_CENTER || s == CPGPU_ADV_CENTER4;
}

int adv_slot(int which) {
  switch (which) {
    case CPGPU_F_U: return 0;
    case CPGPU_F_V: return 1;
    case CPGPU_F_UF: return 2;
    case CPGPU_F_VF: return 3;
    case CPGPU_F_PHI: return 4;
    default: return -1;
  }
}

template <int SCH>
void adv_sweep(cpgpu_advection* A, hipStream_t st, int nti, int ntj) {
  const Axes a = A->axes();
  for (int D = 0; D < nti + ntj - 1; ++D) {
    const int lo = std::max(0, D - (ntj - 1)), hi = std::min(D, nti - 1);
    hipLaunchKernelGGL((k_adv_sweep<SCH>), dim3(hi - lo + 1), dim3(64), 0, st, a, A->F[2].p,
                       A->F[3].p, A->F[4].p, A->dt, D, nti, ntj);
  }
}

}  // namespace

extern "C" {

int cpgpu_advection_create(cpgpu_ctx* ctx, const cpgpu_advection_desc* d,
                           cpgpu_advection** out) {
  if (!ctx || !d || !out) return CPGPU_EINVAL;
  *out = nullptr;
  const cpgpu_grid2d& g = d->grid;
  if (g.nx < 1 || g.ny < 1 || g.gl < 2 || !g.cx || !g.sx || !g.cy || !g.sy)
    return set_error(ctx, CPGPU_EINVAL, "Advection needs gl >= 2 and axis arrays");
  if (!adv_scheme_ok(d->scheme) || !(d->dt > 0.0))
    return set_error(ctx, CPGPU_EINVAL, "bad advection scheme or dt");
  const int* types[3] = {g.bc_type, d->bcu_type, d->bcv_type};
  for (int k = 0; k < 3; ++k)
    for (int s = 0; s < 4; ++s)
      if (types[k][s] != CPGPU_BC_DIRICHLET && types[k][s] != CPGPU_BC_NEUMANN)
        return set_error(ctx, CPGPU_EINVAL, "bad advection boundary type");
  CPGPU_HIP(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<cpgpu_advection> A(new cpgpu_advection());
  A->ctx = ctx;
  A->nx = g.nx; A->ny = g.ny; A->gl = g.gl;
  A->NX = g.nx + 2 * g.gl; A->NY = g.ny + 2 * g.gl;
  A->NT = (size_t)A->NX * A->NY;
  const size_t nbc = 2 * (size_t)A->NY + 2 * (size_t)A->NX;
  for (int k = 0; k < 5; ++k) {
    if (A->F[k].alloc(A->NT)) return set_error(ctx, CPGPU_ENOMEM, "advection fields");
    CPGPU_HIP(ctx, hipMemsetAsync(A->F[k].p, 0, sizeof(double) * A->NT, ctx->stream));
  }
  if (A->cx.alloc(A->NX) || A->sx.alloc(A->NX) || A->cy.alloc(A->NY) || A->sy.alloc(A->NY) ||
      A->bc_val[0].alloc(nbc) || A->bc_val[1].alloc(nbc) || A->bc_val[2].alloc(nbc))
    return set_error(ctx, CPGPU_ENOMEM, "advection buffers");
  std::vector<double> zero(nbc, 0.0);
  const double* vals[3] = {g.bc_val, d->bcu_val, d->bcv_val};
  int rc;
  if ((rc = ns_upload(ctx, A->cx.p, g.cx, sizeof(double) * A->NX)) ||
      (rc = ns_upload(ctx, A->sx.p, g.sx, sizeof(double) * A->NX)) ||
      (rc = ns_upload(ctx, A->cy.p, g.cy, sizeof(double) * A->NY)) ||
      (rc = ns_upload(ctx, A->sy.p, g.sy, sizeof(double) * A->NY)))
    return rc;
  for (int k = 0; k < 3; ++k) {
    if ((rc = ns_upload(ctx, A->bc_val[k].p, vals[k] ? vals[k] : zero.data(),
                        sizeof(double) * nbc)))
      return rc;
    for (int s = 0; s < 4; ++s) A->bc_type[k][s] = types[k][s];
  }
  A->scheme = d->scheme;
  A->dt = d->dt;
  CPGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = A.release();
  return CPGPU_OK;
}

int cpgpu_advection_destroy(cpgpu_advection* A) {
  if (!A) return CPGPU_EINVAL;
  (void)hipStreamSynchronize(A->ctx->stream);
  delete A;
  return CPGPU_OK;
}

int cpgpu_advection_set_field(cpgpu_advection* A, int which, const double* field) {
  const int k = adv_slot(which);
  if (!A || !field || k < 0) return CPGPU_EINVAL;
  CPGPU_HIP(A->ctx, hipSetDevice(A->ctx->device));
  return ns_upload(A->ctx, A->F[k].p, field, sizeof(double) * A->NT);
}

int cpgpu_advection_get_field(cpgpu_advection* A, int which, double* field) {
  const int k = adv_slot(which);
  if (!A || !field || k < 0) return CPGPU_EINVAL;
  CPGPU_HIP(A->ctx, hipSetDevice(A->ctx->device));
  CPGPU_HIP(A->ctx, hipMemcpyAsync(field, A->F[k].p, sizeof(double) * A->NT,
                                   hipMemcpyDeviceToHost, A->ctx->stream));
  CPGPU_HIP(A->ctx, hipStreamSynchronize(A->ctx->stream));
  return CPGPU_OK;
}

int cpgpu_advection_set_scheme(cpgpu_advection* A, int scheme) {
  if (!A || !adv_scheme_ok(scheme)) return CPGPU_EINVAL;
  A->scheme = scheme;
  return CPGPU_OK;
}

int cpgpu_advection_initial(cpgpu_advection* A) {
  if (!A) return CPGPU_EINVAL;
  CPGPU_HIP(A->ctx, hipSetDevice(A->ctx->device));
  hipStream_t st = A->ctx->stream;
  const Axes a = A->axes();
  hipLaunchKernelGGL(k_ghost_edges, dim3(64, 2), dim3(256), 0, st, a, A->F[0].p, A->bcdev(1),
                     A->F[1].p, A->bcdev(2));
  hipLaunchKernelGGL(k_ghost_corners, dim3(1, 2), dim3(256), 0, st, a, A->F[0].p, A->bcdev(1),
                     A->F[1].p, A->bcdev(2));
  const dim3 b2(64, 4), g2((A->ny + 63) / 64, (A->nx + 3) / 4);
  hipLaunchKernelGGL(k_adv_faces, g2, b2, 0, st, a, A->F[0].p, A->F[1].p, A->F[2].p, A->F[3].p);
  CPGPU_HIP(A->ctx, hipGetLastError());
  CPGPU_HIP(A->ctx, hipStreamSynchronize(st));
  return CPGPU_OK;
}

int cpgpu_advection_step(cpgpu_advection* A, int nsteps) {
  if (!A || nsteps < 0) return CPGPU_EINVAL;
  CPGPU_HIP(A->ctx, hipSetDevice(A->ctx->device));
  hipStream_t st = A->ctx->stream;
  const Axes a = A->axes();
  const int nti = (A->nx + 15) / 16, ntj = (A->ny + 15) / 16;
  for (int s = 0; s < nsteps; ++s) {
    // run_one_step: apply_bc of u, v (their faces are kept) and phi, then
    // the in-place sweep
    hipLaunchKernelGGL(k_ghost_edges, dim3(64, 2), dim3(256), 0, st, a, A->F[0].p,
                       A->bcdev(1), A->F[1].p, A->bcdev(2));
    hipLaunchKernelGGL(k_ghost_corners, dim3(1, 2), dim3(256), 0, st, a, A->F[0].p,
                       A->bcdev(1), A->F[1].p, A->bcdev(2));
    hipLaunchKernelGGL(k_ghost_edges, dim3(64, 1), dim3(256), 0, st, a, A->F[4].p,
                       A->bcdev(0), A->F[4].p, A->bcdev(0));
    hipLaunchKernelGGL(k_ghost_corners, dim3(1, 1), dim3(256), 0, st, a, A->F[4].p,
                       A->bcdev(0), A->F[4].p, A->bcdev(0));
    switch (A->scheme) {
      case CPGPU_LIM_VANLEER: adv_sweep<CPGPU_LIM_VANLEER>(A, st, nti, ntj); break;
      case CPGPU_LIM_SUPERBEE: adv_sweep<CPGPU_LIM_SUPERBEE>(A, st, nti, ntj); break;
      case CPGPU_LIM_WAHYD: adv_sweep<CPGPU_LIM_WAHYD>(A, st, nti, ntj); break;
      case CPGPU_LIM_QUICK: adv_sweep<CPGPU_LIM_QUICK>(A, st, nti, ntj); break;
      case CPGPU_LIM_CUI: adv_sweep<CPGPU_LIM_CUI>(A, st, nti, ntj); break;
      case CPGPU_LIM_UPWIND2: adv_sweep<CPGPU_LIM_UPWIND2>(A, st, nti, ntj); break;
      case CPGPU_ADV_UPWIND1: adv_sweep<CPGPU_ADV_UPWIND1>(A, st, nti, ntj); break;
      case CPGPU_ADV_CENTER: adv_sweep<CPGPU_ADV_CENTER>(A, st, nti, ntj); break;
      default: adv_sweep<CPGPU_ADV_CENTER4>(A, st, nti, ntj); break;
    }
    CPGPU_HIP(A->ctx, hipGetLastError());
    A->t += A->dt;
    ++A->steps;
  }
  CPGPU_HIP(A->ctx, hipStreamSynchronize(st));
  return CPGPU_OK;
}

int cpgpu_advection_time(cpgpu_advection* A, double* t, long long* steps) {
  if (!A || !t || !steps) return CPGPU_EINVAL;
  *t = A->t;
  *steps = A->steps;
  return CPGPU_OK;
}

}  // extern "C"
