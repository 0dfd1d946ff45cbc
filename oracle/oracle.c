/*
 * CPU oracle for the structured-grid projection hot path of Carpios.
 * TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker, never by the product library.
 *
 * Parity pins (what checks each function; the tests live in tests/):
 *   orc_axis_uniform ........ closed form of a uniform axis; ghost symmetry
 *   orc_dot / orc_nrm2_* .... exact small cases (nrm2([3,4]) = 3 in the
 *                             reference mode), golden values from the
 *                             reference's own BLAS_Level1 built in oracle/_ref
 *   orc_poisson_assemble .... closed-form eigenvector sin(pi x)sin(pi y)
 *                             (Dirichlet), A*1 = 0 (all-Neumann), symmetry,
 *                             SpMV + solvers checked against oracle/_ref
 *   orc_cg/jpcg/bicgstab/jacobi  golden iteration counts and residual
 *                             histories from the reference's lib/algebra
 *                             built in oracle/_ref; dense Gaussian
 *                             elimination on tiny grids
 *   Neumann ghost constant .. a linear phi with nonzero Neumann data is
 *                             reproduced exactly by the assembled rows
 *                             (a flipped sign fails)
 *   TVD / diffusion / C2F / Div / Grad / F2C  exactness on constant and
 *                             linear fields, limiter special values,
 *                             Div(Grad phi) == -A phi / vol consistency,
 *                             hand-computed per-cell upwind values on a
 *                             cell whose face velocities differ in sign
 *                             (face-centric upwinding fails)
 *   orc_ns_step .............. discrete divergence after projection at the
 *                             solver tolerance; Taylor-Green vortex with
 *                             nu = mu/rho = 1: second-order u and the
 *                             pressure (a flipped advection sign or dif*rho
 *                             fails)
 *   orc_ns_kim_step .......... Taylor-Green decay; predictor rows against
 *                             the all-Neumann Helmholtz matrix built by
 *                             Kronecker products; a hand-computed AB2
 *                             source on a wall cell at step 1
 *   orc_set_plant ............ switches on one deliberate mistake so the
 *                             tests can show each pin fails on it
 *   orc_jpcg_mixed ........... true FP64 residual reaches tol; solution
 *                             matches orc_jpcg to the conditioning bound
 *   orc_error_norms .......... closed forms: a constant error c gives
 *                             (|c|, |c| / sqrt(V), |c|), a single-cell
 *                             error e gives (|e| v / V, |e| sqrt(v) / V,
 *                             |e|) with V the domain volume; scaling
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define FIDX(g, i, j) (((i) + (g)->gl) * ((g)->ny + 2 * (g)->gl) + ((j) + (g)->gl))

/* Planted-mistake switch (ORC_PLANT_*, tests only; 0 = faithful). */
static int g_plant = 0;
void orc_set_plant(int plant) { g_plant = plant; }

/* Threads of the CPU baseline (orc_set_threads; default 1).  With one
 * thread every routine runs its serial loop, in the reference's order.
 * With more, the row loops of the matrix product, the vector updates and
 * the cell loops of the NS predictor split over threads (each value is
 * computed as before), and dots and norms sum fixed contiguous chunks in
 * parallel, then the chunk sums in order: deterministic for a given thread
 * count, but not the serial order, so parity checks run on one thread. */
static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }
#define ORC_PAR _Pragma("omp parallel for if (g_threads > 1) num_threads(g_threads) schedule(static)")

/* sum over i of f(x[i], y[i]) in g_threads contiguous chunks
 * (mode 0: x*y, 1: x*x, 2: x*x for x != 0) */
static double par_sum(int n, const double *x, const double *y, int mode) {
  double part[256];
  int T = g_threads < 256 ? g_threads : 256, t;
  double s = 0.0;
#pragma omp parallel for num_threads(T) schedule(static, 1)
  for (t = 0; t < T; ++t) {
    long lo = (long)n * t / T, hi = (long)n * (t + 1) / T, i;
    double a = 0.0;
    for (i = lo; i < hi; ++i) {
      if (mode == 0) a = a + x[i] * y[i];
      else if (mode == 1 || x[i] != 0.0) a += x[i] * x[i];
    }
    part[t] = a;
  }
  for (t = 0; t < T; ++t) s += part[t];
  return s;
}

/* lib/structure/s_grid.hpp:372-395 (_set_uniform_grid): sizes are l/n,
 * centres by accumulation csum + 0.5*cs, then shifted by min - sum(ghost
 * sizes). */
void orc_axis_uniform(int n, int gl, double vmin, double vmax, double *c,
                      double *s) {
  int ng = n + 2 * gl;
  double l = vmax - vmin;
  double ds = l / n;
  double csum = 0.0, llg = 0.0, sh;
  int ii;
  for (ii = 0; ii < ng; ++ii) s[ii] = ds;
  for (ii = 0; ii < ng; ++ii) {
    c[ii] = csum + 0.5 * s[ii];
    csum += s[ii];
  }
  for (ii = 0; ii < gl; ++ii) llg += s[ii];
  sh = vmin - llg;
  for (ii = 0; ii < ng; ++ii) c[ii] += sh;
}

/* lib/structure/s_grid.hpp:397-434 (_set_non_uniform_grid): interior sizes
 * from the caller, each ghost layer repeats the nearest interior size, then
 * the centres exactly as for the uniform axis. */
void orc_axis_sizes(int n, int gl, double vmin, const double *sizes, double *c,
                    double *s) {
  int ng = n + 2 * gl, ii;
  double csum = 0.0, llg = 0.0, sh;
  for (ii = 0; ii < ng; ++ii) {
    if (ii < gl) s[ii] = sizes[0];
    else if (ii - gl >= n) s[ii] = sizes[n - 1];
    else s[ii] = sizes[ii - gl];
  }
  for (ii = 0; ii < ng; ++ii) {
    c[ii] = csum + 0.5 * s[ii];
    csum += s[ii];
  }
  for (ii = 0; ii < gl; ++ii) llg += s[ii];
  sh = vmin - llg;
  for (ii = 0; ii < ng; ++ii) c[ii] += sh;
}

/* lib/algebra/BLAS_Level1.hpp:500-544: the 5-way unrolled loop still adds
 * left to right, so it is a sequential sum. */
double orc_dot(int n, const double *x, const double *y) {
  double d = 0.0;
  int i;
  if (g_threads > 1) return par_sum(n, x, y, 0);
  for (i = 0; i < n; ++i) d = d + x[i] * y[i];
  return d;
}

/* lib/algebra/BLAS_Level1.hpp:556-577: sums i < n-1 only (drops the last
 * element) and skips exact zeros; n == 1 returns |x0|. */
double orc_nrm2_ref(int n, const double *x) {
  double norm = 0.0;
  int i;
  if (n <= 0) return 0.0;
  if (n == 1) return fabs(x[0]);
  if (g_threads > 1) return sqrt(par_sum(n - 1, x, NULL, 2));
  for (i = 0; i < n - 1; ++i)
    if (x[i] != 0.0) norm += x[i] * x[i];
  return sqrt(norm);
}

double orc_nrm2_true(int n, const double *x) {
  double norm = 0.0;
  int i;
  if (g_threads > 1) return sqrt(par_sum(n, x, NULL, 1));
  for (i = 0; i < n; ++i) norm += x[i] * x[i];
  return sqrt(norm);
}

static double nrm(int mode, int n, const double *x) {
  return mode == ORC_NORM_REF ? orc_nrm2_ref(n, x) : orc_nrm2_true(n, x);
}

/* lib/algebra/matrix_SparCompRow.hpp:236-249: res[i] starts at 0 and adds
 * x[col]*val in CSR order. */
void orc_csr_mult(const orc_csr *A, const double *x, double *y) {
  int i;
  ORC_PAR
  for (i = 0; i < A->n; ++i) {
    double acc = 0.0;
    int k;
    for (k = A->row_ptr[i]; k < A->row_ptr[i + 1]; ++k)
      acc += x[A->col[k]] * A->val[k];
    y[i] = acc;
  }
}

int orc_poisson_nnz(int nx, int ny) {
  return 5 * nx * ny - 2 * nx - 2 * ny;
}

static double bcval(const orc_grid *g, const orc_bc *bc, int seg, int line) {
  int NX = g->nx + 2 * g->gl, NY = g->ny + 2 * g->gl, off;
  if (!bc->val) return 0.0;
  if (seg == 0) off = 0;
  else if (seg == 1) off = NY;
  else if (seg == 2) off = 2 * NY;
  else off = 2 * NY + NX;
  return bc->val[off + line];
}

/* axis helpers: c_(d, idx) and s_(d, idx) of lib/structure/s_grid.hpp */
static double cax(const orc_grid *g, int d, int idx) {
  return d == 0 ? g->cx[idx + g->gl] : g->cy[idx + g->gl];
}
static double sax(const orc_grid *g, int d, int idx) {
  return d == 0 ? g->sx[idx + g->gl] : g->sy[idx + g->gl];
}

/* _coe_other of StencilPoisson_ (s_poisson.hpp:98-104 with the default
 * lambdas :115-135): -betaf*a/d, a = product of the other cell sizes (the z
 * size is 1 in 2-D), d = |c - c_nb| from the accumulated centres. */
static double coe_other(const orc_grid *g, double beta, int i, int j, int d,
                        int ori /* -1 or +1 */) {
  double a, dist, c, co;
  if (d == 0) {
    a = sax(g, 1, j) * 1.0;
    c = cax(g, 0, i);
    co = cax(g, 0, i + ori);
  } else {
    a = 1.0 * sax(g, 0, i);
    c = cax(g, 1, j);
    co = cax(g, 1, j + ori);
  }
  dist = fabs(c - co);
  return -beta * a / dist;
}

typedef struct {
  int di, dj; /* neighbour offset */
  double coe;
} term_t;

/* One matrix row: StencilPoisson_::coe_row (s_poisson.hpp:79-97) followed by
 * GhostRegular_::substitute (s_ghost.hpp:334-357) and concise
 * (polynomial.hpp:317-320).  Terms come out in Index_compare_ order
 * (s_define.hpp:227-241): (i-1,j) (i,j-1) (i,j) (i,j+1) (i+1,j). */
static int build_row(const orc_grid *g, double beta, const double *src_field,
                     double src_uniform, const orc_bc *bc, int i, int j,
                     term_t *t, double *cst, int helm, double alpha) {
  int nx = g->nx, ny = g->ny, d, n = 0, k;
  double coec = 0.0, cm[2], cp[2], v, src;
  double constant;
  /* neighbour coefficients */
  for (d = 0; d < 2; ++d) {
    cp[d] = coe_other(g, beta, i, j, d, +1);
    cm[d] = coe_other(g, beta, i, j, d, -1);
  }
  /* coe(index, 0, _C_): coec += -coem; coec += -coep; per dimension */
  for (d = 0; d < 2; ++d) {
    double coep = -cp[d];
    double coem = -cm[d];
    coec += coem;
    coec += coep;
  }
  /* src(index, 0, _C_): v *= s(index, dim=0) for each d (uses dim, not d) */
  v = 1.0;
  for (d = 0; d < 2; ++d) v *= sax(g, 0, i);
  src = src_field ? src_field[FIDX(g, i, j)] : src_uniform;
  constant = v * src;
  if (helm) {
    /* StencilHelmholtz_::coe_row (s_poisson.hpp:230-259): v is the product
     * of the sizes over d (not dim 0 twice), the centre term gets v*alpha
     * before the ghost substitutions and the constant is v*source */
    v = 1.0;
    for (d = 0; d < 2; ++d) v *= sax(g, d, d == 0 ? i : j);
    constant = v * src;
    coec = coec + v * alpha;
  }
  /* map-ordered terms, ghost ones flagged by coordinates out of range */
  t[n].di = -1; t[n].dj = 0; t[n].coe = cm[0]; n++;
  t[n].di = 0; t[n].dj = -1; t[n].coe = cm[1]; n++;
  t[n].di = 0; t[n].dj = 0; t[n].coe = coec; n++;
  t[n].di = 0; t[n].dj = 1; t[n].coe = cp[1]; n++;
  t[n].di = 1; t[n].dj = 0; t[n].coe = cp[0]; n++;
  /* substitute ghosts in map order (s_ghost.hpp:334-357, :140-179) */
  for (k = 0; k < 5; ++k) {
    int gi = i + t[k].di, gj = j + t[k].dj, a, seg, ori;
    double coe = t[k].coe, c0, c1, cf, dl, d0, d1, val;
    if (gi >= 0 && gi < nx && gj >= 0 && gj < ny) continue;
    if (coe != 0.0) {
      if (gi < 0) { a = 0; seg = 0; ori = -1; }
      else if (gi >= nx) { a = 0; seg = 1; ori = 1; }
      else if (gj < 0) { a = 1; seg = 2; ori = -1; }
      else { a = 1; seg = 3; ori = 1; }
      c0 = cax(g, a, a == 0 ? gi : gj);
      c1 = cax(g, a, a == 0 ? i : j);
      /* f_(a, ori, oriidx): centre +- half size */
      cf = ori > 0 ? c1 + sax(g, a, a == 0 ? i : j) * 0.5
                   : c1 - sax(g, a, a == 0 ? i : j) * 0.5;
      dl = fabs(c1 - c0);
      d0 = fabs(cf - c0);
      d1 = fabs(c1 - cf);
      val = bcval(g, bc, seg, (a == 0 ? j : i) + g->gl);
      if (bc->type[seg] == ORC_BC1) {
        t[2].coe = t[2].coe + (-coe * d0 / d1);
        constant = constant + (coe * dl * val / d1);
      } else {
        t[2].coe = t[2].coe + coe;
        if (g_plant == ORC_PLANT_NEUMANN_SIGN)
          constant = constant + (val * dl * coe);
        else
          constant = constant + (-val * dl * coe);
      }
    }
    t[k].coe = 0.0;
    t[k].di = 99; /* erased */
  }
  *cst = constant;
  return n;
}

/* Solver_::_build_matrix (s_solver.hpp:94-161): non-constant terms become
 * CSR entries in map order, b = -constant ("negative added here"); zero
 * coefficients were trimmed by concise. */
static int assemble_rows(const orc_grid *g, double beta,
                         const double *src_field, double src_uniform,
                         const orc_bc *bc, int *row_ptr, int *col, double *val,
                         double *b, int helm, double alpha) {
  int nx = g->nx, ny = g->ny, i, j, k, nz = 0;
  term_t t[5];
  double cst;
  if (nx < 1 || ny < 1 || g->gl < 1) return -1;
  row_ptr[0] = 0;
  for (j = 0; j < ny; ++j) {
    for (i = 0; i < nx; ++i) {
      int row = i + nx * j;
      build_row(g, beta, src_field, src_uniform, bc, i, j, t, &cst, helm, alpha);
      for (k = 0; k < 5; ++k) {
        if (t[k].di == 99 || t[k].coe == 0.0) continue;
        col[nz] = (i + t[k].di) + nx * (j + t[k].dj);
        val[nz] = t[k].coe;
        nz++;
      }
      b[row] = cst == 0.0 ? 0.0 : -cst;
      row_ptr[row + 1] = nz;
    }
  }
  return nz;
}

/* Right-hand side only: the constant column of the same rows. */
static void rhs_rows(const orc_grid *g, double beta, const double *src_field,
                     double src_uniform, const orc_bc *bc, double *b, int helm,
                     double alpha) {
  int i, j;
  term_t t[5];
  double cst;
  for (j = 0; j < g->ny; ++j) {
    for (i = 0; i < g->nx; ++i) {
      build_row(g, beta, src_field, src_uniform, bc, i, j, t, &cst, helm, alpha);
      b[i + g->nx * j] = cst == 0.0 ? 0.0 : -cst;
    }
  }
}

int orc_poisson_assemble(const orc_grid *g, double beta,
                         const double *src_field, double src_uniform,
                         const orc_bc *bc, int *row_ptr, int *col, double *val,
                         double *b) {
  return assemble_rows(g, beta, src_field, src_uniform, bc, row_ptr, col, val,
                       b, 0, 0.0);
}

void orc_poisson_rhs(const orc_grid *g, double beta, const double *src_field,
                     double src_uniform, const orc_bc *bc, double *b) {
  rhs_rows(g, beta, src_field, src_uniform, bc, b, 0, 0.0);
}

/* Poisson_ with set_uniform_alpha (s_poisson.hpp:461-467, :576-592):
 * StencilHelmholtz_ rows with uniform alpha and beta. */
int orc_helmholtz_assemble(const orc_grid *g, double alpha, double beta,
                           const double *src_field, double src_uniform,
                           const orc_bc *bc, int *row_ptr, int *col,
                           double *val, double *b) {
  return assemble_rows(g, beta, src_field, src_uniform, bc, row_ptr, col, val,
                       b, 1, alpha);
}

void orc_helmholtz_rhs(const orc_grid *g, double alpha, double beta,
                       const double *src_field, double src_uniform,
                       const orc_bc *bc, double *b) {
  rhs_rows(g, beta, src_field, src_uniform, bc, b, 1, alpha);
}

void orc_field_to_x(const orc_grid *g, const double *field, double *x) {
  int i, j;
  for (j = 0; j < g->ny; ++j)
    for (i = 0; i < g->nx; ++i) x[i + g->nx * j] = field[FIDX(g, i, j)];
}

void orc_x_to_field(const orc_grid *g, const double *x, double *field) {
  int i, j;
  for (j = 0; j < g->ny; ++j)
    for (i = 0; i < g->nx; ++i) field[FIDX(g, i, j)] = x[i + g->nx * j];
}

static void hist_push(double *hist, int cap, int *len, double v) {
  if (hist && *len < cap) hist[*len] = v;
  (*len)++;
}

/* DiaPre(A, 1) (preconditioner.hpp:182-230): 1/a_ii; zero diagonal is a
 * failure (the reference calls exit(0); the oracle returns -10). */
static int diapre(const orc_csr *A, double *d) {
  int i, k;
  for (i = 0; i < A->n; ++i) {
    d[i] = 0.0;
    for (k = A->row_ptr[i]; k < A->row_ptr[i + 1]; ++k) {
      if (A->col[k] == i) {
        if (A->val[k] == 0.0) return -10;
        d[i] = 1. / A->val[k];
        break;
      }
    }
    if (d[i] == 0.0) return -10;
  }
  return 0;
}

/* r = b - A*x through a temporary (array_list.hpp:820-825) */
static void residual(const orc_csr *A, const double *x, const double *b,
                     double *r, double *tmp) {
  int i;
  orc_csr_mult(A, x, tmp);
  ORC_PAR
  for (i = 0; i < A->n; ++i) r[i] = b[i] - tmp[i];
}

/* Shared body of CG (solver_matrix.hpp:264-316) and IC_CG
 * (solver_matrix.hpp:318-374) with ICPre replaced by DiaPre(A,1): the
 * Jacobi-preconditioned CG the projection configs name. */
static int cg_body(const orc_csr *A, double *x, const double *b, int *max_iter,
                   double *tol, int norm_mode, double *hist, int hist_cap,
                   int *hist_len, int precond) {
  int n = A->n, i, k, code = 1;
  double *r, *z, *p, *q, *d = NULL;
  double normb, resid, rho = 0.0, rho_1 = 0.0, alpha, beta, pq;
  *hist_len = 0;
  if (n <= 0) return -1;
  r = (double *)malloc(sizeof(double) * n);
  z = (double *)malloc(sizeof(double) * n);
  p = (double *)malloc(sizeof(double) * n);
  q = (double *)malloc(sizeof(double) * n);
  if (precond) {
    d = (double *)malloc(sizeof(double) * n);
    if (diapre(A, d) != 0) {
      code = -10;
      goto done;
    }
  }
  normb = nrm(norm_mode, n, b);
  residual(A, x, b, r, q);
  if (normb == 0.0) normb = 1;
  if ((resid = nrm(norm_mode, n, r) / normb) <= *tol) {
    *tol = resid;
    *max_iter = 0;
    code = 0;
    goto done;
  }
  for (i = 1; i <= *max_iter; ++i) {
    if (precond) {
      ORC_PAR
      for (k = 0; k < n; ++k) z[k] = r[k] * d[k];
    } else {
      memcpy(z, r, sizeof(double) * n);
    }
    rho = orc_dot(n, r, z);
    if (i == 1) {
      memcpy(p, z, sizeof(double) * n);
    } else {
      beta = rho / rho_1;
      ORC_PAR
      for (k = 0; k < n; ++k) p[k] = z[k] + beta * p[k];
    }
    orc_csr_mult(A, p, q);
    pq = orc_dot(n, p, q);
    alpha = rho / pq;
    ORC_PAR
    for (k = 0; k < n; ++k) x[k] += alpha * p[k];
    ORC_PAR
    for (k = 0; k < n; ++k) r[k] -= alpha * q[k];
    resid = nrm(norm_mode, n, r) / normb;
    hist_push(hist, hist_cap, hist_len, resid);
    if (resid <= *tol) {
      *tol = resid;
      *max_iter = i;
      code = 0;
      goto done;
    }
    rho_1 = rho;
  }
  *tol = resid;
  code = 1;
done:
  free(r);
  free(z);
  free(p);
  free(q);
  free(d);
  return code;
}

int orc_cg(const orc_csr *A, double *x, const double *b, int *max_iter,
           double *tol, int norm_mode, double *hist, int hist_cap,
           int *hist_len) {
  return cg_body(A, x, b, max_iter, tol, norm_mode, hist, hist_cap, hist_len,
                 0);
}

int orc_jpcg(const orc_csr *A, double *x, const double *b, int *max_iter,
             double *tol, int norm_mode, double *hist, int hist_cap,
             int *hist_len) {
  return cg_body(A, x, b, max_iter, tol, norm_mode, hist, hist_cap, hist_len,
                 1);
}

/* Dia_BiCGSTAB, solver_matrix.hpp:599-672.  Note the strict '<' tests and
 * the half-step exit that leaves max_iter unchanged (:642-647). */
int orc_dia_bicgstab(const orc_csr *A, double *x, const double *b,
                     int *max_iter, double *tol, int norm_mode, double *hist,
                     int hist_cap, int *hist_len) {
  int n = A->n, i, k, code = 1;
  double *d, *r, *rt, *p, *ph, *s, *sh, *t, *v;
  double normb, resid, rho_1 = 0, rho_2 = 0, alpha = 0, beta, omega = 0;
  *hist_len = 0;
  if (n <= 0) return -1;
  d = (double *)malloc(sizeof(double) * n);
  r = (double *)malloc(sizeof(double) * n);
  rt = (double *)malloc(sizeof(double) * n);
  p = (double *)calloc(n, sizeof(double));
  ph = (double *)malloc(sizeof(double) * n);
  s = (double *)malloc(sizeof(double) * n);
  sh = (double *)malloc(sizeof(double) * n);
  t = (double *)malloc(sizeof(double) * n);
  v = (double *)calloc(n, sizeof(double));
  if (diapre(A, d) != 0) {
    code = -10;
    goto done;
  }
  normb = nrm(norm_mode, n, b);
  residual(A, x, b, r, t);
  memcpy(rt, r, sizeof(double) * n);
  if (normb == 0.0) normb = 1;
  if ((resid = nrm(norm_mode, n, r) / normb) <= *tol) {
    *tol = resid;
    *max_iter = 0;
    hist_push(hist, hist_cap, hist_len, resid);
    code = 0;
    goto done;
  }
  for (i = 1; i <= *max_iter; i++) {
    rho_1 = orc_dot(n, rt, r);
    if (rho_1 == 0) {
      *tol = nrm(norm_mode, n, r) / normb;
      code = 2;
      goto done;
    }
    if (i == 1) {
      memcpy(p, r, sizeof(double) * n);
    } else {
      beta = (rho_1 / rho_2) * (alpha / omega);
      for (k = 0; k < n; ++k) p[k] = r[k] + beta * (p[k] - omega * v[k]);
    }
    for (k = 0; k < n; ++k) ph[k] = p[k] * d[k];
    orc_csr_mult(A, ph, v);
    alpha = rho_1 / orc_dot(n, rt, v);
    for (k = 0; k < n; ++k) s[k] = r[k] - alpha * v[k];
    if ((resid = nrm(norm_mode, n, s) / normb) < *tol) {
      for (k = 0; k < n; ++k) x[k] += alpha * ph[k];
      *tol = resid;
      hist_push(hist, hist_cap, hist_len, resid);
      code = 0;
      goto done;
    }
    for (k = 0; k < n; ++k) sh[k] = s[k] * d[k];
    orc_csr_mult(A, sh, t);
    omega = orc_dot(n, t, s) / orc_dot(n, t, t);
    for (k = 0; k < n; ++k) x[k] += alpha * ph[k] + omega * sh[k];
    for (k = 0; k < n; ++k) r[k] = s[k] - omega * t[k];
    rho_2 = rho_1;
    resid = nrm(norm_mode, n, r) / normb;
    hist_push(hist, hist_cap, hist_len, resid);
    if (resid < *tol) {
      *tol = resid;
      *max_iter = i;
      code = 0;
      goto done;
    }
    if (omega == 0) {
      *tol = nrm(norm_mode, n, r) / normb;
      code = 3;
      goto done;
    }
  }
  *tol = resid;
  code = 1;
done:
  free(d); free(r); free(rt); free(p); free(ph);
  free(s); free(sh); free(t); free(v);
  return code;
}

/* Solver_Jacobi_::_solve, solver_matrix.hpp:842-909: T = -D^{-1}(L+U),
 * C = D^{-1} b; newx = T*x + C.  On convergence it returns WITHOUT copying
 * newx into x (:898-903), so x holds the previous iterate. */
int orc_jacobi(const orc_csr *A, double *x, const double *b, int *max_iter,
               double *tol, int norm_mode, double *hist, int hist_cap,
               int *hist_len) {
  int n = A->n, i, k, code = 1, nnz = A->row_ptr[n];
  double *tv, *c, *newx, *r, *tmp, normb, resid = 0;
  orc_csr T;
  *hist_len = 0;
  if (n <= 0) return -1;
  tv = (double *)malloc(sizeof(double) * nnz);
  c = (double *)malloc(sizeof(double) * n);
  newx = (double *)malloc(sizeof(double) * n);
  r = (double *)malloc(sizeof(double) * n);
  tmp = (double *)malloc(sizeof(double) * n);
  memcpy(tv, A->val, sizeof(double) * nnz);
  for (i = 0; i < n; ++i) c[i] = 1.0 / ORC_SMALL;
  normb = nrm(norm_mode, n, b);
  residual(A, x, b, r, tmp);
  if (normb == 0.0) normb = 1;
  if ((resid = nrm(norm_mode, n, r) / normb) <= *tol) {
    *tol = resid;
    *max_iter = 0;
    code = 0;
    goto done;
  }
  for (i = 0; i < n; ++i) {
    double dv = 0.0;
    int flag = -1;
    for (k = A->row_ptr[i]; k < A->row_ptr[i + 1]; ++k) {
      if (A->col[k] == i) {
        flag = k;
        dv = tv[k];
        tv[k] = 0;
        break;
      }
    }
    for (k = A->row_ptr[i]; k < A->row_ptr[i + 1]; ++k) {
      if (k == flag)
        c[i] = b[A->col[k]] / dv;
      else
        tv[k] = -tv[k] / dv;
    }
  }
  T.n = n;
  T.row_ptr = A->row_ptr;
  T.col = A->col;
  T.val = tv;
  for (i = 1; i <= *max_iter; ++i) {
    orc_csr_mult(&T, x, newx);
    for (k = 0; k < n; ++k) newx[k] = newx[k] + c[k];
    residual(A, newx, b, r, tmp);
    resid = nrm(norm_mode, n, r) / normb;
    hist_push(hist, hist_cap, hist_len, resid);
    if (resid <= *tol) {
      *tol = resid;
      *max_iter = i;
      code = 0;
      goto done;
    }
    memcpy(x, newx, sizeof(double) * n);
  }
  *tol = resid;
  code = 1;
done:
  free(tv); free(c); free(newx); free(r); free(tmp);
  return code;
}

void orc_project_mean(int n, double *x) {
  double s = 0.0, m;
  int i;
  for (i = 0; i < n; ++i) s += x[i];
  m = s / n;
  for (i = 0; i < n; ++i) x[i] -= m;
}

/* FP32-inner Jacobi-PCG inside FP64 iterative refinement.  No reference
 * counterpart: the outer loop recomputes r = b - A x in FP64, solves
 * A e = r in FP32 (matrix, vectors and accumulation all FP32) to a relative
 * inner tolerance, and updates x += e in FP64. */
int orc_jpcg_mixed(const orc_csr *A, double *x, const double *b, double tol,
                   double inner_tol, int max_outer, int max_inner,
                   int project_mean, int *outer, int *inner, double *resid) {
  int n = A->n, nnz = A->row_ptr[n], it, k, o, code = 1;
  double *r = (double *)malloc(sizeof(double) * n);
  double *tmp = (double *)malloc(sizeof(double) * n);
  float *av = (float *)malloc(sizeof(float) * nnz);
  float *dg = (float *)malloc(sizeof(float) * n);
  float *rf = (float *)malloc(sizeof(float) * n);
  float *ef = (float *)malloc(sizeof(float) * n);
  float *pf = (float *)malloc(sizeof(float) * n);
  float *qf = (float *)malloc(sizeof(float) * n);
  double normb, res;
  double *dd = (double *)malloc(sizeof(double) * n);
  *outer = 0;
  *inner = 0;
  if (diapre(A, dd) != 0) {
    code = -10;
    goto done;
  }
  for (k = 0; k < nnz; ++k) av[k] = (float)A->val[k];
  for (k = 0; k < n; ++k) dg[k] = (float)dd[k];
  normb = orc_nrm2_true(n, b);
  if (normb == 0.0) normb = 1.0;
  for (o = 0; o <= max_outer; ++o) {
    float rho, rho1 = 0.f, alpha, beta, pq, nr0, nr;
    residual(A, x, b, r, tmp);
    if (project_mean) orc_project_mean(n, r);
    res = orc_nrm2_true(n, r) / normb;
    *resid = res;
    if (res <= tol) {
      code = 0;
      break;
    }
    if (o == max_outer) break;
    (*outer)++;
    for (k = 0; k < n; ++k) {
      rf[k] = (float)r[k];
      ef[k] = 0.f;
    }
    nr0 = 0.f;
    for (k = 0; k < n; ++k) nr0 += rf[k] * rf[k];
    nr0 = sqrtf(nr0);
    for (it = 1; it <= max_inner; ++it) {
      int i, kk;
      rho = 0.f;
      for (k = 0; k < n; ++k) rho += rf[k] * (rf[k] * dg[k]);
      if (it == 1) {
        for (k = 0; k < n; ++k) pf[k] = rf[k] * dg[k];
      } else {
        beta = rho / rho1;
        for (k = 0; k < n; ++k) pf[k] = rf[k] * dg[k] + beta * pf[k];
      }
      for (i = 0; i < n; ++i) {
        float acc = 0.f;
        for (kk = A->row_ptr[i]; kk < A->row_ptr[i + 1]; ++kk)
          acc += pf[A->col[kk]] * av[kk];
        qf[i] = acc;
      }
      pq = 0.f;
      for (k = 0; k < n; ++k) pq += pf[k] * qf[k];
      alpha = rho / pq;
      for (k = 0; k < n; ++k) {
        ef[k] += alpha * pf[k];
        rf[k] -= alpha * qf[k];
      }
      (*inner)++;
      nr = 0.f;
      for (k = 0; k < n; ++k) nr += rf[k] * rf[k];
      nr = sqrtf(nr);
      if (nr <= inner_tol * nr0) break;
      rho1 = rho;
    }
    for (k = 0; k < n; ++k) x[k] += (double)ef[k];
    if (project_mean) orc_project_mean(n, x);
  }
done:
  free(r); free(tmp); free(av); free(dg); free(rf); free(ef); free(pf);
  free(qf); free(dd);
  return code;
}

/* ------------------------------------------------------------------ */
/* Equation_::apply_bc (s_equation.hpp:340-378): sweep all IJK with i fastest;
 * for every ghost: segment/orientation/original index from the first axis
 * out of range (s_ghost.hpp:218-320); BC1 linear interpolation through the
 * face value and the original cell (Interp_::Linear, arithmetic.hpp:
 * 422-431); BC2 ghost = original - val*dl. */
void orc_apply_bc(const orc_grid *g, double *f, const orc_bc *bc) {
  int nx = g->nx, ny = g->ny, gl = g->gl, I, J;
  for (J = 0; J < ny + 2 * gl; ++J) {
    for (I = 0; I < nx + 2 * gl; ++I) {
      int i = I - gl, j = J - gl, a, seg, ori, oi, oj, line;
      double xg, xf, xo, val, dl;
      if (i >= 0 && i < nx && j >= 0 && j < ny) continue;
      if (i < 0) { a = 0; seg = 0; ori = -1; oi = 0; oj = j; }
      else if (i >= nx) { a = 0; seg = 1; ori = 1; oi = nx - 1; oj = j; }
      else if (j < 0) { a = 1; seg = 2; ori = -1; oi = i; oj = 0; }
      else { a = 1; seg = 3; ori = 1; oi = i; oj = ny - 1; }
      line = (a == 0 ? j : i) + gl;
      val = bcval(g, bc, seg, line);
      xg = cax(g, a, a == 0 ? i : j);
      xo = cax(g, a, a == 0 ? oi : oj);
      xf = ori > 0 ? xo + sax(g, a, a == 0 ? oi : oj) * 0.5
                   : xo - sax(g, a, a == 0 ? oi : oj) * 0.5;
      dl = fabs(xg - xo);
      if (bc->type[seg] == ORC_BC1) {
        double u = (xg - xf) / (xo - xf);
        f[FIDX(g, i, j)] = (val * (1.0 - u)) + (f[FIDX(g, oi, oj)] * u);
      } else {
        double po = f[FIDX(g, oi, oj)];
        f[FIDX(g, i, j)] = po - val * dl;
      }
    }
  }
}

/* std::min / std::max semantics (the first argument wins ties and NaNs) */
static double stdmin(double a, double b) { return (b < a) ? b : a; }
static double stdmax(double a, double b) { return (a < b) ? b : a; }

static double limiter(int lim, double r, double R) {
  switch (lim) {
    case ORC_LIM_VANLEER: /* s_operation.hpp:501-503 */
      return 0.5 * R * (r + fabs(r)) / (R - 1 + r);
    case ORC_LIM_SUPERBEE: /* :505-507 */
      return stdmax(stdmax(0.0, stdmin(R * r, 1.0)), stdmin(r, R));
    case ORC_LIM_WAHYD: /* :509-515 */
      if (r <= 1.0) return 0.5 * R * (r + fabs(r)) / (R - 1 + r);
      return stdmin((r + R * r * fabs(r)) / (R + r * r), R);
    case ORC_LIM_QUICK: /* :537-539 */
      return 0.75 * r + 0.25;
    case ORC_LIM_CUI: /* :540-542 */
      return 2.0 / 3.0 * r + 1.0 / 3.0;
    default: /* upwind2, :598-600 */
      return r;
  }
}

/* VdotNabla_TVD (s_operation.hpp:549-596) with _RCD/_rCD (:484-499): the
 * upwind side is chosen once per cell from the averaged face velocity and
 * used for both faces. */
double orc_tvd_cell(const orc_grid *g, const double *uf, const double *vf,
                    const double *phi, int i, int j, int lim) {
  double vdd[2] = {0.0, 0.0}, res = 0.0;
  int d;
  for (d = 0; d < 2; ++d) {
    const double *vface = d == 0 ? uf : vf;
    int ei = d == 0 ? 1 : 0, ej = d == 0 ? 0 : 1;
    int Cm, Um, Dm, Cp, Up, Dp; /* offsets along the axis */
    double veo, fp, fm, R, r, sU, sC, sD, vU, vC, vD, vfp, vfm;
    int c = d == 0 ? i : j;
    vfp = vface[FIDX(g, i, j)];
    vfm = vface[FIDX(g, i - ei, j - ej)];
    veo = (vfp + vfm) * 0.5;
    if (veo >= 0.0) {
      Dm = 0; Cm = -1; Um = -2;
      Dp = 1; Cp = 0; Up = -1;
    } else {
      Dm = -1; Cm = 0; Um = 1;
      Dp = 0; Cp = 1; Up = 2;
    }
    if (g_plant == ORC_PLANT_FACE_UPWIND) {
      /* planted: each face upwinded by its own velocity */
      if (vfm >= 0.0) { Dm = 0; Cm = -1; Um = -2; }
      else { Dm = -1; Cm = 0; Um = 1; }
      if (vfp >= 0.0) { Dp = 1; Cp = 0; Up = -1; }
      else { Dp = 0; Cp = 1; Up = 2; }
    }
#define VAL(o) phi[FIDX(g, i + ei * (o), j + ej * (o))]
#define SZ(o) sax(g, d, c + (o))
    sC = SZ(Cp); sD = SZ(Dp); sU = SZ(Up);
    vC = VAL(Cp); vD = VAL(Dp); vU = VAL(Up);
    R = (sD + sC) / sC;
    r = (vC - vU) * (sD + sC) / (vD - vC + ORC_SMALL) / (sC + sU);
    fp = vC + limiter(lim, r, R) / R * (vD - vC);
    sC = SZ(Cm); sD = SZ(Dm); sU = SZ(Um);
    vC = VAL(Cm); vD = VAL(Dm); vU = VAL(Um);
    R = (sD + sC) / sC;
    r = (vC - vU) * (sD + sC) / (vD - vC + ORC_SMALL) / (sC + sU);
    fm = vC + limiter(lim, r, R) / R * (vD - vC);
#undef VAL
#undef SZ
    vdd[d] = (vfp * fp - vfm * fm) / sax(g, d, c);
  }
  for (d = 0; d < 2; ++d) res += vdd[d];
  return res;
}

/* VdotNabla_center (s_operation.hpp:442-459) */
double orc_center_cell(const orc_grid *g, const double *uf, const double *vf,
                       const double *phi, int i, int j) {
  double t[2], res = 0.0;
  int d;
  for (d = 0; d < 2; ++d) {
    const double *f = d == 0 ? uf : vf;
    int ei = d == 0 ? 1 : 0, ej = d == 0 ? 0 : 1, k = d == 0 ? i : j;
    double veo = (f[FIDX(g, i, j)] + f[FIDX(g, i - ei, j - ej)]) * 0.5;
    t[d] = veo * (phi[FIDX(g, i + ei, j + ej)] - phi[FIDX(g, i - ei, j - ej)]) /
           (cax(g, d, k + 1) - cax(g, d, k - 1));
  }
  for (d = 0; d < 2; ++d) res += t[d];
  return res;
}

/* VdotNabla_center4 (s_operation.hpp:461-480): the averaged velocity is
 * formed and not applied, so the term is the fourth-order gradient alone */
double orc_center4_cell(const orc_grid *g, const double *phi, int i, int j) {
  double t[2], res = 0.0;
  int d;
  for (d = 0; d < 2; ++d) {
    int ei = d == 0 ? 1 : 0, ej = d == 0 ? 0 : 1, k = d == 0 ? i : j;
    double inn = (phi[FIDX(g, i + ei, j + ej)] - phi[FIDX(g, i - ei, j - ej)]) /
                 (cax(g, d, k + 1) - cax(g, d, k - 1));
    double out = (phi[FIDX(g, i + 2 * ei, j + 2 * ej)] -
                  phi[FIDX(g, i - 2 * ei, j - 2 * ej)]) /
                 (cax(g, d, k + 2) - cax(g, d, k - 2));
    t[d] = (4.0 * inn - out) / 3.0;
  }
  for (d = 0; d < 2; ++d) res += t[d];
  return res;
}

/* the advection term of Advection_'s scheme (set_advection_scheme,
 * s_advection.hpp:140-182): ORC_ADV_UPWIND1 / CENTER / CENTER4, or a TVD
 * limiter id */
double orc_adv_cell(const orc_grid *g, const double *uf, const double *vf,
                    const double *phi, int i, int j, int scheme) {
  if (scheme == ORC_ADV_UPWIND1) return orc_upwind1_cell(g, uf, vf, phi, i, j);
  if (scheme == ORC_ADV_CENTER) return orc_center_cell(g, uf, vf, phi, i, j);
  if (scheme == ORC_ADV_CENTER4) return orc_center4_cell(g, phi, i, j);
  return orc_tvd_cell(g, uf, vf, phi, i, j, scheme);
}

/* Advection_::run_one_step (s_advection.hpp:185-212): phi's ghosts from its
 * boundary, then every cell in Grid_::begin_ijk order (i fastest,
 * Ijk_::_incr, s_define.hpp:394-410) overwritten in place by
 * phi - dt * adv, so later cells read the new values of earlier ones.  The
 * centre velocities' apply_bc of the step leaves the faces unchanged. */
void orc_advection_step(const orc_grid *g, const double *uf, const double *vf,
                        double *phi, const orc_bc *bc, int scheme, double dt) {
  int i, j;
  const double *src = phi;
  double *old = NULL;
  orc_apply_bc(g, phi, bc);
  if (g_plant == ORC_PLANT_ADV_TWO_BUFFER) {
    /* planted: every cell reads the values from before the sweep */
    size_t nt = (size_t)(g->nx + 2 * g->gl) * (g->ny + 2 * g->gl);
    old = (double *)malloc(sizeof(double) * nt);
    memcpy(old, phi, sizeof(double) * nt);
    src = old;
  }
  for (j = 0; j < g->ny; ++j) {
    for (i = 0; i < g->nx; ++i) {
      double adv = orc_adv_cell(g, uf, vf, src, i, j, scheme);
      double phin = phi[FIDX(g, i, j)];
      phi[FIDX(g, i, j)] = phin - dt * adv;
    }
  }
  free(old);
}

/* NablaMuNabla with uniform mu (s_operation.hpp:664-688) */
double orc_diffusion_cell(const orc_grid *g, double mu, const double *phi,
                          int i, int j) {
  double td[2], res = 0.0;
  int d;
  for (d = 0; d < 2; ++d) {
    int ei = d == 0 ? 1 : 0, ej = d == 0 ? 0 : 1, c = d == 0 ? i : j;
    double f = phi[FIDX(g, i, j)];
    double fp = phi[FIDX(g, i + ei, j + ej)];
    double fm = phi[FIDX(g, i - ei, j - ej)];
    double dm = mu * (f - fm) / (cax(g, d, c) - cax(g, d, c - 1));
    double dp = mu * (fp - f) / (cax(g, d, c + 1) - cax(g, d, c));
    td[d] = (dp - dm) / sax(g, d, c);
  }
  for (d = 0; d < 2; ++d) res += td[d];
  return res;
}

/* InterpolateC2F (s_operation.hpp:166-187): uf from uc along x, vf from vc
 * along y; the m-face at index 0 is written into the ghost slot. */
void orc_interp_c2f(const orc_grid *g, const double *uc, const double *vc,
                    double *uf, double *vf) {
  int i, j, d;
  for (j = 0; j < g->ny; ++j) {
    for (i = 0; i < g->nx; ++i) {
      for (d = 0; d < 2; ++d) {
        const double *cd = d == 0 ? uc : vc;
        double *fd = d == 0 ? uf : vf;
        int ei = d == 0 ? 1 : 0, ej = d == 0 ? 0 : 1, c = d == 0 ? i : j;
        double v = cd[FIDX(g, i, j)];
        double hs = sax(g, d, c) * 0.5;
        double hsp = sax(g, d, c + 1) * 0.5;
        fd[FIDX(g, i, j)] = (cd[FIDX(g, i + ei, j + ej)] * hs + v * hsp) /
                            (hs + hsp);
        if (c == 0) {
          double hsm = sax(g, d, c - 1) * 0.5;
          fd[FIDX(g, i - ei, j - ej)] =
              (cd[FIDX(g, i - ei, j - ej)] * hs + v * hsm) / (hs + hsm);
        }
      }
    }
  }
}

/* Div of a face vector (s_operation.hpp:100-120) */
void orc_div(const orc_grid *g, const double *uf, const double *vf,
             double *div) {
  int i, j;
  for (j = 0; j < g->ny; ++j) {
    for (i = 0; i < g->nx; ++i) {
      double dv = 0.0;
      dv += (uf[FIDX(g, i, j)] - uf[FIDX(g, i - 1, j)]) / sax(g, 0, i);
      dv += (vf[FIDX(g, i, j)] - vf[FIDX(g, i, j - 1)]) / sax(g, 1, j);
      div[FIDX(g, i, j)] = dv;
    }
  }
}

/* Grad to faces (s_operation.hpp:46-72); entries it does not write stay 0
 * (a fresh Scalar_ is zero-filled, s_data.hpp:150-153). */
void orc_grad(const orc_grid *g, const double *phi, double *gx, double *gy) {
  int i, j, d, NT = (g->nx + 2 * g->gl) * (g->ny + 2 * g->gl);
  memset(gx, 0, sizeof(double) * NT);
  memset(gy, 0, sizeof(double) * NT);
  for (j = 0; j < g->ny; ++j) {
    for (i = 0; i < g->nx; ++i) {
      for (d = 0; d < 2; ++d) {
        double *gd = d == 0 ? gx : gy;
        int ei = d == 0 ? 1 : 0, ej = d == 0 ? 0 : 1, c = d == 0 ? i : j;
        double vc = phi[FIDX(g, i, j)];
        gd[FIDX(g, i, j)] = (phi[FIDX(g, i + ei, j + ej)] - vc) /
                            (cax(g, d, c + 1) - cax(g, d, c));
        if (c == 0)
          gd[FIDX(g, i - ei, j - ej)] = (phi[FIDX(g, i - ei, j - ej)] - vc) /
                                        (cax(g, d, c - 1) - cax(g, d, c));
      }
    }
  }
}

/* InterpolateF2C (s_operation.hpp:141-164) into fresh zero fields */
void orc_interp_f2c(const orc_grid *g, const double *gx, const double *gy,
                    double *cx, double *cy) {
  int i, j, NT = (g->nx + 2 * g->gl) * (g->ny + 2 * g->gl);
  memset(cx, 0, sizeof(double) * NT);
  memset(cy, 0, sizeof(double) * NT);
  for (j = 0; j < g->ny; ++j) {
    for (i = 0; i < g->nx; ++i) {
      cx[FIDX(g, i, j)] = (gx[FIDX(g, i, j)] + gx[FIDX(g, i - 1, j)]) * 0.5;
      cy[FIDX(g, i, j)] = (gy[FIDX(g, i, j)] + gy[FIDX(g, i, j - 1)]) * 0.5;
    }
  }
}

/* VectorCenter_::max_magnitude (s_vector.hpp:89-103) */
double orc_max_magnitude(const orc_grid *g, const double *u, const double *v) {
  double mx = 0.0;
  int i, j;
  for (j = 0; j < g->ny; ++j) {
    for (i = 0; i < g->nx; ++i) {
      double sum = 0.0, s;
      sum += u[FIDX(g, i, j)] * u[FIDX(g, i, j)];
      sum += v[FIDX(g, i, j)] * v[FIDX(g, i, j)];
      s = sqrt(sum);
      mx = stdmax(mx, s);
    }
  }
  return mx;
}

/* Operation_::Error1, Error2, Errori (s_operation.hpp:869-908) over the
 * normal cells in i-fastest order, volume (1.0*sx)*sy (Grid_::volume,
 * s_grid.hpp:97-104).  Error2 is sqrt(sum err^2 vol) / sum vol as the
 * reference writes it (not the root of the volume mean); Errori keeps the
 * first maximum ('>' test). */
void orc_error_norms(const orc_grid *g, const double *exa, const double *val,
                     double *out) {
  double n1 = 0.0, n2 = 0.0, svol = 0.0, ni = 0.0;
  int i, j;
  for (j = 0; j < g->ny; ++j) {
    for (i = 0; i < g->nx; ++i) {
      double vol = 1.0, err, ae;
      vol *= g->sx[i + g->gl];
      vol *= g->sy[j + g->gl];
      err = val[FIDX(g, i, j)] - exa[FIDX(g, i, j)];
      n1 += fabs(err) * vol;
      n2 += err * err * vol;
      svol += vol;
      ae = fabs(err);
      if (ae > ni) ni = ae;
    }
  }
  out[0] = n1 / svol;
  out[1] = sqrt(n2) / svol;
  out[2] = ni;
}

static double min_size(const double *s, int n) {
  double m = s[0];
  int i;
  for (i = 1; i < n; ++i) m = stdmin(m, s[i]);
  return m;
}

/* NS_::restrict_time (s_ns.hpp:146-169) with TimeRestrict_CourantNumber
 * and TimeRestrict_DiffusionNumber (s_operation.hpp:847-867): dt only
 * shrinks. */
void orc_ns_restrict_time(const orc_grid *g, const orc_ns_params *prm,
                          orc_ns_state *s) {
  int NX = g->nx + 2 * g->gl, NY = g->ny + 2 * g->gl;
  double mx = orc_max_magnitude(g, s->u, s->v);
  double msx = min_size(g->sx, NX), msy = min_size(g->sy, NY);
  double minsize = msx < msy ? msx : msy;
  double rdt = prm->cfl * minsize / mx;
  double rdt2 = s->dt, m;
  if (prm->dnum < 100) {
    double den = 0.0;
    den += 1.0 / (msx * msx);
    den += 1.0 / (msy * msy);
    rdt2 = prm->dnum / prm->dcoef / den;
  }
  m = stdmin(rdt, rdt2);
  if (s->dt > m) s->dt = m;
}

/* NS_explicit_::run_one_step (s_ns_explicit.hpp:118-155) with
 * _veo_star_center (:173-229), followed by Time_::advance (s_time.hpp:
 * 61-64).  The projection is the reference's Poisson_ with an empty
 * BoundaryIndex (all-Neumann, value 0), source = div(u*), warm-started from
 * the previous phi, solved here with Jacobi-PCG. */
/* The projection shared by both steps: ghosts of u*, C2F + Div, the
 * all-Neumann Poisson solve (J-PCG), grad phi, p and the corrections
 * (s_ns_explicit.hpp:134-153, s_ns_kim.hpp:160-203). */
static int ns_project(const orc_grid *g, const orc_csr *A, const orc_bc *bcu,
                      const orc_bc *bcv, const orc_ns_params *prm,
                      orc_ns_state *s, int *iters, double *resid, int kim) {
  int nx = g->nx, ny = g->ny, i, j, NT, n = nx * ny, code, k;
  int NX = nx + 2 * g->gl, NY = ny + 2 * g->gl;
  double dt = s->dt, *x, *b, *gx, *gy, *cgx, *cgy, tol;
  orc_bc bcphi = {{ORC_BC2, ORC_BC2, ORC_BC2, ORC_BC2}, NULL};
  int mi;
  NT = NX * NY;
  orc_apply_bc(g, s->us, bcu);
  orc_apply_bc(g, s->vs, bcv);
  orc_interp_c2f(g, s->us, s->vs, s->usf, s->vsf);
  orc_div(g, s->usf, s->vsf, s->div);
  /* projection: b = -(vol*div) per Poisson_/StencilPoisson_; x = phi */
  x = (double *)malloc(sizeof(double) * n);
  b = (double *)malloc(sizeof(double) * n);
  orc_poisson_rhs(g, 1.0, s->div, 0.0, &bcphi, b);
  orc_field_to_x(g, s->phi, x);
  mi = prm->max_iter;
  tol = prm->tol;
  {
    int hl;
    code = orc_jpcg(A, x, b, &mi, &tol, prm->norm_mode, NULL, 0, &hl);
  }
  *iters = mi;
  *resid = tol;
  orc_x_to_field(g, x, s->phi);
  free(x);
  free(b);
  orc_apply_bc(g, s->phi, &bcphi);
  gx = (double *)malloc(sizeof(double) * NT);
  gy = (double *)malloc(sizeof(double) * NT);
  cgx = (double *)malloc(sizeof(double) * NT);
  cgy = (double *)malloc(sizeof(double) * NT);
  orc_grad(g, s->phi, gx, gy);
  memcpy(s->p, s->phi, sizeof(double) * NT);
  for (j = 0; j < ny; ++j)
    for (i = 0; i < nx; ++i) s->p[FIDX(g, i, j)] = s->p[FIDX(g, i, j)] * (1.0 / dt);
  if (kim) {
    /* NS_kim_: p -= (0.5/rho) * NablaMuNabla(mu, phi) over the interior
     * (s_ns_kim.hpp:183-192) */
    for (j = 0; j < ny; ++j)
      for (i = 0; i < nx; ++i)
        s->p[FIDX(g, i, j)] = s->p[FIDX(g, i, j)] -
                              orc_diffusion_cell(g, prm->mu, s->phi, i, j) *
                                  (0.5 / prm->rho);
  }
  for (k = 0; k < NT; ++k) {
    s->usf[k] = s->usf[k] - gx[k];
    s->vsf[k] = s->vsf[k] - gy[k];
  }
  orc_interp_f2c(g, gx, gy, cgx, cgy);
  for (k = 0; k < NT; ++k) {
    s->us[k] = s->us[k] - cgx[k];
    s->vs[k] = s->vs[k] - cgy[k];
  }
  memcpy(s->u, s->us, sizeof(double) * NT);
  memcpy(s->v, s->vs, sizeof(double) * NT);
  memcpy(s->uf, s->usf, sizeof(double) * NT);
  memcpy(s->vf, s->vsf, sizeof(double) * NT);
  free(gx); free(gy); free(cgx); free(cgy);
  s->t += dt;
  return code;
}

int orc_ns_step(const orc_grid *g, const orc_csr *A, const orc_bc *bcu,
                const orc_bc *bcv, const orc_ns_params *prm, orc_ns_state *s,
                int *iters, double *resid) {
  int nx = g->nx, ny = g->ny, i, j, k;
  double dt;
  orc_ns_restrict_time(g, prm, s);
  dt = s->dt;
  /* u* at centres */
  orc_apply_bc(g, s->u, bcu);
  orc_apply_bc(g, s->v, bcv);
  for (k = 0; k < 2; ++k) {
    double *sc = k == 0 ? s->u : s->v;
    double *ss = k == 0 ? s->us : s->vs;
    ORC_PAR
    for (j = 0; j < ny; ++j) {
      int i;
      for (i = 0; i < nx; ++i) {
        double adv = orc_tvd_cell(g, s->uf, s->vf, sc, i, j, prm->limiter);
        double dif = orc_diffusion_cell(g, prm->mu, sc, i, j);
        double src = 0.0;
        if (g_plant == ORC_PLANT_ADV_SIGN) adv = -adv;
        if (g_plant == ORC_PLANT_DIF_RHO)
          ss[FIDX(g, i, j)] =
              (-adv + (dif + src) * prm->rho) * dt + sc[FIDX(g, i, j)];
        else
          ss[FIDX(g, i, j)] =
              (-adv + (dif + src) / prm->rho) * dt + sc[FIDX(g, i, j)];
      }
    }
  }
  return ns_project(g, A, bcu, bcv, prm, s, iters, resid, 0);
}

/* VdotNabla_upwind1 at one cell (s_operation.hpp:400-422), the NS_ default
 * advection (s_ns.hpp:225). */
double orc_upwind1_cell(const orc_grid *g, const double *uf, const double *vf,
                        const double *phi, int i, int j) {
  double t[2] = {0.0, 0.0}, r = 0.0;
  int d;
  for (d = 0; d < 2; ++d) {
    const double *f = d == 0 ? uf : vf;
    int mi = d == 0 ? i - 1 : i, mj = d == 0 ? j : j - 1;
    int pi = d == 0 ? i + 1 : i, pj = d == 0 ? j : j + 1;
    int c = d == 0 ? i : j;
    double veo = (f[FIDX(g, i, j)] + f[FIDX(g, mi, mj)]) * 0.5;
    if (veo > 0.0)
      t[d] = veo * (phi[FIDX(g, i, j)] - phi[FIDX(g, mi, mj)]) /
             (cax(g, d, c) - cax(g, d, c - 1));
    else if (veo < 0.0)
      t[d] = veo * (phi[FIDX(g, pi, pj)] - phi[FIDX(g, i, j)]) /
             (cax(g, d, c + 1) - cax(g, d, c));
  }
  r += t[0];
  r += t[1];
  return r;
}

static int solve_by(int method, const orc_csr *A, double *x, const double *b,
                    int *mi, double *tol, int norm_mode) {
  int hl;
  switch (method) {
    case ORC_SOLVE_CG: return orc_cg(A, x, b, mi, tol, norm_mode, NULL, 0, &hl);
    case ORC_SOLVE_JPCG: return orc_jpcg(A, x, b, mi, tol, norm_mode, NULL, 0, &hl);
    case ORC_SOLVE_BICGSTAB:
      return orc_dia_bicgstab(A, x, b, mi, tol, norm_mode, NULL, 0, &hl);
    default: return orc_jacobi(A, x, b, mi, tol, norm_mode, NULL, 0, &hl);
  }
}

/* NS_kim_::run_one_step + Time_::advance (s_ns_kim.hpp:130-205, :222-287).
 * No restrict_time (commented out there).  The step first copies u (ghosts
 * included, Operation_::Copy sweeps IJK, s_operation.hpp:818-845) into um
 * and uf into umf (:132-133), then apply_bc(u) refreshes u's ghosts
 * (:225-227).  um therefore keeps the ghost values u carried out of the
 * previous step: the velocity walls applied to u* before its correction
 * (the correction's InterpolateF2C leaves ghosts at 0, s_operation.hpp:
 * 141-164).  The predictor's source per component is
 * (adv - (0.5*dif + 0)/rho)*dt - u with adv = upwind1(uf, u) at step 0 and
 * 1.5*upwind1(uf, u) - 0.5*upwind1(umf, um) after, umf == uf.  Each
 * component then solves the Helmholtz problem alpha = 1,
 * beta = dt*0.5/rho (mu does not enter, :267-272) from the current u* as the
 * first guess.  The predictor is its own Poisson_ (:105) with its own empty
 * BoundaryIndex (s_equation.hpp:104), so its rows see the default boundary
 * condition, BC2 with value 0 (s_boundary.hpp:29-36, :195-207), on every
 * wall: H[k] is that matrix, assembled by the caller; the velocity walls
 * reach u* only through apply_bc(us) afterwards (:285-287). */
int orc_ns_kim_step(const orc_grid *g, const orc_csr *A, const orc_csr *Hu,
                    const orc_csr *Hv, const orc_bc *bcu, const orc_bc *bcv,
                    const orc_ns_params *prm, const orc_kim_params *kp,
                    int step, orc_ns_state *s, int *iters, double *resid) {
  int nx = g->nx, ny = g->ny, i, j, k, n = nx * ny, code;
  int NT = (nx + 2 * g->gl) * (ny + 2 * g->gl);
  double dt = s->dt, beta = dt * 0.5 / prm->rho;
  double *sscr = (double *)calloc((size_t)NT, sizeof(double));
  double *um = (double *)malloc(sizeof(double) * NT);
  double *vm = (double *)malloc(sizeof(double) * NT);
  double *x = (double *)malloc(sizeof(double) * n);
  double *b = (double *)malloc(sizeof(double) * n);
  orc_bc bcpred = {{ORC_BC2, ORC_BC2, ORC_BC2, ORC_BC2}, NULL};
  memcpy(um, s->u, sizeof(double) * NT);
  memcpy(vm, s->v, sizeof(double) * NT);
  orc_apply_bc(g, s->u, bcu);
  orc_apply_bc(g, s->v, bcv);
  if (g_plant == ORC_PLANT_KIM_AB2_GHOSTS) {
    memcpy(um, s->u, sizeof(double) * NT);
    memcpy(vm, s->v, sizeof(double) * NT);
  }
  for (k = 0; k < 2; ++k) {
    double *sc = k == 0 ? s->u : s->v;
    double *sm = k == 0 ? um : vm;
    double *ss = k == 0 ? s->us : s->vs;
    int mi = kp->max_iter;
    double tol = kp->tol;
    for (j = 0; j < ny; ++j) {
      for (i = 0; i < nx; ++i) {
        double advn = orc_upwind1_cell(g, s->uf, s->vf, sc, i, j);
        double advnm1 = orc_upwind1_cell(g, s->uf, s->vf, sm, i, j);
        double adv = 1.5 * advn - 0.5 * advnm1;
        double dif = orc_diffusion_cell(g, prm->mu, sc, i, j);
        double src = 0.0;
        double cur = sc[FIDX(g, i, j)];
        sscr[FIDX(g, i, j)] =
            step == 0 ? (advn - (0.5 * dif + src) / prm->rho) * dt - cur
                      : (adv - (0.5 * dif + src) / prm->rho) * dt - cur;
      }
    }
    if (s->kim_src) memcpy(s->kim_src + (size_t)k * NT, sscr, sizeof(double) * NT);
    orc_helmholtz_rhs(g, 1.0, beta, sscr, 0.0,
                      g_plant == ORC_PLANT_KIM_PRED_WALLS ? (k == 0 ? bcu : bcv)
                                                          : &bcpred,
                      b);
    orc_field_to_x(g, ss, x);
    code = solve_by(kp->method, k == 0 ? Hu : Hv, x, b, &mi, &tol,
                    prm->norm_mode);
    if (iters) iters[1 + k] = mi;
    (void)code;
    orc_x_to_field(g, x, ss);
  }
  free(sscr);
  free(um);
  free(vm);
  free(x);
  free(b);
  code = ns_project(g, A, bcu, bcv, prm, s, iters, resid, 1);
  memcpy(s->us, s->u, sizeof(double) * NT);
  memcpy(s->vs, s->v, sizeof(double) * NT);
  return code;
}
