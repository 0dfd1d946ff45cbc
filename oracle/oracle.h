/*
 * CPU oracle for the structured-grid projection hot path of Carpios.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm, used by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg as the checker.  The product path (libcpgpu) never links,
 * loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Arithmetic is written in the reference's
 * evaluation order and built with -ffp-contract=off so results are bitwise
 * those of a non-FMA build of the reference.
 *
 * Layout conventions (reference lib/structure/s_data.hpp:134-137,
 * lib/algebra/space.hpp:185-210): a 2-D cell field is a row-major
 * (nx+2gl) x (ny+2gl) double buffer, j (y) contiguous; cell (i,j) with
 * -gl <= i < nx+gl lives at ((i+gl)*(ny+2gl) + (j+gl)).  Face fields share the
 * buffer shape: face component d at index (i,j) is the face between (i,j) and
 * (i,j)+e_d (reference lib/structure/s_operation.hpp:166-187).
 *
 * Unknown numbering for the assembled system is i-fastest lexicographic
 * (k = i + nx*j), the iteration order of Ijk_ (s_define.hpp:394-410); the
 * reference's OrderRegular_ is absent from the reference tree, so this order
 * is an assumption (it fixes which element nrm2 drops and the dot order).
 */
#ifndef CARPIOS_ORACLE_H
#define CARPIOS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_SMALL 1e-10 /* lib/type_define.hpp:16 */

/* boundary types, lib/structure/s_boundary.hpp:16-18 */
#define ORC_BC1 1 /* Dirichlet */
#define ORC_BC2 2 /* Neumann */

/* norm modes for the Krylov stopping test */
#define ORC_NORM_REF 0  /* reference nrm2: drops the last element, skips zeros */
#define ORC_NORM_TRUE 1 /* plain Euclidean norm */

/* limiter ids, lib/structure/s_operation.hpp:501-546 */
#define ORC_LIM_VANLEER 0
#define ORC_LIM_SUPERBEE 1
#define ORC_LIM_WAHYD 2
#define ORC_LIM_QUICK 3
#define ORC_LIM_CUI 4
#define ORC_LIM_UPWIND2 5

typedef struct {
  int nx, ny, gl;
  const double *cx, *sx; /* length nx+2gl: centre coordinate and size per i */
  const double *cy, *sy; /* length ny+2gl */
} orc_grid;

typedef struct {
  int n;
  const int *row_ptr;
  const int *col;
  const double *val;
} orc_csr;

/* Boundary description for one scalar.  type[s] for segments
 * 0:x-, 1:x+, 2:y-, 3:y+ (s_ghost.hpp:218-232).  val holds the boundary
 * value sampled per ghost line: segments 0,1 have ny+2gl entries indexed by
 * J=j+gl, segments 2,3 have nx+2gl entries indexed by I=i+gl, concatenated
 * as [seg0 | seg1 | seg2 | seg3].  NULL val means all zeros. */
typedef struct {
  int type[4];
  const double *val;
} orc_bc;

/* ---- grid ---- */
void orc_axis_uniform(int n, int gl, double vmin, double vmax, double *c,
                      double *s);
/* Grid_::_set_non_uniform_grid: n interior sizes, ghosts repeat the ends */
void orc_axis_sizes(int n, int gl, double vmin, const double *sizes, double *c,
                    double *s);

/* ---- BLAS-1 ---- */
double orc_dot(int n, const double *x, const double *y);
double orc_nrm2_ref(int n, const double *x);
double orc_nrm2_true(int n, const double *x);
void orc_csr_mult(const orc_csr *A, const double *x, double *y);

/* ---- Poisson assembly ---- */
int orc_poisson_nnz(int nx, int ny);
int orc_poisson_assemble(const orc_grid *g, double beta,
                         const double *src_field, double src_uniform,
                         const orc_bc *bc, int *row_ptr, int *col, double *val,
                         double *b);
void orc_poisson_rhs(const orc_grid *g, double beta, const double *src_field,
                     double src_uniform, const orc_bc *bc, double *b);
int orc_helmholtz_assemble(const orc_grid *g, double alpha, double beta,
                           const double *src_field, double src_uniform,
                           const orc_bc *bc, int *row_ptr, int *col,
                           double *val, double *b);
void orc_helmholtz_rhs(const orc_grid *g, double alpha, double beta,
                       const double *src_field, double src_uniform,
                       const orc_bc *bc, double *b);
void orc_field_to_x(const orc_grid *g, const double *field, double *x);
void orc_x_to_field(const orc_grid *g, const double *x, double *field);

/* ---- Krylov / stationary solvers (CSR) ----
 * All return the reference codes: 0 converged, 1 max_iter, 2/3 BiCGSTAB
 * breakdowns; -1 invalid argument; -10 zero diagonal (the reference exits).
 * max_iter and tol are in/out exactly as in the reference.  hist receives
 * the residual history (capacity hist_cap), *hist_len its length. */
int orc_cg(const orc_csr *A, double *x, const double *b, int *max_iter,
           double *tol, int norm_mode, double *hist, int hist_cap,
           int *hist_len);
int orc_jpcg(const orc_csr *A, double *x, const double *b, int *max_iter,
             double *tol, int norm_mode, double *hist, int hist_cap,
             int *hist_len);
int orc_dia_bicgstab(const orc_csr *A, double *x, const double *b,
                     int *max_iter, double *tol, int norm_mode, double *hist,
                     int hist_cap, int *hist_len);
int orc_jacobi(const orc_csr *A, double *x, const double *b, int *max_iter,
               double *tol, int norm_mode, double *hist, int hist_cap,
               int *hist_len);
/* FP32-inner J-PCG wrapped in FP64 iterative refinement (no reference
 * counterpart).  Returns outer iterations in *outer, total inner in *inner. */
int orc_jpcg_mixed(const orc_csr *A, double *x, const double *b, double tol,
                   double inner_tol, int max_outer, int max_inner,
                   int project_mean, int *outer, int *inner, double *resid);
void orc_project_mean(int n, double *x);

/* ---- structured operations (lib/structure/s_operation.hpp) ---- */
void orc_apply_bc(const orc_grid *g, double *f, const orc_bc *bc);
double orc_tvd_cell(const orc_grid *g, const double *uf, const double *vf,
                    const double *phi, int i, int j, int limiter);
double orc_diffusion_cell(const orc_grid *g, double mu, const double *phi,
                          int i, int j);
void orc_interp_c2f(const orc_grid *g, const double *uc, const double *vc,
                    double *uf, double *vf);
void orc_div(const orc_grid *g, const double *uf, const double *vf,
             double *div);
void orc_grad(const orc_grid *g, const double *phi, double *gx, double *gy);
void orc_interp_f2c(const orc_grid *g, const double *gx, const double *gy,
                    double *cx, double *cy);
double orc_max_magnitude(const orc_grid *g, const double *u, const double *v);
/* Operation_::Error1 / Error2 / Errori of val against exa (both ghost-padded
 * fields of g): out = {Error1, Error2, Errori}. */
void orc_error_norms(const orc_grid *g, const double *exa, const double *val,
                     double *out);

/* ---- NS_explicit_ projection step ---- */
typedef struct {
  double rho, mu;              /* uniform rho and mu */
  double cfl;                  /* CFL number */
  double dnum, dcoef;          /* diffusion number / coefficient */
  int limiter;                 /* ORC_LIM_* */
  int max_iter;                /* projection solver limits */
  double tol;
  int norm_mode;
} orc_ns_params;

typedef struct {
  double *u, *v;     /* centre velocity */
  double *uf, *vf;   /* face velocity */
  double *us, *vs;   /* centre velocity star */
  double *usf, *vsf; /* face velocity star */
  double *p, *phi;   /* pressure and projection potential */
  double *div;       /* divergence scratch */
  double dt, t;
  double *kim_src;   /* NULL, or 2 fields: NS_kim_'s predictor sources of
                        u and v from the last step (sscr, s_ns_kim.hpp:262) */
} orc_ns_state;

/* One NS_explicit_::run_one_step followed by Time_::advance.  A is the
 * all-Neumann projection matrix (assembled once by the caller with beta=1),
 * iters receives the solver's iteration count.  Returns the solver code. */
int orc_ns_step(const orc_grid *g, const orc_csr *A, const orc_bc *bcu,
                const orc_bc *bcv, const orc_ns_params *prm, orc_ns_state *s,
                int *iters, double *resid);
void orc_ns_restrict_time(const orc_grid *g, const orc_ns_params *prm,
                          orc_ns_state *s);

/* ---- NS_kim_ semi-implicit step ---- */
#define ORC_SOLVE_CG 0
#define ORC_SOLVE_JPCG 1
#define ORC_SOLVE_BICGSTAB 2
#define ORC_SOLVE_JACOBI 3 /* Solver_Jacobi_, the predictor's default */
typedef struct {
  int method;   /* ORC_SOLVE_* of the Helmholtz predictor */
  int max_iter; /* Solver_ defaults: 1000 and 1e-4 (s_solver.hpp:73-74) */
  double tol;
} orc_kim_params;
/* Advection_ (s_advection.hpp:20-217): schemes beyond the TVD limiter ids */
#define ORC_ADV_UPWIND1 10
#define ORC_ADV_CENTER 11
#define ORC_ADV_CENTER4 12
double orc_center_cell(const orc_grid *g, const double *uf, const double *vf,
                       const double *phi, int i, int j);
double orc_center4_cell(const orc_grid *g, const double *phi, int i, int j);
double orc_adv_cell(const orc_grid *g, const double *uf, const double *vf,
                    const double *phi, int i, int j, int scheme);
void orc_advection_step(const orc_grid *g, const double *uf, const double *vf,
                        double *phi, const orc_bc *bc, int scheme, double dt);
double orc_upwind1_cell(const orc_grid *g, const double *uf, const double *vf,
                        const double *phi, int i, int j);
/* iters[0] projection, iters[1..2] the u and v Helmholtz solves */
int orc_ns_kim_step(const orc_grid *g, const orc_csr *A, const orc_csr *Hu,
                    const orc_csr *Hv, const orc_bc *bcu, const orc_bc *bcv,
                    const orc_ns_params *prm, const orc_kim_params *kp,
                    int step, orc_ns_state *s, int *iters, double *resid);

/* ---- planted mistakes (tests only) ----
 * A test switches one deliberate error on to show that a pin fails when the
 * restatement is wrong; 0 (the default) is the faithful restatement.
 *   1  the advection term enters u* with the wrong sign
 *   2  the diffusion term is multiplied by rho instead of divided
 *   3  TVD picks the upwind side per face from the face velocity
 *   4  the Neumann ghost constant enters a Poisson row with the wrong sign
 *   5  NS_kim_'s predictor rows take the velocity walls, not BC2 value 0
 *   6  NS_kim_'s AB2 term reads u with refreshed ghosts for both levels */
#define ORC_PLANT_ADV_SIGN 1
#define ORC_PLANT_DIF_RHO 2
#define ORC_PLANT_FACE_UPWIND 3
#define ORC_PLANT_NEUMANN_SIGN 4
#define ORC_PLANT_KIM_PRED_WALLS 5
#define ORC_PLANT_KIM_AB2_GHOSTS 6
#define ORC_PLANT_ADV_TWO_BUFFER 7
void orc_set_plant(int plant);
/* threads of the CPU baseline (1: the serial reference order) */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
