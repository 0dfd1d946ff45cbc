/*
 * cpgpu.h -- C-ABI of the MI355X (gfx950) structured-grid projection path of
 * Carpios: matrix-free 5-point Poisson solves and the NS_explicit_ projection
 * step, with fields resident in HBM.
 *
 * Plain C types only.  Every entry point returns an int status: the
 * reference's solver codes where a solver runs (0 converged, 1 max_iter,
 * 2 rho breakdown, 3 omega breakdown; lib/algebra/solver_matrix.hpp:
 * 599-672) and negative CPGPU_E* codes for invalid arguments and HIP
 * failures.  Nothing throws across the ABI and nothing calls exit().  The
 * library never frees or retains caller pointers.  A context owns one HIP
 * stream; all work is stream-ordered on it and every reduction is
 * deterministic (run-to-run bitwise identical).
 *
 * Field layout (lib/structure/s_data.hpp:134-137): a cell field is a
 * row-major (nx+2gl) x (ny+2gl) double buffer, j contiguous, cell (i,j) at
 * ((i+gl)*(ny+2gl) + (j+gl)) -- the buffer of a structure::Scalar_<2>.
 */
#ifndef CPGPU_H
#define CPGPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define CPGPU_OK 0
#define CPGPU_MAXITER 1
#define CPGPU_BREAKDOWN_RHO 2
#define CPGPU_BREAKDOWN_OMEGA 3
#define CPGPU_EINVAL (-1)
#define CPGPU_EHIP (-2)
#define CPGPU_EUNSUPPORTED (-3)
#define CPGPU_ENOMEM (-4)
#define CPGPU_EZERODIAG (-10) /* DiaPre zero diagonal: the reference exit()s */

/* ---- boundary types, lib/structure/s_boundary.hpp:16-18 ---- */
#define CPGPU_BC_DIRICHLET 1 /* BC1 */
#define CPGPU_BC_NEUMANN 2   /* BC2 (also the reference default, s_boundary.hpp:29-36) */

/* ---- solver methods ---- */
#define CPGPU_CG 0        /* carpio::CG, solver_matrix.hpp:264-316 */
#define CPGPU_JPCG 1      /* IC_CG loop (:318-374) with DiaPre(A,1) (preconditioner.hpp:182-262) */
#define CPGPU_BICGSTAB 2  /* carpio::Dia_BiCGSTAB, solver_matrix.hpp:599-672 */
#define CPGPU_JACOBI 3    /* carpio::Solver_Jacobi_, solver_matrix.hpp:820-910 */
#define CPGPU_MIXED 4     /* FP32 J-PCG with FP64 residual replacement (new) */
#define CPGPU_MIXED_RESTART 5 /* FP32 J-PCG restarted inside FP64 refinement (new) */
#define CPGPU_FCG_UNET 6  /* FP64 flexible CG with the U-Net preconditioner,
                             J-PCG fallback (new; cpgpu_poisson_set_unet) */

/* ---- stopping-norm modes ---- */
#define CPGPU_NORM_REF 0  /* carpio::nrm2: drops the last unknown (BLAS_Level1.hpp:569) */
#define CPGPU_NORM_TRUE 1 /* plain Euclidean norm */

/* ---- option flags for solves ---- */
#define CPGPU_OPT_PROJECT_MEAN 1 /* all-Neumann: zero-mean b and the update (new) */

typedef struct cpgpu_ctx cpgpu_ctx;
typedef struct cpgpu_poisson cpgpu_poisson;

/* Matrix-free description of structure::StencilPoisson_<2> +
 * GhostRegular_<2> (s_poisson.hpp:60-181, s_ghost.hpp:140-357) on a
 * structure::Grid_<2> (s_grid.hpp:10-440).  cx/sx (cy/sy) are the grid's
 * per-axis centre and size arrays including ghosts (nx+2gl, ny+2gl
 * entries, Grid_::_c and Grid_::_cs).  bc_type[4] for segments x-, x+, y-,
 * y+ (s_ghost.hpp:218-232); bc_val is NULL (all zero) or the packed
 * per-line boundary values [x- (ny+2gl) | x+ (ny+2gl) | y- (nx+2gl) |
 * y+ (nx+2gl)] indexed by the ghost-inclusive line index, sampled where the
 * reference's BoundaryCondition::get_val is evaluated (s_ghost.hpp:161-165). */
typedef struct {
  int nx, ny, gl;
  const double *cx, *sx;
  const double *cy, *sy;
  double beta; /* uniform beta (Poisson_::set_uniform_beta, s_poisson.hpp:454-458) */
  int bc_type[4];
  const double *bc_val;
} cpgpu_grid2d;

/* ---- grid helper ---- */
/* Centre and size arrays (n+2gl entries, ghosts included) of one uniform
 * axis of a structure::Grid_ (Grid_::_set_uniform_grid, s_grid.hpp:
 * 372-395): sizes (vmax-vmin)/n, centres accumulated cell by cell. */
int cpgpu_grid_uniform_axis(int n, int gl, double vmin, double vmax, double *c,
                            double *s);
/* The same arrays of an axis given its n interior cell sizes
 * (Grid_::_set_non_uniform_grid, s_grid.hpp:397-434): each ghost layer
 * repeats the nearest interior size, centres accumulated from vmin.  Grids
 * whose sizes vary take per-face coefficient arrays (FX, FY, centre
 * coefficients: 3 N doubles more) and solve with CPGPU_JPCG or CPGPU_CG;
 * the projection step (cpgpu_ns_*) runs on them. */
int cpgpu_grid_axis_sizes(int n, int gl, double vmin, const double *sizes,
                          double *c, double *s);

/* ---- context ---- */
/* Replaces: nothing in the reference (it is single-threaded host code); owns
 * the device, one stream and the scratch of the solvers. */
int cpgpu_create(int device, cpgpu_ctx **out);
int cpgpu_destroy(cpgpu_ctx *ctx);
const char *cpgpu_last_error(const cpgpu_ctx *ctx);
int cpgpu_synchronize(cpgpu_ctx *ctx);
/* The context's HIP stream as an opaque pointer (hipStream_t). */
void *cpgpu_stream(cpgpu_ctx *ctx);
/* Measured device copy bandwidth (no reference counterpart): reps copies of
 * `bytes` between two fresh buffers by a 16-byte-per-lane grid-stride
 * kernel at 2, 4, 8 and 32 blocks per CU, timed with events; the fastest
 * grid counts: *gbs = 2 * bytes / time (read + write), *us = time per copy. */
int cpgpu_copy_bandwidth(cpgpu_ctx *ctx, size_t bytes, int reps, double *gbs,
                         double *us);

/* ---- Poisson problem (device resident) ----
 * Replaces the assembly + solve of structure::Solver_<2>::solve()
 * (s_solver.hpp:94-171, :261-291): the operator is built once from the
 * descriptor and applied matrix-free. */
int cpgpu_poisson_create(cpgpu_ctx *ctx, const cpgpu_grid2d *g,
                         cpgpu_poisson **out);
int cpgpu_poisson_destroy(cpgpu_poisson *P);

/* b = -(vol*src + ghost constants), as Solver_::_build_matrix
 * (s_solver.hpp:122) builds it; src is a host cell field or NULL for the
 * uniform value (Poisson_::set_uniform_source, s_poisson.hpp:487-491). */
int cpgpu_poisson_set_source(cpgpu_poisson *P, const double *src_field,
                             double src_uniform);
/* Same with a device cell field (interior-only nx*ny, j contiguous). */
int cpgpu_poisson_set_source_device(cpgpu_poisson *P, const double *d_src);
/* Synthetic source (no reference counterpart): a sum of nblobs Gaussian
 * blobs A_k exp(-|x-c_k|^2/(2 w_k^2)) at the cell centres with c_k uniform
 * in [0,1)^2, w_k uniform in [wmin,wmax), A_k standard normal, drawn from
 * xoshiro256** seeded by SplitMix64(seed); the cell mean is subtracted. */
int cpgpu_poisson_set_source_blobs(cpgpu_poisson *P, unsigned long long seed,
                                   int nblobs, double wmin, double wmax);
/* The blob parameters of that generator: out[4k..4k+3] = cx, cy, w, A. */
int cpgpu_blob_params(unsigned long long seed, int nblobs, double wmin,
                      double wmax, double *out);
/* Replaces Poisson_::set_uniform_alpha (s_poisson.hpp:461-467): the rows
 * become StencilHelmholtz_'s (s_poisson.hpp:185-260): the centre gets
 * v*alpha before the ghost substitutions and the source weight is v, with
 * v the product of the cell sizes.  Uniform cells only (else
 * CPGPU_EUNSUPPORTED).  May follow solves: the captured iteration graphs
 * are dropped, and b is rebuilt when its source was a host field, a
 * uniform value or the blob generator; after cpgpu_poisson_set_source_device
 * it returns CPGPU_EINVAL (set alpha first, then the device source).  A
 * batch (cpgpu_batch_create) over this problem must be created after it. */
int cpgpu_poisson_set_alpha(cpgpu_poisson *P, double alpha);
/* New boundary values (packed as cpgpu_grid2d.bc_val) for time-dependent
 * boundary conditions: the reference samples BoundaryCondition::get_val at
 * each assembly (s_ghost.hpp:161-165).  Takes effect at the next source. */
int cpgpu_poisson_set_bc_values(cpgpu_poisson *P, const double *bc_val);
/* Warm start / result: host cell fields (interior read or written). */
int cpgpu_poisson_set_phi(cpgpu_poisson *P, const double *phi_field);
int cpgpu_poisson_get_phi(cpgpu_poisson *P, double *phi_field);
/* Device-side fill of phi (a zero initial guess without a host transfer). */
int cpgpu_poisson_fill_phi(cpgpu_poisson *P, double value);

/* Solve A phi = b from the current phi.  max_iter and tol are in/out with the
 * reference's meaning (on return: iterations used and the final relative
 * residual).  hist (capacity hist_cap, may be NULL) receives the residual
 * history; *hist_len its full length.  opts: CPGPU_OPT_* flags. */
int cpgpu_poisson_solve(cpgpu_poisson *P, int method, int *max_iter,
                        double *tol, int norm_mode, int opts, double *hist,
                        int hist_cap, int *hist_len);

/* Mixed solve controls (CPGPU_MIXED): inner relative tolerance and the
 * maximum number of refinement sweeps. */
int cpgpu_poisson_set_mixed(cpgpu_poisson *P, double inner_tol, int max_outer);

/* Direction-kernel variant of the Krylov iteration per precision: 0 LDS
 * tile, 1..4 column march with 4, 8, 16, 32 rows per wavefront, 5..6 the
 * aligned column march (64 aligned columns per wavefront) with 8, 16 rows,
 * 7..10 the streaming aligned march (loads of the next 4 rows in flight
 * while the current ones are computed) over chunks of 32, 64 rows with the
 * XCD-aware unit order and 64, 128 rows in plain order; 11..14 the
 * block-wide streaming march (a block's four wavefronts march together over
 * 256 16-byte vectors of a row, wave-boundary neighbours through LDS) over
 * chunks of 16, 32, 64 rows in XCD order and 32 rows in plain order, 15..16
 * the same over 32 rows (XCD, plain order) with non-temporal stores of p, q
 * and of the update's x, r (the streaming march of 9 when ny is not a
 * multiple of 2 (FP64) / 4 (FP32)); 5..16 are paired with the
 * two-cells-per-lane update kernel when ny is even. */
int cpgpu_poisson_set_variant(cpgpu_poisson *P, int v64, int v32);
/* Autotuner: times reps launches of every variant on this problem (FP64 and
 * FP32), keeps the fastest and reports it with its time per launch (us).
 * phi is preserved. */
int cpgpu_poisson_autotune(cpgpu_poisson *P, int reps, int *v64, int *v32,
                           float *us64, float *us32);

/* GPU time per launch (us) of the current direction and update kernels of
 * the Krylov iteration (precision 32 or 64): reps launches of each are
 * captured in a hipGraph and replayed between two events.  phi is kept. */
int cpgpu_poisson_kernel_times(cpgpu_poisson *P, int precision, int reps,
                               double *us_dir, double *us_upd);
/* Krylov iterations per launched batch (even, >= 2) and whether batches
 * are replayed hipGraphs (default 16, graphs on).  use_graphs: 0 direct
 * launches, 1 replayed Krylov batches, 2 also each flexible-CG iteration
 * (preconditioner cycle + Krylov kernels) replayed as one captured graph.
 * Not default: at 1024^2 it measured 256 us per iteration against 251 us
 * launched directly -- the device, not the launching host, bounds an
 * iteration. */
int cpgpu_poisson_set_batch(cpgpu_poisson *P, int iters, int use_graphs);

/* y = A x on interior host fields (tests of the operator). */
int cpgpu_poisson_apply(cpgpu_poisson *P, const double *x_field,
                        double *y_field);
/* True FP64 relative residual ||b - A phi|| / ||b|| of the current phi. */
int cpgpu_poisson_residual(cpgpu_poisson *P, double *rel);
/* Operation_<2>::Error1 / Error2 / Errori (s_operation.hpp:869-908) of the
 * current phi against a host ghost-padded exact field, over the normal
 * cells: out[0] = sum|e|v / sum v, out[1] = sqrt(sum e^2 v) / sum v (the
 * reference's normalisation), out[2] = max|e|, v = sx_i * sy_j. */
int cpgpu_poisson_error_norms(cpgpu_poisson *P, const double *exact_field,
                              double *out);

/* Device pointers of the problem's interior-only vectors (nx*ny doubles,
 * j contiguous): which = 0 phi, 1 b, 2 last residual.  For device-resident
 * callers. */
double *cpgpu_poisson_device_vec(cpgpu_poisson *P, int which);
/* Host copy of one of those vectors into the interior of a cell field. */
int cpgpu_poisson_get_vec(cpgpu_poisson *P, int which, double *field);

/* One-shot host-in/host-out solve, the drop-in body of a
 * structure::Solver_<2>::solve() (s_solver.hpp:261-291). */
int cpgpu_poisson2d_solve(cpgpu_ctx *ctx, const cpgpu_grid2d *g,
                          const double *src_field, double src_uniform,
                          double *phi_field, int method, int *max_iter,
                          double *tol, int norm_mode, double *hist,
                          int hist_cap, int *hist_len);

/* Scratch-free variant of cpgpu_poisson2d_solve, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  d_src (interior-only nx*ny device array, j contiguous, or
 * NULL for src_uniform) and d_phi (the same layout: initial guess in,
 * solution out) are the caller's device arrays; every other vector,
 * coefficient array, reduction partial and the solver state live in one
 * static device array per GPU of 10.8 MB, so grids are limited to
 * 262144 cells (512^2) and 4096 cells per side (larger: CPGPU_EINVAL).
 * method CPGPU_JPCG or CPGPU_CG; max_iter, tol and the codes as
 * cpgpu_poisson2d_solve (results bitwise equal to it).  No hipMalloc,
 * hipFree or other allocation on any call; it synchronizes the context's
 * stream.  All calls on one GPU share that array: calls that overlap on one
 * GPU (several contexts or host threads) race. */
int cpgpu_poisson2d_solve_static(cpgpu_ctx *ctx, const cpgpu_grid2d *g,
                                 const double *d_src, double src_uniform,
                                 double *d_phi, int method, int *max_iter,
                                 double *tol, int norm_mode);

/* ---- batched solves over many problems of uneven sizes ----
 * No counterpart in the reference, which solves one problem at a time
 * (solver_matrix.hpp:264-374, the algorithm batched here).  A batch is a
 * view over existing problems: their operators, right-hand sides and phi
 * (the initial guess, overwritten by the solution) are used in place. */
typedef struct cpgpu_batch cpgpu_batch;

/* Batch of n problems of one context (not all-Neumann: singular problems
 * are rejected).  The problems must outlive the batch. */
int cpgpu_batch_create(cpgpu_ctx *ctx, int n, cpgpu_poisson *const *problems,
                       cpgpu_batch **out);
int cpgpu_batch_destroy(cpgpu_batch *B);
/* Tiles (16 x 64 cells) per block; 1 gives each problem the block count of
 * its single-problem solve, so results are bitwise those of CPGPU_JPCG with
 * direction variant 0. */
int cpgpu_batch_set_tiles(cpgpu_batch *B, int tiles_per_block);
/* Graph-captured iteration batches on (default) or plain launches. */
int cpgpu_batch_set_graphs(cpgpu_batch *B, int use_graphs);
/* on != 0: the CPGPU_JPCG pass of cpgpu_batch_solve runs as ONE persistent
 * cooperative launch: set-up, all iterations and the per-problem stops
 * happen on the device, blocks synchronised by a grid barrier; results are
 * bitwise those of the launched batches.  CPGPU_MIXED and the batched passes
 * of CPGPU_FCG_UNET keep launched batches.  No reference counterpart (the
 * reference solves one problem at a time). */
int cpgpu_batch_set_persistent(cpgpu_batch *B, int on);
/* Solves every problem: method CPGPU_JPCG (FP64, the reference's J-PCG test
 * on the true relative residual every iteration), CPGPU_MIXED (FP32
 * iterations, FP64 residual replacement and test every 16 iterations), or
 * CPGPU_FCG_UNET: the members without a U-Net attached as preconditioner
 * (cpgpu_poisson_set_unet, CPGPU_UNET_PRECOND) by the batched CPGPU_MIXED
 * iteration, then the others by flexible CG one after another, smallest
 * first; a member whose flexible CG stalls or takes more than 150
 * iterations is finished, from its iterate, by a second batched pass.  Per problem (arrays of n, each may be NULL): iterations,
 * final relative residual (true norm), status code, and the time from the
 * start of the solve to the problem's stop in ms (device clock for batched
 * members, host clock for flexible-CG ones).  Returns CPGPU_OK when all
 * converged, else CPGPU_MAXITER. */
int cpgpu_batch_solve(cpgpu_batch *B, int method, int max_iter, double tol,
                      int *iters, double *resid, int *codes,
                      double *latency_ms);
/* Times `reps` graph batches of `method` for 1, 2, 4 and 8 tiles per block
 * (ms_per_iter[4], may be NULL), keeps the fastest; phi is restored. */
int cpgpu_batch_autotune(cpgpu_batch *B, int method, int reps, int *best_tpb,
                         float *ms_per_iter);
/* Random problem set (sizes, wall types, blob seeds) from one seeded
 * generator: log2 nx ~ U[lo, hi], log2 ny ~ U[max(lo, log2 nx - 2),
 * min(hi, log2 nx + 2)], both rounded (aspect <= 4); each wall Dirichlet or
 * Neumann with probability 1/2, redrawn while all four are Neumann;
 * bc_type[4 p + s] for segments x-, x+, y-, y+. */
int cpgpu_batch_draw(unsigned long long seed, int n, int lo, int hi, int *nx,
                     int *ny, int *bc_type, unsigned long long *blob_seed);

/* ---- NS_explicit_ projection step (device resident) ----
 * Replaces structure::NS_explicit_<2>::run_one_step + Time_::advance
 * (s_ns_explicit.hpp:118-155, s_time.hpp:61-64) for a uniform grid, uniform
 * rho and mu, zero source, the TVD advection schemes of
 * NS_::set_advection_scheme (s_ns.hpp:171-211) and a Krylov projection. */
typedef struct cpgpu_ns cpgpu_ns;

/* limiter ids (s_operation.hpp:501-546) */
#define CPGPU_LIM_VANLEER 0
#define CPGPU_LIM_SUPERBEE 1
#define CPGPU_LIM_WAHYD 2
#define CPGPU_LIM_QUICK 3
#define CPGPU_LIM_CUI 4
#define CPGPU_LIM_UPWIND2 5

/* field ids of cpgpu_ns_set_field / cpgpu_ns_get_field: the reference's
 * "u","v","uf","vf","us","vs","usf","vsf","p","phi" scalars and the
 * divergence of u*f */
#define CPGPU_F_U 0
#define CPGPU_F_V 1
#define CPGPU_F_UF 2
#define CPGPU_F_VF 3
#define CPGPU_F_US 4
#define CPGPU_F_VS 5
#define CPGPU_F_USF 6
#define CPGPU_F_VSF 7
#define CPGPU_F_P 8
#define CPGPU_F_PHI 9
#define CPGPU_F_DIV 10

typedef struct {
  cpgpu_grid2d grid;   /* axes (gl >= 2 for TVD); bc/beta fields unused */
  int bcu_type[4];     /* velocity boundary types per segment */
  const double *bcu_val; /* packed per ghost line, as cpgpu_grid2d.bc_val */
  int bcv_type[4];
  const double *bcv_val;
  double rho, mu;      /* NS_::set_uniform_rho / set_uniform_mu */
  double cfl;          /* NS_::set_CFL_number */
  double dnum, dcoef;  /* NS_::set_diffusion_number (no limit when >= 100) */
  int limiter;         /* CPGPU_LIM_* */
  int solver;          /* CPGPU_JPCG, CPGPU_CG or CPGPU_MIXED */
  int max_iter;        /* projection solver limits */
  double tol;
  int norm_mode;
  double dt;           /* initial time step (Time_ dt) */
} cpgpu_ns_desc;

int cpgpu_ns_create(cpgpu_ctx *ctx, const cpgpu_ns_desc *d, cpgpu_ns **out);
int cpgpu_ns_destroy(cpgpu_ns *S);
/* Host ghost-padded fields in/out (CPGPU_F_*). */
int cpgpu_ns_set_field(cpgpu_ns *S, int which, const double *field);
int cpgpu_ns_get_field(cpgpu_ns *S, int which, double *field);
/* Error1 / Error2 / Errori (s_operation.hpp:869-908) of the device field
 * `which` against a host ghost-padded exact field (the order and meaning
 * of cpgpu_poisson_error_norms). */
int cpgpu_ns_error_norms(cpgpu_ns *S, int which, const double *exact_field,
                         double *out);
/* The steady-state test of EventStopScalarError_ (s_event.hpp:520-566) on
 * the device: Error1/2/i of field `which` against its snapshot from the
 * previous call (out = -1, -1, -1 on the first call), then the snapshot
 * is replaced by the current field. */
int cpgpu_ns_field_change(cpgpu_ns *S, int which, double *out);
/* nsteps projection steps; per-step projection iterations and final
 * residuals into iters/resid (may be NULL).  Returns the last solver code. */
int cpgpu_ns_step(cpgpu_ns *S, int nsteps, int *iters, double *resid);
int cpgpu_ns_time(cpgpu_ns *S, double *t, double *dt);
/* NS_kim_ mode (s_ns_kim.hpp:130-287): when enable is set, cpgpu_ns_step
 * runs NS_kim_::run_one_step + Time_::advance: no time-step restriction,
 * the upwind advection of the NS_ default (VdotNabla_upwind1,
 * s_operation.hpp:400-422) in the Adams-Bashforth form of the reference, a
 * Helmholtz predictor per component (alpha 1, beta dt*0.5/rho) solved by
 * helm_solver (CPGPU_JACOBI, the reference's default, or CPGPU_CG /
 * CPGPU_JPCG / CPGPU_BICGSTAB) within helm_max_iter / helm_tol (Solver_
 * defaults 1000 and 1e-4, s_solver.hpp:73-74), then the projection with
 * p = phi/dt - (0.5/rho) mu lap(phi).  The step counter restarts at 0. */
int cpgpu_ns_set_kim(cpgpu_ns *S, int enable, int helm_solver,
                     int helm_max_iter, double helm_tol);
/* Iterations of the last step's u and v Helmholtz solves. */
int cpgpu_ns_kim_iters(cpgpu_ns *S, int *iters);
/* New velocity boundary values of component comp (0 u, 1 v), packed as
 * cpgpu_grid2d.bc_val, for time-dependent BoundaryCondition functions
 * (sampled by the caller where the reference calls get_val). */
int cpgpu_ns_set_bc_values(cpgpu_ns *S, int comp, const double *bc_val);
/* The projection problem (all-Neumann Poisson) of the step. */
cpgpu_poisson *cpgpu_ns_projection(cpgpu_ns *S);
/* Restore the time state (Time_ t and dt) of a saved step sequence. */
int cpgpu_ns_set_time(cpgpu_ns *S, double t, double dt);
/* Projection solver of the following steps: the reference's name switch
 * NS_::set_projection_solver (s_ns.hpp:120-130) over CPGPU_JPCG, CPGPU_CG,
 * CPGPU_MIXED and CPGPU_FCG_UNET (a U-Net attached to the projection
 * problem with cpgpu_poisson_set_unet). */
int cpgpu_ns_set_solver(cpgpu_ns *S, int solver, int max_iter, double tol,
                        int norm_mode);

/* ---- Advection_ scalar driver (device resident) ----
 * Replaces structure::Advection_<2> (lib/structure/s_advection.hpp:20-217):
 * explicit Euler steps of phi_t + v . grad phi = 0 with face velocities
 * fixed at initial().  run_one_step (:185-212) overwrites phi in place in
 * Grid_::begin_ijk order (i fastest), so a cell reads the new values of
 * the cells before it; the device reproduces that order exactly by
 * sweeping anti-diagonals of 16 x 16 tiles. */
typedef struct cpgpu_advection cpgpu_advection;

/* schemes of Advection_::set_advection_scheme (s_advection.hpp:140-182):
 * the TVD family by limiter id (CPGPU_LIM_VANLEER .. CPGPU_LIM_UPWIND2:
 * "VanLeer", "superbee", "WAHYD", "QUICK", "CUI", "upwind2") and */
#define CPGPU_ADV_UPWIND1 10 /* "upwind1", VdotNabla_upwind1, the default */
#define CPGPU_ADV_CENTER 11  /* "center", VdotNabla_center */
#define CPGPU_ADV_CENTER4 12 /* "center4", VdotNabla_center4 (the reference
                                leaves the velocity out of this term) */

typedef struct {
  cpgpu_grid2d grid;       /* axes (gl >= 2); bc_type / bc_val: phi's boundary
                              (add_bc of "phi"); beta and source unused */
  int bcu_type[4];         /* centre velocity boundaries ("u", "v") */
  const double *bcu_val;   /* packed per ghost line, as cpgpu_grid2d.bc_val */
  int bcv_type[4];
  const double *bcv_val;
  int scheme;              /* CPGPU_ADV_* or a CPGPU_LIM_* id */
  double dt;               /* Time_ dt (set_time) */
} cpgpu_advection_desc;

int cpgpu_advection_create(cpgpu_ctx *ctx, const cpgpu_advection_desc *d,
                           cpgpu_advection **out);
int cpgpu_advection_destroy(cpgpu_advection *A);
/* Host ghost-padded fields in/out: CPGPU_F_U, CPGPU_F_V (centre velocities,
 * initial_CS), CPGPU_F_UF, CPGPU_F_VF (face velocities), CPGPU_F_PHI. */
int cpgpu_advection_set_field(cpgpu_advection *A, int which, const double *field);
int cpgpu_advection_get_field(cpgpu_advection *A, int which, double *field);
int cpgpu_advection_set_scheme(cpgpu_advection *A, int scheme);
/* Advection_::initial (s_advection.hpp:103-135): the centre velocities'
 * ghosts from their boundaries, then the faces by InterpolateC2F. */
int cpgpu_advection_initial(cpgpu_advection *A);
/* nsteps of run_one_step: apply_bc of u, v and phi, then the in-place sweep
 * phi <- phi - dt * VdotNabla(v, phi). */
int cpgpu_advection_step(cpgpu_advection *A, int nsteps);
/* time t = steps * dt and the step count */
int cpgpu_advection_time(cpgpu_advection *A, double *t, long long *steps);

/* ---- learned preconditioner: a BF16 U-Net on MFMA ----
 * No counterpart in the reference: its Poisson solvers are the CSR Krylov
 * methods of algebra/solver_matrix.hpp (Jacobi, CG, BiCGSTAB, IC-CG ...).
 * This is the extension those solvers are selected beside, by name, in
 * Poisson_::_init_solver (s_poisson.hpp:357-404).
 *
 * Network: 4 levels of 16-32-64-128 channels, two 3x3 convolutions + ReLU
 * per level, 4x4 average pooling (stride 4), 4x4 stride-4 transposed
 * convolutions, skip connections by concatenation, 1x1 linear output.
 * Activations are NHWC BF16, accumulation FP32, master parameters FP32.
 * A field is H = nx rows of W = ny contiguous values (the layout of
 * cpgpu_poisson_device_vec); H and W must be multiples of 64. */
typedef struct cpgpu_unet cpgpu_unet;

/* Kaiming-normal weights (fan-in; gain sqrt(2), 1 for the output layer),
 * zero biases, from xoshiro256** seeded through SplitMix64 by seed. */
int cpgpu_unet_create(cpgpu_ctx *ctx, unsigned long long seed,
                      cpgpu_unet **out);
int cpgpu_unet_destroy(cpgpu_unet *U);
/* Launch shape of the 16-input-channel 3x3 convolutions: 4 or 8 pixel rows
 * per wavefront (blocks of 16 x 16 or 32 x 16 pixels), and with persist > 0
 * a persistent grid of `persist` blocks per compute unit that walks 16 x 16
 * tiles, staging the next tile while it computes one (rows 4).  Default
 * rows 4, persist 4. */
int cpgpu_unet_set_conv_shape(cpgpu_unet *U, int rows, int persist);
/* Level-0 fusion of an applied network (a preconditioner or warm start, no
 * backward pass): level0 = 1 (default) runs the finest level in two kernels
 * (first layer + convolution + pooling; up-convolution + both decoder
 * convolutions + output layer) whose inner activations stay in LDS, with
 * the same per-value arithmetic as the layer-per-kernel path (level0 = 0).
 * Needs the persistent shape (persist > 0).  No reference counterpart. */
int cpgpu_unet_set_fusion(cpgpu_unet *U, int level0);
long long cpgpu_unet_num_params(const cpgpu_unet *U);
/* Flat FP32 parameter vector (grads != 0: the last computed gradients);
 * per layer: weights [tap][cout][cin], then biases [cout]. */
int cpgpu_unet_get_params(cpgpu_unet *U, float *host, int grads);
int cpgpu_unet_set_params(cpgpu_unet *U, const float *host);
/* The network's second input channel: for each cell the sum over the walls
 * it touches of +1 (Dirichlet) or -1 (Neumann).  bc_type[4] are the types of
 * the x-, x+, y-, y+ segments (as cpgpu_grid2d.bc_type), NULL for none (all
 * zero).  Used by cpgpu_unet_forward / _backward; solves and training take
 * the wall types of their problem. */
int cpgpu_unet_set_walls(cpgpu_unet *U, const int *bc_type);
/* y = net(x) on host arrays of H*W values (no input scaling). */
int cpgpu_unet_forward(cpgpu_unet *U, int H, int W, const float *x, float *y);
/* Gradients of sum(dy * y) after cpgpu_unet_forward into grads (host,
 * num_params values). */
int cpgpu_unet_backward(cpgpu_unet *U, const float *dy, float *grads);
/* Internal tensor of the last forward (grad = 0; BF16 widened to FP32) or
 * backward (grad != 0; FP32 gradient) as H x W x C floats into host (NULL:
 * only dims[3] = {H, W, C}).  id: 0 input, 1 a0, 2 C0 (up-conv | skip), 3
 * d0a, 4 d0, 5 p0, 6 a1, 7 C1, 8 d1a, 9 d1, 10 p1, 11 a2, 12 C2, 13 d2a,
 * 14 d2, 15 p2, 16 a3, 17 b3, 18 output.  For layer-by-layer tests. */
int cpgpu_unet_tensor(cpgpu_unet *U, int id, int grad, float *host, int *dims);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on the last gradients. */
int cpgpu_unet_adam(cpgpu_unet *U, float lr);

/* Training set for the problem P: slot k holds an error e (host, interior
 * H*W doubles) and its residual r = A e, computed on the device.  With
 * smooth > 0 the stored error is e after that many weighted-Jacobi sweeps
 * (e <- e - omega D^-1 A e), the error the network meets inside the
 * preconditioner cycle of cpgpu_poisson_set_fcg.  Slots grow on demand;
 * all samples share P's grid. */
int cpgpu_unet_set_sample(cpgpu_unet *U, cpgpu_poisson *P, int slot,
                          const double *e, int smooth, double omega);
int cpgpu_unet_num_samples(const cpgpu_unet *U);
/* Drops every sample (training then moves on to another problem: the
 * samples and the problem passed to cpgpu_unet_train must share one
 * operator); the parameters and the Adam state are kept. */
int cpgpu_unet_clear_samples(cpgpu_unet *U);
/* steps Adam updates, each on the mean energy-norm loss of batch samples
 * drawn with a seeded generator:  L = (s y - e)' A (s y - e) / e' A e,
 * y = net(r / s), s = rms(r).  losses (may be NULL) receives each step's
 * mean loss. */
int cpgpu_unet_train(cpgpu_unet *U, cpgpu_poisson *P, int steps, int batch,
                     float lr, unsigned long long seed, float *losses);
/* Loss of every sample (num_samples values) without updating. */
int cpgpu_unet_eval(cpgpu_unet *U, cpgpu_poisson *P, float *losses);

/* Attach a network to a problem (NULL detaches).  flags: CPGPU_UNET_WARM
 * starts every solve from phi = s net(b / s); CPGPU_UNET_PRECOND makes
 * CPGPU_FCG_UNET use it as the flexible-CG preconditioner. */
#define CPGPU_UNET_WARM 1
#define CPGPU_UNET_PRECOND 2
int cpgpu_poisson_set_unet(cpgpu_poisson *P, cpgpu_unet *U, int flags);
/* Preconditioner cycle of CPGPU_FCG_UNET: z = M(r) is `pre` weighted-
 * Jacobi sweeps (weight omega) from zero, plus the network applied to the
 * remaining residual, followed by `post` sweeps -- a two-level cycle with
 * the network as the coarse correction (default 2, 2, 0.8). */
int cpgpu_poisson_set_fcg(cpgpu_poisson *P, int pre, int post, double omega);

/* ---- general CSR matrices (the algebra solvers' MatrixSCR_) ----
 * Replaces the free functions of lib/algebra/solver_matrix.hpp on a
 * MatrixSCR_<double> (matrix_SparCompRow.hpp): int F(const MatrixSCR_& A,
 * ArrayListV& x, const ArrayListV& b, int& max_iter, Float& tol,
 * std::list<Float>& lresid) for F = CG (:264-316), IC_CG's loop with
 * DiaPre(A,1) (:318-374), Dia_BiCGSTAB (:599-672) and Solver_Jacobi_
 * (:820-910), and MatrixSCR_::operator* (:236-249).  The matrix is copied to
 * the device once; row_ptr has n + 1 entries starting at 0, col_ind and val
 * row_ptr[n] entries (0-based columns). */
typedef struct cpgpu_csr cpgpu_csr;
int cpgpu_csr_create(cpgpu_ctx *ctx, int n, const int *row_ptr, const int *col_ind,
                     const double *val, cpgpu_csr **out);
int cpgpu_csr_destroy(cpgpu_csr *A);
/* y = A x (host vectors of n doubles), each row summed in CSR order. */
int cpgpu_csr_mult(cpgpu_csr *A, const double *x, double *y);
/* method: CPGPU_CG, CPGPU_JPCG, CPGPU_BICGSTAB or CPGPU_JACOBI (other
 * methods: CPGPU_EUNSUPPORTED).  x (host, n doubles) is the start on entry
 * and the result on exit; max_iter and tol in/out as the reference; hist
 * receives the residual after each iteration (hist_len = their count, hist
 * may be NULL).  Returns 0 converged, 1 max_iter, 2 / 3 BiCGSTAB breakdowns,
 * CPGPU_EZERODIAG when DiaPre meets a zero or missing diagonal entry. */
int cpgpu_csr_solve(cpgpu_csr *A, int method, double *x, const double *b, int *max_iter,
                    double *tol, int norm_mode, double *hist, int hist_cap, int *hist_len);

/* ---- kernel timing (HIP events on the context stream) ----
 * When enabled, every launch of the named kernel family is bracketed by
 * events; cpgpu_timer_read returns launches and summed milliseconds.
 * family: 0 = stencil apply (p update + A p + p.q), 1 = vector update
 * (x, r update + dots), 2 = the U-Net's 16 -> 16 3x3 convolutions on the
 * finest level that store their output, 3 = the same convolution when its
 * epilogue applies the 1 x 1 output layer instead (a network only applied),
 * 4 = the NS_explicit_ u* kernel (TVD advection + diffusion of both
 * components), 5 = C2F + Div of u*, 6 = grad phi + corrections + p,
 * 7 = the fused u* + C2F + Div kernel, 8 = the fused level-0 encoder of an
 * applied U-Net, 9 = its fused level-0 decoder.
 * Timed launches carry their own start/stop timestamps (no extra packets
 * between kernels). */
int cpgpu_timer_enable(cpgpu_ctx *ctx, int family, int on);
int cpgpu_timer_read(cpgpu_ctx *ctx, int family, int *launches, double *ms);

#ifdef __cplusplus
}
#endif
#endif
