// C++ host side of the structured-grid path, for the reference's structure
// layer: what a structure::Solver_GPU_<2> (a Solver_ subclass selected by
// name in Poisson_::_init_solver, s_poisson.hpp:357-404) does in solve(),
// and what an NS_explicit_<2> subclass does in run_one_step
// (s_ns_explicit.hpp:118-155), written against the grid's per-axis arrays
// (Grid_::_c, Grid_::_cs) and Scalar_ buffers (s_data.hpp:134-137:
// (nx+2gl) x (ny+2gl) doubles, j contiguous, &scalar.VAL(0, 0) is the
// first), so it compiles without the reference headers.  C++11, header
// only; links libcpgpu.so.
#pragma once

#include <string>
#include <vector>

#include "../../include/cpgpu.h"

namespace cpgpu_bind {

// Centre and size arrays of one axis, ghosts included.
struct Axis {
  std::vector<double> c, s;
  int n = 0, gl = 0;
};

// A uniform axis as structure::Grid_ builds it (s_grid.hpp:372-395).
inline Axis uniform_axis(int n, int gl, double lo, double hi) {
  Axis a;
  a.n = n;
  a.gl = gl;
  a.c.assign(n + 2 * gl, 0.0);
  a.s.assign(n + 2 * gl, 0.0);
  cpgpu_grid_uniform_axis(n, gl, lo, hi, a.c.data(), a.s.data());
  return a;
}

// The boundary conditions of one variable as the kernels read them: per
// segment (x-, x+, y-, y+; Ghost_::which_boudary_seg, s_ghost.hpp:218-232)
// its type, BC1 or BC2 (s_boundary.hpp:16-18), and its value on every
// ghost-inclusive line, sampled by the caller where
// BoundaryCondition::get_val is evaluated (s_ghost.hpp:161-165).
struct Walls {
  int type[4] = {CPGPU_BC_NEUMANN, CPGPU_BC_NEUMANN, CPGPU_BC_NEUMANN, CPGPU_BC_NEUMANN};
  std::vector<double> val;  // [x- (ny+2gl) | x+ (ny+2gl) | y- (nx+2gl) | y+ (nx+2gl)]
};

// One value per segment (a BoundaryCondition whose function is constant).
inline Walls constant_walls(const Axis& x, const Axis& y, const int type[4],
                            const double value[4]) {
  Walls w;
  const size_t NX = x.c.size(), NY = y.c.size();
  for (int s = 0; s < 4; ++s) w.type[s] = type[s];
  w.val.insert(w.val.end(), NY, value[0]);
  w.val.insert(w.val.end(), NY, value[1]);
  w.val.insert(w.val.end(), NX, value[2]);
  w.val.insert(w.val.end(), NX, value[3]);
  return w;
}

inline cpgpu_grid2d descriptor(const Axis& x, const Axis& y, double beta, const Walls& w) {
  cpgpu_grid2d d;
  d.nx = x.n;
  d.ny = y.n;
  d.gl = x.gl;
  d.cx = x.c.data();
  d.sx = x.s.data();
  d.cy = y.c.data();
  d.sy = y.s.data();
  d.beta = beta;
  for (int s = 0; s < 4; ++s) d.bc_type[s] = w.type[s];
  d.bc_val = w.val.empty() ? nullptr : w.val.data();
  return d;
}

// One context per host thread: the reference is single-threaded and the
// solver objects it creates per step are short-lived.
inline cpgpu_ctx* thread_context(int device = 0) {
  static thread_local cpgpu_ctx* ctx = nullptr;
  if (!ctx && cpgpu_create(device, &ctx) != CPGPU_OK) ctx = nullptr;
  return ctx;
}

// Solver_GPU_::solve(): the Poisson_ rows of the grid, beta, the source
// (a Scalar_ buffer or the uniform value) and the walls of "phi", solved
// from the warm start in phi; max_iter and tol in/out as the reference's
// solvers take them (solver_matrix.hpp:264-374).  Returns the solver code
// (negative: CPGPU_E*, message in error()).
inline int solve_poisson(const Axis& x, const Axis& y, double beta, const Walls& w,
                         const double* src_scalar, double src_uniform, double* phi_scalar,
                         int method, int& max_iter, double& tol,
                         int norm_mode = CPGPU_NORM_REF, std::string* error = nullptr) {
  cpgpu_ctx* ctx = thread_context();
  if (!ctx) return CPGPU_EHIP;
  const cpgpu_grid2d d = descriptor(x, y, beta, w);
  const int code = cpgpu_poisson2d_solve(ctx, &d, src_scalar, src_uniform, phi_scalar, method,
                                         &max_iter, &tol, norm_mode, nullptr, 0, nullptr);
  if (code < 0 && error) *error = cpgpu_last_error(ctx);
  return code;
}

// The device-resident NS_explicit_ step: the fields stay in HBM between
// steps; Scalar_ buffers cross to the host only when the caller asks.
class ExplicitStep {
 public:
  // u and v walls, rho and mu (NS_::set_uniform_rho / set_uniform_mu), the
  // CFL and diffusion numbers (s_ns.hpp:132-144), the TVD scheme
  // (CPGPU_LIM_*), the projection solver and its limits, the first dt
  ExplicitStep(const Axis& x, const Axis& y, const Walls& wu, const Walls& wv, double rho,
               double mu, double cfl, double dnum, double dcoef, int limiter, int solver,
               int max_iter, double tol, int norm_mode, double dt)
      : x_(x), y_(y), wu_(wu), wv_(wv) {
    cpgpu_ns_desc d;
    d.grid = descriptor(x_, y_, 1.0, Walls());
    for (int s = 0; s < 4; ++s) {
      d.bcu_type[s] = wu_.type[s];
      d.bcv_type[s] = wv_.type[s];
    }
    d.bcu_val = wu_.val.data();
    d.bcv_val = wv_.val.data();
    d.rho = rho;
    d.mu = mu;
    d.cfl = cfl;
    d.dnum = dnum;
    d.dcoef = dcoef;
    d.limiter = limiter;
    d.solver = solver;
    d.max_iter = max_iter;
    d.tol = tol;
    d.norm_mode = norm_mode;
    d.dt = dt;
    cpgpu_ctx* ctx = thread_context();
    rc_ = ctx ? cpgpu_ns_create(ctx, &d, &S_) : CPGPU_EHIP;
  }
  ~ExplicitStep() {
    if (S_) cpgpu_ns_destroy(S_);
  }
  ExplicitStep(const ExplicitStep&) = delete;
  ExplicitStep& operator=(const ExplicitStep&) = delete;
  int status() const { return rc_; }
  // run_one_step + Time_::advance, n times; per-step projection iterations
  int step(int n, std::vector<int>* iters = nullptr) {
    std::vector<int> it(n > 0 ? n : 1);
    std::vector<double> res(n > 0 ? n : 1);
    const int code = cpgpu_ns_step(S_, n, it.data(), res.data());
    if (iters) iters->assign(it.begin(), it.begin() + n);
    return code;
  }
  // CPGPU_F_* field into / out of a Scalar_ buffer
  int get(int which, double* scalar) { return cpgpu_ns_get_field(S_, which, scalar); }
  int set(int which, const double* scalar) { return cpgpu_ns_set_field(S_, which, scalar); }
  int time(double& t, double& dt) { return cpgpu_ns_time(S_, &t, &dt); }

 private:
  Axis x_, y_;
  Walls wu_, wv_;
  cpgpu_ns* S_ = nullptr;
  int rc_ = CPGPU_OK;
};

}  // namespace cpgpu_bind
