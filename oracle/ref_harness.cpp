// Test-only harness around the reference's own lib/algebra sparse solvers.
//
// Built by oracle/Makefile straight from the reference tree (-I<ref>/lib) into
// oracle/_ref/libcarpios_ref.so; nothing from the reference is copied here.
// Used only by tests/ (and tests/golden/make_golden.py) to pin the C oracle
// against the reference's arithmetic: SpMV (matrix_SparCompRow.hpp:236-249),
// dot/nrm2 (BLAS_Level1.hpp:500-577), DiaPre (preconditioner.hpp:182-262) and
// the Krylov / stationary solvers of solver_matrix.hpp.
#include <list>

#include "algebra/solver_matrix.hpp"

using carpio::ArrayListV;
using carpio::MatrixSCR_;
using carpio::St;
using carpio::Float;

namespace {

typedef MatrixSCR_<double> Mat;
typedef ArrayListV<double> Arr;

Mat make_csr(int n, const int* rp, const int* col, const double* val) {
  int nnz = rp[n];
  Arr v(nnz);
  ArrayListV<St> r(n + 1), c(nnz);
  for (int i = 0; i <= n; ++i) r[i] = St(rp[i]);
  for (int k = 0; k < nnz; ++k) {
    c[k] = St(col[k]);
    v[k] = val[k];
  }
  return Mat(St(n), St(n), St(nnz), v, r, c);
}

Arr make_vec(int n, const double* p) {
  Arr a(n);
  for (int i = 0; i < n; ++i) a[i] = p[i];
  return a;
}

// Jacobi-preconditioned CG composed from reference parts: the loop of
// IC_CG (solver_matrix.hpp:318-374) with its ICPre replaced by DiaPre(A,1)
// (preconditioner.hpp:182-262).  No reference function has this exact body.
int composed_jpcg(const Mat& A, Arr& x, const Arr& b, int& max_iter,
                  Float& tol, std::list<Float>& lresid) {
  Float resid;
  Arr p, z, q;
  Float alpha, beta, rho, rho_1 = 0;
  carpio::DiaPre<double> preA(A, 1);
  Float normb = carpio::nrm2(b);
  Arr r = b - A * x;
  if (normb == 0.0) normb = 1;
  if ((resid = carpio::nrm2(r) / normb) <= tol) {
    tol = resid;
    max_iter = 0;
    return 0;
  }
  for (int i = 1; i <= max_iter; i++) {
    z = preA.solve(r);
    rho = carpio::dot(r, z);
    if (i == 1) {
      p = z;
    } else {
      beta = rho / rho_1;
      p = z + beta * p;
    }
    q = A * p;
    alpha = rho / carpio::dot(p, q);
    x += alpha * p;
    r -= alpha * q;
    resid = carpio::nrm2(r) / normb;
    lresid.push_back(resid);
    if (resid <= tol) {
      tol = resid;
      max_iter = i;
      return 0;
    }
    rho_1 = rho;
  }
  tol = resid;
  return 1;
}

}  // namespace

extern "C" {

double ref_dot(int n, const double* x, const double* y) {
  Arr a = make_vec(n, x), b = make_vec(n, y);
  return carpio::dot(a, b);
}

double ref_nrm2(int n, const double* x) {
  Arr a = make_vec(n, x);
  return carpio::nrm2(a);
}

void ref_spmv(int n, const int* rp, const int* col, const double* val,
              const double* x, double* y) {
  Mat A = make_csr(n, rp, col, val);
  Arr xv = make_vec(n, x);
  Arr yv = A * xv;
  for (int i = 0; i < n; ++i) y[i] = yv[i];
}

// method: 0 CG, 1 Jacobi-PCG (composed), 2 Dia_BiCGSTAB, 3 Jacobi (free
// function), 4 Solver_Jacobi_ class, 5 IC_CG, 6 IC_CGS, 7 IC_BiCGSTAB
int ref_solve(int method, int n, const int* rp, const int* col,
              const double* val, double* x, const double* b, int* max_iter,
              double* tol, double* hist, int hist_cap, int* hist_len) {
  Mat A = make_csr(n, rp, col, val);
  Arr xv = make_vec(n, x), bv = make_vec(n, b);
  std::list<Float> lr;
  int code = -1;
  int mi = *max_iter;
  Float t = *tol;
  switch (method) {
    case 0: code = carpio::CG(A, xv, bv, mi, t, lr); break;
    case 1: code = composed_jpcg(A, xv, bv, mi, t, lr); break;
    case 2: code = carpio::Dia_BiCGSTAB(A, xv, bv, mi, t, lr); break;
    case 3: code = carpio::Jacobi(A, xv, bv, mi, t, lr); break;
    case 4: {
      carpio::Solver_Jacobi_<double> s(mi, t);
      code = s.solve(A, xv, bv);
      mi = s.num_iter();
      t = s.residual();
      Arr ra = s.get_residual_array();
      for (St i = 0; i < ra.size(); ++i) lr.push_back(ra[i]);
      break;
    }
    case 5: code = carpio::IC_CG(A, xv, bv, mi, t, lr); break;
    case 6: code = carpio::IC_CGS(A, xv, bv, mi, t, lr); break;
    case 7: code = carpio::IC_BiCGSTAB(A, xv, bv, mi, t, lr); break;
    default: return -1;
  }
  for (int i = 0; i < n; ++i) x[i] = xv[i];
  *max_iter = mi;
  *tol = t;
  int k = 0;
  for (Float v : lr) {
    if (hist && k < hist_cap) hist[k] = v;
    ++k;
  }
  *hist_len = k;
  return code;
}

}  // extern "C"
