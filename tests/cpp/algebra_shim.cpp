// The reference's own lib/algebra solvers against their GPU counterparts
// (carpios_amd/binding/algebra_gpu.hpp) called with the reference's own
// types: a MatrixSCR_ and ArrayListV built from a CSR file, each method run
// by both from x = 0.  Built by oracle/Makefile from the reference tree
// (-I<reference>/lib) into oracle/_ref/algebra_shim; nothing of the
// reference is copied.  Prints one line per method:
//   <method> ref <code> <max_iter> <tol> <len> gpu <code> <max_iter> <tol> <len> dx <rel L2>
// Input file: n nnz, row_ptr (n+1), col (nnz), val (nnz), b (n), whitespace
// separated.
#include <cmath>
#include <cstdio>
#include <list>
#include <vector>

#include "algebra/solver_matrix.hpp"
#include "../../carpios_amd/binding/algebra_gpu.hpp"

using carpio::ArrayListV;
using carpio::Float;
using carpio::MatrixSCR_;
using carpio::St;

typedef MatrixSCR_<double> Mat;
typedef ArrayListV<double> Arr;

static double rel_l2(const Arr& a, const Arr& b) {
  double d = 0.0, n = 0.0;
  for (St i = 0; i < a.size(); ++i) {
    d += (a[i] - b[i]) * (a[i] - b[i]);
    n += b[i] * b[i];
  }
  return n > 0 ? std::sqrt(d / n) : std::sqrt(d);
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: algebra_shim <csr file> <max_iter> <tol>\n");
    return 1;
  }
  FILE* f = std::fopen(argv[1], "r");
  if (!f) return 2;
  int n = 0, nz = 0;
  if (std::fscanf(f, "%d %d", &n, &nz) != 2 || n < 1 || nz < 1) return 3;
  ArrayListV<St> rp(n + 1), ci(nz);
  Arr val(nz), b(n);
  for (int i = 0; i <= n; ++i) {
    unsigned v;
    if (std::fscanf(f, "%u", &v) != 1) return 3;
    rp[i] = v;
  }
  for (int k = 0; k < nz; ++k) {
    unsigned v;
    if (std::fscanf(f, "%u", &v) != 1) return 3;
    ci[k] = v;
  }
  for (int k = 0; k < nz; ++k)
    if (std::fscanf(f, "%lf", &val[k]) != 1) return 3;
  for (int i = 0; i < n; ++i)
    if (std::fscanf(f, "%lf", &b[i]) != 1) return 3;
  std::fclose(f);
  const St sn = (St)n, snz = (St)nz;
  const Mat A(sn, sn, snz, val, rp, ci);
  const int max_iter = std::atoi(argv[2]);
  const double tol = std::atof(argv[3]);
  for (int m = 0; m < 3; ++m) {
    const char* name = m == 0 ? "CG" : (m == 1 ? "Dia_BiCGSTAB" : "Solver_Jacobi_");
    Arr xr(n, 0.0), xg(n, 0.0);
    int mr = max_iter, mg = max_iter;
    Float tr = tol, tg = tol;
    std::list<Float> lr, lg;
    int cr = 0, cg = 0;
    if (m == 0) {
      cr = carpio::CG(A, xr, b, mr, tr, lr);
      cg = carpio::GPU_CG(A, xg, b, mg, tg, lg);
    } else if (m == 1) {
      cr = carpio::Dia_BiCGSTAB(A, xr, b, mr, tr, lr);
      cg = carpio::GPU_Dia_BiCGSTAB(A, xg, b, mg, tg, lg);
    } else {
      carpio::Solver_Jacobi_<double> sr(max_iter, tol);
      carpio::Solver_GPU_Jacobi_<double> sg(max_iter, tol);
      cr = sr.solve(A, xr, b);
      cg = sg.solve(A, xg, b);
      mr = sr.num_iter();
      mg = sg.num_iter();
      tr = sr.residual();
      tg = sg.residual();
      lr.assign((size_t)sr.get_residual_array().size(), 0.0);
      lg.assign((size_t)sg.get_residual_array().size(), 0.0);
    }
    std::printf("%s ref %d %d %.17g %zu gpu %d %d %.17g %zu dx %.3e\n", name, cr, mr, tr,
                lr.size(), cg, mg, tg, lg.size(), rel_l2(xg, xr));
  }
  return 0;
}
