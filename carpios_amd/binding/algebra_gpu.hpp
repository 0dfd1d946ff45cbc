// C++ host side of the general CSR path: drop-in counterparts of the free
// solver functions of lib/algebra/solver_matrix.hpp and of Solver_Jacobi_
// (int F(const MatrixSCR_& A, ArrayListV& x, const ArrayListV& b,
// int& max_iter, Float& tol, std::list<Float>& lresid)), running on the GPU
// through libcpgpu.so's cpgpu_csr_* entry points.  Include after the
// reference's "algebra/solver_matrix.hpp" (it uses MatrixSCR_, ArrayListV,
// Solver_ and Float); C++11, header only.
//   GPU_CG            carpio::CG            (solver_matrix.hpp:264-316)
//   GPU_JPCG          IC_CG's loop with DiaPre(A,1) (:318-374,
//                     preconditioner.hpp:182-262)
//   GPU_Dia_BiCGSTAB  carpio::Dia_BiCGSTAB  (:599-672)
//   Solver_GPU_Jacobi_ carpio::Solver_Jacobi_ (:820-910)
// Same status codes (0 converged, 1 max_iter, 2 / 3 BiCGSTAB breakdowns),
// the same in/out max_iter and tol, lresid = the residual after each
// iteration.  A zero or missing diagonal returns CPGPU_EZERODIAG where
// DiaPre calls exit(0).  The matrix is copied to the device on each call.
#pragma once

#include <algorithm>
#include <list>
#include <type_traits>
#include <vector>

#include "../../include/cpgpu.h"

namespace carpio {
namespace gpu_detail {

inline cpgpu_ctx* context() {
  static thread_local cpgpu_ctx* ctx = nullptr;
  if (!ctx && cpgpu_create(0, &ctx) != CPGPU_OK) ctx = nullptr;
  return ctx;
}

template <class VALUE>
int csr_solve(int method, const MatrixSCR_<VALUE>& A, ArrayListV<VALUE>& x,
              const ArrayListV<VALUE>& b, int& max_iter, Float& tol, std::list<Float>& lresid) {
  static_assert(std::is_same<VALUE, double>::value, "the device path is FP64");
  cpgpu_ctx* ctx = context();
  if (!ctx) return CPGPU_EHIP;
  const int n = (int)A.size1(), nz = (int)A.NumNonzeros();
  std::vector<int> rp(n + 1), ci(nz);
  for (int i = 0; i <= n; ++i) rp[i] = (int)A.row_ptr(i);
  for (int k = 0; k < nz; ++k) ci[k] = (int)A.col_ind(k);
  std::vector<double> val(nz), hx(n), hb(n);
  for (int k = 0; k < nz; ++k) val[k] = A.val(k);
  for (int i = 0; i < n; ++i) {
    hx[i] = x[i];
    hb[i] = b[i];
  }
  cpgpu_csr* M = nullptr;
  int code = cpgpu_csr_create(ctx, n, rp.data(), ci.data(), val.data(), &M);
  if (code != CPGPU_OK) return code;
  std::vector<double> hist((size_t)std::max(max_iter, 0) + 2);
  int hl = 0;
  code = cpgpu_csr_solve(M, method, hx.data(), hb.data(), &max_iter, &tol, CPGPU_NORM_REF,
                         hist.data(), (int)hist.size(), &hl);
  cpgpu_csr_destroy(M);
  if (code < 0) return code;
  for (int i = 0; i < n; ++i) x[i] = hx[i];
  lresid.assign(hist.begin(), hist.begin() + std::min<int>(hl, (int)hist.size()));
  return code;
}

}  // namespace gpu_detail

template <class VALUE>
int GPU_CG(const MatrixSCR_<VALUE>& A, ArrayListV<VALUE>& x, const ArrayListV<VALUE>& b,
           int& max_iter, Float& tol, std::list<Float>& lresid) {
  return gpu_detail::csr_solve(CPGPU_CG, A, x, b, max_iter, tol, lresid);
}

template <class VALUE>
int GPU_JPCG(const MatrixSCR_<VALUE>& A, ArrayListV<VALUE>& x, const ArrayListV<VALUE>& b,
             int& max_iter, Float& tol, std::list<Float>& lresid) {
  return gpu_detail::csr_solve(CPGPU_JPCG, A, x, b, max_iter, tol, lresid);
}

template <class VALUE>
int GPU_Dia_BiCGSTAB(const MatrixSCR_<VALUE>& A, ArrayListV<VALUE>& x,
                     const ArrayListV<VALUE>& b, int& max_iter, Float& tol,
                     std::list<Float>& lresid) {
  return gpu_detail::csr_solve(CPGPU_BICGSTAB, A, x, b, max_iter, tol, lresid);
}

// Solver_Jacobi_ with the same members: returns the previous iterate on
// convergence (solver_matrix.hpp:898-903), as the reference does.
template <class VALUE>
class Solver_GPU_Jacobi_ : public Solver_<VALUE> {
 public:
  typedef MatrixSCR_<VALUE> MatSCR;
  typedef ArrayListV<VALUE> Arr;
  Solver_GPU_Jacobi_(int max_iter = 100, VALUE tol = 1e-3) : Solver_<VALUE>(max_iter, tol) {}
  int solve(const MatSCR& A, Arr& x, const Arr& b) {
    int mi = this->_max_iter;
    Float tol = this->_tol;
    std::list<Float> lr;
    const int code = gpu_detail::csr_solve(CPGPU_JACOBI, A, x, b, mi, tol, lr);
    if (code < 0) return code;
    this->_num_iter = mi;
    this->_resid = tol;
    this->_lresid.insert(this->_lresid.end(), lr.begin(), lr.end());
    return code;
  }
};

}  // namespace carpio
