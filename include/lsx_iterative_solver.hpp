// lsx_iterative_solver.hpp — header-only C++ mirror of the reference solver interface over the
// lsx C ABI.  LsxIterativeSolver keeps the reference's public contract
// (src/Base/Solvers/iterative_solver_interface.h:31-73): set(...) stores the matrix and builds the
// preconditioner, solve() updates the solution in place and returns the iteration count,
// rebuildPreconditioner() rebuilds after a geometry change, and the members eps, norm_r and it
// have the reference meaning (absolute 2-norm tolerance, final residual norm, iterations).
// Failures throw lsx_exception with the library's message (the reference aborted instead).
#ifndef LSX_ITERATIVE_SOLVER_HPP_
#define LSX_ITERATIVE_SOLVER_HPP_

#include <stdexcept>
#include <string>

#include "lsx.h"

namespace lsx_cpp {

struct lsx_exception : std::runtime_error {
  int status;
  lsx_exception(int s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void check(int status) {
  if (status != LSX_OK) throw lsx_exception(status, lsx_last_error());
}

// Matrix in the reference SparseMatrix layout (sparse_matrix.h:220-228).
struct CsrView {
  int N, M, K;              // m_N, m_M, m_K
  const int* Number;        // row pointers, N*M*K + 1
  const int* Portrait;      // column indices
  const double* Cell;       // values
  long elemCount;           // nnz
  int mem;                  // LSX_MEM_HOST or LSX_MEM_DEVICE
};

class LsxIterativeSolver {
 public:
  explicit LsxIterativeSolver(const lsx_options& o, void* stream = nullptr) : opt_(o), stream_(stream) {}
  virtual ~LsxIterativeSolver() {
    if (h_) lsx_destroy(h_);
  }
  LsxIterativeSolver(const LsxIterativeSolver&) = delete;
  LsxIterativeSolver& operator=(const LsxIterativeSolver&) = delete;

  // IterativeSolverInterface::set(A, rhs, sol, task)
  void set(const CsrView& A, double* rightM, double* sol) {
    if (h_) lsx_destroy(h_);
    h_ = nullptr;
    check(lsx_create(&h_, &opt_, stream_));
    A_ = A;
    rightM_ = rightM;
    sol_ = sol;
    eps = opt_.eps;
    check(lsx_set_csr(h_, A.N, A.M, A.K, A.Number, A.Portrait, A.Cell, A.elemCount, A.mem));
  }
  // IterativeSolverInterface::solve()
  int solve() {
    check(lsx_set_eps(h_, eps));
    check(lsx_solve(h_, rightM_, sol_, A_.mem, &it, &norm_r));
    return it;
  }
  // IterativeSolverInterface::rebuildPreconditioner(): the matrix values changed in place
  void rebuildPreconditioner() {
    check(lsx_update_values(h_, A_.Cell, A_.mem));
    check(lsx_rebuild(h_));
  }

  double eps = 0.0;     ///< absolute residual tolerance
  double norm_r = 0.0;  ///< residual norm after solve()
  int it = 1;           ///< iterations of the last solve()

 private:
  lsx_options opt_;
  void* stream_;
  lsx_solver* h_ = nullptr;
  CsrView A_{};
  double *rightM_ = nullptr, *sol_ = nullptr;
};

// MGFGMRES / MGBiCGStab of the reference (mg_fgmres.h:33-52, mg_bicgstab.h:33-52)
inline lsx_options mg_options(int krylov, int restart, int max_it, double eps, int n_post, int precision) {
  lsx_options o;
  lsx_options_default(&o);
  o.krylov = krylov;
  o.restart = restart;
  o.max_it = max_it;
  o.eps = eps;
  o.n_post = n_post;
  o.precision = precision;
  if (precision != LSX_PREC_FP64) {  // both mixed modes run the V-cycle
    o.mg = LSX_MG_VCYCLE;
    o.n_pre = 2;
    o.n_post = 2;
  }
  return o;
}

}  // namespace lsx_cpp

#endif  // LSX_ITERATIVE_SOLVER_HPP_
