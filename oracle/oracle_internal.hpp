// ORACLE — test infrastructure only (see oracle_solver.cpp header).
#pragma once

#include <vector>

namespace lsxo {

// Constants of the reference, src/Base/ls_stag_real_constants.h:27-33, 62
constexpr double PLUS_TINY = 1e-16;
constexpr double EQ_DOUBLE = 1e-5;
constexpr double THETA_EQ_ZERO = 0.05;
constexpr double EQ_TINY2_DOUBLE = 1e-5;
constexpr double EQ_TINY_DOUBLE = 1e-12;
constexpr double EQ_TINY3_DOUBLE = 1e-13;
constexpr double ONE_THIRD = 1.0 / 3.0;
constexpr double DEFAULT_OMEGA_3D = 0.805;

// CSR with int32 row pointers and columns (sparse_matrix.h:220-228: Number/Portrait/Cell)
struct Csr {
  long rows = 0, cols = 0;
  std::vector<int> rowptr, col;
  std::vector<double> val;
};

double dot(long n, const double* a, const double* b);
double norm2(long n, const double* a);
void spmv(const Csr& A, const double* x, double* y);
void residual(const Csr& A, const double* x, const double* b, double* r);
double residual_norm(const Csr& A, const double* x, const double* b, double* r);
Csr spgemm(const Csr& L, const Csr& U, int cols);
Csr build_P(int NN, int MM, int KK);
Csr build_R(int NN, int MM, int KK);
int count_levels(int N, int M, int K);

struct Adlj {
  int N = 0, M = 0, K = 0;
  // index 0: x lines (natural order), 1: y lines (y-order), 2: z lines (z-order)
  std::vector<double> c[3], dg[3], e[3];
  std::vector<int> yho, yvo, zho, zvo;
  void build(const Csr& A, int N, int M, int K);
  static void tri_lu(std::vector<double>& d, const std::vector<double>& c, std::vector<double>& e);
  void tdma(int dir, const double* rhs, double* sol, double* aux) const;
};

struct DenseLU {
  int n = 0;
  double tiny = 0.0;
  std::vector<double> scale, L, U;
  std::vector<int> perm;
  void factor(const Csr& A);
  void solve(const double* b, double* x) const;
};

struct Preconditioner {
  virtual ~Preconditioner() {}
  virtual void apply(const double* b, double* x) = 0;
};

struct Identity : Preconditioner {
  long n = 0;
  void apply(const double* b, double* x) override {
    for (long i = 0; i < n; ++i) x[i] = b[i];
  }
};

struct Level {
  int N = 0, M = 0, K = 0;
  Csr A, P, R;
  Adlj adlj;
};

struct Multigrid : Preconditioner {
  std::vector<Level> levels;  // levels[0] = fine; levels.back() = coarsest (LU)
  DenseLU lu;
  int n_post = 3;
  double omega = DEFAULT_OMEGA_3D;
  std::vector<double> c, ch, cHH, q;
  std::vector<std::vector<double>> r;
  // single: one level only (no transfers, no coarsest LU): the apply is n_post ADLJ sweeps of the
  // fine level from a zero correction, for grids of any dimensions
  void set(const Csr& A0, int N, int M, int K, int n_post, double omega, bool single = false);
  void apply(const double* b, double* x) override;
};

int fgmres_solve(const Csr& A, Preconditioner& prec, const double* b, double* x, double eps, int max_it, int restart,
                 double* norm_out);
int bicgstab_solve(const Csr& A, Preconditioner& prec, const double* b, double* x, double eps, int max_it,
                   double* norm_out);

}  // namespace lsxo
