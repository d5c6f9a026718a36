// ORACLE — test infrastructure only.
//
// CPU restatement of the reference pressure-solver algorithm (vvp-cfd/LS-STAG-3D_parallel),
// used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
// It is never linked into, loaded by or called from the product library.
//
// Every function cites the reference file:line whose behaviour it restates.  The reference
// sources do not compile unmodified (MSVC-permissive two-phase lookup inside its kernel
// macros), so this restatement is pinned by closed forms and brute-force checks in
// tests/test_oracle_*.py instead of by a build of the reference.
//
// Determinism: the reference's own reductions have no fixed order (OpenMP reductions on the CPU,
// atomicAdd on the GPU).  Every reduction here follows one fixed order, the one the HIP library
// uses (device_order_sum below), so results do not depend on the thread count and the Krylov
// scalars of the two implementations can be compared bit for bit.

#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace lsxo {

// ---------------------------------------------------------------------------------------
// BLAS-1 with deterministic reductions
// ---------------------------------------------------------------------------------------
// Fixed summation order of the HIP reductions: nb = clamp(ceil(n / 1024), 1, 1024) contiguous
// chunks of ceil(n / nb) elements; in a chunk, lane t of 256 adds elements t, t + 256, ... left to
// right; each 64-lane wave folds its lanes pairwise (l += l + 32, then + 16, ..., + 1); lane 0 of
// the 4 waves are added in wave order.  The nb chunk sums are then reduced the same way, lane t
// adding the chunk sums [t p, t p + p) with p = ceil(nb / 256).
static double wave_fold(double* v) {
  for (int off = 32; off > 0; off >>= 1)
    for (int l = 0; l < off; ++l) v[l] = v[l] + v[l + off];
  return v[0];
}

static double block_fold(double* lane) {  // 256 lane values
  double s = 0.0;
  for (int w = 0; w < 4; ++w) s += wave_fold(lane + 64 * w);
  return s;
}

template <class F>
static double device_order_sum(long n, F term) {
  long nb = (n + 1023) / 1024;
  if (nb > 1024) nb = 1024;
  if (nb < 1) nb = 1;
  const long per = (n + nb - 1) / nb;
  std::vector<double> part(nb, 0.0);
#pragma omp parallel for schedule(static)
  for (long b = 0; b < nb; ++b) {
    double lane[256];
    const long lo = per * b, hi = std::min(n, lo + per);
    for (int t = 0; t < 256; ++t) {
      double acc = 0.0;
      for (long i = lo + t; i < hi; i += 256) acc += term(i);
      lane[t] = acc;
    }
    part[b] = block_fold(lane);
  }
  const long pp = (nb + 255) / 256;
  double lane[256];
  for (int t = 0; t < 256; ++t) {
    double acc = 0.0;
    for (long i = t * pp; i < t * pp + pp && i < nb; ++i) acc += part[i];
    lane[t] = acc;
  }
  return block_fold(lane);
}

template <class F>
static double blocked_sum(long n, F term) {
  return device_order_sum(n, term);
}

double dot(long n, const double* a, const double* b) {
  // mesh_function.h:150-153 (dot) / :305-309 (operator*)
  return blocked_sum(n, [&](long i) { return a[i] * b[i]; });
}

double norm2(long n, const double* a) {
  // mesh_function.h:141-145 + getNorm :223-228
  return std::sqrt(blocked_sum(n, [&](long i) { return a[i] * a[i]; }));
}

// ---------------------------------------------------------------------------------------
// CSR products — real_valued_mesh_function.h:81-88 (mvProd), :195-202 (mvProdDiff),
// :203-211 (mvRes)
// ---------------------------------------------------------------------------------------
void spmv(const Csr& A, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < A.rows; ++i) {
    double aux = 0.0;
    for (int j = A.rowptr[i]; j < A.rowptr[i + 1]; ++j) aux += A.val[j] * x[A.col[j]];
    y[i] = aux;
  }
}

void residual(const Csr& A, const double* x, const double* b, double* r) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < A.rows; ++i) {
    double aux = b[i];
    for (int j = A.rowptr[i]; j < A.rowptr[i + 1]; ++j) aux -= A.val[j] * x[A.col[j]];
    r[i] = aux;
  }
}

double residual_norm(const Csr& A, const double* x, const double* b, double* r) {
  residual(A, x, b, r);
  return norm2(A.rows, r);
}

// ---------------------------------------------------------------------------------------
// Sparse product with the reference's per-row insertion order:
// product_sparse_matrix.h:45-104 (numSpMM symbolic count, spMM numeric insertion sort).
// For output entry (i, c) the terms are added in the order they are met: rows h of row i of
// the left factor in CSR order, then the right factor's row h in CSR order.
// ---------------------------------------------------------------------------------------
Csr spgemm(const Csr& L, const Csr& U, int cols) {
  Csr C;
  C.rows = L.rows;
  C.cols = cols;
  C.rowptr.assign(L.rows + 1, 0);
  std::vector<int> mark(cols, -1);
  long total = 0;
  for (long i = 0; i < L.rows; ++i) {
    for (int h = L.rowptr[i]; h < L.rowptr[i + 1]; ++h) {
      const int nl = L.col[h];
      for (int k = U.rowptr[nl]; k < U.rowptr[nl + 1]; ++k) {
        const int up = U.col[k];
        if (mark[up] != (int)i) {
          mark[up] = (int)i;
          ++total;
        }
      }
    }
    C.rowptr[i + 1] = (int)total;
  }
  C.col.assign(total, -1);
  C.val.assign(total, 0.0);
#pragma omp parallel for schedule(dynamic, 256)
  for (long i = 0; i < L.rows; ++i) {
    const int start = C.rowptr[i];
    int count = 0;
    for (int h = L.rowptr[i]; h < L.rowptr[i + 1]; ++h) {
      const double a = L.val[h];
      const int nl = L.col[h];
      for (int k = U.rowptr[nl]; k < U.rowptr[nl + 1]; ++k) {
        const int colk = U.col[k];
        const double v = a * U.val[k];
        int pos = start;
        const int shift = start + count;
        while (pos < shift && C.col[pos] < colk) ++pos;
        if (pos < shift && C.col[pos] == colk) {
          C.val[pos] += v;
        } else {
          for (int q = shift; q > pos; --q) {
            C.col[q] = C.col[q - 1];
            C.val[q] = C.val[q - 1];
          }
          C.col[pos] = colk;
          C.val[pos] = v;
          ++count;
        }
      }
    }
  }
  return C;
}

// ---------------------------------------------------------------------------------------
// Even-even-even transfer operators.
// P: cell-centred trilinear prolongation, per-axis weights 3/4 (parent) and 1/4 (the
//    neighbouring coarse cell on the child's side), weight 1 along an axis where the fine
//    cell touches the domain boundary (transfer_operators_even_even_even.h:47-86 with
//    transfer_operators_interface.h:338-428).
// R: full restriction, 1/8 on each of the 8 children, children ordered
//    (x,y,z),(x+1,y,z),(x,y+1,z),(x+1,y+1,z), then the z+1 layer
//    (transfer_operators_even_even_even.h:87-93, PortraitWidth8_1xy; buildFullR
//    transfer_operators_interface.h:579-587).
// ---------------------------------------------------------------------------------------
static void axis_weights(int f, int nf, int* c, double* w, int* cnt) {
  const int p = f >> 1;
  if (f == 0 || f == nf - 1) {
    c[0] = p;
    w[0] = 1.0;
    *cnt = 1;
    return;
  }
  if (f & 1) {  // right child: neighbour is p+1
    c[0] = p;
    w[0] = 0.75;
    c[1] = p + 1;
    w[1] = 0.25;
  } else {  // left child: neighbour is p-1
    c[0] = p - 1;
    w[0] = 0.25;
    c[1] = p;
    w[1] = 0.75;
  }
  *cnt = 2;
}

// Odd-odd-odd transfer operators (transfer_operators_odd_odd_odd.h:44-100 with
// transfer_operators_interface.h:48-57, 131-154, 181-198, 203-235): an axis with an odd number
// 2n - 1 of fine cells coarsens to n cells whose centres sit on the even fine cells; P weights 1
// on a coincident fine cell and 1/2 on each of the two coarse neighbours of an odd fine cell, each
// row's entries ordered z, then y, then x, coarse index ascending.  R = P^T
// (transfer_operators_interface.h:572-578, the transposition of transposed_sparse_matrix.h:47-68:
// each R row lists its fine cells in ascending order).  The mixed-parity classes build R the same
// way but keep the cell-centred 3/4, 1/4 prolongation on their even axes, which widens the coarse
// operator to five cells along those axes; they are not restated.
static void axis_weights_odd(int f, int* c, double* w, int* cnt) {
  if ((f & 1) == 0) {
    c[0] = f >> 1;
    w[0] = 1.0;
    *cnt = 1;
  } else {
    c[0] = (f - 1) >> 1;
    w[0] = 0.5;
    c[1] = (f + 1) >> 1;
    w[1] = 0.5;
    *cnt = 2;
  }
}

int coarse_count(int n) { return (n & 1) ? (n + 1) / 2 : n / 2; }  // coarse_grid.h:60-67

Csr build_P_odd(int NN, int MM, int KK) {
  const int N = coarse_count(NN), M = coarse_count(MM), K = coarse_count(KK);
  Csr P;
  P.rows = (long)NN * MM * KK;
  P.cols = (long)N * M * K;
  P.rowptr.assign(P.rows + 1, 0);
  P.col.reserve(P.rows * 8);
  P.val.reserve(P.rows * 8);
  long r = 0;
  for (int z = 0; z < KK; ++z)
    for (int y = 0; y < MM; ++y)
      for (int x = 0; x < NN; ++x, ++r) {
        int cx[2], cy[2], cz[2], nx, ny, nz;
        double wx[2], wy[2], wz[2];
        axis_weights_odd(x, cx, wx, &nx);
        axis_weights_odd(y, cy, wy, &ny);
        axis_weights_odd(z, cz, wz, &nz);
        for (int c = 0; c < nz; ++c)
          for (int b = 0; b < ny; ++b)
            for (int a = 0; a < nx; ++a) {
              P.col.push_back(cx[a] + N * cy[b] + N * M * cz[c]);
              P.val.push_back(wx[a] * wy[b] * wz[c]);
            }
        P.rowptr[r + 1] = (int)P.col.size();
      }
  return P;
}

Csr transpose(const Csr& P) {
  Csr R;
  R.rows = P.cols;
  R.cols = P.rows;
  R.rowptr.assign(R.rows + 1, 0);
  for (int c : P.col) ++R.rowptr[c + 1];
  for (long i = 0; i < R.rows; ++i) R.rowptr[i + 1] += R.rowptr[i];
  R.col.assign(P.col.size(), 0);
  R.val.assign(P.col.size(), 0.0);
  std::vector<int> fill(R.rowptr.begin(), R.rowptr.end() - 1);
  for (long i = 0; i < P.rows; ++i)
    for (int j = P.rowptr[i]; j < P.rowptr[i + 1]; ++j) {
      const int k = fill[P.col[j]]++;
      R.col[k] = (int)i;
      R.val[k] = P.val[j];
    }
  return R;
}

Csr build_P(int NN, int MM, int KK) {
  const int N = NN / 2, M = MM / 2, K = KK / 2;
  Csr P;
  P.rows = (long)NN * MM * KK;
  P.cols = (long)N * M * K;
  P.rowptr.assign(P.rows + 1, 0);
  P.col.reserve(P.rows * 8);
  P.val.reserve(P.rows * 8);
  long r = 0;
  for (int z = 0; z < KK; ++z)
    for (int y = 0; y < MM; ++y)
      for (int x = 0; x < NN; ++x, ++r) {
        int cx[2], cy[2], cz[2], nx, ny, nz;
        double wx[2], wy[2], wz[2];
        axis_weights(x, NN, cx, wx, &nx);
        axis_weights(y, MM, cy, wy, &ny);
        axis_weights(z, KK, cz, wz, &nz);
        for (int c = 0; c < nz; ++c)
          for (int b = 0; b < ny; ++b)
            for (int a = 0; a < nx; ++a) {
              P.col.push_back(cx[a] + N * cy[b] + N * M * cz[c]);
              P.val.push_back(wx[a] * wy[b] * wz[c]);
            }
        P.rowptr[r + 1] = (int)P.col.size();
      }
  return P;
}

Csr build_R(int NN, int MM, int KK) {
  const int N = NN / 2, M = MM / 2, K = KK / 2;
  const long NMf = (long)NN * MM;
  Csr R;
  R.rows = (long)N * M * K;
  R.cols = (long)NN * MM * KK;
  R.rowptr.resize(R.rows + 1);
  R.col.resize(R.rows * 8);
  R.val.assign(R.rows * 8, 0.125);
  long r = 0;
  for (int z = 0; z < K; ++z)
    for (int y = 0; y < M; ++y)
      for (int x = 0; x < N; ++x, ++r) {
        const long a = 2 * x + (long)NN * 2 * y + NMf * 2 * z;
        const long ch[8] = {a, a + 1, a + NN, a + NN + 1, a + NMf, a + NMf + 1, a + NMf + NN, a + NMf + NN + 1};
        R.rowptr[r] = (int)(r * 8);
        for (int q = 0; q < 8; ++q) R.col[r * 8 + q] = (int)ch[q];
      }
  R.rowptr[R.rows] = (int)(R.rows * 8);
  return R;
}

// ---------------------------------------------------------------------------------------
// ADLJ smoother — smoother_adlj.h:40-86 (tridiagonal extraction), tridiagonal_matrix.h:45-67
// (triLU), smoother_adlj_x.h:38-48, smoother_adlj_y.h:36-65, smoother_adlj_z.h:34-58.
// ---------------------------------------------------------------------------------------
void Adlj::build(const Csr& A, int N_, int M_, int K_) {
  N = N_;
  M = M_;
  K = K_;
  const long n = (long)N * M * K;
  const long NM = (long)N * M, NK = (long)N * K;
  for (int d = 0; d < 3; ++d) {
    c[d].assign(n > 0 ? n - 1 : 0, 0.0);
    dg[d].assign(n, 0.0);
    e[d].assign(n > 0 ? n - 1 : 0, 0.0);
  }
  // y-order: index y + M*x + NM*z ; z-order: index z + K*x + NK*y (initOrder / initOrderZ)
  yho.resize(n);
  yvo.resize(n);
  zho.resize(n);
  zvo.resize(n);
  for (int z = 0; z < K; ++z)
    for (int y = 0; y < M; ++y)
      for (int x = 0; x < N; ++x) {
        const long nat = x + (long)N * y + NM * z;
        const long yi = y + (long)M * x + NM * z;
        const long zi = z + (long)K * x + NK * y;
        yho[yi] = (int)nat;
        yvo[nat] = (int)yi;
        zho[zi] = (int)nat;
        zvo[nat] = (int)zi;
      }
  for (long i = 0; i < n; ++i) {
    const long num = yvo[i], numZ = zvo[i];
    const long iw = i - 1, is = i - N, ib = i - NM, it = i + NM, in = i + N, ie = i + 1;
    for (int j = A.rowptr[i]; j < A.rowptr[i + 1]; ++j) {
      const long ap = A.col[j];
      const double v = A.val[j];
      if (ap == ib)
        c[2][numZ - 1] = v;
      else if (ap == is)
        c[1][num - 1] = v;
      else if (ap == iw)
        c[0][iw] = v;
      else if (ap == i) {
        dg[2][numZ] = v;
        dg[1][num] = v;
        dg[0][i] = v;
      } else if (ap == ie)
        e[0][i] = v;
      else if (ap == in)
        e[1][num] = v;
      else if (ap == it)
        e[2][numZ] = v;
    }
  }
  for (int d = 0; d < 3; ++d) tri_lu(dg[d], c[d], e[d]);
}

void Adlj::tri_lu(std::vector<double>& d, const std::vector<double>& cc, std::vector<double>& ee) {
  const long n = (long)d.size();
  if (!n) return;
  double maxA = 0.0;
  for (long i = 0; i < n; ++i) maxA = std::max(maxA, std::fabs(d[i]));
  double tiny = EQ_TINY2_DOUBLE * maxA;
  if (tiny < EQ_TINY_DOUBLE) tiny = EQ_TINY_DOUBLE;
  double aux = d[0];
  d[0] = (std::fabs(aux) < tiny) ? 0.0 : (1.0 / aux);
  for (long i = 1, j = 0; i < n; ++i, ++j) {
    ee[j] *= d[j];
    aux = d[i] - cc[j] * ee[j];
    d[i] = (std::fabs(aux) < tiny) ? 0.0 : (1.0 / aux);
  }
}

// c = T_dir^{-1} rhs (solution written in natural order), reference TDMA kernels.
void Adlj::tdma(int dir, const double* rhs, double* sol, double* aux) const {
  const long NM = (long)N * M, NK = (long)N * K;
  if (dir == 0) {  // smoother_adlj_x.h:38-48, 2D launch (K, M)
    const double* cc = c[0].data();
    const double* d = dg[0].data();
    const double* ee = e[0].data();
#pragma omp parallel for collapse(2) schedule(static)
    for (int kk = 0; kk < K; ++kk)
      for (int jj = 0; jj < M; ++jj) {
        const long num = (long)jj * N + kk * NM;
        sol[num] = rhs[num] * d[num];
        for (long i2 = num, j2 = i2 + 1, end = i2 + N; j2 < end; ++j2, ++i2) sol[j2] = (rhs[j2] - cc[i2] * sol[i2]) * d[j2];
        for (long j2 = num + N - 2; j2 >= num; --j2) sol[j2] -= ee[j2] * sol[j2 + 1];
      }
    return;
  }
  // y: launch (K, N), arrays in y-order, stride NM per plane, line length M
  // z: launch (M, N), arrays in z-order, stride NK per y, line length K; first rhs element
  //    is read at index i*NK + j exactly as the reference z-call does (it passes NK where
  //    the y kernel expects NM, smoother_adlj_z.h:54-58 with smoother_adlj_y.h:36-50).
  const int di = dir;
  const double* cc = c[di].data();
  const double* d = dg[di].data();
  const double* ee = e[di].data();
  const int* ho = (dir == 1) ? yho.data() : zho.data();
  const int s1 = (dir == 1) ? K : M;
  const long planeStride = (dir == 1) ? NM : NK;
  const int len = (dir == 1) ? M : K;
#pragma omp parallel for collapse(2) schedule(static)
  for (int ii = 0; ii < s1; ++ii)
    for (int jj = 0; jj < N; ++jj) {
      const long is = (long)ii * planeStride;
      const long iY = is + jj;
      const long num = is + (long)jj * len;
      const long last = num + len - 1;
      aux[num] = rhs[iY] * d[num];
      for (long i2 = num, j2 = i2 + 1; j2 <= last; ++j2, ++i2) aux[j2] = (rhs[ho[j2]] - cc[i2] * aux[i2]) * d[j2];
      sol[ho[last]] = aux[last];
      for (long j2 = last - 1; j2 >= num; --j2) sol[ho[j2]] = aux[j2] - ee[j2] * sol[ho[j2 + 1]];
    }
}

// ---------------------------------------------------------------------------------------
// Dense LU with symmetric diagonal scaling and row partial pivoting for the coarsest level —
// lu_decomposition_sparse_matrix.h:43-215 (initScalePerm, calcLU), solve :216-259 and
// decomposition_sparse_matrix_interface.h:43-71.
// ---------------------------------------------------------------------------------------
void DenseLU::factor(const Csr& A) {
  n = (int)A.rows;
  scale.assign(n, 1.0);
  perm.resize(n);
  L.assign((size_t)n * n, 0.0);
  U.assign((size_t)n * n, 0.0);
  auto a_at = [&](int row, int colj) -> double {
    for (int h = A.rowptr[row]; h < A.rowptr[row + 1]; ++h)
      if (A.col[h] == colj) return A.val[h];
    return 0.0;
  };
  for (int i = 0; i < n; ++i) {
    double diag = 0.0;
    for (int k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k)
      if (A.col[k] == i) {
        diag = std::fabs(A.val[k]);
        break;
      }
    scale[i] = (diag > EQ_DOUBLE) ? (1.0 / std::sqrt(diag)) : 1.0;
    perm[i] = i;
  }
  double maxAbs = 0.0, maxVal = 0.0;
  int pivot = 0;
  double si = scale[0];
  for (int i = 0; i < n; ++i) {
    int k = A.rowptr[i];
    const double curA = A.val[k];
    if (A.col[k] == 0) {
      const double absVal = std::fabs(curA * scale[i] * si);
      if (absVal > maxVal) {
        maxVal = absVal;
        pivot = i;
      }
    }
    for (; k < A.rowptr[i + 1]; ++k) maxAbs = std::max(maxAbs, std::fabs(A.val[k]));
  }
  tiny = EQ_TINY3_DOUBLE * maxAbs;
  if (pivot != 0) std::swap(perm[0], perm[pivot]);
  // Row 0 of U (U stored dense row-major, U[i][j] for j >= i)
  int permi = perm[0];
  si = scale[permi];
  for (int j = 0; j < n; ++j) U[j] = a_at(permi, j) * si * scale[j];
  double U00 = U[0];
  if (std::fabs(U00) < tiny) U00 = (U00 < 0) ? -tiny : tiny;
  U[0] = U00;
  const double invU00 = scale[0] / U00;
  for (int k = 1; k < n; ++k) {
    const int pk = perm[k];
    L[(size_t)k * n + 0] = a_at(pk, 0) * scale[pk] * invU00;
  }
  L[0] = 1.0;
  std::vector<double> col(n, 0.0);
  for (int i = 1; i < n; ++i) {
    maxVal = 0.0;
    pivot = i;
    si = scale[i];
    for (int k = i; k < n; ++k) {
      double sum = 0.0;
      for (int t = 0; t < i; ++t) sum += L[(size_t)k * n + t] * U[(size_t)t * n + i];
      const int pk = perm[k];
      const double val = a_at(pk, i) * scale[pk] * si - sum;
      col[k] = val;
      if (std::fabs(val) > maxVal) {
        maxVal = std::fabs(val);
        pivot = k;
      }
    }
    if (pivot != i) {
      std::swap(perm[i], perm[pivot]);
      for (int t = 0; t < i; ++t) std::swap(L[(size_t)i * n + t], L[(size_t)pivot * n + t]);
      std::swap(col[i], col[pivot]);
    }
    permi = perm[i];
    si = scale[permi];
    for (int j = i; j < n; ++j) {
      double sum = 0.0;
      for (int t = 0; t < i; ++t) sum += L[(size_t)i * n + t] * U[(size_t)t * n + j];
      U[(size_t)i * n + j] = a_at(permi, j) * si * scale[j] - sum;
    }
    double Uii = U[(size_t)i * n + i];
    if (std::fabs(Uii) < tiny) Uii = (Uii < 0) ? -tiny : tiny;
    U[(size_t)i * n + i] = Uii;
    const double invUii = 1.0 / Uii;
    for (int k = i + 1; k < n; ++k) L[(size_t)k * n + i] = col[k] * invUii;
    L[(size_t)i * n + i] = 1.0;
  }
}

void DenseLU::solve(const double* b, double* x) const {
  std::vector<double> bs(n), y(n);
  for (int i = 0; i < n; ++i) bs[i] = scale[perm[i]] * b[perm[i]];  // copyCompProdPerm
  for (int i = 0; i < n; ++i) {                                     // directGaussRow
    double aux = bs[i];
    for (int t = 0; t < i; ++t) aux -= L[(size_t)i * n + t] * y[t];
    y[i] = aux;
  }
  for (int i = n - 1; i >= 0; --i) {  // inverseGaussRow (writes into bScaled)
    double aux = y[i];
    for (int j = i + 1; j < n; ++j) aux -= U[(size_t)i * n + j] * bs[j];
    bs[i] = aux / U[(size_t)i * n + i];
  }
  for (int i = 0; i < n; ++i) x[i] = scale[i] * bs[i];  // copyCompProd
}

// ---------------------------------------------------------------------------------------
// Multigrid cascade — multigrid_solver.h:48-101, coarse_grids_hierarchy.h:47-81,
// coarse_grid.h:62-106, smoother_adlj_interface.h:66-82.
// ---------------------------------------------------------------------------------------
int count_levels(int N, int M, int K) {
  double aux = (double)std::min(std::min(N, M), K);
  int num = 0;
  while (aux > 2.0) {
    aux *= 0.5;
    ++num;
  }
  return num;
}

void Multigrid::set(const Csr& A0, int N, int M, int K, int n_post_, double omega_, bool single) {
  n_post = n_post_;
  omega = omega_;
  const int numMG = single ? 0 : count_levels(N, M, K);
  levels.clear();
  levels.resize(numMG + 1);
  levels[0].A = A0;
  levels[0].N = N;
  levels[0].M = M;
  levels[0].K = K;
  for (int l = 0; l < numMG; ++l) {
    Level& f = levels[l];
    const int odd = (f.N & 1) + (f.M & 1) + (f.K & 1);
    if (odd == 0) {
      f.P = build_P(f.N, f.M, f.K);
      f.R = build_R(f.N, f.M, f.K);
    } else if (odd == 3) {
      f.P = build_P_odd(f.N, f.M, f.K);
      f.R = transpose(f.P);
    } else {
      throw std::runtime_error("mixed-parity level dimensions are not supported (even-even-even and odd-odd-odd "
                               "transfers only)");
    }
    Level& cgrid = levels[l + 1];
    cgrid.N = coarse_count(f.N);
    cgrid.M = coarse_count(f.M);
    cgrid.K = coarse_count(f.K);
    Csr V = spgemm(f.R, f.A, f.A.cols);
    cgrid.A = spgemm(V, f.P, f.P.cols);
    f.adlj.build(f.A, f.N, f.M, f.K);
  }
  if (single)
    levels[0].adlj.build(levels[0].A, N, M, K);
  else
    lu.factor(levels[numMG].A);
  const long n0 = A0.rows;
  c.assign(n0, 0.0);
  ch.assign(n0, 0.0);
  cHH.assign(n0, 0.0);
  q.assign(n0, 0.0);
  r.resize(numMG + 1);
  for (int l = 0; l <= numMG; ++l) r[l].assign(levels[l].A.rows, 0.0);
}

void Multigrid::apply(const double* b, double* x) {
  const int numMG = (int)levels.size() - 1;
  std::copy(b, b + levels[0].A.rows, r[0].begin());
  for (int k = 0; k < numMG; ++k) spmv(levels[k].R, r[k].data(), r[k + 1].data());
  if (numMG > 0) lu.solve(r[numMG].data(), cHH.data());
  for (int k = numMG > 0 ? numMG - 1 : 0; k >= 0; --k) {
    Level& L = levels[k];
    const long n = L.A.rows;
    if (numMG > 0) {
      spmv(L.P, cHH.data(), ch.data());
    } else {  // one level: the sweeps start from a zero correction
      std::fill(ch.begin(), ch.begin() + n, 0.0);
    }
    std::copy(ch.begin(), ch.begin() + n, c.begin());
    for (int s = 0; s < n_post; ++s) {
      for (int dir = 2; dir >= 0; --dir) {  // z, y, x  (smoother_adlj.h:121-126)
        spmv(L.A, c.data(), q.data());      // calcRes: q = A c ; r -= q
        double* rk = r[k].data();
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) rk[i] -= q[i];
        L.adlj.tdma(dir, rk, c.data(), q.data());  // aux_sol shares q's storage
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {
          c[i] *= omega;
          ch[i] += c[i];
        }
      }
    }
    std::copy(ch.begin(), ch.begin() + n, cHH.begin());
  }
  std::copy(ch.begin(), ch.begin() + levels[0].A.rows, x);
}

// ---------------------------------------------------------------------------------------
// FGMRES(m) — fgmres_interface.h:59-73 (set), :78-142 (solve), :153-168 (solUpdate)
// ---------------------------------------------------------------------------------------
static void fgmres_sol_update(long n, int last, int lastH, std::vector<double>& y, const std::vector<double>& H,
                              const std::vector<double>& Z, double* sol) {
  int hNum = lastH;
  for (int k = last; k > -1; --k, --hNum) {
    const double curY = y[k] / H[hNum];
    for (int i = k - 1; i > -1; --i) {
      --hNum;
      y[i] -= curY * H[hNum];
    }
    const double* z = Z.data() + (size_t)k * n;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) sol[i] += z[i] * curY;
  }
}

int fgmres_solve(const Csr& A, Preconditioner& prec, const double* b, double* x, double eps, int max_it, int restart,
                 double* norm_out) {
  const long n = A.rows;
  std::vector<double> V((size_t)restart * n), Z((size_t)restart * n), w(n);
  std::vector<double> c(restart), s(restart), y(restart), H(restart * (restart + 1) / 2);
  const int m_size = restart * (restart + 1) / 2;
  int it = 0;
  double norm_r = residual_norm(A, x, b, w.data());
  if (!(norm_r > eps)) {
    *norm_out = norm_r;
    return 0;
  }
  it = 1;
  double normW = norm_r;
  double curH = norm_r;
  for (;;) {
    for (int j = 0, hColStart = 0, hNum = 0; j < restart; ++j, hColStart = hNum, ++it) {
      double* curV = V.data() + (size_t)j * n;
      double* curZ = Z.data() + (size_t)j * n;
      const double inv = 1.0 / normW;
#pragma omp parallel for schedule(static)
      for (long i = 0; i < n; ++i) curV[i] = inv * w[i];
      prec.apply(curV, curZ);
      spmv(A, curZ, w.data());
      for (int i = 0; i <= j; ++i, ++hNum) {
        const double* v2 = V.data() + (size_t)i * n;
        const double h = dot(n, w.data(), v2);
#pragma omp parallel for schedule(static)
        for (long q = 0; q < n; ++q) w[q] -= v2[q] * h;
        H[hNum] = h;
      }
      curH = H[hColStart];
      for (int i = 0, cur = hColStart, nxt = cur + 1; i < j; ++i, cur = nxt, ++nxt) {
        const double cc = c[i], ss = s[i], nextH = H[nxt];
        H[cur] = cc * curH - ss * nextH;
        curH = ss * curH + cc * nextH;
      }
      normW = norm2(n, w.data());
      double curC = curH * curH + normW * normW;
      const double curA = std::sqrt(curC);
      curC = 1.0 / curA;
      const double curS = -normW * curC;
      curC *= curH;
      H[hNum - 1] = curA;
      y[j] = curC * norm_r;
      norm_r *= curS;
      curH = std::fabs(norm_r);
      if ((curH > eps) && (it < max_it)) {
        c[j] = curC;
        s[j] = curS;
      } else {
        norm_r = curH;
        fgmres_sol_update(n, j, hNum - 1, y, H, Z, x);
        *norm_out = norm_r;
        return it;
      }
    }
    fgmres_sol_update(n, restart - 1, m_size - 1, y, H, Z, x);
    residual(A, x, b, w.data());  // w = b - A x, next cycle's first direction
    normW = curH;                 // estimated |gamma|, not recomputed
  }
}

// ---------------------------------------------------------------------------------------
// BiCGStab — bicgstab_interface.h:39-56 (solUpd, pUpd), :76-111 (solve), :124-128
// ---------------------------------------------------------------------------------------
int bicgstab_solve(const Csr& A, Preconditioner& prec, const double* b, double* x, double eps, int max_it,
                   double* norm_out) {
  const long n = A.rows;
  std::vector<double> nu(n), t(n), yv(n), z(n), r(n), p(n), s(n), rr(n);
  int it = 1;
  spmv(A, x, z.data());
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) r[i] = b[i] - z[i];
  double norm_r = norm2(n, r.data());
  if (norm_r > eps) {
    rr = r;
    p = r;
    do {
      prec.apply(p.data(), yv.data());
      spmv(A, yv.data(), nu.data());
      double vs = dot(n, rr.data(), r.data());
      double coef = dot(n, rr.data(), nu.data());
      const double alpha = vs / coef;
#pragma omp parallel for schedule(static)
      for (long i = 0; i < n; ++i) {
        nu[i] *= alpha;
        s[i] = r[i] - nu[i];
      }
      prec.apply(s.data(), z.data());
      spmv(A, z.data(), t.data());
      double w = dot(n, t.data(), s.data());
      coef = dot(n, t.data(), t.data());
      w /= coef;
#pragma omp parallel for schedule(static)
      for (long i = 0; i < n; ++i) {
        yv[i] *= alpha;
        z[i] *= w;
        const double cs = yv[i] + z[i];
        x[i] += cs;
        t[i] *= w;
        r[i] = s[i] - t[i];
      }
      coef = w / alpha;
      vs *= coef;
      double beta = dot(n, rr.data(), r.data());
      beta /= vs;
      norm_r = std::sqrt(blocked_sum(n, [&](long i) {
        const double cv = (p[i] - nu[i] * coef) * beta;
        const double re = r[i];
        p[i] = re + cv;
        return re * re;
      }));
      ++it;
    } while ((norm_r > eps) && (it < max_it));
  }
  *norm_out = norm_r;
  return it;
}

}  // namespace lsxo
