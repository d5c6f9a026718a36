// Geometric V-cycle preconditioner on face-coefficient storage: device kernels.
//
// Two instantiations share the kernels:
//   FP64 : FP64 values and faces on every level (the plain FP64 V-cycle);
//   mixed: FP32 values, FP32 faces on the finest level and FP16 faces (divided by a per-level
//          scale) on coarse levels.
// The coarsest level is always solved by the FP64 dense LU.
#pragma once

#include <hip/hip_runtime.h>

#include "lsx_kernels.hpp"

namespace lsx {
namespace vc {

enum FaceType { FACE_F64 = 0, FACE_F32 = 1, FACE_F16 = 2 };

// One level.  Faces: fe[i] couples cell i with i+1 (x), fn[i] with i+N, ft[i] with i+N*M; a
// face on the domain boundary is 0.  s[i] = 1 / a_p (0 for an empty row).  pin0/pin1 are
// identity rows (the finest level's pinned cells), -1 on coarse levels.
//
// dff: every row other than the pinned ones has a_p = -(sum of its six faces) - sigma, sigma the
// pinned cells' couplings on the sink cells sink0/sink1 of a coarse level (sig0/sig1) and 0
// elsewhere; a row without faces has a_p = 0 (empty) or a_p = dz, the level's one common value
// (the reference's solid cells), and then its own east face is stored as -0.0 (checked and
// marked at set-up in FP64).  The forward line passes and the residual restriction then rebuild
// a_p from the faces they load anyway instead of reading s (4 of 28 bytes per cell of an FP32
// line pass); the other kernels read s.
struct Level {
  Dims d{0, 0, 0};
  int ftype = FACE_F64;
  const void *f_e = nullptr, *f_n = nullptr, *f_t = nullptr;
  const void* s = nullptr;  // value type of the hierarchy
  double fscale = 1.0;      // faces are stored divided by fscale
  int pin0 = -1, pin1 = -1;
  int dff = 0;
  int sink0 = -1, sink1 = -1;
  double sig0 = 0.0, sig1 = 0.0, dz = 0.0;
};

// Set-up of Level::dff on the FP64 faces and diagonal of one level (v: a scratch vector of n
// doubles, partial as for face_absmax, out: one device double):
//   mode 0: out = max |a_p| over the non-pinned rows without faces (the candidate dz);
//   mode 1: out = max over the rows of the relative deviation of a_p from the rule with dz (1 for a
//           row without faces whose a_p is neither 0 nor dz, and with hscale > 0 for a row whose
//           faces all round to 0 when stored as FP16 times hscale);
//   mode 2: marks the rows without faces whose a_p is dz != 0 (fe = -0.0).
void diag_rule(double* fe, const double* fn, const double* ft, const double* dg, Dims d, int pin0, int pin1,
               int sink0, int sink1, double sig0, double sig1, double dz, double hscale, int mode, double* v,
               double* partial, double* out, hipStream_t s);

// The small levels handled by one workgroup in a single launch (k_coarse_cycle): levels with at
// most COARSE_CELLS cells (16^3; a 32^3 level already runs faster on the whole device), down to
// and including the LU level.
constexpr int COARSE_MAX = 8;
constexpr int COARSE_THREADS = 1024;
constexpr long COARSE_CELLS = 4096;
template <class T>
struct CoarseLevels {
  int count = 0;
  Level L[COARSE_MAX];
  T* x[COARSE_MAX] = {};
  T* b[COARSE_MAX] = {};
};

// ---- setup (FP64 temporaries) --------------------------------------------------------------
// faces and diagonal of the finest level from the FP64 7-plane operator [b,s,w,p,e,n,t]; a face
// is read from the row of the non-pinned cell (pinned rows are identity rows)
void faces_from7(const double* c7, Dims d, int pin0, int pin1, double* fe, double* fn, double* ft, double* dg,
                 hipStream_t s);
// coarse faces: half the sum of the four fine faces covering each coarse face
void aggregate(const double* fe, const double* fn, const double* ft, Dims df, double* ce, double* cn, double* ct,
               Dims dc, hipStream_t s);
// Neumann diagonal (-sum of the six faces), plus ex[i] when ex is not null
void diag_from_faces(const double* fe, const double* fn, const double* ft, Dims d, const double* ex, double* dg,
                     hipStream_t s);
// ex = dg + (sum of the six faces): the diagonal beyond the couplings (a mass term)
void excess(const double* fe, const double* fn, const double* ft, const double* dg, Dims d, double* ex, hipStream_t s);
// coarse excess: sum of the fine excess over the children of each coarse cell
void aggregate_cells(const double* ef, Dims df, double* ec, Dims dc, hipStream_t s);
// sigma[q] = sum of the faces of pinned cell q (q = 0, 1) on the finest level
void pin_sigma(const double* fe, const double* fn, const double* ft, Dims d, int pin0, int pin1, double* sigma,
               hipStream_t s);
// dg[c0] -= sigma[0]; dg[c1] -= sigma[1]  (Dirichlet sinks of the pins on a coarse level)
void pin_sink(double* dg, int c0, int c1, const double* sigma, hipStream_t s);
void face_absmax(const double* fe, const double* fn, const double* ft, long n, double* partial, double* out,
                 hipStream_t s);
template <class T>
void inv_diag(const double* dg, long n, T* sv, hipStream_t s);
template <class FT>
void store_faces(const double* fe, const double* fn, const double* ft, long n, double inv_scale, FT* ae, FT* an,
                 FT* at, hipStream_t s);
// coarsest level as the 27-plane FP64 layout of the dense LU kernels
void to27(const double* fe, const double* fn, const double* ft, const double* dg, Dims d, int pin0, int pin1,
          double* c27, hipStream_t s);

// ---- cycle (T = double with FACE_F64 levels, T = float with FACE_F32 / FACE_F16 levels) -------
// one colour of a red-black Gauss-Seidel sweep (zero_guess: 1 x = 0 before this half-sweep,
// 2 first update of this colour after such a start, 0 otherwise)
template <class T>
void rbgs(const Level& L, T* x, const T* b, int color, int zero_guess, hipStream_t s);
// bc = sum over the 8 children of (b - A x)
template <class T>
void resid_restrict(const Level& L, const T* x, const T* b, T* bc, hipStream_t s);
// x_f += P x_c (trilinear)
template <class T>
void prolong_add(const T* xc, Dims dc, T* xf, Dims df, hipStream_t s);
template <class A, class B>
void convert(const A* a, B* b, long n, hipStream_t s);
// the small levels of the cycle in one launch: C.b[0] holds the right-hand side of the first of
// them, C.x[0] receives its correction (lu_*: the dense LU factors of the last level)
template <class T>
void coarse_cycle(const CoarseLevels<T>& C, int n_pre, int n_post, const double* lu_scale, const int* lu_perm,
                  const double* L, const double* U, const double* dinv, hipStream_t s);

// Coarsest level of a capped hierarchy (DENSE_MAX unknowns at most) as a dense FP64 inverse,
// formed at setup by Gauss-Jordan elimination with partial pivoting on the scaled operator:
// inv_c column-major (inv_c[j n + i] = inv(i, j)), inv_r row-major; work holds 2 n^2 + 2 n + 1
// doubles.  dense_apply: x = inv b with inv = inv_c (or inv_r for x = inv^T b); part holds
// n * ceil(n / 128) doubles.  The single-workgroup coarse cycle applies inv_c itself when the
// last level has at most COARSE_DENSE cells.
constexpr int DENSE_MAX = 4096;
constexpr int COARSE_DENSE = 512;
void dense_inverse(const double* fe, const double* fn, const double* ft, const double* dg, Dims d, double* inv_c,
                   double* inv_r, double* work, hipStream_t s);
template <class T>
void dense_apply(const double* inv, int n, const T* b, T* x, double* part, hipStream_t s);

// transposed cycle: one colour of a half-sweep on A^T (the finest level's pinned rows are full
// rows of A^T, other rows drop their couplings to them); r = b - A^T x on the whole level;
// bc = P^T r (trilinear weights); x_f += R^T x_c (each child gets its parent's value)
template <class T>
void rbgs_tr(const Level& L, T* x, const T* b, int color, int zero_guess, hipStream_t s);
template <class T>
void resid_tr(const Level& L, const T* x, const T* b, T* r, hipStream_t s);
template <class T>
void restrict_pt(const T* r, Dims df, T* bc, hipStream_t s);
template <class T>
void prolong_inject_add(const T* xc, Dims dc, T* xf, Dims df, hipStream_t s);

template <>
void rbgs_tr<double>(const Level& L, double* x, const double* b, int color, int zero_guess, hipStream_t s);
template <>
void rbgs_tr<float>(const Level& L, float* x, const float* b, int color, int zero_guess, hipStream_t s);
template <>
void resid_tr<double>(const Level& L, const double* x, const double* b, double* r, hipStream_t s);
template <>
void resid_tr<float>(const Level& L, const float* x, const float* b, float* r, hipStream_t s);

// line relaxation: every line of direction dir (0 x, 1 y, 2 z) solved exactly against x_in's
// values across lines (x_in unread when zero); trans: the same on A^T (the finest level's pinned
// rows are full rows, other rows drop their couplings to them)
template <class T>
void line_pass(const Level& L, int dir, const T* x_in, T* x_out, const T* b, int zero, int trans, hipStream_t s);
template <>
void line_pass<double>(const Level& L, int dir, const double* x_in, double* x_out, const double* b,
                       int zero, int trans, hipStream_t s);
template <>
void line_pass<float>(const Level& L, int dir, const float* x_in, float* x_out, const float* b,
                      int zero, int trans, hipStream_t s);
template <class T>
void copy_vec(const T* a, T* b, long n, hipStream_t s);
// one damped Jacobi sweep x_out = x_in + omega D^-1 (b - A x_in) (zero: x_in taken as 0, not
// read; trans: the same with A^T); empty rows keep their value
template <class T>
void jacobi(const Level& L, const T* x_in, T* x_out, const T* b, T omega, int zero, int trans, hipStream_t s);
template <>
void jacobi<double>(const Level& L, const double* x_in, double* x_out, const double* b, double omega, int zero,
                    int trans, hipStream_t s);
template <>
void jacobi<float>(const Level& L, const float* x_in, float* x_out, const float* b, float omega, int zero, int trans,
                   hipStream_t s);
// one full red-black sweep x_in -> x_out in a single pass (zero: x_in taken as 0, not read);
// bit-identical to the red and black half-sweeps in place
template <class T>
void rb_sweep(const Level& L, const T* x_in, T* x_out, const T* b, int zero, hipStream_t s);
template <>
void rb_sweep<double>(const Level& L, const double* x_in, double* x_out, const double* b, int zero, hipStream_t s);
template <>
void rb_sweep<float>(const Level& L, const float* x_in, float* x_out, const float* b, int zero, hipStream_t s);
// x = 0 on the empty rows of level L (whose 1/a_p, of type T, is stored as 0)
template <class T>
void zero_empty(const Level& L, double* x, hipStream_t s);

template <>
void rbgs<double>(const Level& L, double* x, const double* b, int color, int zero_guess, hipStream_t s);
template <>
void rbgs<float>(const Level& L, float* x, const float* b, int color, int zero_guess, hipStream_t s);
template <>
void resid_restrict<double>(const Level& L, const double* x, const double* b, double* bc, hipStream_t s);
template <>
void resid_restrict<float>(const Level& L, const float* x, const float* b, float* bc, hipStream_t s);

}  // namespace vc
}  // namespace lsx
