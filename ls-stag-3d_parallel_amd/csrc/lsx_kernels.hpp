// Host-side launchers of the HIP kernels (lsx_kernels.hip).  All launches are asynchronous
// on the given stream; sizes are checked by the callers in lsx_solver.cpp.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lsx {

struct Dims {
  int N, M, K;
  __host__ __device__ long n() const { return (long)N * M * K; }
};

// Operator storage: npts SoA coefficient planes of n values each.
//  npts == 7 : [b, s, w, p, e, n, t]                       (fine level, reference row order)
//  npts == 27: index (dx+1) + 3(dy+1) + 9(dz+1)            (Galerkin coarse levels)
enum StencilMode { ST_PROD = 0, ST_RESID = 1, ST_SUBR = 2 };

// Identity-like rows of a symmetric fine operator (the pinned cells), at most CAP.
struct SymExc {
  static constexpr int CAP = 8;
  int count = 0;
  int row[CAP] = {-1, -1, -1, -1, -1, -1, -1, -1};
};

// symmetric fine storage: planes [a_p, a_e, a_n, a_t]; exc[i] = 1 for identity-like rows;
// asym[0] set when the operator is not symmetric off those rows
void sym_check(const double* coef7, Dims d, unsigned char* exc, int* asym, hipStream_t s);
void sym_build(const double* coef7, const unsigned char* exc, Dims d, double* s4, hipStream_t s);
void stencil_apply_sym(const double* s4, const SymExc& ex, Dims d, const double* x, const double* b, double* y,
                       int mode, double* partial, hipStream_t s);

// the same with the transposed operator A^T (identity-like rows become full rows of A^T)
void stencil_apply_sym_t(const double* s4, const SymExc& ex, Dims d, const double* x, const double* b, double* y,
                         int mode, double* partial, hipStream_t s);

// Deterministic reductions: partial sums per block into `partial`, then finalize sums them in
// a fixed order into out[out_index] (device memory).
constexpr int RED_THREADS = 256;
constexpr int CGS_MAX = 32;
int reduction_blocks(long n);

// ---- setup --------------------------------------------------------------------------------
// CSR (device) -> 7 stencil planes; err_flag gets bit 1 if a column is not a 7-point neighbour of
// its row, bit 2 if a row names a column twice.
void csr_to_stencil7(const int* rowptr, const int* col, const double* val, Dims d, double* coef7, int* err_flag,
                     hipStream_t s);
// Galerkin coarse operator A_c = R A_f P (even-even-even transfers), 27 planes on the coarse grid.
void galerkin(const double* coef_f, int npts_f, Dims df, double* coef27_c, Dims dc, hipStream_t s);
// max |a_i| of an array into out[0]
void absmax(const double* a, long n, double* partial, double* out, hipStream_t s);
// max |diag| of an operator into out[0]
void diag_absmax(const double* coef, int npts, Dims d, double* partial, double* out, hipStream_t s);
// Thomas LU factors of the line systems of direction dir (0 x, 1 y, 2 z): dinv, eu (scaled upper)
void line_factor(const double* coef, int npts, Dims d, int dir, const double* absmax, double* dinv, double* eu,
                 hipStream_t s);
// Dense LU of the coarsest operator (single thread): scale, perm, L, U (n x n row-major)
void lu_factor(const double* coef27, Dims d, double* scale, int* perm, double* L, double* U, double* col_scratch,
               hipStream_t s);

// ---- solve phase ------------------------------------------------------------------------------
// y = A x (ST_PROD), y = b - A x with partial sums of y^2 (ST_RESID), y -= A x (ST_SUBR)
void stencil_apply(const double* coef, int npts, Dims d, const double* x, const double* b, double* y, int mode,
                   double* partial, hipStream_t s);
void restrict_full(const double* rf, Dims df, double* rc, Dims dc, hipStream_t s);
void prolong_copy(const double* chc, Dims dc, double* chf, double* cf, Dims df, hipStream_t s);
void line_solve(const double* coef, int npts, Dims d, int dir, const double* dinv, const double* eu, const double* r,
                double* c, double* ch, double omega, int z_head_fix, hipStream_t s);
void lu_solve(const double* scale, const int* perm, const double* L, const double* U, int n, const double* b,
              double* x, hipStream_t s);

// ---- BLAS-1 (FP64) ------------------------------------------------------------------------
void copy(const double* a, double* b, long n, hipStream_t s);
void fill_zero(double* b, long n, hipStream_t s);  // a kernel, not a memset: safe under graph capture
void scale_copy(const double* w, const double* normw, double* v, long n, hipStream_t s);  // v = (1/normw) w
// v = (1/normw) w and vf = (float)v in one pass
void scale_copy_f32(const double* w, const double* normw, double* v, float* vf, long n, hipStream_t s);
void dot_partial(const double* a, const double* b, long n, double* partial, hipStream_t s);
void finalize(const double* partial, int nb, double* out, int sqrt_flag, hipStream_t s);
// w -= h*v ; then partial dot(w, vnext) (vnext may be null => partial sum of w^2)
void axpy_dot(double* w, const double* v, const double* h, const double* vnext, long n, double* partial,
              hipStream_t s);
// x += sum_k y_k z_k in the order k = last..0 (coefficients from host array, up to 64)
void sol_update(double* x, const double* Z, long ldz, const double* yk, int count, long n, hipStream_t s);
// classical Gram-Schmidt passes over cnt <= CGS_MAX basis vectors (vector k at V + k*ldv):
// cgs_dot: out[k] = v_k . w (k < cnt);  cgs_update: w -= sum_k h_k v_k, and norm_out[0] = ||w||
// when norm_out is given.  Row chunks of cgs_blocks(n) blocks (at most CGS_BLOCKS_MAX); partial
// needs CGS_MAX * cgs_blocks(n) doubles.
constexpr int CGS_BLOCKS_MAX = 4096;
int cgs_blocks(long n);
void cgs_dot(const double* V, long ldv, int cnt, const double* w, long n, double* partial, double* out,
             hipStream_t s);
// cgs_dot over FP32 vectors (each widened to FP64 before its product, sums in FP64)
void cgs_dot_f32(const float* V, long ldv, int cnt, const double* w, long n, double* partial, double* out,
                 hipStream_t s);
void cgs_update(const double* V, long ldv, int cnt, const double* h, double* w, long n, double* partial,
                double* norm_out, hipStream_t s);
// the update for a basis stored unnormalised (vector k = s_k times the orthonormal one):
// out = post (w - sum_k (h_k isq_k) v_k) with isq_k = 1 / s_k^2 (out may be w), its FP32 copy in
// out32 when given, and norm_out[0] = ||out|| when given
struct CgsScales {
  double isq[CGS_MAX];
};
void cgs_update_scaled(const double* V, long ldv, int cnt, const double* h, const CgsScales& sc, double post,
                       const double* w, double* out, float* out32, long n, double* partial, double* norm_out,
                       hipStream_t s);
// mixed precision: y = A x with FP32 x (fine 7-point operator); x += sum_k y_k z_k with FP32 z
void stencil7_prod_f32x(const double* coef, Dims d, const float* x, double* y, hipStream_t s);
void stencil_prod_sym_f32x(const double* s4, const SymExc& ex, Dims d, const float* x, double* y, hipStream_t s);
void stencil_prod_sym_t_f32x(const double* s4, const SymExc& ex, Dims d, const float* x, double* y, hipStream_t s);
void sol_update_f32(double* x, const float* Z, long ldz, const double* yk, int count, long n, hipStream_t s);
// BiCGStab fused updates (bicgstab_interface.h:39-56)
void bicg_s_update(double* nu, const double* r, double* sv, double alpha, long n, hipStream_t s);
void bicg_sol_update(const double* sv, double* y, double* z, double* x, double* t, double* r, double alpha, double w,
                     long n, hipStream_t s);
void bicg_p_update(double* p, const double* nu, const double* r, double coef, double beta, long n, double* partial,
                   hipStream_t s);
void sub_vec(const double* b, const double* z, double* r, long n, hipStream_t s);  // r = b - z

}  // namespace lsx
