// FP64 kernels of the pressure solve (reference-semantics path).
//
// Built with -ffp-contract=off: every product and sum is rounded separately, in the order the
// reference writes it, so the GPU path and a scalar CPU evaluation of the same formulas agree
// to the last bit wherever the summation order is the same (stencil rows, transfers, line
// solves, Galerkin products).  Reductions are two-stage and deterministic (fixed partition,
// fixed combination order), but their order differs from a sequential sum.
//
// Index convention (reference ls_stag_uniform_mesh.h:174-185): num = x + N*y + N*M*z.

#include "lsx_kernels.hpp"

#include <cmath>

namespace lsx {

namespace {

constexpr double PLUS_TINY = 1e-16;        // ls_stag_real_constants.h:27
constexpr double EQ_DOUBLE = 1e-5;         // :28
constexpr double EQ_TINY2_DOUBLE = 1e-5;   // :30
constexpr double EQ_TINY_DOUBLE = 1e-12;   // :31
constexpr double EQ_TINY3_DOUBLE = 1e-13;  // :32

inline int div_up(long a, long b) { return (int)((a + b - 1) / b); }

// ---- deterministic block reduction ------------------------------------------------------------
template <int BS>
__device__ double block_sum(double v) {
  __shared__ double lds[BS / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < BS / 64; ++w) s += lds[w];
  __syncthreads();
  return s;  // valid in thread 0
}

template <int BS>
__device__ double block_max(double v) {
  __shared__ double lds[BS / 64];
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < BS / 64; ++w) s = fmax(s, lds[w]);
  __syncthreads();
  return s;
}

// contiguous chunk of [0, n) owned by block b of nb
__device__ inline void chunk_of(long n, int b, int nb, long* lo, long* hi) {
  const long per = (n + nb - 1) / nb;
  *lo = per * b;
  long h = *lo + per;
  *hi = h < n ? h : n;
}

// ---- stencil geometry helpers -------------------------------------------------------------
// One thread per cell for the plain stencil products: 256-thread blocks of 64 x by 4 y cells,
// blocks x fastest, then y, then z (tile_blocks of them).  The cell of this thread, false past the
// grid's edge.  No per-cell division, and each row's loads in flight together (measured 139 us
// against 156 us for 4096-cell chunks walked by 256 threads at 256^3).
__device__ inline bool tile_cell(int N, int M, int* x, int* y, int* z) {
  const int gx = (N + 63) / 64, gy = (M + 3) / 4;
  const int bx = (int)(blockIdx.x % gx), r = (int)(blockIdx.x / gx);
  *x = bx * 64 + (int)(threadIdx.x & 63);
  *y = (r % gy) * 4 + (int)(threadIdx.x >> 6);
  *z = r / gy;
  return *x < N && *y < M;
}

__device__ inline void decode(int i, int N, int M, int* x, int* y, int* z) {
  const int NM = N * M;
  *z = i / NM;
  const int r = i - *z * NM;
  *y = r / N;
  *x = r - *y * N;
}

__device__ inline int slot_sub(int npts, int dir) {  // coupling to the previous cell of a line
  if (npts == 7) return dir == 0 ? 2 : (dir == 1 ? 1 : 0);
  return dir == 0 ? 12 : (dir == 1 ? 10 : 4);
}
__device__ inline int slot_super(int npts, int dir) {  // coupling to the next cell of a line
  if (npts == 7) return dir == 0 ? 4 : (dir == 1 ? 5 : 6);
  return dir == 0 ? 14 : (dir == 1 ? 16 : 22);
}
__device__ inline int slot_diag(int npts) { return npts == 7 ? 3 : 13; }

// Row i of the operator applied to x, accumulated in ascending column order (the CSR order
// of the reference matrices: [b,s,w,p,e,n,t] on the fine level, sorted columns on Galerkin
// levels).  mode: 0 -> acc starts at 0 and adds; 1 -> acc starts at b and subtracts.
template <int NPTS>
__device__ inline double row_apply(const double* __restrict__ coef, long n, int i, int x, int y, int z, int N, int M,
                                   int K, const double* __restrict__ xv, double acc, bool subtract) {
  const int NM = N * M;
  if (NPTS == 7) {
    const int off[7] = {-NM, -N, -1, 0, 1, N, NM};
    const bool ok[7] = {z > 0, y > 0, x > 0, true, x < N - 1, y < M - 1, z < K - 1};
#pragma unroll
    for (int q = 0; q < 7; ++q)
      if (ok[q]) {
        const double t = coef[q * n + i] * xv[i + off[q]];
        acc = subtract ? acc - t : acc + t;
      }
  } else {
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
          const int zz = z + dz, yy = y + dy, xx = x + dx;
          if (zz < 0 || zz >= K || yy < 0 || yy >= M || xx < 0 || xx >= N) continue;
          const int q = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
          const double t = coef[q * n + i] * xv[i + dx + N * dy + NM * dz];
          acc = subtract ? acc - t : acc + t;
        }
  }
  return acc;
}

// ---- kernels: setup ----------------------------------------------------------------------------
__global__ void k_csr_to_stencil7(const int* __restrict__ rowptr, const int* __restrict__ col,
                                  const double* __restrict__ val, int N, int M, int K, double* __restrict__ coef,
                                  int* err) {
  const long n = (long)N * M * K;
  const int NM = N * M;
  for (long li = blockIdx.x * (long)blockDim.x + threadIdx.x; li < n; li += (long)gridDim.x * blockDim.x) {
    const int i = (int)li;
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    double a[7] = {0, 0, 0, 0, 0, 0, 0};
    unsigned seen = 0;
    for (int j = rowptr[i]; j < rowptr[i + 1]; ++j) {
      const int off = col[j] - i;
      int q = -1;
      if (off == -NM && z > 0) q = 0;
      else if (off == -N && y > 0) q = 1;
      else if (off == -1 && x > 0) q = 2;
      else if (off == 0) q = 3;
      else if (off == 1 && x < N - 1) q = 4;
      else if (off == N && y < M - 1) q = 5;
      else if (off == NM && z < K - 1) q = 6;
      if (q < 0) {
        atomicOr(err, 1);
      } else if (seen & (1u << q)) {  // a repeated column would drop a term
        atomicOr(err, 2);
      } else {
        seen |= 1u << q;
        a[q] = val[j];
      }
    }
#pragma unroll
    for (int q = 0; q < 7; ++q) coef[q * n + i] = a[q];
  }
}

// Galerkin product A_c = R (A_f P): coarse_grid.h:103-104 with product_sparse_matrix.h:72-104.
// Terms reach each entry in the reference order: V = R A_f accumulates over the 8 children in
// R's column order, then A_c accumulates V's fine columns in ascending order.
__device__ inline void axis_w(int f, int nf, int* c, double* w, int* cnt) {
  const int p = f >> 1;
  if (f == 0 || f == nf - 1) {
    c[0] = p; w[0] = 1.0; *cnt = 1;
  } else if (f & 1) {
    c[0] = p; w[0] = 0.75; c[1] = p + 1; w[1] = 0.25; *cnt = 2;
  } else {
    c[0] = p - 1; w[0] = 0.25; c[1] = p; w[1] = 0.75; *cnt = 2;
  }
}

template <int NPTS>
__global__ void k_galerkin(const double* __restrict__ cf, int Nf, int Mf, int Kf, double* __restrict__ cc, int Nc,
                           int Mc, int Kc) {
  const long nc = (long)Nc * Mc * Kc, nf = (long)Nf * Mf * Kf;
  const int NMf = Nf * Mf;
  for (long li = blockIdx.x * (long)blockDim.x + threadIdx.x; li < nc; li += (long)gridDim.x * blockDim.x) {
    const int I = (int)li;
    int X, Y, Z;
    decode(I, Nc, Mc, &X, &Y, &Z);
    double V[64];
#pragma unroll
    for (int q = 0; q < 64; ++q) V[q] = 0.0;
    for (int ch = 0; ch < 8; ++ch) {
      const int hx = 2 * X + (ch & 1), hy = 2 * Y + ((ch >> 1) & 1), hz = 2 * Z + (ch >> 2);
      const int h = hx + Nf * hy + NMf * hz;
      if (NPTS == 7) {
        const int dxs[7] = {0, 0, -1, 0, 1, 0, 0}, dys[7] = {0, -1, 0, 0, 0, 1, 0}, dzs[7] = {-1, 0, 0, 0, 0, 0, 1};
        for (int q = 0; q < 7; ++q) {
          const int fx = hx + dxs[q], fy = hy + dys[q], fz = hz + dzs[q];
          if (fx < 0 || fx >= Nf || fy < 0 || fy >= Mf || fz < 0 || fz >= Kf) continue;
          const int b = (fx - 2 * X + 1) + 4 * (fy - 2 * Y + 1) + 16 * (fz - 2 * Z + 1);
          V[b] = V[b] + 0.125 * cf[q * nf + h];
        }
      } else {
        for (int q = 0; q < 27; ++q) {
          const int dx = q % 3 - 1, dy = (q / 3) % 3 - 1, dz = q / 9 - 1;
          const int fx = hx + dx, fy = hy + dy, fz = hz + dz;
          if (fx < 0 || fx >= Nf || fy < 0 || fy >= Mf || fz < 0 || fz >= Kf) continue;
          const int b = (fx - 2 * X + 1) + 4 * (fy - 2 * Y + 1) + 16 * (fz - 2 * Z + 1);
          V[b] = V[b] + 0.125 * cf[q * nf + h];
        }
      }
    }
    double A[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) A[q] = 0.0;
    for (int b = 0; b < 64; ++b) {
      const int fx = 2 * X - 1 + (b & 3), fy = 2 * Y - 1 + ((b >> 2) & 3), fz = 2 * Z - 1 + (b >> 4);
      if (fx < 0 || fx >= Nf || fy < 0 || fy >= Mf || fz < 0 || fz >= Kf) continue;
      int cx[2], cy[2], cz[2], nx, ny, nz;
      double wx[2], wy[2], wz[2];
      axis_w(fx, Nf, cx, wx, &nx);
      axis_w(fy, Mf, cy, wy, &ny);
      axis_w(fz, Kf, cz, wz, &nz);
      const double v = V[b];
      for (int c3 = 0; c3 < nz; ++c3)
        for (int c2 = 0; c2 < ny; ++c2)
          for (int c1 = 0; c1 < nx; ++c1) {
            const double w = wx[c1] * wy[c2] * wz[c3];
            const int q = (cx[c1] - X + 1) + 3 * (cy[c2] - Y + 1) + 9 * (cz[c3] - Z + 1);
            A[q] = A[q] + v * w;
          }
    }
#pragma unroll
    for (int q = 0; q < 27; ++q) cc[q * nc + I] = A[q];
  }
}

// ---- odd-odd-odd transfers (transfer_operators_odd_odd_odd.h:44-100, R = P^T by
// transfer_operators_interface.h:572-578 and transposed_sparse_matrix.h:47-68) --------------------
// An odd axis of 2n - 1 fine cells coarsens to n cells centred on the even fine cells: P weights
// 1 on a coincident fine cell, 1/2 on each coarse neighbour of an odd fine cell; a row's entries
// run z, then y, then x, coarse index ascending.  R row C lists fine cells 2C - 1, 2C, 2C + 1 per
// axis (weights 1/2, 1, 1/2, those inside the level) in ascending fine index.
__device__ inline void axis_w_odd(int f, int* c, double* w, int* cnt) {
  if ((f & 1) == 0) {
    c[0] = f >> 1; w[0] = 1.0; *cnt = 1;
  } else {
    c[0] = (f - 1) >> 1; w[0] = 0.5; c[1] = (f + 1) >> 1; w[1] = 0.5; *cnt = 2;
  }
}
__device__ inline int axis_r_odd(int C, int nf, int* f, double* w) {
  int n = 0;
  if (2 * C - 1 >= 0) { f[n] = 2 * C - 1; w[n] = 0.5; ++n; }
  f[n] = 2 * C; w[n] = 1.0; ++n;
  if (2 * C + 1 < nf) { f[n] = 2 * C + 1; w[n] = 0.5; ++n; }
  return n;
}

// Galerkin product A_c = (R A_f) P for odd-odd-odd levels, terms in the order of
// product_sparse_matrix.h:72-104: V = R A_f accumulates over R's fine cells in ascending order,
// each with its A_f row in ascending column order; A_c then accumulates V's fine columns in
// ascending order (V's window: fine cells 2C - 2 .. 2C + 2 per axis).
template <int NPTS>
__global__ void k_galerkin_odd(const double* __restrict__ cf, int Nf, int Mf, int Kf, double* __restrict__ cc,
                               int Nc, int Mc, int Kc) {
  const long nc = (long)Nc * Mc * Kc, nf = (long)Nf * Mf * Kf;
  const int NMf = Nf * Mf;
  for (long li = blockIdx.x * (long)blockDim.x + threadIdx.x; li < nc; li += (long)gridDim.x * blockDim.x) {
    const int I = (int)li;
    int X, Y, Z;
    decode(I, Nc, Mc, &X, &Y, &Z);
    double V[125];
    for (int q = 0; q < 125; ++q) V[q] = 0.0;
    int rx[3], ry[3], rz[3];
    double wx[3], wy[3], wz[3];
    const int nx = axis_r_odd(X, Nf, rx, wx), ny = axis_r_odd(Y, Mf, ry, wy), nz = axis_r_odd(Z, Kf, rz, wz);
    for (int c3 = 0; c3 < nz; ++c3)
      for (int c2 = 0; c2 < ny; ++c2)
        for (int c1 = 0; c1 < nx; ++c1) {
          const double w = wx[c1] * wy[c2] * wz[c3];
          const int hx = rx[c1], hy = ry[c2], hz = rz[c3];
          const int h = hx + Nf * hy + NMf * hz;
          for (int q = 0; q < NPTS; ++q) {
            int dx, dy, dz;
            if (NPTS == 7) {
              const int dxs[7] = {0, 0, -1, 0, 1, 0, 0}, dys[7] = {0, -1, 0, 0, 0, 1, 0}, dzs[7] = {-1, 0, 0, 0, 0, 0, 1};
              dx = dxs[q];
              dy = dys[q];
              dz = dzs[q];
            } else {
              dx = q % 3 - 1;
              dy = (q / 3) % 3 - 1;
              dz = q / 9 - 1;
            }
            const int fx = hx + dx, fy = hy + dy, fz = hz + dz;
            if (fx < 0 || fx >= Nf || fy < 0 || fy >= Mf || fz < 0 || fz >= Kf) continue;
            const int b = (fx - 2 * X + 2) + 5 * (fy - 2 * Y + 2) + 25 * (fz - 2 * Z + 2);
            V[b] = V[b] + w * cf[q * nf + h];
          }
        }
    double A[27];
    for (int q = 0; q < 27; ++q) A[q] = 0.0;
    for (int b = 0; b < 125; ++b) {
      const int fx = 2 * X - 2 + b % 5, fy = 2 * Y - 2 + (b / 5) % 5, fz = 2 * Z - 2 + b / 25;
      if (fx < 0 || fx >= Nf || fy < 0 || fy >= Mf || fz < 0 || fz >= Kf) continue;
      int cx[2], cy[2], cz[2], mx, my, mz;
      double px[2], py[2], pz[2];
      axis_w_odd(fx, cx, px, &mx);
      axis_w_odd(fy, cy, py, &my);
      axis_w_odd(fz, cz, pz, &mz);
      const double v = V[b];
      for (int c3 = 0; c3 < mz; ++c3)
        for (int c2 = 0; c2 < my; ++c2)
          for (int c1 = 0; c1 < mx; ++c1) {
            const double w = px[c1] * py[c2] * pz[c3];
            const int q = (cx[c1] - X + 1) + 3 * (cy[c2] - Y + 1) + 9 * (cz[c3] - Z + 1);
            A[q] = A[q] + v * w;
          }
    }
    for (int q = 0; q < 27; ++q) cc[q * nc + I] = A[q];
  }
}

// r_c = R r_f on odd-odd-odd levels (R = P^T, fine cells in ascending order)
__global__ void k_restrict_odd(const double* __restrict__ rf, int Nf, int Mf, int Kf, double* __restrict__ rc, int Nc,
                               int Mc, int Kc) {
  const long nc = (long)Nc * Mc * Kc;
  const int NMf = Nf * Mf;
  for (long li = blockIdx.x * (long)blockDim.x + threadIdx.x; li < nc; li += (long)gridDim.x * blockDim.x) {
    int X, Y, Z;
    decode((int)li, Nc, Mc, &X, &Y, &Z);
    int rx[3], ry[3], rz[3];
    double wx[3], wy[3], wz[3];
    const int nx = axis_r_odd(X, Nf, rx, wx), ny = axis_r_odd(Y, Mf, ry, wy), nz = axis_r_odd(Z, Kf, rz, wz);
    double acc = 0.0;
    for (int c3 = 0; c3 < nz; ++c3)
      for (int c2 = 0; c2 < ny; ++c2)
        for (int c1 = 0; c1 < nx; ++c1) {
          const double w = wx[c1] * wy[c2] * wz[c3];
          acc = acc + w * rf[rx[c1] + Nf * ry[c2] + NMf * rz[c3]];
        }
    rc[li] = acc;
  }
}

// ch_f = P ch_c, c_f = ch_f on odd-odd-odd levels
__global__ void k_prolong_odd(const double* __restrict__ chc, int Nc, int Mc, double* __restrict__ chf,
                              double* __restrict__ cf, int Nf, int Mf, int Kf) {
  const long nf = (long)Nf * Mf * Kf;
  const int NMc = Nc * Mc;
  for (long li = blockIdx.x * (long)blockDim.x + threadIdx.x; li < nf; li += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode((int)li, Nf, Mf, &x, &y, &z);
    int cx[2], cy[2], cz[2], nx, ny, nz;
    double wx[2], wy[2], wz[2];
    axis_w_odd(x, cx, wx, &nx);
    axis_w_odd(y, cy, wy, &ny);
    axis_w_odd(z, cz, wz, &nz);
    double acc = 0.0;
    for (int c3 = 0; c3 < nz; ++c3)
      for (int c2 = 0; c2 < ny; ++c2)
        for (int c1 = 0; c1 < nx; ++c1) {
          const double w = wx[c1] * wy[c2] * wz[c3];
          acc = acc + w * chc[cx[c1] + Nc * cy[c2] + NMc * cz[c3]];
        }
    chf[li] = acc;
    cf[li] = acc;
  }
}

template <int BS>
__global__ void k_diag_absmax_partial(const double* __restrict__ diag, long n, double* partial) {
  long lo, hi;
  chunk_of(n, blockIdx.x, gridDim.x, &lo, &hi);
  double m = 0.0;
  for (long i = lo + threadIdx.x; i < hi; i += BS) m = fmax(m, fabs(diag[i]));
  m = block_max<BS>(m);
  if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

template <int BS>
__global__ void k_max_finalize(const double* partial, int nb, double* out) {
  double m = 0.0;
  for (int i = threadIdx.x; i < nb; i += BS) m = fmax(m, partial[i]);
  m = block_max<BS>(m);
  if (threadIdx.x == 0) out[0] = m;
}

// Line decomposition of direction dir: tridiagonal_matrix.h:45-67 (triLU) applied to the
// lines of smoother_adlj.h:40-86.  One thread per line; the zero-pivot threshold is global
// (max |diag| over the level), pivots below it are set to 0 which decouples the row.
__device__ inline void line_geometry(int t, int dir, int N, int M, int K, int* start, int* stride, int* len) {
  if (dir == 0) {  // lines (y, z)
    *start = t * N;
    *stride = 1;
    *len = N;
  } else if (dir == 1) {  // lines (x, z)
    const int x = t % N, z = t / N;
    *start = x + N * M * z;
    *stride = N;
    *len = M;
  } else {  // lines (x, y)
    *start = t;
    *stride = N * M;
    *len = K;
  }
}

__global__ void k_line_factor(const double* __restrict__ coef, int npts, int N, int M, int K, int dir,
                              const double* absmax, double* __restrict__ dinv, double* __restrict__ eu) {
  const long n = (long)N * M * K;
  const int nlines = dir == 0 ? M * K : (dir == 1 ? N * K : N * M);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlines) return;
  double tiny = EQ_TINY2_DOUBLE * absmax[0];
  if (tiny < EQ_TINY_DOUBLE) tiny = EQ_TINY_DOUBLE;
  int start, stride, len;
  line_geometry(t, dir, N, M, K, &start, &stride, &len);
  const double* dg = coef + slot_diag(npts) * n;
  const double* sb = coef + slot_sub(npts, dir) * n;
  const double* sp = coef + slot_super(npts, dir) * n;
  int i = start;
  double aux = dg[i];
  double dprev = (fabs(aux) < tiny) ? 0.0 : (1.0 / aux);
  dinv[i] = dprev;
  for (int k = 1; k < len; ++k) {
    const int ip = i;
    i += stride;
    const double e = sp[ip] * dprev;
    eu[ip] = e;
    aux = dg[i] - sb[i] * e;
    dprev = (fabs(aux) < tiny) ? 0.0 : (1.0 / aux);
    dinv[i] = dprev;
  }
  eu[i] = 0.0;
}

// Dense LU of the coarsest level: lu_decomposition_sparse_matrix.h:60-215 (initScalePerm,
// calcLU).  A single thread: the coarsest level has 8-16 unknowns.
__device__ inline double dense_at(const double* coef, int n, int N, int M, int K, int i, int j) {
  int xi, yi, zi, xj, yj, zj;
  decode(i, N, M, &xi, &yi, &zi);
  decode(j, N, M, &xj, &yj, &zj);
  const int dx = xj - xi, dy = yj - yi, dz = zj - zi;
  if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return 0.0;
  return coef[((dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)) * (long)n + i];
}

__global__ void k_lu_factor(const double* __restrict__ coef, int N, int M, int K, double* scale, int* perm, double* L,
                            double* U, double* colv) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int n = N * M * K;
  for (int i = 0; i < n; ++i) {
    const double diag = fabs(dense_at(coef, n, N, M, K, i, i));
    scale[i] = (diag > EQ_DOUBLE) ? (1.0 / sqrt(diag)) : 1.0;
    perm[i] = i;
  }
  for (int q = 0; q < n * n; ++q) {
    L[q] = 0.0;
    U[q] = 0.0;
  }
  double maxAbs = 0.0, maxVal = 0.0;
  int pivot = 0;
  double si = scale[0];
  for (int i = 0; i < n; ++i) {
    const double a0 = dense_at(coef, n, N, M, K, i, 0);
    const double absVal = fabs(a0 * scale[i] * si);
    if (absVal > maxVal) {
      maxVal = absVal;
      pivot = i;
    }
    for (int j = 0; j < n; ++j) maxAbs = fmax(maxAbs, fabs(dense_at(coef, n, N, M, K, i, j)));
  }
  const double tiny = EQ_TINY3_DOUBLE * maxAbs;
  if (pivot != 0) {
    const int tmp = perm[0];
    perm[0] = perm[pivot];
    perm[pivot] = tmp;
  }
  int permi = perm[0];
  si = scale[permi];
  for (int j = 0; j < n; ++j) U[j] = dense_at(coef, n, N, M, K, permi, j) * si * scale[j];
  double U00 = U[0];
  if (fabs(U00) < tiny) U00 = (U00 < 0) ? -tiny : tiny;
  U[0] = U00;
  const double invU00 = scale[0] / U00;
  for (int k = 1; k < n; ++k) {
    const int pk = perm[k];
    L[k * n] = dense_at(coef, n, N, M, K, pk, 0) * scale[pk] * invU00;
  }
  L[0] = 1.0;
  for (int i = 1; i < n; ++i) {
    maxVal = 0.0;
    pivot = i;
    si = scale[i];
    for (int k = i; k < n; ++k) {
      double sum = 0.0;
      for (int t = 0; t < i; ++t) sum += L[k * n + t] * U[t * n + i];
      const int pk = perm[k];
      const double val = dense_at(coef, n, N, M, K, pk, i) * scale[pk] * si - sum;
      colv[k] = val;
      if (fabs(val) > maxVal) {
        maxVal = fabs(val);
        pivot = k;
      }
    }
    if (pivot != i) {
      int tmp = perm[i];
      perm[i] = perm[pivot];
      perm[pivot] = tmp;
      for (int t = 0; t < i; ++t) {
        const double d = L[i * n + t];
        L[i * n + t] = L[pivot * n + t];
        L[pivot * n + t] = d;
      }
      const double v = colv[i];
      colv[i] = colv[pivot];
      colv[pivot] = v;
    }
    permi = perm[i];
    si = scale[permi];
    for (int j = i; j < n; ++j) {
      double sum = 0.0;
      for (int t = 0; t < i; ++t) sum += L[i * n + t] * U[t * n + j];
      U[i * n + j] = dense_at(coef, n, N, M, K, permi, j) * si * scale[j] - sum;
    }
    double Uii = U[i * n + i];
    if (fabs(Uii) < tiny) Uii = (Uii < 0) ? -tiny : tiny;
    U[i * n + i] = Uii;
    const double invUii = 1.0 / Uii;
    for (int k = i + 1; k < n; ++k) L[k * n + i] = colv[k] * invUii;
    L[i * n + i] = 1.0;
  }
}

// ---- symmetric storage of the fine operator ----------------------------------------------------
// The reference pressure operator is symmetric except its identity (pinned) rows, whose
// neighbours still couple to them.  A row is an exception when its own off-diagonals are all 0
// while a neighbour's coupling to it is not; each face then takes the coupled side's value.
__device__ inline bool row_decoupled(const double* c7, long n, long i) {
  return c7[0 * n + i] == 0.0 && c7[1 * n + i] == 0.0 && c7[2 * n + i] == 0.0 && c7[4 * n + i] == 0.0 &&
         c7[5 * n + i] == 0.0 && c7[6 * n + i] == 0.0;
}

__global__ void k_sym_check(const double* __restrict__ c7, int N, int M, int K, unsigned char* __restrict__ exc,
                            int* __restrict__ asym) {
  const long n = (long)N * M * K, NM = (long)N * M;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode((int)i, N, M, &x, &y, &z);
    unsigned char e = 0;
    const bool di = row_decoupled(c7, n, i);
    // faces towards lower neighbours: (this row's w/s/b) versus (neighbour's e/n/t)
    const long nb[3] = {i - 1, i - N, i - NM};
    const bool has[3] = {x > 0, y > 0, z > 0};
    const int lo_plane[3] = {2, 1, 0}, hi_plane[3] = {4, 5, 6};
    for (int d = 0; d < 3; ++d) {
      if (!has[d]) continue;
      const double mine = c7[lo_plane[d] * n + i], theirs = c7[hi_plane[d] * n + nb[d]];
      if (mine == theirs) continue;
      if (di && mine == 0.0) e = 1;                                   // this row is an identity-like row
      else if (mine != 0.0 && theirs == 0.0 && row_decoupled(c7, n, nb[d])) continue;  // flagged by that row
      else atomicOr(asym, 1);
    }
    const long up[3] = {i + 1, i + N, i + NM};
    const bool hasu[3] = {x < N - 1, y < M - 1, z < K - 1};
    for (int d = 0; d < 3; ++d) {
      if (!hasu[d]) continue;
      const double mine = c7[hi_plane[d] * n + i], theirs = c7[lo_plane[d] * n + up[d]];
      if (mine != theirs && di && mine == 0.0) e = 1;
    }
    exc[i] = e;
  }
}

__global__ void k_sym_build(const double* __restrict__ c7, const unsigned char* __restrict__ exc, int N, int M, int K,
                            double* __restrict__ s4) {
  const long n = (long)N * M * K, NM = (long)N * M;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode((int)i, N, M, &x, &y, &z);
    s4[0 * n + i] = c7[3 * n + i];
    s4[1 * n + i] = x < N - 1 ? (exc[i] ? c7[2 * n + i + 1] : c7[4 * n + i]) : 0.0;
    s4[2 * n + i] = y < M - 1 ? (exc[i] ? c7[1 * n + i + N] : c7[5 * n + i]) : 0.0;
    s4[3 * n + i] = z < K - 1 ? (exc[i] ? c7[0 * n + i + NM] : c7[6 * n + i]) : 0.0;
  }
}

// Row i of the symmetric fine operator, terms in the reference CSR order [b,s,w,p,e,n,t]
// (identical values and order to the 7-plane row, hence bit-identical results).  Every load is
// issued unconditionally (absent neighbours read the cell itself) and absent terms leave the sum
// unchanged, so the seven loads of a row are in flight together.
template <class XT>
__device__ inline double row_apply_sym(const double* __restrict__ s4, long n, int i, int x, int y, int z, int N,
                                       int M, int K, const XT* __restrict__ xv, double acc, bool subtract,
                                       const SymExc& ex) {
  const int NM = N * M;
  bool exc_row = false;
#pragma unroll
  for (int q = 0; q < SymExc::CAP; ++q) exc_row |= (q < ex.count) && (ex.row[q] == i);
  const bool hb = !exc_row && z > 0, hs = !exc_row && y > 0, hw = !exc_row && x > 0;
  const bool he = !exc_row && x < N - 1, hn = !exc_row && y < M - 1, ht = !exc_row && z < K - 1;
  const int ib = z > 0 ? i - NM : i, is = y > 0 ? i - N : i, iw = x > 0 ? i - 1 : i;
  const int ie = x < N - 1 ? i + 1 : i, in = y < M - 1 ? i + N : i, it = z < K - 1 ? i + NM : i;
  const double cb = s4[3 * n + ib], cs = s4[2 * n + is], cw = s4[1 * n + iw], cp = s4[i];
  const double ce = s4[1 * n + i], cn = s4[2 * n + i], ct = s4[3 * n + i];
  const double xb = (double)xv[ib], xs = (double)xv[is], xw = (double)xv[iw], xp = (double)xv[i];
  const double xe = (double)xv[ie], xn = (double)xv[in], xt = (double)xv[it];
  auto add = [&](bool h, double c, double xj) {
    const double t = c * xj;
    const double a = subtract ? acc - t : acc + t;
    acc = h ? a : acc;
  };
  add(hb, cb, xb);
  add(hs, cs, xs);
  add(hw, cw, xw);
  add(true, cp, xp);
  add(he, ce, xe);
  add(hn, cn, xn);
  add(ht, ct, xt);
  return acc;
}

// Row i of the transposed operator: A^T(i, j) = A(j, i).  An identity-like row of A becomes a
// full row of A^T (its own faces, which hold the neighbours' couplings to it), and every other
// row drops its couplings to those cells, whose rows in A do not couple back.
__device__ inline bool is_exc(const SymExc& ex, long j) {
  bool e = false;
#pragma unroll
  for (int q = 0; q < SymExc::CAP; ++q) e |= (q < ex.count) && (ex.row[q] == j);
  return e;
}

template <class XT>
__device__ inline double row_apply_sym_t(const double* __restrict__ s4, long n, int i, int x, int y, int z, int N,
                                         int M, int K, const XT* __restrict__ xv, double acc, bool subtract,
                                         const SymExc& ex) {
  const int NM = N * M;
  const bool exc_row = is_exc(ex, i);
  double t;
  auto term = [&](double c, long j) {
    if (!exc_row && is_exc(ex, j)) return;
    t = c * (double)xv[j];
    acc = subtract ? acc - t : acc + t;
  };
  if (z > 0) term(s4[3 * n + i - NM], i - NM);
  if (y > 0) term(s4[2 * n + i - N], i - N);
  if (x > 0) term(s4[1 * n + i - 1], i - 1);
  t = s4[i] * (double)xv[i];
  acc = subtract ? acc - t : acc + t;
  if (x < N - 1) term(s4[1 * n + i], i + 1);
  if (y < M - 1) term(s4[2 * n + i], i + N);
  if (z < K - 1) term(s4[3 * n + i], i + NM);
  return acc;
}

template <int MODE, int BS>
__global__ void k_stencil_sym_t(const double* __restrict__ s4, int N, int M, int K, SymExc ex,
                                const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ y,
                                double* __restrict__ partial) {
  const long n = (long)N * M * K;
  long lo, hi;
  chunk_of(n, blockIdx.x, gridDim.x, &lo, &hi);
  double ss = 0.0;
  for (long li = lo + threadIdx.x; li < hi; li += BS) {
    const int i = (int)li;
    int xx, yy, zz;
    decode(i, N, M, &xx, &yy, &zz);
    if (MODE == ST_PROD) {
      y[i] = row_apply_sym_t(s4, n, i, xx, yy, zz, N, M, K, x, 0.0, false, ex);
    } else if (MODE == ST_RESID) {
      const double r = row_apply_sym_t(s4, n, i, xx, yy, zz, N, M, K, x, b[i], true, ex);
      y[i] = r;
      ss += r * r;
    } else {
      const double q = row_apply_sym_t(s4, n, i, xx, yy, zz, N, M, K, x, 0.0, false, ex);
      y[i] = y[i] - q;
    }
  }
  if (MODE == ST_RESID) {
    ss = block_sum<BS>(ss);
    if (threadIdx.x == 0) partial[blockIdx.x] = ss;
  }
}

template <int MODE, int BS>
__global__ void k_stencil_sym(const double* __restrict__ s4, int N, int M, int K, SymExc ex,
                              const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ y,
                              double* __restrict__ partial) {
  const long n = (long)N * M * K;
  long lo, hi;
  chunk_of(n, blockIdx.x, gridDim.x, &lo, &hi);
  double ss = 0.0;
  for (long li = lo + threadIdx.x; li < hi; li += BS) {
    const int i = (int)li;
    int xx, yy, zz;
    decode(i, N, M, &xx, &yy, &zz);
    if (MODE == ST_PROD) {
      y[i] = row_apply_sym(s4, n, i, xx, yy, zz, N, M, K, x, 0.0, false, ex);
    } else if (MODE == ST_RESID) {
      const double r = row_apply_sym(s4, n, i, xx, yy, zz, N, M, K, x, b[i], true, ex);
      y[i] = r;
      ss += r * r;
    } else {
      const double q = row_apply_sym(s4, n, i, xx, yy, zz, N, M, K, x, 0.0, false, ex);
      y[i] = y[i] - q;
    }
  }
  if (MODE == ST_RESID) {
    ss = block_sum<BS>(ss);
    if (threadIdx.x == 0) partial[blockIdx.x] = ss;
  }
}

// y = A^T x with an FP32 x
__global__ void k_stencil_sym_t_f32x(const double* __restrict__ s4, int N, int M, int K, SymExc ex,
                                     const float* __restrict__ x, double* __restrict__ y) {
  const long n = (long)N * M * K;
  int xx, yy, zz;
  if (!tile_cell(N, M, &xx, &yy, &zz)) return;
  const int i = xx + N * (yy + M * zz);
  y[i] = row_apply_sym_t(s4, n, i, xx, yy, zz, N, M, K, x, 0.0, false, ex);
}

// y = A x with an FP32 x (mixed-precision Krylov products)
__global__ void k_stencil_sym_f32x(const double* __restrict__ s4, int N, int M, int K, SymExc ex,
                                   const float* __restrict__ x, double* __restrict__ y) {
  const long n = (long)N * M * K;
  int xx, yy, zz;
  if (!tile_cell(N, M, &xx, &yy, &zz)) return;
  const int i = xx + N * (yy + M * zz);
  y[i] = row_apply_sym(s4, n, i, xx, yy, zz, N, M, K, x, 0.0, false, ex);
}

// ---- kernels: solve phase ---------------------------------------------------------------------
template <int NPTS, int MODE, int BS>
__global__ void k_stencil(const double* __restrict__ coef, int N, int M, int K, const double* __restrict__ x,
                          const double* __restrict__ b, double* __restrict__ y, double* __restrict__ partial) {
  const long n = (long)N * M * K;
  long lo, hi;
  chunk_of(n, blockIdx.x, gridDim.x, &lo, &hi);
  double ss = 0.0;
  for (long li = lo + threadIdx.x; li < hi; li += BS) {
    const int i = (int)li;
    int xx, yy, zz;
    decode(i, N, M, &xx, &yy, &zz);
    if (MODE == ST_PROD) {
      y[i] = row_apply<NPTS>(coef, n, i, xx, yy, zz, N, M, K, x, 0.0, false);
    } else if (MODE == ST_RESID) {
      const double r = row_apply<NPTS>(coef, n, i, xx, yy, zz, N, M, K, x, b[i], true);
      y[i] = r;
      ss += r * r;
    } else {
      const double q = row_apply<NPTS>(coef, n, i, xx, yy, zz, N, M, K, x, 0.0, false);
      y[i] = y[i] - q;
    }
  }
  if (MODE == ST_RESID) {
    ss = block_sum<BS>(ss);
    if (threadIdx.x == 0) partial[blockIdx.x] = ss;
  }
}

// r_c = R r_f: each coarse cell sums 1/8 of its 8 children (R row order a,b,c,d,aa,bb,cc,dd)
__global__ void k_restrict(const double* __restrict__ rf, int Nf, int Mf, double* __restrict__ rc, int Nc, int Mc,
                           int Kc) {
  const long nc = (long)Nc * Mc * Kc;
  const int NMf = Nf * Mf;
  for (long li = blockIdx.x * (long)blockDim.x + threadIdx.x; li < nc; li += (long)gridDim.x * blockDim.x) {
    int X, Y, Z;
    decode((int)li, Nc, Mc, &X, &Y, &Z);
    const int a = 2 * X + Nf * 2 * Y + NMf * 2 * Z;
    double acc = 0.0;
    acc = acc + 0.125 * rf[a];
    acc = acc + 0.125 * rf[a + 1];
    acc = acc + 0.125 * rf[a + Nf];
    acc = acc + 0.125 * rf[a + Nf + 1];
    acc = acc + 0.125 * rf[a + NMf];
    acc = acc + 0.125 * rf[a + NMf + 1];
    acc = acc + 0.125 * rf[a + NMf + Nf];
    acc = acc + 0.125 * rf[a + NMf + Nf + 1];
    rc[li] = acc;
  }
}

// ch_f = P ch_c (trilinear), c_f = ch_f
__global__ void k_prolong(const double* __restrict__ chc, int Nc, int Mc, double* __restrict__ chf,
                          double* __restrict__ cf, int Nf, int Mf, int Kf) {
  const long nf = (long)Nf * Mf * Kf;
  const int NMc = Nc * Mc;
  for (long li = blockIdx.x * (long)blockDim.x + threadIdx.x; li < nf; li += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode((int)li, Nf, Mf, &x, &y, &z);
    int cx[2], cy[2], cz[2], nx, ny, nz;
    double wx[2], wy[2], wz[2];
    axis_w(x, Nf, cx, wx, &nx);
    axis_w(y, Mf, cy, wy, &ny);
    axis_w(z, Kf, cz, wz, &nz);
    double acc = 0.0;
    for (int c3 = 0; c3 < nz; ++c3)
      for (int c2 = 0; c2 < ny; ++c2)
        for (int c1 = 0; c1 < nx; ++c1) {
          const double w = wx[c1] * wy[c2] * wz[c3];
          acc = acc + w * chc[cx[c1] + Nc * cy[c2] + NMc * cz[c3]];
        }
    chf[li] = acc;
    cf[li] = acc;
  }
}

// c = omega * T_dir^{-1} r ; ch += c (smoother_adlj_interface.h:73-82 with the TDMA kernels
// smoother_adlj_x.h:38-48, smoother_adlj_y.h:36-50).  One thread per line.  For z lines the
// reference reads the first right-hand-side element at index y*N*K + x (its z call passes
// N*K where the y kernel expects N*M); z_head_fix = 0 reproduces that read.
__global__ void k_line_solve(const double* __restrict__ coef, int npts, int N, int M, int K, int dir,
                             const double* __restrict__ dinv, const double* __restrict__ eu,
                             const double* __restrict__ r, double* __restrict__ c, double* __restrict__ ch,
                             double omega, int z_head_fix) {
  const long n = (long)N * M * K;
  const int nlines = dir == 0 ? M * K : (dir == 1 ? N * K : N * M);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlines) return;
  int start, stride, len;
  line_geometry(t, dir, N, M, K, &start, &stride, &len);
  const double* sb = coef + slot_sub(npts, dir) * n;
  int head = start;
  if (dir == 2 && !z_head_fix) {
    const int x = t % N, y = t / N;
    head = y * N * K + x;
  }
  int i = start;
  double yprev = r[head] * dinv[i];
  c[i] = yprev;
  for (int k = 1; k < len; ++k) {
    i += stride;
    yprev = (r[i] - sb[i] * yprev) * dinv[i];
    c[i] = yprev;
  }
  // backward: x_k = y_k - e_k x_{k+1}; write omega*x and accumulate into ch
  double xnext = yprev;
  double cv = xnext * omega;
  c[i] = cv;
  ch[i] = ch[i] + cv;
  for (int k = len - 2; k >= 0; --k) {
    i -= stride;
    const double xv = c[i] - eu[i] * xnext;
    xnext = xv;
    cv = xv * omega;
    c[i] = cv;
    ch[i] = ch[i] + cv;
  }
}

// lu_decomposition_sparse_matrix.h:216-259 + decomposition_sparse_matrix_interface.h:43-71.
// The factors (n <= 64) and the right-hand side are staged in LDS by the whole workgroup; thread 0
// then runs the triangular solves on LDS operands (the same operations in the same order).
constexpr int LU_SOLVE_THREADS = 256;
__global__ void __launch_bounds__(LU_SOLVE_THREADS) k_lu_solve(const double* scale, const int* perm,
                                                               const double* L, const double* U, int n,
                                                               const double* __restrict__ b, double* __restrict__ x) {
  __shared__ double sL[64 * 64], sU[64 * 64], ssc[64], sb[64], slu_b[64], slu_y[64];
  __shared__ int sp[64];
  for (int k = threadIdx.x; k < n * n; k += blockDim.x) {
    sL[k] = L[k];
    sU[k] = U[k];
  }
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    ssc[k] = scale[k];
    sp[k] = perm[k];
    sb[k] = b[k];
  }
  __syncthreads();
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double* bs = slu_b;  // LDS: as local arrays they would live in scratch memory
  double* yv = slu_y;
  for (int i = 0; i < n; ++i) bs[i] = ssc[sp[i]] * sb[sp[i]];
  for (int i = 0; i < n; ++i) {
    double aux = bs[i];
    for (int t = 0; t < i; ++t) aux -= sL[i * n + t] * yv[t];
    yv[i] = aux;
  }
  for (int i = n - 1; i >= 0; --i) {
    double aux = yv[i];
    for (int j = i + 1; j < n; ++j) aux -= sU[i * n + j] * bs[j];
    bs[i] = aux / sU[i * n + i];
  }
  for (int i = 0; i < n; ++i) x[i] = ssc[i] * bs[i];
}

// ---- kernels: BLAS-1 ---------------------------------------------------------------------------
__global__ void k_copy(const double* __restrict__ a, double* __restrict__ b, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

__global__ void k_fill_zero(double* __restrict__ b, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = 0.0;
}

__global__ void k_scale_copy(const double* __restrict__ w, const double* __restrict__ normw, double* __restrict__ v,
                             long n) {
  const double f = 1.0 / normw[0];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) v[i] = f * w[i];
}

__global__ void k_scale_copy_f32(const double* __restrict__ w, const double* __restrict__ normw, double* __restrict__ v,
                                 float* __restrict__ vf, long n) {
  const double f = 1.0 / normw[0];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double a = f * w[i];
    v[i] = a;
    vf[i] = (float)a;
  }
}

template <int BS>
__global__ void k_dot_partial(const double* __restrict__ a, const double* __restrict__ b, long n, double* partial) {
  long lo, hi;
  chunk_of(n, blockIdx.x, gridDim.x, &lo, &hi);
  double s = 0.0;
  for (long i = lo + threadIdx.x; i < hi; i += BS) s += a[i] * b[i];
  s = block_sum<BS>(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

template <int BS>
__global__ void k_finalize(const double* partial, int nb, double* out, int sqrt_flag) {
  double s = 0.0;
  // each thread sums a contiguous run so the combination order is fixed
  const int per = (nb + BS - 1) / BS;
  const int lo = threadIdx.x * per;
  for (int i = lo; i < lo + per && i < nb; ++i) s += partial[i];
  s = block_sum<BS>(s);
  if (threadIdx.x == 0) out[0] = sqrt_flag ? sqrt(s) : s;
}

template <int BS>
__global__ void k_axpy_dot(double* __restrict__ w, const double* __restrict__ v, const double* __restrict__ h,
                           const double* __restrict__ vnext, long n, double* partial) {
  long lo, hi;
  chunk_of(n, blockIdx.x, gridDim.x, &lo, &hi);
  const double hv = h[0];
  double s = 0.0;
  for (long i = lo + threadIdx.x; i < hi; i += BS) {
    const double aux = v[i] * hv;
    const double wn = w[i] - aux;
    w[i] = wn;
    s += vnext ? wn * vnext[i] : wn * wn;
  }
  s = block_sum<BS>(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// ---- classical Gram-Schmidt passes ------------------------------------------------------------
// V holds cnt <= CGS_MAX basis vectors, vector k at V + k*ldv.  The dot pass runs on a 2D grid
// (row chunks x groups of CGS_GROUP vectors) so each thread keeps only CGS_GROUP accumulators and
// pointers; partial[k * nb + chunk] is summed in a fixed order by k_finalize_multi.  The update
// pass keeps the coefficients in LDS and streams the vectors row by row.
constexpr int CGS_GROUP = 8;

// rows per step of the Gram-Schmidt passes: their loads are in flight together
constexpr int CGS_ROWS = 4;

// acc[k] += sum over the thread's rows i of V_k[i] w[i], rows in increasing order, CGS_ROWS
// rows (i, i + BS, ...) per step so that CGS_ROWS (KC + 1) loads are in flight
template <int BS, int KC, class VT>
__device__ inline void cgs_dot_rows(const VT* __restrict__ Vg, long ldv, const double* __restrict__ w, long lo,
                                    long hi, double (&acc)[CGS_GROUP]) {
  constexpr int R = CGS_ROWS;
  long i = lo + threadIdx.x;
  for (; i + (R - 1) * BS < hi; i += R * BS) {
    double wr[R], v[R][KC];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      wr[r] = w[i + r * BS];
#pragma unroll
      for (int k = 0; k < KC; ++k) v[r][k] = (double)Vg[k * ldv + i + r * BS];
    }
#pragma unroll
    for (int k = 0; k < KC; ++k)
#pragma unroll
      for (int r = 0; r < R; ++r) acc[k] += v[r][k] * wr[r];
  }
  for (; i < hi; i += BS) {
    const double w0 = w[i];
#pragma unroll
    for (int k = 0; k < KC; ++k) acc[k] += (double)Vg[k * ldv + i] * w0;
  }
}

// Block (chunk c, group g) is block c % 8 + 8 (g + ng (c / 8)) of a 1-D grid of 8 ceil(nb / 8) ng
// blocks (those of chunks past nb exit at once): workgroups go to the 8 XCDs round-robin, so the
// ng blocks of one chunk run on one XCD one after the other and read its slice of w from that
// XCD's L2 after the first.
template <int BS, class VT>
__global__ void __launch_bounds__(BS) k_cgs_dot(const VT* __restrict__ V, long ldv, int cnt,
                                                const double* __restrict__ w, long n, int nb, int ng,
                                                double* __restrict__ partial) {
  __shared__ double lds[BS / 64][CGS_GROUP];
  const int b = blockIdx.x, r = b / 8;
  const int chunk = b % 8 + 8 * (r / ng), grp = r % ng;
  if (chunk >= nb) return;  // uniform over the block
  const int k0 = grp * CGS_GROUP;
  const int kc = cnt - k0 < CGS_GROUP ? cnt - k0 : CGS_GROUP;
  long lo, hi;
  chunk_of(n, chunk, nb, &lo, &hi);
  const VT* Vg = V + (long)k0 * ldv;
  double acc[CGS_GROUP];
#pragma unroll
  for (int k = 0; k < CGS_GROUP; ++k) acc[k] = 0.0;
  switch (kc) {  // uniform over the block: every load of a step is issued unconditionally
    case 1: cgs_dot_rows<BS, 1, VT>(Vg, ldv, w, lo, hi, acc); break;
    case 2: cgs_dot_rows<BS, 2, VT>(Vg, ldv, w, lo, hi, acc); break;
    case 3: cgs_dot_rows<BS, 3, VT>(Vg, ldv, w, lo, hi, acc); break;
    case 4: cgs_dot_rows<BS, 4, VT>(Vg, ldv, w, lo, hi, acc); break;
    case 5: cgs_dot_rows<BS, 5, VT>(Vg, ldv, w, lo, hi, acc); break;
    case 6: cgs_dot_rows<BS, 6, VT>(Vg, ldv, w, lo, hi, acc); break;
    case 7: cgs_dot_rows<BS, 7, VT>(Vg, ldv, w, lo, hi, acc); break;
    default: cgs_dot_rows<BS, CGS_GROUP, VT>(Vg, ldv, w, lo, hi, acc); break;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < CGS_GROUP; ++k) {
    double v = acc[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) lds[wave][k] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < kc) {
    double s = 0.0;
    for (int q = 0; q < BS / 64; ++q) s += lds[q][threadIdx.x];
    partial[(long)(k0 + threadIdx.x) * nb + chunk] = s;
  }
}

// w -= sum_k h_k v_k ; with_norm: partial ||w||^2 per block
template <int BS>
__global__ void __launch_bounds__(BS) k_cgs_update(const double* __restrict__ V, long ldv, int cnt,
                                                   const double* __restrict__ h, double* __restrict__ w, long n,
                                                   int with_norm, double* __restrict__ partial) {
  __shared__ double hk[CGS_MAX];
  if ((int)threadIdx.x < cnt) hk[threadIdx.x] = h[threadIdx.x];
  __syncthreads();
  long lo, hi;
  chunk_of(n, blockIdx.x, gridDim.x, &lo, &hi);
  double ss = 0.0;
  constexpr int R = CGS_ROWS;
  long i = lo + threadIdx.x;
  for (; i + (R - 1) * BS < hi; i += R * BS) {  // R rows per step, each updated in the order k = 0, 1, ...
    double wr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) wr[r] = w[i + r * BS];
    const double* vp = V + i;
#pragma unroll 8
    for (int k = 0; k < cnt; ++k)
#pragma unroll
      for (int r = 0; r < R; ++r) wr[r] -= hk[k] * vp[k * ldv + r * BS];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      w[i + r * BS] = wr[r];
      ss += wr[r] * wr[r];
    }
  }
  for (; i < hi; i += BS) {
    double wn = w[i];
    const double* vp = V + i;
#pragma unroll 8
    for (int k = 0; k < cnt; ++k) wn -= hk[k] * vp[k * ldv];
    w[i] = wn;
    ss += wn * wn;
  }
  if (with_norm) {
    ss = block_sum<BS>(ss);
    if (threadIdx.x == 0) partial[blockIdx.x] = ss;
  }
}

// out = post (w - sum_k (h_k isq_k) v_k) (out may alias w), out32 = (float)out when given;
// with_norm: partial ||out||^2 per block
// R rows per step and the vector loop unrolled U times: R U loads in flight (8 x 4 measured
// 342 against 355 us for 4 x 8 at 256^3 with 21 basis vectors)
template <int BS, int R, int U>
__global__ void __launch_bounds__(BS) k_cgs_update_scaled(const double* __restrict__ V, long ldv, int cnt,
                                                          const double* __restrict__ h, CgsScales sc, double post,
                                                          const double* w, double* out, float* __restrict__ out32,
                                                          long n, int with_norm, double* __restrict__ partial) {
  __shared__ double hk[CGS_MAX];
  if ((int)threadIdx.x < cnt) hk[threadIdx.x] = h[threadIdx.x] * sc.isq[threadIdx.x];
  __syncthreads();
  long lo, hi;
  chunk_of(n, blockIdx.x, gridDim.x, &lo, &hi);
  double ss = 0.0;
  long i = lo + threadIdx.x;
  for (; i + (R - 1) * BS < hi; i += R * BS) {
    double wr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) wr[r] = w[i + r * BS];
    const double* vp = V + i;
#pragma unroll U
    for (int k = 0; k < cnt; ++k)
#pragma unroll
      for (int r = 0; r < R; ++r) wr[r] -= hk[k] * vp[k * ldv + r * BS];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double o = wr[r] * post;
      out[i + r * BS] = o;
      if (out32) out32[i + r * BS] = (float)o;
      ss += o * o;
    }
  }
  for (; i < hi; i += BS) {
    double wn = w[i];
    const double* vp = V + i;
#pragma unroll 8
    for (int k = 0; k < cnt; ++k) wn -= hk[k] * vp[k * ldv];
    const double o = wn * post;
    out[i] = o;
    if (out32) out32[i] = (float)o;
    ss += o * o;
  }
  if (with_norm) {
    ss = block_sum<BS>(ss);
    if (threadIdx.x == 0) partial[blockIdx.x] = ss;
  }
}

template <int BS>
__global__ void k_finalize_multi(const double* __restrict__ partial, int nb, double* __restrict__ out,
                                 int sqrt_flag) {
  const double* p = partial + (long)blockIdx.x * nb;
  double s = 0.0;
  const int per = (nb + BS - 1) / BS;
  const int lo = threadIdx.x * per;
  for (int i = lo; i < lo + per && i < nb; ++i) s += p[i];
  s = block_sum<BS>(s);
  if (threadIdx.x == 0) out[blockIdx.x] = sqrt_flag ? sqrt(s) : s;
}

// y = A x on the fine 7-point operator with an FP32 input vector (mixed-precision Krylov basis)
__global__ void k_stencil7_prod_f32x(const double* __restrict__ coef, int N, int M, int K,
                                     const float* __restrict__ x, double* __restrict__ y) {
  const long n = (long)N * M * K;
  const int NM = N * M;
  for (long li = blockIdx.x * (long)blockDim.x + threadIdx.x; li < n; li += (long)gridDim.x * blockDim.x) {
    const int i = (int)li;
    int xx, yy, zz;
    decode(i, N, M, &xx, &yy, &zz);
    double acc = 0.0;
    if (zz > 0) acc = acc + coef[0 * n + i] * (double)x[i - NM];
    if (yy > 0) acc = acc + coef[1 * n + i] * (double)x[i - N];
    if (xx > 0) acc = acc + coef[2 * n + i] * (double)x[i - 1];
    acc = acc + coef[3 * n + i] * (double)x[i];
    if (xx < N - 1) acc = acc + coef[4 * n + i] * (double)x[i + 1];
    if (yy < M - 1) acc = acc + coef[5 * n + i] * (double)x[i + N];
    if (zz < K - 1) acc = acc + coef[6 * n + i] * (double)x[i + NM];
    y[i] = acc;
  }
}

struct YCoef {
  double y[64];
};

// x += sum_k y_k z_k, k = count-1 .. 0 in that order for every row; four rows per step (grid
// stride apart) with the vector loop unrolled four times, so that 16 loads are in flight
template <class ZT>
__global__ void k_sol_update_t(double* __restrict__ x, const ZT* __restrict__ Z, long ldz, YCoef yk, int count,
                               long n) {
  constexpr int R = 4;
  const long st = (long)gridDim.x * blockDim.x;
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  for (; i + (R - 1) * st < n; i += R * st) {
    double aux[R];
#pragma unroll
    for (int r = 0; r < R; ++r) aux[r] = x[i + r * st];
#pragma unroll 4
    for (int k = count - 1; k >= 0; --k)
#pragma unroll
      for (int r = 0; r < R; ++r) aux[r] = aux[r] + (double)Z[k * ldz + i + r * st] * yk.y[k];
#pragma unroll
    for (int r = 0; r < R; ++r) x[i + r * st] = aux[r];
  }
  for (; i < n; i += st) {
    double aux = x[i];
    for (int k = count - 1; k >= 0; --k) aux = aux + (double)Z[k * ldz + i] * yk.y[k];
    x[i] = aux;
  }
}

__global__ void k_bicg_s(double* __restrict__ nu, const double* __restrict__ r, double* __restrict__ sv, double alpha,
                         long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double nv = nu[i] * alpha;
    nu[i] = nv;
    sv[i] = r[i] - nv;
  }
}

__global__ void k_bicg_sol(const double* __restrict__ sv, double* __restrict__ y, double* __restrict__ z,
                           double* __restrict__ x, double* __restrict__ t, double* __restrict__ r, double alpha,
                           double w, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double yi = y[i] * alpha;
    y[i] = yi;
    const double zi = z[i] * w;
    z[i] = zi;
    const double cs = yi + zi;
    x[i] = x[i] + cs;
    const double ti = t[i] * w;
    t[i] = ti;
    r[i] = sv[i] - ti;
  }
}

template <int BS>
__global__ void k_bicg_p(double* __restrict__ p, const double* __restrict__ nu, const double* __restrict__ r,
                         double coef, double beta, long n, double* partial) {
  long lo, hi;
  chunk_of(n, blockIdx.x, gridDim.x, &lo, &hi);
  double s = 0.0;
  for (long i = lo + threadIdx.x; i < hi; i += BS) {
    const double cv = (p[i] - nu[i] * coef) * beta;
    const double re = r[i];
    p[i] = re + cv;
    s += re * re;
  }
  s = block_sum<BS>(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void k_sub(const double* __restrict__ b, const double* __restrict__ z, double* __restrict__ r, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) r[i] = b[i] - z[i];
}

int grid_1d(long n, int bs = 256) {
  long g = (n + bs - 1) / bs;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

// ================================================================================================
// launchers
// ================================================================================================
// blocks of tile_cell
static unsigned tile_blocks(Dims d) { return (unsigned)((long)((d.N + 63) / 64) * ((d.M + 3) / 4) * d.K); }

int reduction_blocks(long n) {
  long b = (n + RED_THREADS * 4 - 1) / (RED_THREADS * 4);
  if (b > 1024) b = 1024;
  if (b < 1) b = 1;
  return (int)b;
}

void csr_to_stencil7(const int* rowptr, const int* col, const double* val, Dims d, double* coef7, int* err_flag,
                     hipStream_t s) {
  k_csr_to_stencil7<<<grid_1d(d.n()), 256, 0, s>>>(rowptr, col, val, d.N, d.M, d.K, coef7, err_flag);
}

void galerkin(const double* coef_f, int npts_f, Dims df, double* coef27_c, Dims dc, hipStream_t s) {
  if (df.N & 1) {  // odd-odd-odd level (the only odd class the hierarchy admits)
    if (npts_f == 7)
      k_galerkin_odd<7><<<grid_1d(dc.n(), 64), 64, 0, s>>>(coef_f, df.N, df.M, df.K, coef27_c, dc.N, dc.M, dc.K);
    else
      k_galerkin_odd<27><<<grid_1d(dc.n(), 64), 64, 0, s>>>(coef_f, df.N, df.M, df.K, coef27_c, dc.N, dc.M, dc.K);
    return;
  }
  if (npts_f == 7)
    k_galerkin<7><<<grid_1d(dc.n(), 64), 64, 0, s>>>(coef_f, df.N, df.M, df.K, coef27_c, dc.N, dc.M, dc.K);
  else
    k_galerkin<27><<<grid_1d(dc.n(), 64), 64, 0, s>>>(coef_f, df.N, df.M, df.K, coef27_c, dc.N, dc.M, dc.K);
}

void diag_absmax(const double* coef, int npts, Dims d, double* partial, double* out, hipStream_t s) {
  const int nb = reduction_blocks(d.n());
  const double* dg = coef + (long)(npts == 7 ? 3 : 13) * d.n();
  k_diag_absmax_partial<RED_THREADS><<<nb, RED_THREADS, 0, s>>>(dg, d.n(), partial);
  k_max_finalize<RED_THREADS><<<1, RED_THREADS, 0, s>>>(partial, nb, out);
}

void absmax(const double* a, long n, double* partial, double* out, hipStream_t s) {
  const int nb = reduction_blocks(n);
  k_diag_absmax_partial<RED_THREADS><<<nb, RED_THREADS, 0, s>>>(a, n, partial);
  k_max_finalize<RED_THREADS><<<1, RED_THREADS, 0, s>>>(partial, nb, out);
}

void line_factor(const double* coef, int npts, Dims d, int dir, const double* absmax, double* dinv, double* eu,
                 hipStream_t s) {
  const int nlines = dir == 0 ? d.M * d.K : (dir == 1 ? d.N * d.K : d.N * d.M);
  k_line_factor<<<div_up(nlines, 64), 64, 0, s>>>(coef, npts, d.N, d.M, d.K, dir, absmax, dinv, eu);
}

void lu_factor(const double* coef27, Dims d, double* scale, int* perm, double* L, double* U, double* col_scratch,
               hipStream_t s) {
  k_lu_factor<<<1, 1, 0, s>>>(coef27, d.N, d.M, d.K, scale, perm, L, U, col_scratch);
}

void stencil_apply(const double* coef, int npts, Dims d, const double* x, const double* b, double* y, int mode,
                   double* partial, hipStream_t s) {
  const int nb = reduction_blocks(d.n());
#define LSX_ST(NP, MD) k_stencil<NP, MD, RED_THREADS><<<nb, RED_THREADS, 0, s>>>(coef, d.N, d.M, d.K, x, b, y, partial)
  if (npts == 7) {
    if (mode == ST_PROD) LSX_ST(7, ST_PROD);
    else if (mode == ST_RESID) LSX_ST(7, ST_RESID);
    else LSX_ST(7, ST_SUBR);
  } else {
    if (mode == ST_PROD) LSX_ST(27, ST_PROD);
    else if (mode == ST_RESID) LSX_ST(27, ST_RESID);
    else LSX_ST(27, ST_SUBR);
  }
#undef LSX_ST
}

void sym_check(const double* coef7, Dims d, unsigned char* exc, int* asym, hipStream_t s) {
  k_sym_check<<<grid_1d(d.n()), 256, 0, s>>>(coef7, d.N, d.M, d.K, exc, asym);
}

void sym_build(const double* coef7, const unsigned char* exc, Dims d, double* s4, hipStream_t s) {
  k_sym_build<<<grid_1d(d.n()), 256, 0, s>>>(coef7, exc, d.N, d.M, d.K, s4);
}

void stencil_apply_sym(const double* s4, const SymExc& ex, Dims d, const double* x, const double* b, double* y,
                       int mode, double* partial, hipStream_t s) {
  const int nb = reduction_blocks(d.n());
  if (mode == ST_PROD)
    k_stencil_sym<ST_PROD, RED_THREADS><<<nb, RED_THREADS, 0, s>>>(s4, d.N, d.M, d.K, ex, x, b, y, partial);
  else if (mode == ST_RESID)
    k_stencil_sym<ST_RESID, RED_THREADS><<<nb, RED_THREADS, 0, s>>>(s4, d.N, d.M, d.K, ex, x, b, y, partial);
  else
    k_stencil_sym<ST_SUBR, RED_THREADS><<<nb, RED_THREADS, 0, s>>>(s4, d.N, d.M, d.K, ex, x, b, y, partial);
}

void restrict_full(const double* rf, Dims df, double* rc, Dims dc, hipStream_t s) {
  if (df.N & 1) {
    k_restrict_odd<<<grid_1d(dc.n()), 256, 0, s>>>(rf, df.N, df.M, df.K, rc, dc.N, dc.M, dc.K);
    return;
  }
  k_restrict<<<grid_1d(dc.n()), 256, 0, s>>>(rf, df.N, df.M, rc, dc.N, dc.M, dc.K);
}

void prolong_copy(const double* chc, Dims dc, double* chf, double* cf, Dims df, hipStream_t s) {
  if (df.N & 1) {
    k_prolong_odd<<<grid_1d(df.n()), 256, 0, s>>>(chc, dc.N, dc.M, chf, cf, df.N, df.M, df.K);
    return;
  }
  k_prolong<<<grid_1d(df.n()), 256, 0, s>>>(chc, dc.N, dc.M, chf, cf, df.N, df.M, df.K);
}

void line_solve(const double* coef, int npts, Dims d, int dir, const double* dinv, const double* eu, const double* r,
                double* c, double* ch, double omega, int z_head_fix, hipStream_t s) {
  const int nlines = dir == 0 ? d.M * d.K : (dir == 1 ? d.N * d.K : d.N * d.M);
  k_line_solve<<<div_up(nlines, 64), 64, 0, s>>>(coef, npts, d.N, d.M, d.K, dir, dinv, eu, r, c, ch, omega,
                                                 z_head_fix);
}

void lu_solve(const double* scale, const int* perm, const double* L, const double* U, int n, const double* b,
              double* x, hipStream_t s) {
  k_lu_solve<<<1, LU_SOLVE_THREADS, 0, s>>>(scale, perm, L, U, n, b, x);
}

void copy(const double* a, double* b, long n, hipStream_t s) { k_copy<<<grid_1d(n), 256, 0, s>>>(a, b, n); }

void scale_copy_f32(const double* w, const double* normw, double* v, float* vf, long n, hipStream_t s) {
  k_scale_copy_f32<<<grid_1d(n), 256, 0, s>>>(w, normw, v, vf, n);
}

void fill_zero(double* b, long n, hipStream_t s) { k_fill_zero<<<grid_1d(n), 256, 0, s>>>(b, n); }

void scale_copy(const double* w, const double* normw, double* v, long n, hipStream_t s) {
  k_scale_copy<<<grid_1d(n), 256, 0, s>>>(w, normw, v, n);
}

void dot_partial(const double* a, const double* b, long n, double* partial, hipStream_t s) {
  k_dot_partial<RED_THREADS><<<reduction_blocks(n), RED_THREADS, 0, s>>>(a, b, n, partial);
}

void finalize(const double* partial, int nb, double* out, int sqrt_flag, hipStream_t s) {
  k_finalize<RED_THREADS><<<1, RED_THREADS, 0, s>>>(partial, nb, out, sqrt_flag);
}

void axpy_dot(double* w, const double* v, const double* h, const double* vnext, long n, double* partial,
              hipStream_t s) {
  k_axpy_dot<RED_THREADS><<<reduction_blocks(n), RED_THREADS, 0, s>>>(w, v, h, vnext, n, partial);
}

void sol_update(double* x, const double* Z, long ldz, const double* yk, int count, long n, hipStream_t s) {
  // chunks of 64 coefficients, highest k first (the reference order k = last..0)
  int hi = count;
  while (hi > 0) {
    const int lo = hi > 64 ? hi - 64 : 0;
    YCoef c;
    for (int k = lo; k < hi; ++k) c.y[k - lo] = yk[k];
    k_sol_update_t<double><<<grid_1d(n), 256, 0, s>>>(x, Z + (long)lo * ldz, ldz, c, hi - lo, n);
    hi = lo;
  }
}

int cgs_blocks(long n) {
  long b = (n + 1023) / 1024;
  if (b > CGS_BLOCKS_MAX) b = CGS_BLOCKS_MAX;
  if (b < 1) b = 1;
  return (int)b;
}

void cgs_dot(const double* V, long ldv, int cnt, const double* w, long n, double* partial, double* out,
             hipStream_t s) {
  const int nb = cgs_blocks(n), ng = (cnt + CGS_GROUP - 1) / CGS_GROUP;
  k_cgs_dot<RED_THREADS, double><<<(nb + 7) / 8 * 8 * ng, RED_THREADS, 0, s>>>(V, ldv, cnt, w, n, nb, ng, partial);
  k_finalize_multi<RED_THREADS><<<cnt, RED_THREADS, 0, s>>>(partial, nb, out, 0);
}

void cgs_dot_f32(const float* V, long ldv, int cnt, const double* w, long n, double* partial, double* out,
                 hipStream_t s) {
  const int nb = cgs_blocks(n), ng = (cnt + CGS_GROUP - 1) / CGS_GROUP;
  k_cgs_dot<RED_THREADS, float><<<(nb + 7) / 8 * 8 * ng, RED_THREADS, 0, s>>>(V, ldv, cnt, w, n, nb, ng, partial);
  k_finalize_multi<RED_THREADS><<<cnt, RED_THREADS, 0, s>>>(partial, nb, out, 0);
}

void cgs_update(const double* V, long ldv, int cnt, const double* h, double* w, long n, double* partial,
                double* norm_out, hipStream_t s) {
  const int nb = cgs_blocks(n);
  k_cgs_update<RED_THREADS><<<nb, RED_THREADS, 0, s>>>(V, ldv, cnt, h, w, n, norm_out != nullptr, partial);
  if (norm_out) k_finalize_multi<RED_THREADS><<<1, RED_THREADS, 0, s>>>(partial, nb, norm_out, 1);
}

void cgs_update_scaled(const double* V, long ldv, int cnt, const double* h, const CgsScales& sc, double post,
                       const double* w, double* out, float* out32, long n, double* partial, double* norm_out,
                       hipStream_t s) {
  const int nb = cgs_blocks(n);
  k_cgs_update_scaled<RED_THREADS, 8, 4><<<nb, RED_THREADS, 0, s>>>(V, ldv, cnt, h, sc, post, w, out, out32, n,
                                                                    norm_out != nullptr, partial);
  if (norm_out) k_finalize_multi<RED_THREADS><<<1, RED_THREADS, 0, s>>>(partial, nb, norm_out, 1);
}

void stencil_apply_sym_t(const double* s4, const SymExc& ex, Dims d, const double* x, const double* b, double* y,
                         int mode, double* partial, hipStream_t s) {
  const int nb = reduction_blocks(d.n());
  if (mode == ST_PROD)
    k_stencil_sym_t<ST_PROD, RED_THREADS><<<nb, RED_THREADS, 0, s>>>(s4, d.N, d.M, d.K, ex, x, b, y, partial);
  else if (mode == ST_RESID)
    k_stencil_sym_t<ST_RESID, RED_THREADS><<<nb, RED_THREADS, 0, s>>>(s4, d.N, d.M, d.K, ex, x, b, y, partial);
  else
    k_stencil_sym_t<ST_SUBR, RED_THREADS><<<nb, RED_THREADS, 0, s>>>(s4, d.N, d.M, d.K, ex, x, b, y, partial);
}

void stencil_prod_sym_t_f32x(const double* s4, const SymExc& ex, Dims d, const float* x, double* y, hipStream_t s) {
  k_stencil_sym_t_f32x<<<tile_blocks(d), 256, 0, s>>>(s4, d.N, d.M, d.K, ex, x, y);
}

void stencil_prod_sym_f32x(const double* s4, const SymExc& ex, Dims d, const float* x, double* y, hipStream_t s) {
  k_stencil_sym_f32x<<<tile_blocks(d), 256, 0, s>>>(s4, d.N, d.M, d.K, ex, x, y);
}

void stencil7_prod_f32x(const double* coef, Dims d, const float* x, double* y, hipStream_t s) {
  k_stencil7_prod_f32x<<<grid_1d(d.n()), 256, 0, s>>>(coef, d.N, d.M, d.K, x, y);
}

void sol_update_f32(double* x, const float* Z, long ldz, const double* yk, int count, long n, hipStream_t s) {
  int hi = count;
  while (hi > 0) {
    const int lo = hi > 64 ? hi - 64 : 0;
    YCoef c;
    for (int k = lo; k < hi; ++k) c.y[k - lo] = yk[k];
    k_sol_update_t<float><<<grid_1d(n), 256, 0, s>>>(x, Z + (long)lo * ldz, ldz, c, hi - lo, n);
    hi = lo;
  }
}

void bicg_s_update(double* nu, const double* r, double* sv, double alpha, long n, hipStream_t s) {
  k_bicg_s<<<grid_1d(n), 256, 0, s>>>(nu, r, sv, alpha, n);
}

void bicg_sol_update(const double* sv, double* y, double* z, double* x, double* t, double* r, double alpha, double w,
                     long n, hipStream_t s) {
  k_bicg_sol<<<grid_1d(n), 256, 0, s>>>(sv, y, z, x, t, r, alpha, w, n);
}

void bicg_p_update(double* p, const double* nu, const double* r, double coef, double beta, long n, double* partial,
                   hipStream_t s) {
  k_bicg_p<RED_THREADS><<<reduction_blocks(n), RED_THREADS, 0, s>>>(p, nu, r, coef, beta, n, partial);
}

void sub_vec(const double* b, const double* z, double* r, long n, hipStream_t s) {
  k_sub<<<grid_1d(n), 256, 0, s>>>(b, z, r, n);
}

}  // namespace lsx
