// Geometric V-cycle preconditioner: red-black Gauss-Seidel on 7-point operators stored as face
// coefficients, full-sum restriction, trilinear prolongation, FP64 dense LU on the coarsest
// level.  It preconditions the FP64 flexible Krylov loop.  Built without FMA contraction so a
// cycle gives the same bits whichever kernels (per-level or fused) carry it out.
//
// Coarse levels are rediscretised: a coarse face is half the sum of the four fine faces it
// covers (area x4, distance x2), the diagonal is the Neumann -sum of faces.  The finest level's
// two pinned cells are identity rows there; a coarse level sees each as a Dirichlet sink on the
// coarse cell that contains it, of conductance equal to the sum of the pinned cell's faces.

#include <hip/hip_fp16.h>

#include "lsx_vcycle.hpp"

namespace lsx {
namespace vc {

namespace {

__device__ inline void decode(long i, int N, int M, int* x, int* y, int* z) {
  const long NM = (long)N * M;
  *z = (int)(i / NM);
  const long r = i - (long)*z * NM;
  *y = (int)(r / N);
  *x = (int)(r - (long)*y * N);
}

inline int grid_1d(long n, int bs = 256) {
  long g = (n + bs - 1) / bs;
  if (g > 16384) g = 16384;
  return (int)(g < 1 ? 1 : g);
}

template <class T, class FT>
__device__ inline T ld(const FT* p, long i) {
  return (T)p[i];
}

// ---- setup -----------------------------------------------------------------------------------
__global__ void k_faces_from7(const double* __restrict__ c7, int N, int M, int K, int pin0, int pin1,
                              double* __restrict__ fe, double* __restrict__ fn, double* __restrict__ ft,
                              double* __restrict__ dg) {
  const long n = (long)N * M * K, NM = (long)N * M;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    const bool pin = (i == pin0) || (i == pin1);
    fe[i] = x < N - 1 ? (pin ? c7[2 * n + i + 1] : c7[4 * n + i]) : 0.0;
    fn[i] = y < M - 1 ? (pin ? c7[1 * n + i + N] : c7[5 * n + i]) : 0.0;
    ft[i] = z < K - 1 ? (pin ? c7[0 * n + i + NM] : c7[6 * n + i]) : 0.0;
    dg[i] = c7[3 * n + i];
  }
}

__global__ void k_aggregate(const double* __restrict__ fe, const double* __restrict__ fn, const double* __restrict__ ft,
                            int Nf, int Mf, double* __restrict__ ce, double* __restrict__ cn, double* __restrict__ ct,
                            int Nc, int Mc, int Kc) {
  const long nc = (long)Nc * Mc * Kc, NMf = (long)Nf * Mf;
  for (long I = blockIdx.x * (long)blockDim.x + threadIdx.x; I < nc; I += (long)gridDim.x * blockDim.x) {
    int X, Y, Z;
    decode(I, Nc, Mc, &X, &Y, &Z);
    const long a = 2L * X + (long)Nf * 2 * Y + NMf * 2 * Z;
    double se = 0, sn = 0, st = 0;
    if (X < Nc - 1) {
      const long f = a + 1;
      se = fe[f] + fe[f + Nf] + fe[f + NMf] + fe[f + Nf + NMf];
    }
    if (Y < Mc - 1) {
      const long f = a + Nf;
      sn = fn[f] + fn[f + 1] + fn[f + NMf] + fn[f + 1 + NMf];
    }
    if (Z < Kc - 1) {
      const long f = a + NMf;
      st = ft[f] + ft[f + 1] + ft[f + Nf] + ft[f + 1 + Nf];
    }
    ce[I] = 0.5 * se;
    cn[I] = 0.5 * sn;
    ct[I] = 0.5 * st;
  }
}

__global__ void k_diag_from_faces(const double* __restrict__ fe, const double* __restrict__ fn,
                                  const double* __restrict__ ft, int N, int M, int K, double* __restrict__ dg) {
  const long n = (long)N * M * K, NM = (long)N * M;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    double d = -fe[i] - fn[i] - ft[i];
    if (x > 0) d -= fe[i - 1];
    if (y > 0) d -= fn[i - N];
    if (z > 0) d -= ft[i - NM];
    dg[i] = d;
  }
}

__global__ void k_pin_sigma(const double* __restrict__ fe, const double* __restrict__ fn, const double* __restrict__ ft,
                            int N, int M, int K, int pin0, int pin1, double* __restrict__ sigma) {
  const int q = threadIdx.x;
  if (blockIdx.x != 0 || q > 1) return;
  const long NM = (long)N * M;
  const long i = q == 0 ? pin0 : pin1;
  double sgm = 0.0;
  if (i >= 0) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    sgm = fe[i] + fn[i] + ft[i];
    if (x > 0) sgm += fe[i - 1];
    if (y > 0) sgm += fn[i - N];
    if (z > 0) sgm += ft[i - NM];
  }
  sigma[q] = sgm;
}

__global__ void k_pin_sink(double* __restrict__ dg, int c0, int c1, const double* __restrict__ sigma) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  if (c0 >= 0) dg[c0] -= sigma[0];
  if (c1 >= 0) dg[c1] -= sigma[1];
}

// inverse diagonal; only empty rows (closed cut cells, coarse cells inside the body) are left
// out.  Tiny cut cells keep their own pivot: a right-hand side that is O(1) there asks for an
// O(1/d) pressure, which the smoother must be able to produce.
template <class T>
__global__ void k_inv_diag(const double* __restrict__ dg, long n, T* __restrict__ s) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double d = dg[i];
    s[i] = d == 0.0 ? (T)0 : (T)(1.0 / d);
  }
}

template <class FT>
__global__ void k_store_faces(const double* __restrict__ fe, const double* __restrict__ fn,
                              const double* __restrict__ ft, long n, double inv_scale, FT* __restrict__ ae,
                              FT* __restrict__ an, FT* __restrict__ at) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    ae[i] = (FT)(fe[i] * inv_scale);
    an[i] = (FT)(fn[i] * inv_scale);
    at[i] = (FT)(ft[i] * inv_scale);
  }
}

__global__ void k_absmax3(const double* __restrict__ a, const double* __restrict__ b, const double* __restrict__ c,
                          long n, double* __restrict__ partial) {
  __shared__ double lds[4];
  double m = 0.0;
  const long per = (n + gridDim.x - 1) / gridDim.x;
  const long lo = per * blockIdx.x, hi = lo + per < n ? lo + per : n;
  for (long i = lo + threadIdx.x; i < hi; i += blockDim.x) m = fmax(m, fmax(fabs(a[i]), fmax(fabs(b[i]), fabs(c[i]))));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
}

__global__ void k_max_reduce(const double* __restrict__ partial, int nb, double* out) {
  __shared__ double lds[4];
  double m = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) m = fmax(m, partial[i]);
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
}

__global__ void k_to27(const double* __restrict__ fe, const double* __restrict__ fn, const double* __restrict__ ft,
                       const double* __restrict__ dg, int N, int M, int K, int pin0, int pin1, double* __restrict__ c27) {
  const long n = (long)N * M * K, NM = (long)N * M;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    for (int q = 0; q < 27; ++q) c27[q * n + i] = 0.0;
    c27[13 * n + i] = dg[i];
    if (i == pin0 || i == pin1) continue;
    if (x < N - 1) c27[14 * n + i] = fe[i];
    if (x > 0) c27[12 * n + i] = fe[i - 1];
    if (y < M - 1) c27[16 * n + i] = fn[i];
    if (y > 0) c27[10 * n + i] = fn[i - N];
    if (z < K - 1) c27[22 * n + i] = ft[i];
    if (z > 0) c27[4 * n + i] = ft[i - NM];
  }
}

// ---- cycle -----------------------------------------------------------------------------------
template <class T, class FT>
__device__ inline T offdiag(const T* __restrict__ x, const FT* __restrict__ ae, const FT* __restrict__ an,
                            const FT* __restrict__ at, long i, int xx, int yy, int zz, int N, int M, int K) {
  const long NM = (long)N * M;
  T acc = 0;
  if (xx > 0) acc += ld<T>(ae, i - 1) * x[i - 1];
  if (xx < N - 1) acc += ld<T>(ae, i) * x[i + 1];
  if (yy > 0) acc += ld<T>(an, i - N) * x[i - N];
  if (yy < M - 1) acc += ld<T>(an, i) * x[i + N];
  if (zz > 0) acc += ld<T>(at, i - NM) * x[i - NM];
  if (zz < K - 1) acc += ld<T>(at, i) * x[i + NM];
  return acc;
}

// one colour of a red-black sweep; work item t (from start, by stride) owns the t-th cell of it.
// zero_guess: 0 normal; 1 x is zero before this half-sweep; 2 first update of this colour after
// a zero start (its values are not read).  An empty row (1/a_p stored as 0) keeps its value, so
// that every step has the form x + W (b - A x) and the cycle's transpose is the transposed cycle.
template <class T, class FT>
__device__ inline void rbgs_body(T* __restrict__ x, const T* __restrict__ b, const FT* __restrict__ ae,
                                 const FT* __restrict__ an, const FT* __restrict__ at, const T* __restrict__ s, T fs,
                                 int N, int M, int K, int color, int pin0, int pin1, int zero_guess, long start,
                                 long stride) {
  const long half_row = (N + 1) / 2;
  const long rows = (long)M * K;
  const long total = half_row * rows;
  for (long t = start; t < total; t += stride) {
    const long row = t / half_row;
    const int yy = (int)(row % M), zz = (int)(row / M);
    const int xx = 2 * (int)(t - row * half_row) + ((yy + zz + color) & 1);
    if (xx >= N) continue;
    const long i = (long)xx + (long)N * row;
    T v;
    if (i == pin0 || i == pin1) {
      v = b[i] * s[i];
    } else if (s[i] == (T)0) {
      v = zero_guess ? (T)0 : x[i];  // empty row: keeps its value (0 from a zero start)
    } else if (zero_guess == 1) {
      v = b[i] * s[i];
    } else {
      v = (b[i] - fs * offdiag<T>(x, ae, an, at, i, xx, yy, zz, N, M, K)) * s[i];
    }
    x[i] = v;
  }
}

template <class T, class FT>
__global__ void k_rbgs(T* __restrict__ x, const T* __restrict__ b, const FT* __restrict__ ae, const FT* __restrict__ an,
                       const FT* __restrict__ at, const T* __restrict__ s, T fs, int N, int M, int K, int color,
                       int pin0, int pin1, int zero_guess) {
  rbgs_body<T, FT>(x, b, ae, an, at, s, fs, N, M, K, color, pin0, pin1, zero_guess,
                   blockIdx.x * (long)blockDim.x + threadIdx.x, (long)gridDim.x * blockDim.x);
}

template <class T, class FT>
__device__ inline void resid_restrict_body(const T* __restrict__ x, const T* __restrict__ b, const FT* __restrict__ ae,
                                           const FT* __restrict__ an, const FT* __restrict__ at,
                                           const T* __restrict__ s, T fs, int N, int M, int K, int pin0, int pin1,
                                           T* __restrict__ bc, long start, long stride) {
  const int Nc = N / 2, Mc = M / 2, Kc = K / 2;
  const long nc = (long)Nc * Mc * Kc, NM = (long)N * M;
  for (long I = start; I < nc; I += stride) {
    int X, Y, Z;
    decode(I, Nc, Mc, &X, &Y, &Z);
    T acc = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int xx = 2 * X + (c & 1), yy = 2 * Y + ((c >> 1) & 1), zz = 2 * Z + (c >> 2);
      const long i = xx + (long)N * yy + NM * zz;
      const T ap = s[i] != (T)0 ? (T)1 / s[i] : (T)0;
      T r = b[i] - ap * x[i];
      if (i != pin0 && i != pin1) r -= fs * offdiag<T>(x, ae, an, at, i, xx, yy, zz, N, M, K);
      acc += r;
    }
    bc[I] = acc;
  }
}

template <class T, class FT>
__global__ void k_resid_restrict(const T* __restrict__ x, const T* __restrict__ b, const FT* __restrict__ ae,
                                 const FT* __restrict__ an, const FT* __restrict__ at, const T* __restrict__ s, T fs,
                                 int N, int M, int K, int pin0, int pin1, T* __restrict__ bc) {
  resid_restrict_body<T, FT>(x, b, ae, an, at, s, fs, N, M, K, pin0, pin1, bc,
                             blockIdx.x * (long)blockDim.x + threadIdx.x, (long)gridDim.x * blockDim.x);
}

__device__ inline void axis_w(int f, int nf, int* c, float* w, int* cnt) {
  const int p = f >> 1;
  if (f == 0 || f == nf - 1) {
    c[0] = p; w[0] = 1.0f; *cnt = 1;
  } else if (f & 1) {
    c[0] = p; w[0] = 0.75f; c[1] = p + 1; w[1] = 0.25f; *cnt = 2;
  } else {
    c[0] = p - 1; w[0] = 0.25f; c[1] = p; w[1] = 0.75f; *cnt = 2;
  }
}

// ---- transposed cycle ------------------------------------------------------------------------
// The transpose of the cycle runs its steps in reverse order, each transposed: red-black
// half-sweeps in the opposite colour order on A^T, restriction by P^T (trilinear weights
// gathered from the 4 x 4 x 4 fine cells around a coarse cell), prolongation by the transposed
// sum restriction (every child receives its parent's value).  The coarse operators are
// symmetric; on the finest level the pinned rows of A^T are full rows and every other row
// drops its coupling to a pinned cell.
template <class T, class FT>
__device__ inline T offdiag_tr(const T* __restrict__ x, const FT* __restrict__ ae, const FT* __restrict__ an,
                               const FT* __restrict__ at, long i, int xx, int yy, int zz, int N, int M, int K,
                               int pin0, int pin1) {
  const long NM = (long)N * M;
  const bool own = i == pin0 || i == pin1;
  T acc = 0;
  auto add = [&](T c, long j) {
    if (own || (j != pin0 && j != pin1)) acc += c * x[j];
  };
  if (xx > 0) add(ld<T>(ae, i - 1), i - 1);
  if (xx < N - 1) add(ld<T>(ae, i), i + 1);
  if (yy > 0) add(ld<T>(an, i - N), i - N);
  if (yy < M - 1) add(ld<T>(an, i), i + N);
  if (zz > 0) add(ld<T>(at, i - NM), i - NM);
  if (zz < K - 1) add(ld<T>(at, i), i + NM);
  return acc;
}

template <class T, class FT>
__global__ void k_rbgs_tr(T* __restrict__ x, const T* __restrict__ b, const FT* __restrict__ ae,
                          const FT* __restrict__ an, const FT* __restrict__ at, const T* __restrict__ s, T fs, int N,
                          int M, int K, int color, int pin0, int pin1, int zero_guess) {
  const long half_row = (N + 1) / 2;
  const long total = half_row * M * K;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const long row = t / half_row;
    const int yy = (int)(row % M), zz = (int)(row / M);
    const int xx = 2 * (int)(t - row * half_row) + ((yy + zz + color) & 1);
    if (xx >= N) continue;
    const long i = (long)xx + (long)N * row;
    if (s[i] == (T)0)
      x[i] = zero_guess ? (T)0 : x[i];
    else
      x[i] = zero_guess == 1
                 ? b[i] * s[i]
                 : (b[i] - fs * offdiag_tr<T, FT>(x, ae, an, at, i, xx, yy, zz, N, M, K, pin0, pin1)) * s[i];
  }
}

template <class T, class FT>
__global__ void k_resid_tr(const T* __restrict__ x, const T* __restrict__ b, const FT* __restrict__ ae,
                           const FT* __restrict__ an, const FT* __restrict__ at, const T* __restrict__ s, T fs, int N,
                           int M, int K, int pin0, int pin1, T* __restrict__ r) {
  const long n = (long)N * M * K;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int xx, yy, zz;
    decode(i, N, M, &xx, &yy, &zz);
    const T ap = s[i] != (T)0 ? (T)1 / s[i] : (T)0;
    r[i] = b[i] - ap * x[i] - fs * offdiag_tr<T, FT>(x, ae, an, at, i, xx, yy, zz, N, M, K, pin0, pin1);
  }
}

// weight of fine cell f in coarse cell C along one axis under the trilinear prolongation
__device__ inline float axis_wt(int f, int nf, int C) {
  int c[2], cnt;
  float w[2];
  axis_w(f, nf, c, w, &cnt);
  float v = 0.0f;
  for (int k = 0; k < cnt; ++k) v = c[k] == C ? w[k] : v;
  return v;
}

template <class T>
__global__ void k_restrict_pt(const T* __restrict__ r, int Nf, int Mf, int Kf, T* __restrict__ bc) {
  const int Nc = Nf / 2, Mc = Mf / 2, Kc = Kf / 2;
  const long nc = (long)Nc * Mc * Kc, NMf = (long)Nf * Mf;
  for (long I = blockIdx.x * (long)blockDim.x + threadIdx.x; I < nc; I += (long)gridDim.x * blockDim.x) {
    int X, Y, Z;
    decode(I, Nc, Mc, &X, &Y, &Z);
    T acc = 0;
    for (int fz = 2 * Z - 1; fz <= 2 * Z + 2; ++fz) {
      if (fz < 0 || fz >= Kf) continue;
      const float wz = axis_wt(fz, Kf, Z);
      if (wz == 0.0f) continue;
      for (int fy = 2 * Y - 1; fy <= 2 * Y + 2; ++fy) {
        if (fy < 0 || fy >= Mf) continue;
        const float wy = axis_wt(fy, Mf, Y);
        if (wy == 0.0f) continue;
        for (int fx = 2 * X - 1; fx <= 2 * X + 2; ++fx) {
          if (fx < 0 || fx >= Nf) continue;
          const float wx = axis_wt(fx, Nf, X);
          if (wx == 0.0f) continue;
          acc += (T)(wx * wy * wz) * r[fx + (long)Nf * fy + NMf * fz];
        }
      }
    }
    bc[I] = acc;
  }
}

template <class T>
__global__ void k_prolong_inject_add(const T* __restrict__ xc, int Nc, int Mc, T* __restrict__ xf, int Nf, int Mf,
                                     int Kf) {
  const long nf = (long)Nf * Mf * Kf, NMc = (long)Nc * Mc;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nf; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(i, Nf, Mf, &x, &y, &z);
    xf[i] += xc[(x >> 1) + (long)Nc * (y >> 1) + NMc * (z >> 1)];
  }
}

template <class T>
__device__ inline void prolong_add_body(const T* __restrict__ xc, int Nc, int Mc, T* __restrict__ xf, int Nf, int Mf,
                                        int Kf, long start, long stride) {
  const long nf = (long)Nf * Mf * Kf;
  const long NMc = (long)Nc * Mc;
  for (long i = start; i < nf; i += stride) {
    int x, y, z;
    decode(i, Nf, Mf, &x, &y, &z);
    int cx[2], cy[2], cz[2], nx, ny, nz;
    float wx[2], wy[2], wz[2];
    axis_w(x, Nf, cx, wx, &nx);
    axis_w(y, Mf, cy, wy, &ny);
    axis_w(z, Kf, cz, wz, &nz);
    T acc = 0;
    for (int c3 = 0; c3 < nz; ++c3)
      for (int c2 = 0; c2 < ny; ++c2)
        for (int c1 = 0; c1 < nx; ++c1)
          acc += (T)(wx[c1] * wy[c2] * wz[c3]) * xc[cx[c1] + Nc * cy[c2] + NMc * cz[c3]];
    xf[i] += acc;
  }
}

template <class T>
__global__ void k_prolong_add(const T* __restrict__ xc, int Nc, int Mc, T* __restrict__ xf, int Nf, int Mf, int Kf) {
  prolong_add_body<T>(xc, Nc, Mc, xf, Nf, Mf, Kf, blockIdx.x * (long)blockDim.x + threadIdx.x,
                      (long)gridDim.x * blockDim.x);
}

// The small levels of the cycle in one workgroup: pre-smoothing and restriction down to the
// coarsest level, its dense LU solve (FP64, thread 0, the factors of lu_factor), then
// prolongation and post-smoothing back up; barriers separate the phases (one launch instead of
// ~10 per level).  Level q of C is levels[first + q]; the last one is the LU level.
template <class T, class FT>
__global__ void __launch_bounds__(COARSE_THREADS) k_coarse_cycle(CoarseLevels<T> C, int n_pre, int n_post,
                                                                const double* __restrict__ lu_scale,
                                                                const int* __restrict__ lu_perm,
                                                                const double* __restrict__ Lf,
                                                                const double* __restrict__ Uf) {
  const long tid = threadIdx.x, nth = blockDim.x;
  const int last = C.count - 1;
  for (int q = 0; q < last; ++q) {
    const Level& L = C.L[q];
    const FT* e = (const FT*)L.f_e;
    const FT* nn = (const FT*)L.f_n;
    const FT* t = (const FT*)L.f_t;
    const T* s = (const T*)L.s;
    for (int sw = 0; sw < n_pre; ++sw) {
      rbgs_body<T, FT>(C.x[q], C.b[q], e, nn, t, s, (T)L.fscale, L.d.N, L.d.M, L.d.K, 0, L.pin0, L.pin1,
                       sw == 0 ? 1 : 0, tid, nth);
      __syncthreads();
      rbgs_body<T, FT>(C.x[q], C.b[q], e, nn, t, s, (T)L.fscale, L.d.N, L.d.M, L.d.K, 1, L.pin0, L.pin1,
                       sw == 0 ? 2 : 0, tid, nth);
      __syncthreads();
    }
    resid_restrict_body<T, FT>(C.x[q], C.b[q], e, nn, t, s, (T)L.fscale, L.d.N, L.d.M, L.d.K, L.pin0, L.pin1,
                               C.b[q + 1], tid, nth);
    __syncthreads();
  }
  if (tid == 0) {
    const int n = (int)C.L[last].d.n();
    const T* bb = C.b[last];
    double bs[64], yv[64];
    for (int i = 0; i < n; ++i) bs[i] = lu_scale[lu_perm[i]] * (double)bb[lu_perm[i]];
    for (int i = 0; i < n; ++i) {
      double aux = bs[i];
      for (int k = 0; k < i; ++k) aux -= Lf[i * n + k] * yv[k];
      yv[i] = aux;
    }
    for (int i = n - 1; i >= 0; --i) {
      double aux = yv[i];
      for (int j = i + 1; j < n; ++j) aux -= Uf[i * n + j] * bs[j];
      bs[i] = aux / Uf[i * n + i];
    }
    for (int i = 0; i < n; ++i) C.x[last][i] = (T)(lu_scale[i] * bs[i]);
  }
  __syncthreads();
  for (int q = last - 1; q >= 0; --q) {
    const Level& L = C.L[q];
    const FT* e = (const FT*)L.f_e;
    const FT* nn = (const FT*)L.f_n;
    const FT* t = (const FT*)L.f_t;
    const T* s = (const T*)L.s;
    prolong_add_body<T>(C.x[q + 1], C.L[q + 1].d.N, C.L[q + 1].d.M, C.x[q], L.d.N, L.d.M, L.d.K, tid, nth);
    __syncthreads();
    for (int sw = 0; sw < n_post; ++sw) {
      rbgs_body<T, FT>(C.x[q], C.b[q], e, nn, t, s, (T)L.fscale, L.d.N, L.d.M, L.d.K, 0, L.pin0, L.pin1, 0, tid, nth);
      __syncthreads();
      rbgs_body<T, FT>(C.x[q], C.b[q], e, nn, t, s, (T)L.fscale, L.d.N, L.d.M, L.d.K, 1, L.pin0, L.pin1, 0, tid, nth);
      __syncthreads();
    }
  }
}

template <class A, class B>
__global__ void k_convert(const A* __restrict__ a, B* __restrict__ b, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = (B)a[i];
}

// ---- fused red-black sweep ---------------------------------------------------------------------
// One full red-black sweep in one pass, x_in -> x_out.  A workgroup owns a 64 x 16 tile of
// (x, y) columns and a run of FZ z planes and marches up in z: at plane p it updates the red
// cells of p over the tile and a one-cell ring around it (from x_in, the old values), then the
// black cells of plane p - 1 over the tile (from the red values of planes p - 2, p - 1, p, all
// new).  The ring's red values are recomputed by the neighbouring workgroups too, so no
// workgroup reads another's results and x_out is the same bit for bit as red and black
// half-sweeps in place: the same formula, the same operand order.  Each array is read about
// once instead of once per colour.
constexpr int FX = 64, FY = 16, FZ = 32;
constexpr int HX = FX + 4, HY = FY + 4;  // x_in staging: tile + 2-cell ring
constexpr int RX = FX + 2, RY = FY + 2;  // red values: tile + 1-cell ring

// per-cell data of one plane: right-hand side, 1/a_p and the six faces (absent ones read the cell
// itself and are never used)
template <class T>
struct RbCell {
  T b, s, fw, fe, fs, fn, fb, ft;
};

template <class T, class FT>
__device__ inline RbCell<T> rb_load(const T* __restrict__ b, const FT* __restrict__ ae, const FT* __restrict__ an,
                                    const FT* __restrict__ at, const T* __restrict__ s, bool ok, int gx, int gy,
                                    int p, int N, int M) {
  RbCell<T> d;
  if (!ok) {
    d.b = d.s = d.fw = d.fe = d.fs = d.fn = d.fb = d.ft = (T)0;
    return d;
  }
  const long NM = (long)N * M;
  const long i = gx + (long)N * gy + NM * p;
  d.b = b[i];
  d.s = s[i];
  d.fw = ld<T>(ae, gx > 0 ? i - 1 : i);
  d.fe = ld<T>(ae, i);
  d.fs = ld<T>(an, gy > 0 ? i - N : i);
  d.fn = ld<T>(an, i);
  d.fb = ld<T>(at, p > 0 ? i - NM : i);
  d.ft = ld<T>(at, i);
  return d;
}

// the red-black update of one cell from its neighbours' values (operand order of offdiag)
template <class T>
__device__ inline T rb_apply(const RbCell<T>& d, T fs, int gx, int gy, int p, int N, int M, int K, T xw, T xe, T xs,
                             T xn, T xb, T xt) {
  T acc = 0;
  if (gx > 0) acc += d.fw * xw;
  if (gx < N - 1) acc += d.fe * xe;
  if (gy > 0) acc += d.fs * xs;
  if (gy < M - 1) acc += d.fn * xn;
  if (p > 0) acc += d.fb * xb;
  if (p < K - 1) acc += d.ft * xt;
  return (d.b - fs * acc) * d.s;
}

// workgroup barrier for LDS traffic only: global loads in flight (the next plane's prefetch)
// are not waited for
__device__ inline void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

template <class T, class FT>
__global__ void __launch_bounds__(FX* FY) k_rb_fused(const T* __restrict__ x_in, T* __restrict__ x_out,
                                                      const T* __restrict__ b, const FT* __restrict__ ae,
                                                      const FT* __restrict__ an, const FT* __restrict__ at,
                                                      const T* __restrict__ s, T fs, int N, int M, int K, int pin0,
                                                      int pin1, int zero) {
  __shared__ T xs_[3][HY][HX];
  __shared__ T rs_[3][RY][RX];
  constexpr int NT = FX * FY;
  const int tid = threadIdx.x;
  const int X0 = blockIdx.x * FX, Y0 = blockIdx.y * FY, Z0 = blockIdx.z * FZ;
  const int Z1 = Z0 + FZ < K ? Z0 + FZ : K;
  const long NM = (long)N * M;
  auto slot = [](int p) { return ((p % 3) + 3) % 3; };
  // this thread's ring cells (tile + 1-cell ring, RX x RY) and staging cells (HX x HY)
  int cgx[2], cgy[2];
  bool cin[2], cint[2];
  for (int k = 0; k < 2; ++k) {
    const int c = tid + k * NT;
    const int rx = c % RX, ry = c / RX;
    cgx[k] = X0 - 1 + rx;
    cgy[k] = Y0 - 1 + ry;
    cin[k] = c < RX * RY && cgx[k] >= 0 && cgx[k] < N && cgy[k] >= 0 && cgy[k] < M;
    cint[k] = cin[k] && rx >= 1 && rx <= FX && ry >= 1 && ry <= FY;
  }
  int sgx[2], sgy[2];
  bool sin_[2];
  for (int k = 0; k < 2; ++k) {
    const int c = tid + k * NT;
    sgx[k] = X0 - 2 + c % HX;
    sgy[k] = Y0 - 2 + c / HX;
    sin_[k] = c < HX * HY && sgx[k] >= 0 && sgx[k] < N && sgy[k] >= 0 && sgy[k] < M;
  }
  auto xload = [&](int k, int q) -> T {
    return (!zero && sin_[k] && q >= 0 && q < K) ? x_in[sgx[k] + (long)N * sgy[k] + NM * q] : (T)0;
  };
  auto xstore = [&](int k, int q, T v) {
    if (sin_[k] && q >= 0 && q < K) xs_[slot(q)][sgy[k] - Y0 + 2][sgx[k] - X0 + 2] = v;
  };
  auto xin = [&](int q, int gx, int gy) -> T { return xs_[slot(q)][gy - Y0 + 2][gx - X0 + 2]; };
  auto red = [&](int q, int gx, int gy) -> T { return rs_[slot(q)][gy - Y0 + 1][gx - X0 + 1]; };
  auto dload = [&](int k, int q) {
    return rb_load<T, FT>(b, ae, an, at, s, cin[k] && q >= 0 && q < K, cgx[k], cgy[k], q, N, M);
  };

  // prologue: x_in planes Z0-2 .. Z0 staged, the data of planes Z0-1 (current) and Z0 (next)
  for (int k = 0; k < 2; ++k) {
    xstore(k, Z0 - 2, xload(k, Z0 - 2));
    xstore(k, Z0 - 1, xload(k, Z0 - 1));
  }
  T xr[2] = {xload(0, Z0), xload(1, Z0)};
  RbCell<T> dp[2], dc[2] = {dload(0, Z0 - 1), dload(1, Z0 - 1)};
  for (int p = Z0 - 1; p <= Z1; ++p) {
    // loads of the next step, in flight while this plane is computed
    const RbCell<T> dn[2] = {dload(0, p + 1), dload(1, p + 1)};
    const T xn2[2] = {xload(0, p + 2), xload(1, p + 2)};
    xstore(0, p + 1, xr[0]);
    xstore(1, p + 1, xr[1]);
    lds_barrier();
    // red cells of plane p over the tile and its ring, from the old values
    if (p >= 0 && p < K) {
      for (int k = 0; k < 2; ++k) {
        const int gx = cgx[k], gy = cgy[k];
        if (!cin[k] || ((gx + gy + p) & 1)) continue;
        const long i = gx + (long)N * gy + NM * p;
        T v;
        if (i == pin0 || i == pin1) {
          v = dc[k].b * dc[k].s;
        } else if (dc[k].s == (T)0) {
          v = zero ? (T)0 : xin(p, gx, gy);
        } else if (zero) {
          v = dc[k].b * dc[k].s;
        } else {
          const T xw = gx > 0 ? xin(p, gx - 1, gy) : (T)0, xe = gx < N - 1 ? xin(p, gx + 1, gy) : (T)0;
          const T xs = gy > 0 ? xin(p, gx, gy - 1) : (T)0, xn = gy < M - 1 ? xin(p, gx, gy + 1) : (T)0;
          const T xb = p > 0 ? xin(p - 1, gx, gy) : (T)0, xt = p < K - 1 ? xin(p + 1, gx, gy) : (T)0;
          v = rb_apply<T>(dc[k], fs, gx, gy, p, N, M, K, xw, xe, xs, xn, xb, xt);
        }
        rs_[slot(p)][gy - Y0 + 1][gx - X0 + 1] = v;
      }
    }
    lds_barrier();
    // plane p - 1 of the tile: black cells from the new red values, red cells as computed
    const int q = p - 1;
    if (q >= Z0 && q < Z1) {
      for (int k = 0; k < 2; ++k) {
        if (!cint[k]) continue;
        const int gx = cgx[k], gy = cgy[k];
        const long i = gx + (long)N * gy + NM * q;
        T v;
        if ((gx + gy + q) & 1) {
          if (i == pin0 || i == pin1) {
            v = dp[k].b * dp[k].s;
          } else if (dp[k].s == (T)0) {
            v = zero ? (T)0 : xin(q, gx, gy);
          } else {
            const T xw = gx > 0 ? red(q, gx - 1, gy) : (T)0, xe = gx < N - 1 ? red(q, gx + 1, gy) : (T)0;
            const T xs = gy > 0 ? red(q, gx, gy - 1) : (T)0, xn = gy < M - 1 ? red(q, gx, gy + 1) : (T)0;
            const T xb = q > 0 ? red(q - 1, gx, gy) : (T)0, xt = q < K - 1 ? red(q + 1, gx, gy) : (T)0;
            v = rb_apply<T>(dp[k], fs, gx, gy, q, N, M, K, xw, xe, xs, xn, xb, xt);
          }
        } else {
          v = red(q, gx, gy);
        }
        x_out[i] = v;
      }
    }
    // the next step overwrites slot (p + 2) % 3 of both rings
    lds_barrier();
    for (int k = 0; k < 2; ++k) {
      dp[k] = dc[k];
      dc[k] = dn[k];
      xr[k] = xn2[k];
    }
  }
}

template <class T>
__global__ void k_zero_empty(const T* __restrict__ s, double* __restrict__ x, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    if (s[i] == (T)0) x[i] = 0.0;
}

// ---- line relaxation ---------------------------------------------------------------------------
// One pass solves every line of direction d (0 x, 1 y, 2 z) exactly, with the couplings across
// lines taken from x_in (line Jacobi): x_out = T_d^{-1} (b - (A - T_d) x_in), T_d the tridiagonal
// part of A along d.  zero: x_in is 0 (first pass of a cycle).  A sweep is the three directions
// in turn; on grids stretched towards the body, where cells are long in one or two directions
// and point smoothers stall, every strong coupling is solved exactly in one of them.
//
// Row of cell i on its line: lower (coupling to the previous cell of the line), diag, upper,
// rhs.  Pinned rows are identity rows; an empty row (1/a_p stored as 0) keeps its value, as in
// the point smoother.
template <class T, class FT>
__device__ inline void line_row(int d, const T* __restrict__ x, const T* __restrict__ b, const FT* __restrict__ ae,
                                const FT* __restrict__ an, const FT* __restrict__ at, const T* __restrict__ s, T fs,
                                long i, int xx, int yy, int zz, int N, int M, int K, int pin0, int pin1, bool zero,
                                bool trans, T& lo, T& dg, T& up, T& rhs) {
  // Every load is issued unconditionally (absent neighbours read the cell itself and are masked
  // out), so that a caller evaluating many rows can have all their loads in flight at once.
  const long NM = (long)N * M;
  const bool own = i == pin0 || i == pin1;
  auto keep = [&](long j) { return !trans || own || (j != pin0 && j != pin1); };
  const bool hw = xx > 0 && keep(i - 1), he = xx < N - 1 && keep(i + 1);
  const bool hs = yy > 0 && keep(i - N), hn = yy < M - 1 && keep(i + N);
  const bool hb = zz > 0 && keep(i - NM), ht = zz < K - 1 && keep(i + NM);
  const long iw = xx > 0 ? i - 1 : i, ie = xx < N - 1 ? i + 1 : i;
  const long is = yy > 0 ? i - N : i, in = yy < M - 1 ? i + N : i;
  const long ib = zz > 0 ? i - NM : i, it = zz < K - 1 ? i + NM : i;
  const T si = s[i], bi = b[i];
  const T fw = fs * ld<T>(ae, iw), fe = fs * ld<T>(ae, i);
  const T fsf = fs * ld<T>(an, is), fn = fs * ld<T>(an, i);
  const T fb = fs * ld<T>(at, ib), ft = fs * ld<T>(at, i);
  const T cw = hw ? fw : (T)0, ce = he ? fe : (T)0, cs = hs ? fsf : (T)0, cn = hn ? fn : (T)0;
  const T cb = hb ? fb : (T)0, ct = ht ? ft : (T)0;
  T off = (T)0, xi = (T)0;
  if (!zero) {  // uniform over the launch
    xi = x[i];
    if (d == 0) off = cs * x[is] + cn * x[in] + cb * x[ib] + ct * x[it];
    else if (d == 1) off = cw * x[iw] + ce * x[ie] + cb * x[ib] + ct * x[it];
    else off = cw * x[iw] + ce * x[ie] + cs * x[is] + cn * x[in];
  }
  lo = d == 0 ? cw : (d == 1 ? cs : cb);
  up = d == 0 ? ce : (d == 1 ? cn : ct);
  dg = si != (T)0 ? (T)1 / si : (T)1;
  rhs = bi - off;
  if (own && !trans) {  // identity row
    lo = up = (T)0;
    dg = (T)1;
    rhs = bi;
  } else if (si == (T)0) {  // empty row: keeps its value
    lo = up = (T)0;
    dg = (T)1;
    rhs = xi;
  }
}

// x lines: one wavefront per line, CPL consecutive cells per lane (coalesced loads), solved by
// parallel cyclic reduction over the 64 * CPL (padded with identity rows) equations.
template <class T>
__device__ inline T wave_get(T v, int src, bool ok) {
  const T r = __shfl(v, ok ? src : 0, 64);
  return r;
}

template <class T, class FT, int CPL>
__global__ void __launch_bounds__(256) k_line_x(const T* __restrict__ x_in, T* __restrict__ x_out,
                                                const T* __restrict__ b, const FT* __restrict__ ae,
                                                const FT* __restrict__ an, const FT* __restrict__ at,
                                                const T* __restrict__ s, T fs, int N, int M, int K, int pin0,
                                                int pin1, int zero, int trans) {
  const int lane = threadIdx.x & 63;
  const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= (long)M * K) return;  // whole wavefronts leave together
  const int yy = (int)(row % M), zz = (int)(row / M);
  constexpr int L = 64 * CPL;
  T a[CPL], d[CPL], c[CPL], r[CPL];
#pragma unroll
  for (int j = 0; j < CPL; ++j) {
    const int k = lane * CPL + j;
    if (k < N) {
      line_row<T, FT>(0, x_in, b, ae, an, at, s, fs, row * N + k, k, yy, zz, N, M, K, pin0, pin1, zero != 0,
                      trans != 0, a[j], d[j], c[j], r[j]);
    } else {
      a[j] = c[j] = r[j] = (T)0;
      d[j] = (T)1;
    }
  }
#pragma unroll
  for (int st = 1; st < L; st <<= 1) {
    T na[CPL], nd[CPL], nc[CPL], nr[CPL];
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
      const int k = lane * CPL + j;
      // equation k - st
      T am, dm, cm, rm;
      if (st < CPL && j - st >= 0) {
        am = a[(j - st) & (CPL - 1)];
        dm = d[(j - st) & (CPL - 1)];
        cm = c[(j - st) & (CPL - 1)];
        rm = r[(j - st) & (CPL - 1)];
      } else {
        const int jj = st < CPL ? ((j - st + CPL) & (CPL - 1)) : j;
        const int src = st < CPL ? lane - 1 : lane - st / CPL;
        const bool ok = k - st >= 0;
        am = wave_get(a[jj], src, ok);
        dm = wave_get(d[jj], src, ok);
        cm = wave_get(c[jj], src, ok);
        rm = wave_get(r[jj], src, ok);
        if (!ok) {
          am = cm = rm = (T)0;
          dm = (T)1;
        }
      }
      // equation k + st
      T ap, dp, cp, rp;
      if (st < CPL && j + st < CPL) {
        ap = a[(j + st) & (CPL - 1)];
        dp = d[(j + st) & (CPL - 1)];
        cp = c[(j + st) & (CPL - 1)];
        rp = r[(j + st) & (CPL - 1)];
      } else {
        const int jj = st < CPL ? ((j + st) & (CPL - 1)) : j;
        const int src = st < CPL ? lane + 1 : lane + st / CPL;
        const bool ok = k + st < L;
        ap = wave_get(a[jj], src, ok);
        dp = wave_get(d[jj], src, ok);
        cp = wave_get(c[jj], src, ok);
        rp = wave_get(r[jj], src, ok);
        if (!ok) {
          ap = cp = rp = (T)0;
          dp = (T)1;
        }
      }
      const T al = dm != (T)0 ? -a[j] / dm : (T)0;
      const T ga = dp != (T)0 ? -c[j] / dp : (T)0;
      na[j] = al * am;
      nc[j] = ga * cp;
      nd[j] = d[j] + al * cm + ga * ap;
      nr[j] = r[j] + al * rm + ga * rp;
    }
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
      a[j] = na[j];
      d[j] = nd[j];
      c[j] = nc[j];
      r[j] = nr[j];
    }
  }
#pragma unroll
  for (int j = 0; j < CPL; ++j) {
    const int k = lane * CPL + j;
    if (k < N) x_out[row * N + k] = d[j] != (T)0 ? r[j] / d[j] : (T)0;
  }
}

// y and z lines: a workgroup holds 64 lines (adjacent x, so every load coalesces) split into S
// segments of SEG rows, one wavefront per segment (S = ceil(len / SEG) <= 16: lines of at most
// 256 cells).  Each thread eliminates its segment with the
// two couplings to the neighbouring segments kept as unknowns, x_k = G_k + HL_k x_L + UR_k x_R
// (x_L the last row of the segment before, x_R the first of the one after; all in registers);
// one thread per line then solves the 2S interface values in order through LDS, and every
// segment writes x_k.  Inputs are read once and x written once, with 64 S threads per 64 lines
// instead of 64.
constexpr int SEG = 16;

template <class T, class FT>
__global__ void __launch_bounds__(1024) k_line_seg(int dir, const T* __restrict__ x_in, T* __restrict__ x_out,
                                                   const T* __restrict__ b, const FT* __restrict__ ae,
                                                   const FT* __restrict__ an, const FT* __restrict__ at,
                                                   const T* __restrict__ s, T fs, int N, int M, int K, int pin0,
                                                   int pin1, int zero, int trans, int S) {
  extern __shared__ unsigned char lds_raw[];
  T* ifc = reinterpret_cast<T*>(lds_raw);  // [8][S][64]: Gs HLs URs Ge HLe URe, then L R
  auto IF = [&](int q, int j, int l) -> T& { return ifc[((long)q * S + j) * 64 + l]; };
  const long NM = (long)N * M;
  const int lane = threadIdx.x & 63, seg = threadIdx.x >> 6;
  // x lines: the 64 lanes take 64 adjacent rows (y); y and z lines: 64 adjacent x
  const int cross = blockIdx.x * 64 + lane, other = blockIdx.y;
  const int len = dir == 0 ? N : (dir == 1 ? M : K);
  const long stride = dir == 0 ? 1L : (dir == 1 ? (long)N : NM);
  const bool live = cross < (dir == 0 ? M : N);
  const long base = dir == 0 ? (long)N * cross + NM * other
                             : (dir == 1 ? cross + NM * other : cross + (long)N * other);
  const int k0 = seg * SEG;
  const int cnt = live ? (len - k0 < SEG ? len - k0 : SEG) : 0;
  T G[SEG], HL[SEG], UR[SEG];
  {
    T cprev = (T)0, gprev = (T)0, hprev = (T)0;
#pragma unroll
    for (int k = 0; k < SEG; ++k) {
      if (k < cnt) {
        const int kk = k0 + k;
        const long i = base + stride * kk;
        const int xx = dir == 0 ? kk : cross;
        const int yy = dir == 0 ? cross : (dir == 1 ? kk : other), zz = dir == 2 ? kk : other;
        T lo, dg, up, rhs;
        line_row<T, FT>(dir, x_in, b, ae, an, at, s, fs, i, xx, yy, zz, N, M, K, pin0, pin1, zero != 0, trans != 0,
                        lo, dg, up, rhs);
        const T w = dg - lo * cprev;
        const T winv = w != (T)0 ? (T)1 / w : (T)0;
        gprev = (rhs - lo * gprev) * winv;
        hprev = (k == 0 ? -lo : -lo * hprev) * winv;
        cprev = up * winv;
        G[k] = gprev;
        HL[k] = hprev;
        UR[k] = cprev;  // eliminated upper coefficient until the backward pass
      }
    }
    // backward: x_k = G_k + HL_k x_L + UR_k x_R
    T Gn = (T)0, Hn = (T)0, Un = (T)0;
#pragma unroll
    for (int k = SEG - 1; k >= 0; --k) {
      if (k < cnt) {
        const T cp = UR[k];
        if (k == cnt - 1) {
          Gn = G[k];
          Hn = HL[k];
          Un = -cp;
        } else {
          Gn = G[k] - cp * Gn;
          Hn = HL[k] - cp * Hn;
          Un = -cp * Un;
        }
        G[k] = Gn;
        HL[k] = Hn;
        UR[k] = Un;
      }
    }
  }
  if (cnt > 0) {
    IF(0, seg, lane) = G[0];
    IF(1, seg, lane) = HL[0];
    IF(2, seg, lane) = UR[0];
    T ge = G[0], he = HL[0], ue = UR[0];
#pragma unroll
    for (int k = 1; k < SEG; ++k)
      if (k == cnt - 1) {
        ge = G[k];
        he = HL[k];
        ue = UR[k];
      }
    IF(3, seg, lane) = ge;
    IF(4, seg, lane) = he;
    IF(5, seg, lane) = ue;
  }
  __syncthreads();
  if (seg == 0 && live) {
    // x_s(j) = gam + del x_s(j+1), x_e(j) = alp + bet x_s(j+1), left to right
    T alp = (T)0, bet = (T)0;
    for (int j = 0; j < S; ++j) {
      const T gs = IF(0, j, lane), hs = IF(1, j, lane), us = IF(2, j, lane);
      const T ge = IF(3, j, lane), he = IF(4, j, lane), ue = IF(5, j, lane);
      const T den = (T)1 - hs * bet;
      const T inv = den != (T)0 ? (T)1 / den : (T)0;
      const T gam = (gs + hs * alp) * inv, del = us * inv;
      const T alp_n = ge + he * (alp + bet * gam);
      const T bet_n = he * bet * del + ue;
      IF(0, j, lane) = gam;
      IF(1, j, lane) = del;
      IF(3, j, lane) = alp_n;
      IF(4, j, lane) = bet_n;
      alp = alp_n;
      bet = bet_n;
    }
    T xs_next = (T)0;  // x_s of the segment after the last one: none
    for (int j = S - 1; j >= 0; --j) {
      const T xe = IF(3, j, lane) + IF(4, j, lane) * xs_next;
      const T xsj = IF(0, j, lane) + IF(1, j, lane) * xs_next;
      IF(7, j, lane) = xs_next;  // R of segment j
      if (j + 1 < S) IF(6, j + 1, lane) = xe;  // L of segment j + 1
      xs_next = xsj;
    }
    IF(6, 0, lane) = (T)0;
  }
  __syncthreads();
  if (cnt > 0) {
    const T xl = IF(6, seg, lane), xr = IF(7, seg, lane);
#pragma unroll
    for (int k = 0; k < SEG; ++k)
      if (k < cnt) x_out[base + stride * (k0 + k)] = G[k] + HL[k] * xl + UR[k] * xr;
  }
}

}  // namespace

// ================================================================================================
void faces_from7(const double* c7, Dims d, int pin0, int pin1, double* fe, double* fn, double* ft, double* dg,
                 hipStream_t s) {
  k_faces_from7<<<grid_1d(d.n()), 256, 0, s>>>(c7, d.N, d.M, d.K, pin0, pin1, fe, fn, ft, dg);
}

void aggregate(const double* fe, const double* fn, const double* ft, Dims df, double* ce, double* cn, double* ct,
               Dims dc, hipStream_t s) {
  k_aggregate<<<grid_1d(dc.n()), 256, 0, s>>>(fe, fn, ft, df.N, df.M, ce, cn, ct, dc.N, dc.M, dc.K);
}

void diag_from_faces(const double* fe, const double* fn, const double* ft, Dims d, double* dg, hipStream_t s) {
  k_diag_from_faces<<<grid_1d(d.n()), 256, 0, s>>>(fe, fn, ft, d.N, d.M, d.K, dg);
}

void pin_sigma(const double* fe, const double* fn, const double* ft, Dims d, int pin0, int pin1, double* sigma,
               hipStream_t s) {
  k_pin_sigma<<<1, 64, 0, s>>>(fe, fn, ft, d.N, d.M, d.K, pin0, pin1, sigma);
}

void pin_sink(double* dg, int c0, int c1, const double* sigma, hipStream_t s) {
  k_pin_sink<<<1, 64, 0, s>>>(dg, c0, c1, sigma);
}

void face_absmax(const double* fe, const double* fn, const double* ft, long n, double* partial, double* out,
                 hipStream_t s) {
  int nb = (int)((n + 1023) / 1024);
  if (nb > 1024) nb = 1024;
  if (nb < 1) nb = 1;
  k_absmax3<<<nb, 256, 0, s>>>(fe, fn, ft, n, partial);
  k_max_reduce<<<1, 256, 0, s>>>(partial, nb, out);
}

template <class T>
void inv_diag(const double* dg, long n, T* sv, hipStream_t s) {
  k_inv_diag<T><<<grid_1d(n), 256, 0, s>>>(dg, n, sv);
}
template void inv_diag<float>(const double*, long, float*, hipStream_t);
template void inv_diag<double>(const double*, long, double*, hipStream_t);

template <class FT>
void store_faces(const double* fe, const double* fn, const double* ft, long n, double inv_scale, FT* ae, FT* an,
                 FT* at, hipStream_t s) {
  k_store_faces<FT><<<grid_1d(n), 256, 0, s>>>(fe, fn, ft, n, inv_scale, ae, an, at);
}
template void store_faces<double>(const double*, const double*, const double*, long, double, double*, double*,
                                  double*, hipStream_t);
template void store_faces<float>(const double*, const double*, const double*, long, double, float*, float*, float*,
                                 hipStream_t);
template void store_faces<_Float16>(const double*, const double*, const double*, long, double, _Float16*, _Float16*,
                                    _Float16*, hipStream_t);

void to27(const double* fe, const double* fn, const double* ft, const double* dg, Dims d, int pin0, int pin1,
          double* c27, hipStream_t s) {
  k_to27<<<grid_1d(d.n()), 256, 0, s>>>(fe, fn, ft, dg, d.N, d.M, d.K, pin0, pin1, c27);
}

template <class T, class FT>
static void rbgs_t(const Level& L, T* x, const T* b, int color, int zero_guess, hipStream_t s) {
  const long total = (long)((L.d.N + 1) / 2) * L.d.M * L.d.K;
  k_rbgs<T, FT><<<grid_1d(total), 256, 0, s>>>(x, b, (const FT*)L.f_e, (const FT*)L.f_n, (const FT*)L.f_t,
                                               (const T*)L.s, (T)L.fscale, L.d.N, L.d.M, L.d.K, color, L.pin0,
                                               L.pin1, zero_guess);
}

template <>
void rbgs<double>(const Level& L, double* x, const double* b, int color, int zero_guess, hipStream_t s) {
  rbgs_t<double, double>(L, x, b, color, zero_guess, s);
}
template <>
void rbgs<float>(const Level& L, float* x, const float* b, int color, int zero_guess, hipStream_t s) {
  if (L.ftype == FACE_F16)
    rbgs_t<float, _Float16>(L, x, b, color, zero_guess, s);
  else
    rbgs_t<float, float>(L, x, b, color, zero_guess, s);
}

template <class T, class FT>
static void resid_restrict_t(const Level& L, const T* x, const T* b, T* bc, hipStream_t s) {
  const long nc = (long)(L.d.N / 2) * (L.d.M / 2) * (L.d.K / 2);
  k_resid_restrict<T, FT><<<grid_1d(nc), 256, 0, s>>>(x, b, (const FT*)L.f_e, (const FT*)L.f_n, (const FT*)L.f_t,
                                                      (const T*)L.s, (T)L.fscale, L.d.N, L.d.M, L.d.K, L.pin0,
                                                      L.pin1, bc);
}

template <>
void resid_restrict<double>(const Level& L, const double* x, const double* b, double* bc, hipStream_t s) {
  resid_restrict_t<double, double>(L, x, b, bc, s);
}
template <>
void resid_restrict<float>(const Level& L, const float* x, const float* b, float* bc, hipStream_t s) {
  if (L.ftype == FACE_F16)
    resid_restrict_t<float, _Float16>(L, x, b, bc, s);
  else
    resid_restrict_t<float, float>(L, x, b, bc, s);
}

template <class T, class FT>
static void rbgs_tr_t(const Level& L, T* x, const T* b, int color, int zero_guess, hipStream_t s) {
  const long total = (long)((L.d.N + 1) / 2) * L.d.M * L.d.K;
  k_rbgs_tr<T, FT><<<grid_1d(total), 256, 0, s>>>(x, b, (const FT*)L.f_e, (const FT*)L.f_n, (const FT*)L.f_t,
                                                  (const T*)L.s, (T)L.fscale, L.d.N, L.d.M, L.d.K, color, L.pin0,
                                                  L.pin1, zero_guess);
}
template <class T, class FT>
static void resid_tr_t(const Level& L, const T* x, const T* b, T* r, hipStream_t s) {
  k_resid_tr<T, FT><<<grid_1d(L.d.n()), 256, 0, s>>>(x, b, (const FT*)L.f_e, (const FT*)L.f_n, (const FT*)L.f_t,
                                                     (const T*)L.s, (T)L.fscale, L.d.N, L.d.M, L.d.K, L.pin0,
                                                     L.pin1, r);
}

template <>
void rbgs_tr<double>(const Level& L, double* x, const double* b, int color, int zero_guess, hipStream_t s) {
  rbgs_tr_t<double, double>(L, x, b, color, zero_guess, s);
}
template <>
void rbgs_tr<float>(const Level& L, float* x, const float* b, int color, int zero_guess, hipStream_t s) {
  if (L.ftype == FACE_F16)
    rbgs_tr_t<float, _Float16>(L, x, b, color, zero_guess, s);
  else
    rbgs_tr_t<float, float>(L, x, b, color, zero_guess, s);
}
template <>
void resid_tr<double>(const Level& L, const double* x, const double* b, double* r, hipStream_t s) {
  resid_tr_t<double, double>(L, x, b, r, s);
}
template <>
void resid_tr<float>(const Level& L, const float* x, const float* b, float* r, hipStream_t s) {
  if (L.ftype == FACE_F16)
    resid_tr_t<float, _Float16>(L, x, b, r, s);
  else
    resid_tr_t<float, float>(L, x, b, r, s);
}

template <class T>
void restrict_pt(const T* r, Dims df, T* bc, hipStream_t s) {
  const long nc = (long)(df.N / 2) * (df.M / 2) * (df.K / 2);
  k_restrict_pt<T><<<grid_1d(nc), 256, 0, s>>>(r, df.N, df.M, df.K, bc);
}
template <class T>
void prolong_inject_add(const T* xc, Dims dc, T* xf, Dims df, hipStream_t s) {
  k_prolong_inject_add<T><<<grid_1d(df.n()), 256, 0, s>>>(xc, dc.N, dc.M, xf, df.N, df.M, df.K);
}
template void restrict_pt<float>(const float*, Dims, float*, hipStream_t);
template void restrict_pt<double>(const double*, Dims, double*, hipStream_t);
template void prolong_inject_add<float>(const float*, Dims, float*, Dims, hipStream_t);
template void prolong_inject_add<double>(const double*, Dims, double*, Dims, hipStream_t);

template <class T, class FT>
static void line_pass_t(const Level& L, int dir, const T* x_in, T* x_out, T* cpv, const T* b, int zero, int trans,
                        hipStream_t s) {
  const Dims& d = L.d;
  const FT* ae = (const FT*)L.f_e;
  const FT* an = (const FT*)L.f_n;
  const FT* at = (const FT*)L.f_t;
  const T* sv = (const T*)L.s;
  const T fs = (T)L.fscale;
  if (dir == 0) {
    const int blocks = (int)(((long)d.M * d.K + 3) / 4);
    if (d.N <= 64)
      k_line_x<T, FT, 1><<<blocks, 256, 0, s>>>(x_in, x_out, b, ae, an, at, sv, fs, d.N, d.M, d.K, L.pin0, L.pin1, zero,
                                                   trans);
    else if (d.N <= 128)
      k_line_x<T, FT, 2><<<blocks, 256, 0, s>>>(x_in, x_out, b, ae, an, at, sv, fs, d.N, d.M, d.K, L.pin0, L.pin1, zero,
                                                   trans);
    else if (d.N <= 256)
      k_line_x<T, FT, 4><<<blocks, 256, 0, s>>>(x_in, x_out, b, ae, an, at, sv, fs, d.N, d.M, d.K, L.pin0, L.pin1, zero,
                                                   trans);
    else
      k_line_x<T, FT, 8><<<blocks, 256, 0, s>>>(x_in, x_out, b, ae, an, at, sv, fs, d.N, d.M, d.K, L.pin0, L.pin1, zero,
                                                   trans);
  } else {
    const int len = dir == 0 ? d.N : (dir == 1 ? d.M : d.K);
    const int S = (len + SEG - 1) / SEG;
    const dim3 grid(((dir == 0 ? d.M : d.N) + 63) / 64, dir == 1 ? d.K : (dir == 0 ? d.K : d.M));
    (void)cpv;
    k_line_seg<T, FT><<<grid, 64 * S, 8 * S * 64 * sizeof(T), s>>>(dir, x_in, x_out, b, ae, an, at, sv, fs, d.N, d.M,
                                                                   d.K, L.pin0, L.pin1, zero, trans, S);
  }
}

template <>
void line_pass<double>(const Level& L, int dir, const double* x_in, double* x_out, double* cpv, const double* b,
                       int zero, int trans, hipStream_t s) {
  line_pass_t<double, double>(L, dir, x_in, x_out, cpv, b, zero, trans, s);
}
template <>
void line_pass<float>(const Level& L, int dir, const float* x_in, float* x_out, float* cpv, const float* b,
                      int zero, int trans, hipStream_t s) {
  if (L.ftype == FACE_F16)
    line_pass_t<float, _Float16>(L, dir, x_in, x_out, cpv, b, zero, trans, s);
  else
    line_pass_t<float, float>(L, dir, x_in, x_out, cpv, b, zero, trans, s);
}

template <class T, class FT>
static void rb_sweep_t(const Level& L, const T* x_in, T* x_out, const T* b, int zero, hipStream_t s) {
  const Dims& d = L.d;
  const dim3 grid((d.N + FX - 1) / FX, (d.M + FY - 1) / FY, (d.K + FZ - 1) / FZ);
  k_rb_fused<T, FT><<<grid, FX * FY, 0, s>>>(x_in, x_out, b, (const FT*)L.f_e, (const FT*)L.f_n, (const FT*)L.f_t,
                                              (const T*)L.s, (T)L.fscale, d.N, d.M, d.K, L.pin0, L.pin1, zero);
}
template <>
void rb_sweep<double>(const Level& L, const double* x_in, double* x_out, const double* b, int zero, hipStream_t s) {
  rb_sweep_t<double, double>(L, x_in, x_out, b, zero, s);
}
template <>
void rb_sweep<float>(const Level& L, const float* x_in, float* x_out, const float* b, int zero, hipStream_t s) {
  if (L.ftype == FACE_F16)
    rb_sweep_t<float, _Float16>(L, x_in, x_out, b, zero, s);
  else
    rb_sweep_t<float, float>(L, x_in, x_out, b, zero, s);
}

template <class T>
void zero_empty(const Level& L, double* x, hipStream_t s) {
  k_zero_empty<T><<<grid_1d(L.d.n()), 256, 0, s>>>((const T*)L.s, x, L.d.n());
}
template void zero_empty<float>(const Level&, double*, hipStream_t);
template void zero_empty<double>(const Level&, double*, hipStream_t);

template <class T>
void copy_vec(const T* a, T* b, long n, hipStream_t s) {
  k_convert<T, T><<<grid_1d(n), 256, 0, s>>>(a, b, n);
}
template void copy_vec<float>(const float*, float*, long, hipStream_t);
template void copy_vec<double>(const double*, double*, long, hipStream_t);

template <class T>
void coarse_cycle(const CoarseLevels<T>& C, int n_pre, int n_post, const double* lu_scale, const int* lu_perm,
                  const double* L, const double* U, hipStream_t s) {
  if (C.count < 2 || C.count > COARSE_MAX) return;
  if (C.L[0].ftype == FACE_F16)
    k_coarse_cycle<T, _Float16><<<1, COARSE_THREADS, 0, s>>>(C, n_pre, n_post, lu_scale, lu_perm, L, U);
  else if (C.L[0].ftype == FACE_F32)
    k_coarse_cycle<T, float><<<1, COARSE_THREADS, 0, s>>>(C, n_pre, n_post, lu_scale, lu_perm, L, U);
  else
    k_coarse_cycle<T, double><<<1, COARSE_THREADS, 0, s>>>(C, n_pre, n_post, lu_scale, lu_perm, L, U);
}

template <class T>
void prolong_add(const T* xc, Dims dc, T* xf, Dims df, hipStream_t s) {
  k_prolong_add<T><<<grid_1d(df.n()), 256, 0, s>>>(xc, dc.N, dc.M, xf, df.N, df.M, df.K);
}
template void prolong_add<float>(const float*, Dims, float*, Dims, hipStream_t);
template void prolong_add<double>(const double*, Dims, double*, Dims, hipStream_t);
template void coarse_cycle<float>(const CoarseLevels<float>&, int, int, const double*, const int*, const double*,
                                  const double*, hipStream_t);
template void coarse_cycle<double>(const CoarseLevels<double>&, int, int, const double*, const int*, const double*,
                                   const double*, hipStream_t);

template <class A, class B>
void convert(const A* a, B* b, long n, hipStream_t s) {
  k_convert<A, B><<<grid_1d(n), 256, 0, s>>>(a, b, n);
}
template void convert<double, float>(const double*, float*, long, hipStream_t);
template void convert<float, double>(const float*, double*, long, hipStream_t);

}  // namespace vc
}  // namespace lsx
