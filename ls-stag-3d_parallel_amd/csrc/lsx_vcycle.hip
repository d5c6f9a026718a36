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

#include <algorithm>
#include <type_traits>

#include "lsx_vcycle.hpp"

namespace lsx {
namespace vc {

namespace {

// cell index -> (x, y, z) in 32-bit arithmetic (every level has fewer than 2^31 cells: the
// CSR boundary indexes rows with 32-bit integers); 64-bit division is a long instruction sequence
__device__ inline void decode(long i, int N, int M, int* x, int* y, int* z) {
  const unsigned ii = (unsigned)i, NM = (unsigned)N * (unsigned)M;
  const unsigned zz = ii / NM, r = ii - zz * NM, yy = r / (unsigned)N;
  *z = (int)zz;
  *y = (int)yy;
  *x = (int)(r - yy * (unsigned)N);
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

// Workgroup b of a 1-D grid of nb runs on XCD b % 8 (the 8 XCDs take workgroups round-robin,
// each XCD with its own L2).  xcd_order(b, nb) renumbers them so that XCD x takes a contiguous
// run of logical blocks: a bijection of [0, nb).  It pays where a block reads halo data that
// its logical neighbours read too (the fused sweep, the y/z line passes, the residual
// restriction); for the plain stencil products and the x-line pass it measured slower and is
// not used.
__device__ inline int xcd_order(int b, int nb) {
  const int q = nb / 8, r = nb % 8, x = b % 8, k = b / 8;
  return x * q + (x < r ? x : r) + k;
}

// a_p of a non-pinned row of a Level::dff level from the sum of its (scaled, present) faces
// (own_e: the row's own east face as stored, -0.0 marking a decoupled row whose a_p is dz)
template <class T>
struct DiagRule {
  int c0, c1;  // sink cells (-1: none)
  T s0, s1, dz;
  __device__ inline T of(T fsum, long i, T own_e) const { return of_marked(fsum, i, __builtin_signbit(own_e)); }
  __device__ inline T of_marked(T fsum, long i, bool marked) const {
    T d = -fsum;
    d = i == c0 ? d - s0 : d;
    d = i == c1 ? d - s1 : d;
    return d == (T)0 && marked ? dz : d;
  }
};
template <class T>
inline DiagRule<T> diag_rule(const Level& L) {
  return DiagRule<T>{L.sink0, L.sink1, (T)L.sig0, (T)L.sig1, (T)L.dz};
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

// Coarse cell (X, Y, Z) covers fine cells 2X, 2X + 1 (and so on per axis); on an axis of odd
// length the last coarse cell has the single child 2X, and the coarse faces sum the fine faces
// that exist (the same terms in the same order as on even levels).
__global__ void k_aggregate(const double* __restrict__ fe, const double* __restrict__ fn, const double* __restrict__ ft,
                            int Nf, int Mf, int Kf, double* __restrict__ ce, double* __restrict__ cn,
                            double* __restrict__ ct, int Nc, int Mc, int Kc) {
  const long nc = (long)Nc * Mc * Kc, NMf = (long)Nf * Mf;
  for (long I = blockIdx.x * (long)blockDim.x + threadIdx.x; I < nc; I += (long)gridDim.x * blockDim.x) {
    int X, Y, Z;
    decode(I, Nc, Mc, &X, &Y, &Z);
    const long a = 2L * X + (long)Nf * 2 * Y + NMf * 2 * Z;
    const bool hx = 2 * X + 1 < Nf, hy = 2 * Y + 1 < Mf, hz = 2 * Z + 1 < Kf;  // second child present
    double se = 0, sn = 0, st = 0;
    if (X < Nc - 1) {
      const long f = a + 1;
      se = fe[f];
      if (hy) se = se + fe[f + Nf];
      if (hz) se = se + fe[f + NMf];
      if (hy && hz) se = se + fe[f + Nf + NMf];
    }
    if (Y < Mc - 1) {
      const long f = a + Nf;
      sn = fn[f];
      if (hx) sn = sn + fn[f + 1];
      if (hz) sn = sn + fn[f + NMf];
      if (hx && hz) sn = sn + fn[f + 1 + NMf];
    }
    if (Z < Kc - 1) {
      const long f = a + NMf;
      st = ft[f];
      if (hx) st = st + ft[f + 1];
      if (hy) st = st + ft[f + Nf];
      if (hx && hy) st = st + ft[f + 1 + Nf];
    }
    ce[I] = 0.5 * se;
    cn[I] = 0.5 * sn;
    ct[I] = 0.5 * st;
  }
}

// dg = -(sum of the six faces), plus ex[i] when ex is given (the aggregated excess of an operator
// with a mass term); ex = null leaves the sum as it is (-0.0 included)
__global__ void k_diag_from_faces(const double* __restrict__ fe, const double* __restrict__ fn,
                                  const double* __restrict__ ft, int N, int M, int K, const double* __restrict__ ex,
                                  double* __restrict__ dg) {
  const long n = (long)N * M * K, NM = (long)N * M;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    double d = -fe[i] - fn[i] - ft[i];
    if (x > 0) d -= fe[i - 1];
    if (y > 0) d -= fn[i - N];
    if (z > 0) d -= ft[i - NM];
    if (ex) d += ex[i];
    dg[i] = d;
  }
}

// excess of the finest level: ex = a_p + (sum of the six faces), the part of the diagonal that
// is not a coupling (the mass and wall terms of a momentum operator)
__global__ void k_excess(const double* __restrict__ fe, const double* __restrict__ fn, const double* __restrict__ ft,
                         const double* __restrict__ dg, int N, int M, int K, double* __restrict__ ex) {
  const long n = (long)N * M * K, NM = (long)N * M;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    double f = fe[i] + fn[i] + ft[i];
    if (x > 0) f += fe[i - 1];
    if (y > 0) f += fn[i - N];
    if (z > 0) f += ft[i - NM];
    ex[i] = dg[i] + f;
  }
}

// coarse excess: the sum over the children present (the diagonal block of P^T A P with the
// piecewise-constant prolongation, less its couplings)
__global__ void k_aggregate_cells(const double* __restrict__ ef, int Nf, int Mf, int Kf, double* __restrict__ ec,
                                  int Nc, int Mc, int Kc) {
  const long nc = (long)Nc * Mc * Kc, NMf = (long)Nf * Mf;
  for (long I = blockIdx.x * (long)blockDim.x + threadIdx.x; I < nc; I += (long)gridDim.x * blockDim.x) {
    int X, Y, Z;
    decode(I, Nc, Mc, &X, &Y, &Z);
    const long a = 2L * X + (long)Nf * 2 * Y + NMf * 2 * Z;
    const bool hx = 2 * X + 1 < Nf, hy = 2 * Y + 1 < Mf, hz = 2 * Z + 1 < Kf;
    double s = 0.0;
    for (int dz = 0; dz <= (int)hz; ++dz)
      for (int dy = 0; dy <= (int)hy; ++dy)
        for (int dx = 0; dx <= (int)hx; ++dx) s += ef[a + dx + (long)Nf * dy + NMf * dz];
    ec[I] = s;
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

// Level::dff set-up.  Row i's rebuilt diagonal: d = -(sum of its present faces) - sigma (sink
// cells); a decoupled row (d = 0) is an empty row (a_p = 0) or, when its own east face is stored
// as -0.0, a row whose a_p is the level's common value dz (the reference's solid cells carry a
// tiny positive diagonal).  mode 0: v_i = |a_p| of the decoupled non-pinned rows (absmax: dz);
// mode 1: v_i = deviation of a_p from the rule (0 on pinned rows; decoupled rows 0 when a_p is 0
// or dz, else 1; with hscale > 0, 1 on a coupled row whose faces all round to 0 as FP16 faces
// stored times hscale); mode 2: fe_i = -0.0 on the decoupled rows with a_p = dz != 0.
__device__ inline double face_sum(const double* __restrict__ fe, const double* __restrict__ fn,
                                  const double* __restrict__ ft, long i, int x, int y, int z, int N, long NM,
                                  double hscale, bool* under) {
  double f = fe[i] + fn[i] + ft[i], h = 0.0;
  auto q = [&](double v) { return (double)(_Float16)(v * hscale); };
  if (hscale > 0.0) h = q(fe[i]) + q(fn[i]) + q(ft[i]);
  if (x > 0) {
    f += fe[i - 1];
    if (hscale > 0.0) h += q(fe[i - 1]);
  }
  if (y > 0) {
    f += fn[i - N];
    if (hscale > 0.0) h += q(fn[i - N]);
  }
  if (z > 0) {
    f += ft[i - NM];
    if (hscale > 0.0) h += q(ft[i - NM]);
  }
  *under = hscale > 0.0 && f != 0.0 && h == 0.0;
  return f;
}

__global__ void k_diag_rule(double* __restrict__ fe, const double* __restrict__ fn, const double* __restrict__ ft,
                            const double* __restrict__ dg, int N, int M, int K, int pin0, int pin1, int c0, int c1,
                            double s0, double s1, double dz, double hscale, int mode, double* __restrict__ v) {
  const long n = (long)N * M * K, NM = (long)N * M;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    bool under;
    double d = -face_sum(fe, fn, ft, i, x, y, z, N, NM, hscale, &under);
    if (i == c0) d -= s0;
    if (i == c1) d -= s1;
    const double a = dg[i];
    const bool pin = i == pin0 || i == pin1;
    if (mode == 0) {
      v[i] = !pin && d == 0.0 ? fabs(a) : 0.0;
    } else if (mode == 1) {
      double e;
      if (pin)
        e = 0.0;
      else if (d == 0.0)
        e = a == 0.0 || a == dz ? 0.0 : 1.0;
      else if (a == 0.0 || under)
        e = 1.0;
      else
        e = fabs(a - d) / fabs(a);
      v[i] = e;
    } else if (!pin && d == 0.0 && dz != 0.0 && a == dz) {
      fe[i] = -0.0;
    }
  }
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
// sum over the present neighbours j of a_ij x_j in the order w, e, s, n, b, t, with every load
// issued unconditionally (absent neighbours read the cell itself) and absent terms leaving the sum
// unchanged, so that the row's loads are in flight together
template <class T, class FT>
__device__ inline T offdiag_u(const T* __restrict__ x, const FT* __restrict__ ae, const FT* __restrict__ an,
                              const FT* __restrict__ at, long i, int xx, int yy, int zz, int N, int M, int K) {
  const long NM = (long)N * M;
  const bool hw = xx > 0, he = xx < N - 1, hs = yy > 0, hn = yy < M - 1, hb = zz > 0, ht = zz < K - 1;
  const long iw = hw ? i - 1 : i, ie = he ? i + 1 : i, is = hs ? i - N : i, in = hn ? i + N : i;
  const long ib = hb ? i - NM : i, it = ht ? i + NM : i;
  const T fw = ld<T>(ae, iw), fe = ld<T>(ae, i), fs_ = ld<T>(an, is), fn = ld<T>(an, i);
  const T fb = ld<T>(at, ib), ft = ld<T>(at, i);
  const T xw = x[iw], xe = x[ie], xs = x[is], xn = x[in], xb = x[ib], xt = x[it];
  T acc = 0;
  acc = hw ? acc + fw * xw : acc;
  acc = he ? acc + fe * xe : acc;
  acc = hs ? acc + fs_ * xs : acc;
  acc = hn ? acc + fn * xn : acc;
  acc = hb ? acc + fb * xb : acc;
  acc = ht ? acc + ft * xt : acc;
  return acc;
}

// offdiag_u with the sum of the present faces and the row's own east face as stored (the inputs
// of DiagRule::of)
template <class T, class FT>
__device__ inline T offdiag_fsum(const T* __restrict__ x, const FT* __restrict__ ae, const FT* __restrict__ an,
                                 const FT* __restrict__ at, long i, int xx, int yy, int zz, int N, int M, int K,
                                 T& fsum, T& own_e) {
  const long NM = (long)N * M;
  const bool hw = xx > 0, he = xx < N - 1, hs = yy > 0, hn = yy < M - 1, hb = zz > 0, ht = zz < K - 1;
  const long iw = hw ? i - 1 : i, ie = he ? i + 1 : i, is = hs ? i - N : i, in = hn ? i + N : i;
  const long ib = hb ? i - NM : i, it = ht ? i + NM : i;
  const T fw0 = ld<T>(ae, iw), fe0 = ld<T>(ae, i), fs0 = ld<T>(an, is), fn0 = ld<T>(an, i);
  const T fb0 = ld<T>(at, ib), ft0 = ld<T>(at, i);
  const T xw = x[iw], xe = x[ie], xs = x[is], xn = x[in], xb = x[ib], xt = x[it];
  const T fw = hw ? fw0 : (T)0, fe = he ? fe0 : (T)0, fs_ = hs ? fs0 : (T)0, fn = hn ? fn0 : (T)0;
  const T fb = hb ? fb0 : (T)0, ft = ht ? ft0 : (T)0;
  fsum = fw + fe + fs_ + fn + fb + ft;
  own_e = fe0;
  return fw * xw + fe * xe + fs_ * xs + fn * xn + fb * xb + ft * xt;
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
  // 32-bit index arithmetic (a level has fewer than 2^31 cells); the neighbours' loads are issued
  // unconditionally (offdiag_u: the same sum in the same order)
  const unsigned half_row = (unsigned)(N + 1) / 2;
  const unsigned total = half_row * (unsigned)M * (unsigned)K;
  for (long t = start; t < (long)total; t += stride) {
    const unsigned tt = (unsigned)t, row = tt / half_row;
    const int yy = (int)(row % (unsigned)M), zz = (int)(row / (unsigned)M);
    const int xx = 2 * (int)(tt - row * half_row) + ((yy + zz + color) & 1);
    if (xx >= N) continue;
    const long i = (long)xx + (long)N * row;
    const T si = s[i], bi = b[i];
    T v;
    if (i == pin0 || i == pin1) {
      v = bi * si;
    } else if (si == (T)0) {
      v = zero_guess ? (T)0 : x[i];  // empty row: keeps its value (0 from a zero start)
    } else if (zero_guess == 1) {
      v = bi * si;
    } else {
      v = (bi - fs * offdiag_u<T, FT>(x, ae, an, at, i, xx, yy, zz, N, M, K)) * si;
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

// bc = sum over the children of (b - A x): one thread per fine cell (64 adjacent x by 2 y by 2 z
// per workgroup, so loads coalesce and each row's loads are in flight together), residuals
// through LDS, and the 8 children of each of the 32 coarse cells summed by one thread in the
// order x, then y, then z fastest-first (the order of the one-workgroup coarse cycle, so the two
// are bit-identical).  On an axis of odd length the last coarse cell has one child: the absent
// one adds 0.  DFF: a_p rebuilt from the faces (Level::dff), s not read; no bitwise twin then.
template <class T, class FT, bool DFF>
__global__ void __launch_bounds__(256) k_resid_restrict_cells(const T* __restrict__ x, const T* __restrict__ b,
                                                              const FT* __restrict__ ae, const FT* __restrict__ an,
                                                              const FT* __restrict__ at, const T* __restrict__ s,
                                                              T fs, int N, int M, int K, int pin0, int pin1,
                                                              T* __restrict__ bc, DiagRule<T> dr) {
  __shared__ T rr[2][2][64];
  const int tx = threadIdx.x & 63, ty = (threadIdx.x >> 6) & 1, tz = threadIdx.x >> 7;
  const int Nc = (N + 1) / 2, Mc = (M + 1) / 2;
  // block (bx, by, bz) of a 1-D grid, bx fastest, in XCD-contiguous order (neighbouring rows and
  // planes read by the same XCD)
  const int gx = (N + 63) / 64;
  const int lb = xcd_order(blockIdx.x, gridDim.x);
  const int bx = lb % gx, by = (lb / gx) % Mc, bz = (lb / gx) / Mc;
  const int xx = bx * 64 + tx, yy = 2 * by + ty, zz = 2 * bz + tz;
  // a thread past the level (x past N, or the absent second child on an odd axis) reads its
  // clamped neighbour and posts 0
  const bool live = xx < N && yy < M && zz < K;
  const int xc = xx < N ? xx : N - 1, yc = yy < M ? yy : M - 1, zc = zz < K ? zz : K - 1;
  const long i = xc + (long)N * yc + (long)N * M * zc;
  const bool pin = i == pin0 || i == pin1;
  T ap, od;
  if constexpr (DFF) {
    T fsum, own_e;
    od = offdiag_fsum<T, FT>(x, ae, an, at, i, xc, yc, zc, N, M, K, fsum, own_e);
    ap = pin ? (T)1 : dr.of(fs * fsum, i, own_e);
  } else {
    const T si = s[i];
    ap = si != (T)0 ? (T)1 / si : (T)0;
    od = offdiag_u<T, FT>(x, ae, an, at, i, xc, yc, zc, N, M, K);
  }
  const T r0 = b[i] - ap * x[i];
  rr[tz][ty][tx] = !live ? (T)0 : (pin ? r0 : r0 - fs * od);
  __syncthreads();
  if (threadIdx.x < 32) {
    const int X = bx * 32 + threadIdx.x;
    if (X < Nc) {
      T acc = 0;
#pragma unroll
      for (int c = 0; c < 8; ++c) acc += rr[c >> 2][(c >> 1) & 1][2 * threadIdx.x + (c & 1)];
      bc[X + (long)Nc * by + (long)Nc * Mc * bz] = acc;
    }
  }
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
  const int Nc = (Nf + 1) / 2, Mc = (Mf + 1) / 2, Kc = (Kf + 1) / 2;
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
__device__ inline void prolong_cell(const T* __restrict__ xc, int Nc, int Mc, T* __restrict__ xf, int Nf, int Mf,
                                    int Kf, long i, int x, int y, int z) {
  const long NMc = (long)Nc * Mc;
  {
    int cx[2], cy[2], cz[2], nx, ny, nz;
    float wx[2], wy[2], wz[2];
    axis_w(x, Nf, cx, wx, &nx);
    axis_w(y, Mf, cy, wy, &ny);
    axis_w(z, Kf, cz, wz, &nz);
    // always two taps per axis (an absent second tap reads the first with weight 0), so that the
    // eight loads are straight-line code
    if (nx < 2) { cx[1] = cx[0]; wx[1] = 0.f; }
    if (ny < 2) { cy[1] = cy[0]; wy[1] = 0.f; }
    if (nz < 2) { cz[1] = cz[0]; wz[1] = 0.f; }
    T v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = xc[cx[c & 1] + Nc * cy[(c >> 1) & 1] + NMc * cz[c >> 2]];
    T acc = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) acc += (T)(wx[c & 1] * wy[(c >> 1) & 1] * wz[c >> 2]) * v[c];
    xf[i] += acc;
  }
}

template <class T>
__device__ inline void prolong_add_body(const T* __restrict__ xc, int Nc, int Mc, T* __restrict__ xf, int Nf, int Mf,
                                        int Kf, long start, long stride) {
  const long nf = (long)Nf * Mf * Kf;
  for (long i = start; i < nf; i += stride) {
    int x, y, z;
    decode(i, Nf, Mf, &x, &y, &z);
    prolong_cell<T>(xc, Nc, Mc, xf, Nf, Mf, Kf, i, x, y, z);
  }
}

// One thread per coarse cell (64 x by 4 y coarse cells per block, blocks x fastest, then y, then
// z) adds the prolongation to its 8 children.  The 27 coarse values around it are loaded once
// (indices clamped to the level) and every child is computed as prolong_cell computes it: the
// same two taps per axis (at the level's boundary the first tap with weight 1 and a second,
// weight 0, that reads the same value), weights multiplied in FP32 and the 8 terms added in the
// same order, so the result is bit-identical with 27 instead of 64 coarse loads per 8 cells.
template <class T>
__global__ void __launch_bounds__(256) k_prolong_add8(const T* __restrict__ xc, int Nc, int Mc, int Kc,
                                                      T* __restrict__ xf) {
  const int gx = (Nc + 63) / 64, gy = (Mc + 3) / 4;
  const int r = (int)(blockIdx.x / gx);
  const int X = (int)(blockIdx.x % gx) * 64 + (int)(threadIdx.x & 63);
  const int Y = (r % gy) * 4 + (int)(threadIdx.x >> 6), Z = r / gy;
  if (X >= Nc || Y >= Mc) return;
  const int Nf = 2 * Nc, Mf = 2 * Mc;
  const long NMc = (long)Nc * Mc, NMf = (long)Nf * Mf;
  T nb[3][3][3];  // [dz + 1][dy + 1][dx + 1]
#pragma unroll
  for (int dz = 0; dz < 3; ++dz)
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) {
        const int cx = min(max(X + dx - 1, 0), Nc - 1), cy = min(max(Y + dy - 1, 0), Mc - 1);
        const int cz = min(max(Z + dz - 1, 0), Kc - 1);
        nb[dz][dy][dx] = xc[cx + Nc * (long)cy + NMc * cz];
      }
  // per axis and child parity i: taps at offsets (i - 1, i) around the coarse cell, weights
  // (0.25, 0.75) / (0.75, 0.25), or (1, 0) on the level's first / last fine cell
  float wx[2][2], wy[2][2], wz[2][2];
  auto wts = [](int P, int n, float (&w)[2][2]) {
    w[0][0] = P == 0 ? 1.0f : 0.25f;
    w[0][1] = P == 0 ? 0.0f : 0.75f;
    w[1][0] = P == n - 1 ? 1.0f : 0.75f;
    w[1][1] = P == n - 1 ? 0.0f : 0.25f;
  };
  wts(X, Nc, wx);
  wts(Y, Mc, wy);
  wts(Z, Kc, wz);
#pragma unroll
  for (int iz = 0; iz < 2; ++iz)
#pragma unroll
    for (int iy = 0; iy < 2; ++iy) {
      const long row = (long)(2 * X) + Nf * (long)(2 * Y + iy) + NMf * (2 * Z + iz);
#pragma unroll
      for (int ix = 0; ix < 2; ++ix) {
        T acc = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const int tx = c & 1, ty = (c >> 1) & 1, tz = c >> 2;
          acc += (T)(wx[ix][tx] * wy[iy][ty] * wz[iz][tz]) * nb[iz + tz][iy + ty][ix + tx];
        }
        xf[row + ix] += acc;
      }
    }
}

// one thread per fine cell: 64 x by 4 y cells per block, blocks x fastest, then y, then z
template <class T>
__global__ void __launch_bounds__(256) k_prolong_add(const T* __restrict__ xc, int Nc, int Mc, T* __restrict__ xf,
                                                     int Nf, int Mf, int Kf) {
  const int gx = (Nf + 63) / 64, gy = (Mf + 3) / 4;
  const int r = (int)(blockIdx.x / gx);
  const int x = (int)(blockIdx.x % gx) * 64 + (int)(threadIdx.x & 63);
  const int y = (r % gy) * 4 + (int)(threadIdx.x >> 6), z = r / gy;
  if (x < Nf && y < Mf) prolong_cell<T>(xc, Nc, Mc, xf, Nf, Mf, Kf, x + (long)Nf * (y + (long)Mf * z), x, y, z);
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
                                                                const double* __restrict__ Uf,
                                                                const double* __restrict__ dinv) {
  const long tid = threadIdx.x, nth = blockDim.x;
  const int last = C.count - 1;
  int n_last = 0;
  T* x_last = nullptr;
  const T* b_last = nullptr;
#pragma unroll
  for (int q = 0; q < COARSE_MAX; ++q)
    if (q == last) {
      n_last = (int)C.L[q].d.n();
      x_last = C.x[q];
      b_last = C.b[q];
    }
  // the LU factors of the last level (at most 64 unknowns) are staged in LDS by the whole
  // workgroup at the start: thread 0's triangular solves then wait on LDS, not on one
  // (by then evicted) global load after another
  __shared__ double sL[64 * 64], sU[64 * 64], ssc[64], slu_b[64], slu_y[64];
  __shared__ int sperm[64];
  __shared__ T sres[COARSE_CELLS];  // residuals of the level being restricted
  if (!dinv) {
    const int n = n_last;
    for (int k = (int)tid; k < n * n; k += (int)nth) {
      sL[k] = Lf[k];
      sU[k] = Uf[k];
    }
    for (int k = (int)tid; k < n; k += (int)nth) {
      ssc[k] = lu_scale[k];
      sperm[k] = lu_perm[k];
    }
  }
  __syncthreads();
  for (int q = 0; q < COARSE_MAX - 1; ++q) {
    if (q < last) {
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
      // residual restriction: every cell's residual into LDS (one thread per cell), then each coarse
      // cell sums its children in the order of k_resid_restrict_cells (the same values; one cell's
      // loads per thread instead of eight keeps the kernel clear of register spills)
      {
        const T* x = C.x[q];
        const T* b = C.b[q];
        const T fs = (T)L.fscale;
        const int N = L.d.N, M = L.d.M, K = L.d.K, n = N * M * K;
        for (int i = (int)tid; i < n; i += (int)nth) {
          int xx, yy, zz;
          decode(i, N, M, &xx, &yy, &zz);
          const T si = s[i];
          const T ap = si != (T)0 ? (T)1 / si : (T)0;
          T r = b[i] - ap * x[i];
          const T od = offdiag_u<T, FT>(x, e, nn, t, i, xx, yy, zz, N, M, K);
          if (i != L.pin0 && i != L.pin1) r -= fs * od;
          sres[i] = r;
        }
        __syncthreads();
        const int Nc = (N + 1) / 2, Mc = (M + 1) / 2, nc = Nc * Mc * ((K + 1) / 2);
        T* bc = C.b[q + 1];
        for (int I = (int)tid; I < nc; I += (int)nth) {
          int X, Y, Z;
          decode(I, Nc, Mc, &X, &Y, &Z);
          T acc = 0;
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            const int cx = 2 * X + (c & 1), cy = 2 * Y + ((c >> 1) & 1), cz = 2 * Z + (c >> 2);
            acc += cx < N && cy < M && cz < K ? sres[cx + N * (cy + M * cz)] : (T)0;  // odd axes: one child
          }
          bc[I] = acc;
        }
        __syncthreads();
      }
    }
  }
  if (dinv) {  // dense inverse of the last level, column-major: x_i = sum_j inv[j n + i] b_j
    const int n = n_last;
    const T* bb = b_last;
    for (int i = (int)tid; i < n; i += (int)nth) {
      double acc = 0.0;
      for (int j = 0; j < n; ++j) acc += dinv[(long)j * n + i] * (double)bb[j];
      x_last[i] = (T)acc;
    }
  } else if (tid == 0) {
    const int n = n_last;
    const T* bb = b_last;
    double* bs = slu_b;  // LDS: as local arrays they would live in scratch memory
    double* yv = slu_y;
    for (int i = 0; i < n; ++i) bs[i] = ssc[sperm[i]] * (double)bb[sperm[i]];
    for (int i = 0; i < n; ++i) {
      double aux = bs[i];
      for (int k = 0; k < i; ++k) aux -= sL[i * n + k] * yv[k];
      yv[i] = aux;
    }
    for (int i = n - 1; i >= 0; --i) {
      double aux = yv[i];
      for (int j = i + 1; j < n; ++j) aux -= sU[i * n + j] * bs[j];
      bs[i] = aux / sU[i * n + i];
    }
    for (int i = 0; i < n; ++i) x_last[i] = (T)(ssc[i] * bs[i]);
  }
  __syncthreads();
  for (int q = COARSE_MAX - 2; q >= 0; --q) {
    if (q < last) {
      const Level& L = C.L[q];
      const FT* e = (const FT*)L.f_e;
      const FT* nn = (const FT*)L.f_n;
      const FT* t = (const FT*)L.f_t;
      const T* s = (const T*)L.s;
      prolong_add_body<T>(C.x[q + 1], C.L[q + 1].d.N, C.L[q + 1].d.M, C.x[q], L.d.N, L.d.M, L.d.K, tid, nth);
      __syncthreads();
      for (int sw = 0; sw < n_post; ++sw) {
        rbgs_body<T, FT>(C.x[q], C.b[q], e, nn, t, s, (T)L.fscale, L.d.N, L.d.M, L.d.K, 0, L.pin0, L.pin1, 0, tid,
                         nth);
        __syncthreads();
        rbgs_body<T, FT>(C.x[q], C.b[q], e, nn, t, s, (T)L.fscale, L.d.N, L.d.M, L.d.K, 1, L.pin0, L.pin1, 0, tid,
                         nth);
        __syncthreads();
      }
    }
  }
}

// ---- coarsest level as a dense inverse ---------------------------------------------------------
// A hierarchy capped above its natural depth ends on a level of up to DENSE_MAX unknowns, too many
// for the single-thread LU.  Its inverse is formed once at setup by Gauss-Jordan elimination with
// partial pivoting on the symmetrically scaled matrix S A S (s_i = 1 / sqrt|a_ii|, the LU's
// scaling) augmented with the identity, W = [S A S | I] (n x 2n, row-major); pivots below
// 1e-13 max|S A S| are clamped to that size as the LU clamps them.  Empty rows (every face 0:
// coarse cells inside the body, whose right-hand side is 0) become identity rows.  A cycle then
// applies the inverse as a dense product in FP64.
__global__ void k_dense_fill(const double* __restrict__ fe, const double* __restrict__ fn,
                             const double* __restrict__ ft, const double* __restrict__ dg, int N, int M, int K,
                             double* __restrict__ sc, double* __restrict__ W) {
  const long n = (long)N * M * K, NM = (long)N * M, w = 2 * n;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n * w; q += (long)gridDim.x * blockDim.x) {
    const long i = q / w, j = q - i * w;
    auto scl = [&](long r) {
      const double d = fabs(dg[r]);
      return d > 1e-5 ? 1.0 / sqrt(d) : 1.0;
    };
    double v = 0.0;
    if (j >= n) {
      v = j - n == i ? 1.0 : 0.0;
    } else if (j == i) {
      v = dg[i] == 0.0 ? 1.0 : scl(i) * dg[i] * scl(i);
    } else {
      int xi, yi, zi;
      decode(i, N, M, &xi, &yi, &zi);
      double a = 0.0;
      if (j == i + 1 && xi < N - 1) a = fe[i];
      if (j == i - 1 && xi > 0) a = fe[j];
      if (j == i + N && yi < M - 1) a = fn[i];
      if (j == i - N && yi > 0) a = fn[j];
      if (j == i + NM && zi < K - 1) a = ft[i];
      if (j == i - NM && zi > 0) a = ft[j];
      v = a == 0.0 ? 0.0 : scl(i) * a * scl(j);
    }
    W[q] = v;
    if (j == 0) sc[i] = scl(i);
  }
}

// max |W| over the left half, into out[0] (one workgroup, setup only)
__global__ void __launch_bounds__(1024) k_dense_absmax(const double* __restrict__ W, int n, double* out) {
  __shared__ double lds[16];
  double m = 0.0;
  const long w = 2L * n;
  for (long q = threadIdx.x; q < (long)n * n; q += blockDim.x) m = fmax(m, fabs(W[(q / n) * w + q % n]));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) r = fmax(r, lds[k]);
    out[0] = r;
  }
}

// pivot step k (one workgroup): the row of largest |W[i][k]|, i >= k (ties: the lowest i), swapped
// into row k, the row scaled by its pivot, and the multipliers f[i] = W[i][k] (f[k] = 0) saved
// for the elimination kernel
__global__ void __launch_bounds__(1024) k_gj_pivot(double* __restrict__ W, int n, int k,
                                                   const double* __restrict__ amax, double* __restrict__ f) {
  __shared__ double bv[1024];
  __shared__ int bi[1024];
  const long w = 2L * n;
  const int tid = threadIdx.x, nth = blockDim.x;
  double best = -1.0;
  int bidx = k;
  for (int i = k + tid; i < n; i += nth) {
    const double v = fabs(W[(long)i * w + k]);
    if (v > best) {
      best = v;
      bidx = i;
    }
  }
  bv[tid] = best;
  bi[tid] = bidx;
  __syncthreads();
  for (int st = nth / 2; st > 0; st >>= 1) {
    if (tid < st) {
      const double o = bv[tid + st];
      const int oi = bi[tid + st];
      if (o > bv[tid] || (o == bv[tid] && oi < bi[tid])) {
        bv[tid] = o;
        bi[tid] = oi;
      }
    }
    __syncthreads();
  }
  const int p = bi[0];
  if (p != k)
    for (long j = tid; j < w; j += nth) {
      const double a = W[(long)k * w + j];
      W[(long)k * w + j] = W[(long)p * w + j];
      W[(long)p * w + j] = a;
    }
  __syncthreads();
  const double tiny = 1e-13 * amax[0];
  double piv = W[(long)k * w + k];
  if (fabs(piv) < tiny) piv = piv < 0.0 ? -tiny : tiny;
  __syncthreads();  // every thread holds the pivot before the row changes
  for (long j = tid; j < w; j += nth) W[(long)k * w + j] = j == k ? 1.0 : W[(long)k * w + j] / piv;
  for (int i = tid; i < n; i += nth) f[i] = i == k ? 0.0 : W[(long)i * w + k];
}

// elimination step k: W[i][j] -= f[i] W[k][j] for every row i != k (row k is only read)
__global__ void k_gj_elim(double* __restrict__ W, int n, int k, const double* __restrict__ f) {
  const long w = 2L * n;
  const int i = blockIdx.y;
  const double fi = f[i];
  if (i == k || fi == 0.0) return;
  for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < w; j += (long)gridDim.x * blockDim.x)
    W[(long)i * w + j] -= fi * W[(long)k * w + j];
}

// inv = S (S A S)^-1 S from the right half of W: column-major copy (inv_c[j n + i] = inv(i, j))
// for the product, row-major copy (inv_r[i n + j]) for the transposed product
__global__ void k_dense_finish(const double* __restrict__ W, const double* __restrict__ sc, int n,
                               double* __restrict__ inv_c, double* __restrict__ inv_r) {
  const long w = 2L * n;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < (long)n * n; q += (long)gridDim.x * blockDim.x) {
    const long i = q / n, j = q - i * n;
    const double v = sc[i] * W[i * w + n + j] * sc[j];
    inv_r[q] = v;
    inv_c[j * n + i] = v;
  }
}

// x = inv b in column chunks: block (r, c) sums columns [c * DCH, (c + 1) * DCH) of rows
// [256 r, 256 r + 256) into part[c][i]; k_dense_sum adds the chunks in order
constexpr int DCH = 128;
template <class T>
__global__ void __launch_bounds__(256) k_dense_part(const double* __restrict__ inv, int n, const T* __restrict__ b,
                                                    double* __restrict__ part) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j0 = blockIdx.y * DCH, j1 = j0 + DCH < n ? j0 + DCH : n;
  if (i >= n) return;
  double acc = 0.0;
  for (int j = j0; j < j1; ++j) acc += inv[(long)j * n + i] * (double)b[j];
  part[(long)blockIdx.y * n + i] = acc;
}

template <class T>
__global__ void k_dense_sum(const double* __restrict__ part, int n, int nch, T* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int c = 0; c < nch; ++c) acc += part[(long)c * n + i];
  x[i] = (T)acc;
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
constexpr int HX = FX + 4;  // x_in staging: tile + 2-cell ring (rows: tile rows + 4)
constexpr int RX = FX + 2;  // red values: tile + 1-cell ring (rows: tile rows + 2)

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

template <class T, class FT, int TY>
__global__ void __launch_bounds__(FX* TY) k_rb_fused(const T* __restrict__ x_in, T* __restrict__ x_out,
                                                      const T* __restrict__ b, const FT* __restrict__ ae,
                                                      const FT* __restrict__ an, const FT* __restrict__ at,
                                                      const T* __restrict__ s, T fs, int N, int M, int K, int pin0,
                                                      int pin1, int zero) {
  constexpr int HY = TY + 4, RY = TY + 2;
  // two ring cells and two staging cells per thread cover the tile's rings
  static_assert(2 * FX * TY >= RX * RY && 2 * FX * TY >= HX * HY, "tile rings exceed two cells per thread");
  __shared__ T xs_[3][HY][HX];
  __shared__ T rs_[3][RY][RX];
  constexpr int NT = FX * TY;
  const int tid = threadIdx.x;
  // tiles in XCD-contiguous order (x fastest, then y, then z): neighbouring tiles, whose ring
  // cells a tile reads, run on the same XCD and share its L2 (measured fetch 26 instead of
  // 50 B/cell)
  const int gxt = (N + FX - 1) / FX, gyt = (M + TY - 1) / TY;
  const int lb = xcd_order(blockIdx.x, gridDim.x);
  const int X0 = (lb % gxt) * FX, Y0 = ((lb / gxt) % gyt) * TY, Z0 = (lb / (gxt * gyt)) * FZ;
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
    cint[k] = cin[k] && rx >= 1 && rx <= FX && ry >= 1 && ry <= TY;
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

// ---- damped Jacobi ---------------------------------------------------------------------------
// One damped Jacobi sweep x_out = x_in + omega D^-1 (b - A x_in): every row of the form
// x + W (b - A x) with W = omega D^-1, so the transposed sweep is the same with A^T (TRANS: the
// finest level's pinned rows are full rows of A^T, other rows drop their couplings to them).
// Empty rows (1/a_p stored as 0) keep their value; ZERO: x_in is 0 and not read.  One thread per
// cell, every load of a row issued before the update.
template <class T, class FT, bool ZERO, bool TRANS>
__global__ void __launch_bounds__(256) k_jacobi(const T* __restrict__ x_in, T* __restrict__ x_out,
                                                const T* __restrict__ b, const FT* __restrict__ ae,
                                                const FT* __restrict__ an, const FT* __restrict__ at,
                                                const T* __restrict__ s, T fs, T omega, int N, int M, int K,
                                                int pin0, int pin1) {
  // one thread per cell: 64 x by 4 y cells per block, blocks x fastest, then y, then z
  const int gx = (N + 63) / 64, gy = (M + 3) / 4;
  const int r = (int)(blockIdx.x / gx);
  const int xx = (int)(blockIdx.x % gx) * 64 + (int)(threadIdx.x & 63);
  const int yy = (r % gy) * 4 + (int)(threadIdx.x >> 6), zz = r / gy;
  if (xx < N && yy < M) {
    const long i = xx + (long)N * (yy + (long)M * zz);
    const T si = s[i], bi = b[i];
    T xi = (T)0, od = (T)0;
    if constexpr (!ZERO) {
      xi = x_in[i];
      if constexpr (TRANS) {
        od = offdiag_tr<T, FT>(x_in, ae, an, at, i, xx, yy, zz, N, M, K, pin0, pin1);
      } else {
        const T o = offdiag_u<T, FT>(x_in, ae, an, at, i, xx, yy, zz, N, M, K);
        od = (i == pin0 || i == pin1) ? (T)0 : o;  // pinned rows of A are identity rows
      }
    }
    x_out[i] = si == (T)0 ? xi : xi + omega * ((bi - fs * od) * si - xi);
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
// 1 / v (0 for v = 0): the FP32 hardware reciprocal (1 ulp), FP64 division
__device__ inline float recip0(float v) { return v != 0.f ? __builtin_amdgcn_rcpf(v) : 0.f; }
__device__ inline double recip0(double v) { return v != 0.0 ? 1.0 / v : 0.0; }

// Row of cell i on its line of direction DIR: lower (coupling to the previous cell of the line),
// diag, upper, rhs.  Pinned rows are identity rows; an empty row (1/a_p stored as 0) keeps its
// value, as in the point smoother.  ZERO: x taken as 0 (not read).  DFF (forward rows only,
// trans false): a_p rebuilt from the faces by dr, s not read.
template <class T, class FT, int DIR, bool ZERO, bool DFF = false>
__device__ inline void line_row(const T* __restrict__ x, const T* __restrict__ b, const FT* __restrict__ ae,
                                const FT* __restrict__ an, const FT* __restrict__ at, const T* __restrict__ s, T fs,
                                long i, int xx, int yy, int zz, int N, int M, int K, int pin0, int pin1, bool trans,
                                T& lo, T& dg, T& up, T& rhs, DiagRule<T> dr = DiagRule<T>{-1, -1, (T)0, (T)0}) {
  #pragma clang fp contract(fast)  // line solves: no bitwise twin to match
  // Straight-line code: every load is issued unconditionally (absent neighbours read the cell
  // itself and are masked out), so that a caller evaluating many rows can have all their loads
  // in flight at once.
  const long NM = (long)N * M;
  const bool own = i == pin0 || i == pin1;
  auto keep = [&](long j) { return !trans || own || (j != pin0 && j != pin1); };
  const bool hw = xx > 0 && keep(i - 1), he = xx < N - 1 && keep(i + 1);
  const bool hs = yy > 0 && keep(i - N), hn = yy < M - 1 && keep(i + N);
  const bool hb = zz > 0 && keep(i - NM), ht = zz < K - 1 && keep(i + NM);
  const long iw = xx > 0 ? i - 1 : i, ie = xx < N - 1 ? i + 1 : i;
  const long is = yy > 0 ? i - N : i, in = yy < M - 1 ? i + N : i;
  const long ib = zz > 0 ? i - NM : i, it = zz < K - 1 ? i + NM : i;
  T si = (T)0;
  if constexpr (!DFF) si = s[i];
  const T bi = b[i];
  const T fw = fs * ld<T>(ae, iw), fe = fs * ld<T>(ae, i);
  const T fsf = fs * ld<T>(an, is), fn = fs * ld<T>(an, i);
  const T fb = fs * ld<T>(at, ib), ft = fs * ld<T>(at, i);
  const T cw = hw ? fw : (T)0, ce = he ? fe : (T)0, cs = hs ? fsf : (T)0, cn = hn ? fn : (T)0;
  const T cb = hb ? fb : (T)0, ct = ht ? ft : (T)0;
  T off = (T)0, xi = (T)0;
  if constexpr (!ZERO) {
    xi = x[i];
    if constexpr (DIR == 0) off = cs * x[is] + cn * x[in] + cb * x[ib] + ct * x[it];
    else if constexpr (DIR == 1) off = cw * x[iw] + ce * x[ie] + cb * x[ib] + ct * x[it];
    else off = cw * x[iw] + ce * x[ie] + cs * x[is] + cn * x[in];
  }
  T ap;
  if constexpr (DFF)
    ap = dr.of(cw + ce + cs + cn + cb + ct, i, fe);
  else
    ap = recip0(si);
  const bool ident = own && !trans, empty = DFF ? ap == (T)0 : si == (T)0;
  const bool cut = ident || empty;  // identity row (pinned) or kept value (empty)
  const T l0 = DIR == 0 ? cw : (DIR == 1 ? cs : cb);
  const T u0 = DIR == 0 ? ce : (DIR == 1 ? cn : ct);
  lo = cut ? (T)0 : l0;
  up = cut ? (T)0 : u0;
  dg = cut ? (T)1 : ap;
  rhs = ident ? bi : (empty ? xi : bi - off);
}

// x lines: one wavefront per line, CPL consecutive cells per lane (coalesced loads).  Partition
// method: each lane eliminates its first CPL - 1 cells with two unknowns kept, the last cell of
// the lane before (x_L) and its own last cell (x_E): x_k = G_k + H_k x_L + Q_k x_E.  The rows of
// the 64 last cells then form a tridiagonal system in x_E, one unknown per lane, solved by
// parallel cyclic reduction with cross-lane shuffles (6 steps); every lane substitutes back.
// value of v in lane src of this lane's group of W lanes (src relative to the group)
template <int W = 64, class T>
__device__ inline T wave_get(T v, int src, bool ok) {
  const T r = __shfl(v, ok ? src : 0, W);
  return r;
}

// CPL consecutive elements of p from index i (i a multiple of CPL) as one wide load
template <class E, int C>
__device__ inline void ld_pack(const E* __restrict__ p, long i, E (&out)[C]) {
  struct alignas(sizeof(E) * C) Pack {
    E v[C];
  };
  const Pack pk = *reinterpret_cast<const Pack*>(p + i);
#pragma unroll
  for (int j = 0; j < C; ++j) out[j] = pk.v[j];
}

// The rows of cells i0 .. i0 + CPL - 1 of one x line (i0 = row start + CPL * lane, N a multiple
// of CPL): the same rows as line_row, each array read with one wide load per lane.  The lines'
// neighbours in y and z exist or not for the whole row; pinned cells are checked per cell.
template <class T, class FT, int CPL, bool ZERO, bool DFF>
__device__ inline void line_rows_x(const T* __restrict__ x, const T* __restrict__ b, const FT* __restrict__ ae,
                                   const FT* __restrict__ an, const FT* __restrict__ at, const T* __restrict__ s,
                                   T fs, long i0, int k0, int yy, int zz, int N, int M, int K, int pin0, int pin1,
                                   bool trans, T (&lo)[CPL], T (&dg)[CPL], T (&up)[CPL], T (&rh)[CPL],
                                   DiagRule<T> dr) {
  #pragma clang fp contract(fast)  // line solves: no bitwise twin to match
  const long NM = (long)N * M;
  const bool ys = yy > 0, yn = yy < M - 1, zb = zz > 0, zt = zz < K - 1;
  const long is0 = ys ? i0 - N : i0, in0 = yn ? i0 + N : i0, ib0 = zb ? i0 - NM : i0, it0 = zt ? i0 + NM : i0;
  T sv[CPL], bv[CPL];
  FT fev[CPL], fnv[CPL], fsv[CPL], ftv[CPL], fbv[CPL];
  if constexpr (!DFF) ld_pack<T, CPL>(s, i0, sv);
  ld_pack<T, CPL>(b, i0, bv);
  ld_pack<FT, CPL>(ae, i0, fev);
  ld_pack<FT, CPL>(an, i0, fnv);
  ld_pack<FT, CPL>(an, is0, fsv);
  ld_pack<FT, CPL>(at, i0, ftv);
  ld_pack<FT, CPL>(at, ib0, fbv);
  const FT fw0 = ae[k0 > 0 ? i0 - 1 : i0];
  T xv[CPL], xs[CPL], xn[CPL], xb[CPL], xt[CPL];
  if constexpr (!ZERO) {
    ld_pack<T, CPL>(x, i0, xv);
    ld_pack<T, CPL>(x, is0, xs);
    ld_pack<T, CPL>(x, in0, xn);
    ld_pack<T, CPL>(x, ib0, xb);
    ld_pack<T, CPL>(x, it0, xt);
  }
  auto pinned = [&](long j) { return j == pin0 || j == pin1; };
#pragma unroll
  for (int j = 0; j < CPL; ++j) {
    const long i = i0 + j;
    const int k = k0 + j;
    const bool own = pinned(i);
    auto keep = [&](long q) { return !trans || own || !pinned(q); };
    const bool hw = k > 0 && keep(i - 1), he = k < N - 1 && keep(i + 1);
    const bool hs = ys && keep(i - N), hn = yn && keep(i + N);
    const bool hb = zb && keep(i - NM), ht = zt && keep(i + NM);
    const T fw = fs * (T)(j == 0 ? fw0 : fev[j > 0 ? j - 1 : 0]), fe = fs * (T)fev[j];
    const T fsf = fs * (T)fsv[j], fn = fs * (T)fnv[j];
    const T fb = fs * (T)fbv[j], ft = fs * (T)ftv[j];
    const T cw = hw ? fw : (T)0, ce = he ? fe : (T)0, cs = hs ? fsf : (T)0, cn = hn ? fn : (T)0;
    const T cb = hb ? fb : (T)0, ct = ht ? ft : (T)0;
    T off = (T)0, xi = (T)0;
    if constexpr (!ZERO) {
      xi = xv[j];
      off = cs * xs[j] + cn * xn[j] + cb * xb[j] + ct * xt[j];
    }
    const T bi = bv[j];
    T ap;
    bool empty;
    if constexpr (DFF) {
      ap = dr.of(cw + ce + cs + cn + cb + ct, i, fe);
      empty = ap == (T)0;
    } else {
      ap = recip0(sv[j]);
      empty = sv[j] == (T)0;
    }
    const bool ident = own && !trans, cut = ident || empty;
    lo[j] = cut ? (T)0 : cw;
    up[j] = cut ? (T)0 : ce;
    dg[j] = cut ? (T)1 : ap;
    rh[j] = ident ? bi : (empty ? xi : bi - off);
  }
}

// LANES = 64: one wavefront per line; LANES = 32: two lines per wavefront (lines of at most
// 32 CPL cells: shorter cyclic reduction, wider loads per lane)
template <class T, class FT, int CPL, bool ZERO, bool TRANS, bool DFF, int LANES = 64>
__global__ void __launch_bounds__(256) k_line_x(const T* __restrict__ x_in, T* __restrict__ x_out,
                                                const T* __restrict__ b, const FT* __restrict__ ae,
                                                const FT* __restrict__ an, const FT* __restrict__ at,
                                                const T* __restrict__ s, T fs, int N, int M, int K, int pin0,
                                                int pin1, DiagRule<T> dr) {
  #pragma clang fp contract(fast)  // line solves: no bitwise twin to match
  static_assert(!(TRANS && DFF), "a_p from the faces is a forward-operator rule");
  const int lane = threadIdx.x & (LANES - 1);
  const long row_t = (long)blockIdx.x * (256 / LANES) + (threadIdx.x / LANES);
  if constexpr (LANES == 64)
    if (row_t >= (long)M * K) return;  // whole wavefronts leave together
  const bool live = row_t < (long)M * K;  // two-line wavefronts: a line past the end repeats the last
  const long row = live ? row_t : (long)M * K - 1;
  const int yy = (int)(row % M), zz = (int)(row / M);
  T lo[CPL], dg[CPL], up[CPL], rh[CPL];
  if (N % CPL == 0 && lane * CPL < N) {
    line_rows_x<T, FT, CPL, ZERO, DFF>(x_in, b, ae, an, at, s, fs, row * N + lane * CPL, lane * CPL, yy, zz, N, M,
                                       K, pin0, pin1, TRANS, lo, dg, up, rh, dr);
  } else {
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
      const int k = lane * CPL + j;
      if (k < N) {
        line_row<T, FT, 0, ZERO, DFF>(x_in, b, ae, an, at, s, fs, row * N + k, k, yy, zz, N, M, K, pin0, pin1,
                                      TRANS, lo[j], dg[j], up[j], rh[j], dr);
      } else {  // padding: decoupled identity rows
        lo[j] = up[j] = rh[j] = (T)0;
        dg[j] = (T)1;
      }
    }
  }
  // local elimination: forward x_j = g_j + h_j x_L - c_j x_{j+1}, backward with x_{CPL-1} = x_E
  T G[CPL], H[CPL], Q[CPL];
  {
    T cp = (T)0, gp = (T)0, hp = (T)1;  // x_{-1} = x_L
#pragma unroll
    for (int j = 0; j < CPL - 1; ++j) {
      const T winv = recip0(dg[j] - lo[j] * cp);
      gp = (rh[j] - lo[j] * gp) * winv;
      hp = -lo[j] * hp * winv;
      cp = up[j] * winv;
      G[j] = gp;
      H[j] = hp;
      Q[j] = cp;
    }
    T Gn = (T)0, Hn = (T)0, Qn = (T)1;  // cell CPL-1 is x_E
#pragma unroll
    for (int j = CPL - 2; j >= 0; --j) {
      const T c = Q[j];
      Gn = G[j] - c * Gn;
      Hn = H[j] - c * Hn;
      Qn = -c * Qn;
      G[j] = Gn;
      H[j] = Hn;
      Q[j] = Qn;
    }
  }
  // the cell before the last in terms of (x_L, x_E), and the first cell, which the lane before needs
  const T Gb = CPL > 1 ? G[CPL > 1 ? CPL - 2 : 0] : (T)0;
  const T Hb = CPL > 1 ? H[CPL > 1 ? CPL - 2 : 0] : (T)1;
  const T Qb = CPL > 1 ? Q[CPL > 1 ? CPL - 2 : 0] : (T)0;
  const T Gf = CPL > 1 ? G[0] : (T)0, Hf = CPL > 1 ? H[0] : (T)0, Qf = CPL > 1 ? Q[0] : (T)1;
  const bool nxt = lane < LANES - 1;
  // every lane takes part in the shuffles (a lane outside a branch supplies no value)
  const T G1v = wave_get<LANES>(Gf, lane + 1, nxt), H1v = wave_get<LANES>(Hf, lane + 1, nxt);
  const T Q1v = wave_get<LANES>(Qf, lane + 1, nxt);
  const T G1 = nxt ? G1v : (T)0, H1 = nxt ? H1v : (T)0, Q1 = nxt ? Q1v : (T)0;
  // row of the last cell: lo x_{CPL-2} + dg x_E + up x_0(next lane)
  const T le = lo[CPL - 1], ue = up[CPL - 1];
  T a = le * Hb;
  T d = dg[CPL - 1] + le * Qb + ue * H1;
  T c = ue * Q1;
  T r = rh[CPL - 1] - le * Gb - ue * G1;
#pragma unroll
  for (int st = 1; st < LANES; st <<= 1) {
    const bool okm = lane - st >= 0, okp = lane + st < LANES;
    T am = wave_get<LANES>(a, lane - st, okm), dm = wave_get<LANES>(d, lane - st, okm);
    T cm = wave_get<LANES>(c, lane - st, okm), rm = wave_get<LANES>(r, lane - st, okm);
    T ap = wave_get<LANES>(a, lane + st, okp), dp = wave_get<LANES>(d, lane + st, okp);
    T cp = wave_get<LANES>(c, lane + st, okp), rp = wave_get<LANES>(r, lane + st, okp);
    if (!okm) {
      am = cm = rm = (T)0;
      dm = (T)1;
    }
    if (!okp) {
      ap = cp = rp = (T)0;
      dp = (T)1;
    }
    const T al = -a * recip0(dm), ga = -c * recip0(dp);
    a = al * am;
    c = ga * cp;
    d = d + al * cm + ga * ap;
    r = r + al * rm + ga * rp;
  }
  const T xe = r * recip0(d);
  const T xl0 = wave_get<LANES>(xe, lane - 1, lane > 0);
  const T xl = lane > 0 ? xl0 : (T)0;
#pragma unroll
  for (int j = 0; j < CPL; ++j) {
    const int k = lane * CPL + j;
    const T v = j == CPL - 1 ? xe : G[j] + H[j] * xl + Q[j] * xe;
    if (live && k < N) x_out[row * N + k] = v;
  }
}

// y and z lines: a workgroup holds LPW lines (adjacent x, LPW = 128 bytes / value size: 32 FP32
// or 16 FP64 lines, so that every row load is one full 128-byte line) split into S segments of
// SEG = 8 rows; thread t takes line t % LPW and segment t / LPW (S = ceil(len / SEG) <= 32: lines
// of at most 256 cells).  The partition method of the x lines with segments in place of lanes:
// each thread eliminates its segment's first cnt - 1 rows keeping x_L (last cell of the segment
// before) and x_E (its own last cell) as unknowns, x_k = G_k + H_k x_L + Q_k x_E; the S last-cell
// rows of a line form a tridiagonal system, solved by cyclic reduction (log2 S steps); every
// segment then writes x_k.  Inputs are read once and x written once.  The FP64 workgroups are half
// the size of the FP32 ones (512 threads), so two of them share a CU at the FP64 kernel's register
// count.
constexpr int SEG = 8;
template <class T>
constexpr int lpw() {
  return 128 / (int)sizeof(T);
}

// element i of p addressed by a 32-bit byte offset (i * sizeof(E) < 2^32), so that loads take a
// uniform base and one 32-bit offset register
template <class E>
__device__ inline E ldo(const E* __restrict__ p, unsigned i) {
  return *reinterpret_cast<const E*>(reinterpret_cast<const char*>(p) + i * (unsigned)sizeof(E));
}
// the same load marked non-temporal (read once by this workgroup only)
template <class E>
__device__ inline E ldo_nt(const E* __restrict__ p, unsigned i) {
  return __builtin_nontemporal_load(reinterpret_cast<const E*>(reinterpret_cast<const char*>(p) + i * (unsigned)sizeof(E)));
}

// Row of cell i of a y (DIR 1) or z (DIR 2) line of the forward operator, with the line's
// per-thread constants: the x neighbours (A) and the other cross direction (B) exist or not for
// the whole line (offsets of absent ones are 0: the cell itself, masked).  Same values and
// operand order as line_row.  DFF: a_p rebuilt from the faces by dr, s not read.
template <class T, class FT, int DIR, bool ZERO, bool DFF>
struct SegLine {
  const T* x;
  const T* b;
  const T* s;
  const FT* ae;
  const FT* fB;  // faces of the B direction (z faces for y lines, y faces for z lines)
  const FT* fL;  // faces along the line
  T fs;
  int strideL, oAl, oAu, oBl, oBu, len, pin0, pin1;
  // x neighbours come from the neighbouring lanes of the LPW-lane group (lane l32 holds cell
  // x0 + l32 of the row); only the group's west and east halo cells are loaded, at offsets
  // oW0 / oE0 from the cell (the same address for the whole group)
  static constexpr int LPW = lpw<T>();
  int l32, oW0, oE0;
  bool hAl, hAu, hBl, hBu;
  DiagRule<T> dr;
  // a face in operator units: FP16 faces are stored divided by fs; FP32 and FP64 faces are
  // stored as they are (fs = 1)
  __device__ inline T scl(T f) const {
    if constexpr (std::is_same<FT, _Float16>::value)
      return fs * f;
    else
      return f;
  }
  __device__ inline void row(unsigned i, int kk, bool first, T& fl_carry, T& lo, T& dg, T& up, T& rh) const {
  #pragma clang fp contract(fast)  // line solves: no bitwise twin to match
    const bool hl = kk > 0, hu = kk < len - 1;
    T sv = (T)0;
    if constexpr (!DFF) sv = ldo(s, i);
    const T bv = ldo_nt(b, i);
    const T fau = (T)ldo_nt(ae, i);
    const bool marked = DFF && __builtin_signbit(fau);
    const T fbl = (T)ldo(fB, i + oBl), fbu = (T)ldo(fB, i);
    const T flu = (T)ldo_nt(fL, i);
    const T fll = first ? (T)ldo(fL, hl ? i - strideL : i) : fl_carry;  // upper face of the row before
    fl_carry = flu;
    T off = (T)0, xi = (T)0, fal = (T)0;
    if constexpr (!ZERO) {
      xi = ldo(x, i);
      const T xbl = ldo(x, i + oBl), xbu = ldo(x, i + oBu);
      const T xw0 = ldo(x, i + oW0), xe0 = ldo(x, i + oE0), fw0 = (T)ldo(ae, i + oW0);
      const T xws = __shfl_up(xi, 1, LPW), xes = __shfl_down(xi, 1, LPW), fws = __shfl_up(fau, 1, LPW);
      const T xal = l32 == 0 ? xw0 : xws, xau = l32 == LPW - 1 ? xe0 : xes;
      fal = l32 == 0 ? fw0 : fws;
      const T cAl = hAl ? scl(fal) : (T)0, cAu = hAu ? scl(fau) : (T)0;
      const T cBl = hBl ? scl(fbl) : (T)0, cBu = hBu ? scl(fbu) : (T)0;
      off = cAl * xal + cAu * xau + cBl * xbl + cBu * xbu;
    } else if constexpr (DFF) {
      const T fw0 = (T)ldo(ae, i + oW0), fws = __shfl_up(fau, 1, LPW);
      fal = l32 == 0 ? fw0 : fws;
    }
    T ap;
    bool empty;
    if constexpr (DFF) {
      const T fsum = (hAl ? fal : (T)0) + (hAu ? fau : (T)0) + (hBl ? fbl : (T)0) + (hBu ? fbu : (T)0) +
                     (hl ? fll : (T)0) + (hu ? flu : (T)0);
      ap = dr.of_marked(scl(fsum), (long)i, marked);
      empty = ap == (T)0;
    } else {
      ap = recip0(sv);
      empty = sv == (T)0;
    }
    const bool own = (int)i == pin0 || (int)i == pin1, cut = own || empty;
    lo = cut || !hl ? (T)0 : scl(fll);
    up = cut || !hu ? (T)0 : scl(flu);
    dg = cut ? (T)1 : ap;
    rh = own ? bv : (empty ? xi : bv - off);
  }
};

template <class T, class FT, int DIR, bool ZERO, bool TRANS, bool DFF>
__global__ void __launch_bounds__(1024) k_line_seg(const T* __restrict__ x_in, T* __restrict__ x_out,
                                                   const T* __restrict__ b, const FT* __restrict__ ae,
                                                   const FT* __restrict__ an, const FT* __restrict__ at,
                                                   const T* __restrict__ s, T fs, int N, int M, int K, int pin0,
                                                   int pin1, int S, DiagRule<T> dr) {
  #pragma clang fp contract(fast)  // line solves: no bitwise twin to match
  static_assert(!(TRANS && DFF), "a_p from the faces is a forward-operator rule");
  constexpr int LPW = lpw<T>();
  extern __shared__ unsigned char lds_raw[];
  // [8][S][LPW + 1]: a d c r of the reduced rows, G H Q of first cells, x_E; rows padded by one
  // element so that both the (segment, line) and the (line, segment) thread mappings read
  // consecutive banks
  T* ifc = reinterpret_cast<T*>(lds_raw);
  auto IF = [&](int q, int j, int l) -> T& { return ifc[((long)q * S + j) * (LPW + 1) + l]; };
  const long NM = (long)N * M;
  const int lane = threadIdx.x % LPW, seg = threadIdx.x / LPW;
  // Line groups in XCD-contiguous order, columns fastest: the halo columns and neighbouring
  // planes a workgroup reads are read by workgroups of the same XCD at about the same time
  // (measured fetch 26 instead of 35 B/cell, 24 B/cell compulsory reads; 9 % less time).  y lines:
  // LPW adjacent x of one (y = 0, z) row; z lines: LPW consecutive cells p = x + N y of the (x, y)
  // plane, so that every row load is one aligned 128-byte line whatever N is (the same groups as
  // x runs of one row when LPW divides N)
  const int gxb = (N + LPW - 1) / LPW;
  const int lb = xcd_order(blockIdx.x, gridDim.x);
  bool live;
  int xx, other;
  long p0 = 0, pl = 0;
  if constexpr (DIR == 1) {
    const int xr = (lb % gxb) * LPW + lane;
    other = lb / gxb;
    live = xr < N;
    xx = live ? xr : N - 1;  // lanes past the last line repeat it and store nothing
  } else {
    p0 = (long)lb * LPW;
    pl = p0 + LPW - 1 < NM - 1 ? p0 + LPW - 1 : NM - 1;
    const long pr = p0 + lane;
    live = pr < NM;
    const long pc = live ? pr : NM - 1;
    xx = (int)(pc % N);
    other = (int)(pc / N);
  }
  const int len = DIR == 1 ? M : K;
  const long stride = DIR == 1 ? (long)N : NM;
  const long base = DIR == 1 ? xx + NM * other : xx + (long)N * other;
  const int k0 = seg * SEG;
  const int cnt = len - k0 < SEG ? len - k0 : SEG;  // >= 1
  // Straight-line code over the SEG rows (rows past the segment's end re-read its last row and
  // are masked), so that the loads of all of them are in flight together.
  T G[SEG], H[SEG], Q[SEG];
  T le = (T)0, de = (T)1, ue = (T)0, re = (T)0;  // row of the last cell
  T Gb = (T)0, Hb = (T)1, Qb = (T)0;              // the cell before the last, in (x_L, x_E)
  T Gf = (T)0, Hf = (T)0, Qf = (T)1;              // the first cell, in (x_L, x_E)
  // the elimination of the segment; FULL: a segment of SEG rows (every segment of a line whose
  // length is a multiple of SEG), whose masks are then known at compile time
  auto eliminate = [&](auto full_tag) {
    constexpr bool FULL = decltype(full_tag)::value;
    const int nrow = FULL ? SEG : cnt;
    {
      T cp = (T)0, gp = (T)0, hp = (T)1;  // x_{-1} = x_L
      // x_k = g_k + h_k x_L - c_k x_{k+1} for the rows before the last, each row eliminated as soon
      // as it is built
      auto elim = [&](int k, T lo, T dg, T up, T rh) {
        const bool el = k < nrow - 1, last = k == nrow - 1;
        const T winv = recip0(dg - lo * cp);
        const T gn = (rh - lo * gp) * winv, hn = -lo * hp * winv, cn = up * winv;
        gp = el ? gn : gp;
        hp = el ? hn : hp;
        cp = el ? cn : cp;
        G[k] = gp;
        H[k] = hp;
        Q[k] = cp;
        le = last ? lo : le;
        de = last ? dg : de;
        ue = last ? up : ue;
        re = last ? rh : re;
      };
      if constexpr (!TRANS) {
        SegLine<T, FT, DIR, ZERO, DFF> Ln;
        Ln.dr = dr;
        const int nB = DIR == 1 ? K : M, dB = DIR == 1 ? N * M : N;
        Ln.x = x_in;
        Ln.b = b;
        Ln.s = s;
        Ln.ae = ae;
        Ln.fB = DIR == 1 ? at : an;
        Ln.fL = DIR == 1 ? an : at;
        Ln.fs = fs;
        Ln.strideL = (int)stride;
        Ln.len = len;
        Ln.pin0 = pin0;
        Ln.pin1 = pin1;
        Ln.hAl = xx > 0;
        Ln.hAu = xx < N - 1;
        Ln.hBl = other > 0;
        Ln.hBu = other < nB - 1;
        Ln.oAl = Ln.hAl ? -1 : 0;
        Ln.oAu = Ln.hAu ? 1 : 0;
        Ln.oBl = Ln.hBl ? -dB : 0;
        Ln.oBu = Ln.hBu ? dB : 0;
        Ln.l32 = lane;
        if constexpr (DIR == 1) {
          const int x0 = (lb % gxb) * LPW, xl = x0 + LPW - 1 < N - 1 ? x0 + LPW - 1 : N - 1;
          Ln.oW0 = (x0 - xx) + (x0 > 0 ? -1 : 0);
          Ln.oE0 = (xl - xx) + (xl < N - 1 ? 1 : 0);
        } else {
          // the group's west and east halo cells p0 - 1 and pl + 1 (read only by a first or last
          // lane that has that x neighbour, which then lies in its row)
          Ln.oW0 = (int)(p0 - base) + (p0 > 0 ? -1 : 0);
          Ln.oE0 = (int)(pl - base) + (pl < NM - 1 ? 1 : 0);
        }
        T carry = (T)0;
  #pragma unroll
        for (int k = 0; k < SEG; ++k) {
          const int kk = k0 + (k < nrow ? k : nrow - 1);
          T lo, dg, up, rh;
          Ln.row((unsigned)(base + stride * kk), kk, k == 0, carry, lo, dg, up, rh);
          elim(k, lo, dg, up, rh);
        }
      } else {
  #pragma unroll
        for (int k = 0; k < SEG; ++k) {
          const int kk = k0 + (k < nrow ? k : nrow - 1);
          const int yy = DIR == 1 ? kk : other, zz = DIR == 1 ? other : kk;
          T lo, dg, up, rh;
          line_row<T, FT, DIR, ZERO>(x_in, b, ae, an, at, s, fs, base + stride * kk, xx, yy, zz, N, M, K, pin0, pin1,
                                     true, lo, dg, up, rh);
          elim(k, lo, dg, up, rh);
        }
      }
    }
    {
      T Gn = (T)0, Hn = (T)0, Qn = (T)1;  // x_{nrow-1} = x_E
  #pragma unroll
      for (int k = SEG - 2; k >= 0; --k) {
        const bool act = k < nrow - 1;
        const T c = Q[k];
        const T gn = G[k] - c * Gn, hn = H[k] - c * Hn, qn = -c * Qn;
        Gn = act ? gn : Gn;
        Hn = act ? hn : Hn;
        Qn = act ? qn : Qn;
        G[k] = Gn;
        H[k] = Hn;
        Q[k] = Qn;
        const bool bl = k == nrow - 2;
        Gb = bl ? Gn : Gb;
        Hb = bl ? Hn : Hb;
        Qb = bl ? Qn : Qb;
      }
      if (nrow > 1) {
        Gf = Gn;
        Hf = Hn;
        Qf = Qn;
      }
    }
  };
  if (cnt == SEG)
    eliminate(std::true_type{});
  else
    eliminate(std::false_type{});
  if ((S & (S - 1)) == 0) {
    // S a power of two (<= 32): the reduced systems are solved inside wavefronts.  Each thread
    // posts its partial last-cell row and its first cell; after one barrier thread t takes
    // segment t % S of line t / S, so the S rows of a line sit in S adjacent lanes of one
    // wavefront and the cyclic reduction runs on cross-lane shuffles with no further barrier.
    IF(0, seg, lane) = le * Hb;         // a
    IF(1, seg, lane) = de + le * Qb;    // d without the next segment's term
    IF(2, seg, lane) = re - le * Gb;    // r without the next segment's term
    IF(3, seg, lane) = ue;
    IF(4, seg, lane) = Gf;
    IF(5, seg, lane) = Hf;
    IF(6, seg, lane) = Qf;
    __syncthreads();
    const int tl = (int)threadIdx.x / S, ts = (int)threadIdx.x % S;  // line, segment
    const int w64 = (int)threadIdx.x & 63;                         // lane in the wavefront
    const bool nx = ts + 1 < S;
    const T u0 = IF(3, ts, tl);
    const T G1 = nx ? IF(4, ts + 1, tl) : (T)0;
    const T H1 = nx ? IF(5, ts + 1, tl) : (T)0;
    const T Q1 = nx ? IF(6, ts + 1, tl) : (T)0;
    // row of the last cell: lo x_before + dg x_E + up x_first(next segment)
    T a = IF(0, ts, tl);
    T d = IF(1, ts, tl) + u0 * H1;
    T c = u0 * Q1;
    T r = IF(2, ts, tl) - u0 * G1;
    for (int st = 1; st < S; st <<= 1) {
      const bool okm = ts - st >= 0, okp = ts + st < S;
      const int sm = okm ? w64 - st : w64, sp = okp ? w64 + st : w64;
      const T am0 = __shfl(a, sm, 64), dm0 = __shfl(d, sm, 64), cm0 = __shfl(c, sm, 64), rm0 = __shfl(r, sm, 64);
      const T ap0 = __shfl(a, sp, 64), dp0 = __shfl(d, sp, 64), cp0 = __shfl(c, sp, 64), rp0 = __shfl(r, sp, 64);
      const T am = okm ? am0 : (T)0, dm = okm ? dm0 : (T)1, cm = okm ? cm0 : (T)0, rm = okm ? rm0 : (T)0;
      const T ap = okp ? ap0 : (T)0, dp = okp ? dp0 : (T)1, cq = okp ? cp0 : (T)0, rp = okp ? rp0 : (T)0;
      const T al = -a * recip0(dm), ga = -c * recip0(dp);
      a = al * am;
      c = ga * cq;
      d = d + al * cm + ga * ap;
      r = r + al * rm + ga * rp;
    }
    IF(7, ts, tl) = r * recip0(d);
    __syncthreads();
  } else {
    // general S: cyclic reduction across the wavefronts through LDS, one barrier pair per step
    IF(4, seg, lane) = Gf;
    IF(5, seg, lane) = Hf;
    IF(6, seg, lane) = Qf;
    __syncthreads();
    const bool nxt = seg + 1 < S;
    const T G1 = nxt ? IF(4, seg + 1, lane) : (T)0;
    const T H1 = nxt ? IF(5, seg + 1, lane) : (T)0;
    const T Q1 = nxt ? IF(6, seg + 1, lane) : (T)0;
    // row of the last cell: lo x_before + dg x_E + up x_first(next segment)
    T a = le * Hb;
    T d = de + le * Qb + ue * H1;
    T c = ue * Q1;
    T r = re - le * Gb - ue * G1;
    for (int st = 1; st < S; st <<= 1) {  // S is uniform over the workgroup
      IF(0, seg, lane) = a;
      IF(1, seg, lane) = d;
      IF(2, seg, lane) = c;
      IF(3, seg, lane) = r;
      __syncthreads();
      const bool okm = seg - st >= 0, okp = seg + st < S;
      const T am = okm ? IF(0, seg - st, lane) : (T)0, dm = okm ? IF(1, seg - st, lane) : (T)1;
      const T cm = okm ? IF(2, seg - st, lane) : (T)0, rm = okm ? IF(3, seg - st, lane) : (T)0;
      const T ap = okp ? IF(0, seg + st, lane) : (T)0, dp = okp ? IF(1, seg + st, lane) : (T)1;
      const T cq = okp ? IF(2, seg + st, lane) : (T)0, rp = okp ? IF(3, seg + st, lane) : (T)0;
      const T al = -a * recip0(dm), ga = -c * recip0(dp);
      a = al * am;
      c = ga * cq;
      d = d + al * cm + ga * ap;
      r = r + al * rm + ga * rp;
      __syncthreads();
    }
    IF(7, seg, lane) = r * recip0(d);
    __syncthreads();
  }
  const T xe = IF(7, seg, lane);
  const T xl = seg > 0 ? IF(7, seg - 1, lane) : (T)0;
  if (live) {
#pragma unroll
    for (int k = 0; k < SEG; ++k)
      if (k < cnt) {
        const unsigned i = (unsigned)(base + stride * (k0 + k));
        __builtin_nontemporal_store(k == cnt - 1 ? xe : G[k] + H[k] * xl + Q[k] * xe,
                                    reinterpret_cast<T*>(reinterpret_cast<char*>(x_out) + i * (unsigned)sizeof(T)));
      }
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
  k_aggregate<<<grid_1d(dc.n()), 256, 0, s>>>(fe, fn, ft, df.N, df.M, df.K, ce, cn, ct, dc.N, dc.M, dc.K);
}

void diag_from_faces(const double* fe, const double* fn, const double* ft, Dims d, const double* ex, double* dg,
                     hipStream_t s) {
  k_diag_from_faces<<<grid_1d(d.n()), 256, 0, s>>>(fe, fn, ft, d.N, d.M, d.K, ex, dg);
}

void excess(const double* fe, const double* fn, const double* ft, const double* dg, Dims d, double* ex,
            hipStream_t s) {
  k_excess<<<grid_1d(d.n()), 256, 0, s>>>(fe, fn, ft, dg, d.N, d.M, d.K, ex);
}

void aggregate_cells(const double* ef, Dims df, double* ec, Dims dc, hipStream_t s) {
  k_aggregate_cells<<<grid_1d(dc.n()), 256, 0, s>>>(ef, df.N, df.M, df.K, ec, dc.N, dc.M, dc.K);
}

void pin_sigma(const double* fe, const double* fn, const double* ft, Dims d, int pin0, int pin1, double* sigma,
               hipStream_t s) {
  k_pin_sigma<<<1, 64, 0, s>>>(fe, fn, ft, d.N, d.M, d.K, pin0, pin1, sigma);
}

void pin_sink(double* dg, int c0, int c1, const double* sigma, hipStream_t s) {
  k_pin_sink<<<1, 64, 0, s>>>(dg, c0, c1, sigma);
}

void diag_rule(double* fe, const double* fn, const double* ft, const double* dg, Dims d, int pin0, int pin1,
               int sink0, int sink1, double sig0, double sig1, double dz, double hscale, int mode, double* v,
               double* partial, double* out, hipStream_t s) {
  k_diag_rule<<<grid_1d(d.n()), 256, 0, s>>>(fe, fn, ft, dg, d.N, d.M, d.K, pin0, pin1, sink0, sink1, sig0, sig1, dz,
                                             hscale, mode, v);
  if (mode != 2) face_absmax(v, v, v, d.n(), partial, out, s);
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
static void jacobi_t(const Level& L, const T* x_in, T* x_out, const T* b, T omega, int zero, int trans,
                     hipStream_t s) {
  const FT* ae = (const FT*)L.f_e;
  const FT* an = (const FT*)L.f_n;
  const FT* at = (const FT*)L.f_t;
  const T* sv = (const T*)L.s;
  const T fs = (T)L.fscale;
  const Dims& d = L.d;
#define LSX_JACOBI(Z, TR)                                                                                     \
  k_jacobi<T, FT, Z, TR><<<(unsigned)((long)((d.N + 63) / 64) * ((d.M + 3) / 4) * d.K), 256, 0, s>>>(             \
      x_in, x_out, b, ae, an, at, sv, fs, omega, d.N, d.M, d.K,                                                \
                                                        L.pin0, L.pin1)
  if (zero)
    LSX_JACOBI(true, false);  // zero start: the transposed first sweep is the same
  else if (trans)
    LSX_JACOBI(false, true);
  else
    LSX_JACOBI(false, false);
#undef LSX_JACOBI
}

template <>
void jacobi<double>(const Level& L, const double* x_in, double* x_out, const double* b, double omega, int zero,
                    int trans, hipStream_t s) {
  jacobi_t<double, double>(L, x_in, x_out, b, omega, zero, trans, s);
}
template <>
void jacobi<float>(const Level& L, const float* x_in, float* x_out, const float* b, float omega, int zero, int trans,
                   hipStream_t s) {
  if (L.ftype == FACE_F16)
    jacobi_t<float, _Float16>(L, x_in, x_out, b, omega, zero, trans, s);
  else
    jacobi_t<float, float>(L, x_in, x_out, b, omega, zero, trans, s);
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
  const long blocks = (long)((L.d.N + 63) / 64) * ((L.d.M + 1) / 2) * ((L.d.K + 1) / 2);
#define LSX_RESID(DFF)                                                                                          \
  k_resid_restrict_cells<T, FT, DFF><<<(unsigned)blocks, 256, 0, s>>>(                                          \
      x, b, (const FT*)L.f_e, (const FT*)L.f_n, (const FT*)L.f_t, (const T*)L.s, (T)L.fscale, L.d.N, L.d.M, L.d.K, \
      L.pin0, L.pin1, bc, diag_rule<T>(L))
  if (L.dff)
    LSX_RESID(true);
  else
    LSX_RESID(false);
#undef LSX_RESID
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
  const long nc = (long)((df.N + 1) / 2) * ((df.M + 1) / 2) * ((df.K + 1) / 2);
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

template <class T, class FT, bool ZERO>
static void line_pass_z(const Level& L, int dir, const T* x_in, T* x_out, const T* b, int trans, hipStream_t s) {
  const Dims& d = L.d;
  const FT* ae = (const FT*)L.f_e;
  const FT* an = (const FT*)L.f_n;
  const FT* at = (const FT*)L.f_t;
  const T* sv = (const T*)L.s;
  const T fs = (T)L.fscale;
  const DiagRule<T> dr = diag_rule<T>(L);
  if (dir == 0) {
    // the FP64 x-line pass reads 1/a_p even on a Level::dff level: rebuilding a_p there measured
    // slower at 256^3 (226 against 195 us), the FP32 one faster (80 against 99 us)
#define LSX_LINE_X(CPL, LANES)                                                                                 \
  do {                                                                                                         \
    const int blocks = (int)(((long)d.M * d.K + 256 / LANES - 1) / (256 / LANES));                             \
    if (trans)                                                                                                 \
      k_line_x<T, FT, CPL, ZERO, true, false, LANES><<<blocks, 256, 0, s>>>(x_in, x_out, b, ae, an, at, sv, fs, \
                                                                            d.N, d.M, d.K, L.pin0, L.pin1, dr); \
    else if (L.dff && sizeof(T) == 4)                                                                          \
      k_line_x<T, FT, CPL, ZERO, false, true, LANES><<<blocks, 256, 0, s>>>(x_in, x_out, b, ae, an, at, sv, fs, \
                                                                            d.N, d.M, d.K, L.pin0, L.pin1, dr); \
    else                                                                                                       \
      k_line_x<T, FT, CPL, ZERO, false, false, LANES><<<blocks, 256, 0, s>>>(x_in, x_out, b, ae, an, at, sv,    \
                                                                             fs, d.N, d.M, d.K, L.pin0, L.pin1, \
                                                                             dr);                               \
  } while (0)
    // lines of up to 128 cells: two per wavefront, four cells per lane
    if (d.N <= 64)
      LSX_LINE_X(1, 64);
    else if (d.N <= 128)
      LSX_LINE_X(4, 32);
    else if (d.N <= 256)
      LSX_LINE_X(4, 64);
    else
      LSX_LINE_X(8, 64);
#undef LSX_LINE_X
  } else {
    const int len = dir == 1 ? d.M : d.K;
    const int S = (len + SEG - 1) / SEG;
    constexpr int LPW = lpw<T>();
    const dim3 grid(dir == 1 ? ((d.N + LPW - 1) / LPW) * d.K : (unsigned)((d.N * (long)d.M + LPW - 1) / LPW));
    const size_t lds = 8 * (size_t)S * (LPW + 1) * sizeof(T);
#define LSX_LINE_SEG(DIR, TR, DFF)                                                                             \
  k_line_seg<T, FT, DIR, ZERO, TR, DFF><<<grid, LPW * S, lds, s>>>(x_in, x_out, b, ae, an, at, sv, fs, d.N, d.M,  \
                                                                   d.K, L.pin0, L.pin1, S, dr)
    if (dir == 1) {
      if (trans)
        LSX_LINE_SEG(1, true, false);
      else if (L.dff)
        LSX_LINE_SEG(1, false, true);
      else
        LSX_LINE_SEG(1, false, false);
    } else {
      if (trans)
        LSX_LINE_SEG(2, true, false);
      else if (L.dff)
        LSX_LINE_SEG(2, false, true);
      else
        LSX_LINE_SEG(2, false, false);
    }
#undef LSX_LINE_SEG
  }
}

template <class T, class FT>
static void line_pass_t(const Level& L, int dir, const T* x_in, T* x_out, const T* b, int zero, int trans,
                        hipStream_t s) {
  if (zero)
    line_pass_z<T, FT, true>(L, dir, x_in, x_out, b, trans, s);
  else
    line_pass_z<T, FT, false>(L, dir, x_in, x_out, b, trans, s);
}

template <>
void line_pass<double>(const Level& L, int dir, const double* x_in, double* x_out, const double* b,
                       int zero, int trans, hipStream_t s) {
  line_pass_t<double, double>(L, dir, x_in, x_out, b, zero, trans, s);
}
template <>
void line_pass<float>(const Level& L, int dir, const float* x_in, float* x_out, const float* b,
                      int zero, int trans, hipStream_t s) {
  if (L.ftype == FACE_F16)
    line_pass_t<float, _Float16>(L, dir, x_in, x_out, b, zero, trans, s);
  else
    line_pass_t<float, float>(L, dir, x_in, x_out, b, zero, trans, s);
}

template <class T, class FT>
static void rb_sweep_t(const Level& L, const T* x_in, T* x_out, const T* b, int zero, hipStream_t s) {
  const Dims& d = L.d;
  // FP64 cells run 512-thread workgroups (8 rows): at 1024 threads a thread has 128 VGPRs and
  // the three planes of FP64 cell data spilled
  constexpr int TY = sizeof(T) == 4 ? FY : FY / 2;
  const dim3 grid(((d.N + FX - 1) / FX) * ((d.M + TY - 1) / TY) * ((d.K + FZ - 1) / FZ));
  k_rb_fused<T, FT, TY><<<grid, FX * TY, 0, s>>>(x_in, x_out, b, (const FT*)L.f_e, (const FT*)L.f_n, (const FT*)L.f_t,
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
                  const double* L, const double* U, const double* dinv, hipStream_t s) {
  if (C.count < 2 || C.count > COARSE_MAX) return;
  if (C.L[0].ftype == FACE_F16)
    k_coarse_cycle<T, _Float16><<<1, COARSE_THREADS, 0, s>>>(C, n_pre, n_post, lu_scale, lu_perm, L, U, dinv);
  else if (C.L[0].ftype == FACE_F32)
    k_coarse_cycle<T, float><<<1, COARSE_THREADS, 0, s>>>(C, n_pre, n_post, lu_scale, lu_perm, L, U, dinv);
  else
    k_coarse_cycle<T, double><<<1, COARSE_THREADS, 0, s>>>(C, n_pre, n_post, lu_scale, lu_perm, L, U, dinv);
}

void dense_inverse(const double* fe, const double* fn, const double* ft, const double* dg, Dims d, double* inv_c,
                   double* inv_r, double* work, hipStream_t s) {
  const int n = (int)d.n();
  double* W = work;                    // n x 2n
  double* sc = W + 2L * n * n;         // n
  double* f = sc + n;                  // n
  double* amax = f + n;                // 1
  const long tot = 2L * n * n;
  k_dense_fill<<<(unsigned)std::min<long>((tot + 255) / 256, 65536), 256, 0, s>>>(fe, fn, ft, dg, d.N, d.M, d.K, sc, W);
  k_dense_absmax<<<1, 1024, 0, s>>>(W, n, amax);
  const dim3 eg((unsigned)((2L * n + 255) / 256), (unsigned)n);
  for (int k = 0; k < n; ++k) {
    k_gj_pivot<<<1, 1024, 0, s>>>(W, n, k, amax, f);
    k_gj_elim<<<eg, 256, 0, s>>>(W, n, k, f);
  }
  k_dense_finish<<<(unsigned)std::min<long>(((long)n * n + 255) / 256, 65536), 256, 0, s>>>(W, sc, n, inv_c, inv_r);
}

template <class T>
void dense_apply(const double* inv, int n, const T* b, T* x, double* part, hipStream_t s) {
  const int nch = (n + DCH - 1) / DCH;
  k_dense_part<T><<<dim3((unsigned)((n + 255) / 256), (unsigned)nch), 256, 0, s>>>(inv, n, b, part);
  k_dense_sum<T><<<(n + 255) / 256, 256, 0, s>>>(part, n, nch, x);
}
template void dense_apply<float>(const double*, int, const float*, float*, double*, hipStream_t);
template void dense_apply<double>(const double*, int, const double*, double*, double*, hipStream_t);

template <class T>
void prolong_add(const T* xc, Dims dc, T* xf, Dims df, hipStream_t s) {
  if (df.N == 2 * dc.N && df.M == 2 * dc.M && df.K == 2 * dc.K) {
    k_prolong_add8<T><<<(unsigned)((long)((dc.N + 63) / 64) * ((dc.M + 3) / 4) * dc.K), 256, 0, s>>>(xc, dc.N, dc.M,
                                                                                                dc.K, xf);
    return;
  }
  k_prolong_add<T><<<(unsigned)((long)((df.N + 63) / 64) * ((df.M + 3) / 4) * df.K), 256, 0, s>>>(xc, dc.N, dc.M, xf,
                                                                                             df.N, df.M, df.K);
}
template void prolong_add<float>(const float*, Dims, float*, Dims, hipStream_t);
template void prolong_add<double>(const double*, Dims, double*, Dims, hipStream_t);
template void coarse_cycle<float>(const CoarseLevels<float>&, int, int, const double*, const int*, const double*,
                                  const double*, const double*, hipStream_t);
template void coarse_cycle<double>(const CoarseLevels<double>&, int, int, const double*, const int*, const double*,
                                   const double*, const double*, hipStream_t);

template <class A, class B>
void convert(const A* a, B* b, long n, hipStream_t s) {
  k_convert<A, B><<<grid_1d(n), 256, 0, s>>>(a, b, n);
}
template void convert<double, float>(const double*, float*, long, hipStream_t);
template void convert<float, double>(const float*, double*, long, hipStream_t);

}  // namespace vc
}  // namespace lsx
