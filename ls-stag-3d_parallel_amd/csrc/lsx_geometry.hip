// Cut-cell geometry of the pressure grid and the pressure operator's coefficients, on the GPU.
//
// One thread per node/cell.  The level-set oscillation fix depends on the order nodes are
// visited (reference ls_stag_uniform_mesh.h:2906-2916 inside the serial sweep
// flow_task_description.h:102-137).  A node can only change there if its level set is positive
// and one of its upper neighbours is negative, so a flag pass marks those few nodes and
// repeated parallel passes over them reach the ordered sweep's result (k_osc_iter); every other
// node keeps its value and all later stages read final values only.
//
// Built with -ffp-contract=off: every formula is evaluated with the operation order of the
// reference expressions it restates.

#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "lsx_geometry.hpp"

namespace lsx {
namespace geo {

namespace {

constexpr double PLUS_TINY = 1e-16;
constexpr double EQ_DOUBLE = 1e-5;
constexpr double THETA_EQ_ZERO = 0.05;
constexpr double EQ_TINY_DOUBLE = 1e-12;
constexpr double ONE_THIRD = 1.0 / 3.0;
constexpr int FLUID_SIDE = 1111;

__device__ inline void decode(long i, int N, int M, int* x, int* y, int* z) {
  const long NM = (long)N * M;
  *z = (int)(i / NM);
  const long r = i - (long)*z * NM;
  *y = (int)(r / N);
  *x = (int)(r - (long)*y * N);
}

__device__ inline bool in_box(int x, int y, const Box& b) { return x >= b.hs && x < b.hf && y >= b.vs && y < b.vf; }

// ---- six-patch bicubic Bezier body ----------------------------------------------------------------
// Patches on the six cube faces: face f has normal axis BZ_A[f] with sign BZ_S[f], u along BZ_U[f]
// and v along BZ_V[f] (S_u x S_v outwards); control point (i, j) of face f, coordinate c, at
// ctrl[(f * 16 + i * 4 + j) * 3 + c].  phi = +distance inside, -distance outside: per patch 12
// Newton steps from the cube-sphere inverse of the point (a coordinate at its bound with the
// gradient pointing out holds there while the other takes a 1D step; Gauss-Newton where the
// Hessian is not positive definite), smallest distance wins, side from (x - S) . (S_u x S_v).
__constant__ int BZ_A[6] = {0, 0, 1, 1, 2, 2};
__constant__ double BZ_S[6] = {1.0, -1.0, 1.0, -1.0, 1.0, -1.0};
__constant__ int BZ_U[6] = {1, 2, 2, 0, 0, 1};
__constant__ int BZ_V[6] = {2, 1, 0, 2, 1, 0};

__device__ inline void bz_basis(double t, double* b, double* db, double* ddb) {
  const double s = 1.0 - t;
  b[0] = s * s * s;
  b[1] = 3.0 * t * s * s;
  b[2] = 3.0 * t * t * s;
  b[3] = t * t * t;
  db[0] = -3.0 * s * s;
  db[1] = 3.0 * s * s - 6.0 * t * s;
  db[2] = 6.0 * t * s - 3.0 * t * t;
  db[3] = 3.0 * t * t;
  ddb[0] = 6.0 * s;
  ddb[1] = 18.0 * t - 12.0;
  ddb[2] = 6.0 - 18.0 * t;
  ddb[3] = 6.0 * t;
}

// S, S_u and S_v of patch P at (u, v): each row i of the control net is first reduced along v
// (q_i = sum_j B_j(v) P_ij and its v derivative), then the rows along u
__device__ inline void bz_eval1(const double* __restrict__ P, double u, double v, double* S, double* Su, double* Sv) {
  double bu[4], dbu[4], ddbu[4], bv[4], dbv[4], ddbv[4];
  bz_basis(u, bu, dbu, ddbu);
  bz_basis(v, bv, dbv, ddbv);
  for (int c = 0; c < 3; ++c) S[c] = Su[c] = Sv[c] = 0.0;
  for (int i = 0; i < 4; ++i)
    for (int c = 0; c < 3; ++c) {
      double q = 0.0, dq = 0.0;
      for (int j = 0; j < 4; ++j) {
        const double p = P[(i * 4 + j) * 3 + c];
        q += bv[j] * p;
        dq += dbv[j] * p;
      }
      S[c] += bu[i] * q;
      Su[c] += dbu[i] * q;
      Sv[c] += bu[i] * dq;
    }
}

// r . S_uu, r . S_uv and r . S_vv of patch P at (u, v), through w_ij = P_ij . r
__device__ inline void bz_eval2(const double* __restrict__ P, double u, double v, const double* r, double* ruu,
                                double* ruv, double* rvv) {
  double bu[4], dbu[4], ddbu[4], bv[4], dbv[4], ddbv[4];
  bz_basis(u, bu, dbu, ddbu);
  bz_basis(v, bv, dbv, ddbv);
  double a = 0.0, b = 0.0, c = 0.0;
  for (int i = 0; i < 4; ++i) {
    double q = 0.0, dq = 0.0, ddq = 0.0;
    for (int j = 0; j < 4; ++j) {
      const double* pij = P + (i * 4 + j) * 3;
      const double w = pij[0] * r[0] + pij[1] * r[1] + pij[2] * r[2];
      q += bv[j] * w;
      dq += dbv[j] * w;
      ddq += ddbv[j] * w;
    }
    a += ddbu[i] * q;
    b += dbu[i] * dq;
    c += bu[i] * ddq;
  }
  *ruu = a;
  *ruv = b;
  *rvv = c;
}

__device__ inline double bz_dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ inline double bz_clamp(double t) { return t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t); }

// closest point of the winning patch: phi = sgn * |X - S(u, v)| with r = X - S
struct BzHit {
  int f;
  double u, v, r[3], dist, sgn;
};

// Newton search for the closest point of patch f (control values P) to X (d: X in the body's
// normalised coordinates, for the start); returns |X - S|^2 with (u, v), r = X - S and
// side = r . (S_u x S_v) there.  A step that leaves (u, v) unchanged ends the search: the steps
// after it would repeat it exactly.
__device__ inline double bz_patch(const double* __restrict__ P, int f, const double* X, const double* d, double& u,
                                  double& v, double* r, double& side) {
  const double da = fabs(d[BZ_A[f]]);
  u = 0.5;
  v = 0.5;
  if (da > 0.0) {
    u = bz_clamp(0.5 * (d[BZ_U[f]] / da + 1.0));
    v = bz_clamp(0.5 * (d[BZ_V[f]] / da + 1.0));
  }
  double S[3], Su[3], Sv[3];
  for (int it = 0; it < 12; ++it) {
    bz_eval1(P, u, v, S, Su, Sv);
    double q[3];
    for (int c = 0; c < 3; ++c) q[c] = S[c] - X[c];
    const double gu = bz_dot(Su, q), gv = bz_dot(Sv, q);
    const double juu = bz_dot(Su, Su), juv = bz_dot(Su, Sv), jvv = bz_dot(Sv, Sv);
    double ruu, ruv, rvv;
    bz_eval2(P, u, v, q, &ruu, &ruv, &rvv);
    double huu = juu + ruu, huv = juv + ruv, hvv = jvv + rvv;
    double det = huu * hvv - huv * huv;
    if (!(det > 0.0 && huu > 0.0)) {
      huu = juu;
      huv = juv;
      hvv = jvv;
      det = huu * hvv - huv * huv;
    }
    const bool hold_u = (u >= 1.0 && gu < 0.0) || (u <= 0.0 && gu > 0.0);
    const bool hold_v = (v >= 1.0 && gv < 0.0) || (v <= 0.0 && gv > 0.0);
    double du = 0.0, dv = 0.0;
    if (hold_u && hold_v) break;
    if (hold_u) {
      dv = -gv / (hvv > 0.0 ? hvv : jvv);
    } else if (hold_v) {
      du = -gu / (huu > 0.0 ? huu : juu);
    } else {
      if (!(det > 0.0)) break;
      du = -(hvv * gu - huv * gv) / det;
      dv = -(huu * gv - huv * gu) / det;
    }
    const double un = bz_clamp(u + du), vn = bz_clamp(v + dv);
    if (un == u && vn == v) break;
    u = un;
    v = vn;
  }
  bz_eval1(P, u, v, S, Su, Sv);
  for (int c = 0; c < 3; ++c) r[c] = X[c] - S[c];
  const double nrm[3] = {Su[1] * Sv[2] - Su[2] * Sv[1], Su[2] * Sv[0] - Su[0] * Sv[2], Su[0] * Sv[1] - Su[1] * Sv[0]};
  side = bz_dot(r, nrm);
  return bz_dot(r, r);
}

// squared distance from X to the bounding box of patch P's control points (the patch lies in
// their convex hull)
__device__ inline double bz_box_dist2(const double* __restrict__ P, const double* X) {
  double acc = 0.0;
  for (int c = 0; c < 3; ++c) {
    double lo = P[c], hi = P[c];
    for (int k = 1; k < 16; ++k) {
      lo = fmin(lo, P[3 * k + c]);
      hi = fmax(hi, P[3 * k + c]);
    }
    const double e = X[c] < lo ? lo - X[c] : (X[c] > hi ? X[c] - hi : 0.0);
    acc += e * e;
  }
  return acc;
}

// The patch with the smallest distance wins, the lowest index among equal distances.  Each point
// first searches its home patch (the cube face its normalised direction points to); another
// patch is searched only if its control-point box comes within the best distance so far plus a
// margin of 1e-12, far above rounding (a patch further away cannot win or tie).  The loops run
// over the patches in order so that the control values stay uniform over the wavefront.
__device__ double bezier_closest(const double* __restrict__ ctrl, const double* c0, const double* a0, double x,
                                 double y, double z, BzHit& hit) {
  const double X[3] = {x, y, z};
  double d[3];
  for (int c = 0; c < 3; ++c) d[c] = (X[c] - c0[c]) / a0[c];
  int ax = 0;
  for (int c = 1; c < 3; ++c)
    if (fabs(d[c]) > fabs(d[ax])) ax = c;
  const int home = 2 * ax + (d[ax] < 0.0 ? 1 : 0);  // faces 2a, 2a + 1: normal axis a, sign + / -
  double best = 0.0, side = 0.0;
  // pass 0 searches the home patch, pass 1 the others (one call site: one copy of the search)
  for (int k = 0; k < 12; ++k) {
    const int f = k < 6 ? k : k - 6;
    const double* P = ctrl + f * 48;
    if (k < 6) {
      if (f != home) continue;
    } else {
      if (f == home) continue;
      const double m = sqrt(bz_box_dist2(P, X)) - 1e-12;
      if (m > 0.0 && m * m > best) continue;
    }
    double u, v, r[3], sd;
    const double dist2 = bz_patch(P, f, X, d, u, v, r, sd);
    if (k < 6 || dist2 < best || (dist2 == best && f < hit.f)) {
      best = dist2;
      side = sd;
      hit.f = f;
      hit.u = u;
      hit.v = v;
      for (int c = 0; c < 3; ++c) hit.r[c] = r[c];
    }
  }
  const double dist = sqrt(best);
  hit.dist = dist;
  hit.sgn = side > 0.0 ? -1.0 : 1.0;
  return side > 0.0 ? -dist : dist;
}

__device__ inline double bezier_phi(const double* __restrict__ ctrl, const double* c0, const double* a0, double x,
                                    double y, double z) {
  BzHit hit;
  return bezier_closest(ctrl, c0, a0, x, y, z, hit);
}

// ---- level set -------------------------------------------------------------------------------
__global__ void k_level_set(const double* __restrict__ cx, const double* __restrict__ cy, const double* __restrict__ cz,
                            int N, int M, int K, Box box, Body body, double* __restrict__ phi) {
  const long n = (long)N * M * K;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    double v = 0.0;
    if (in_box(x, y, box)) {
      if (body.kind == BODY_SPHERE) {
        const double qx = cx[x] - body.c[0], qy = cy[y] - body.c[1], qz = cz[z] - body.c[2];
        v = body.a[0] - sqrt(qx * qx + qy * qy + qz * qz);
      } else if (body.kind == BODY_BEZIER6) {
        v = bezier_phi(body.ctrl, body.c, body.a, cx[x], cy[y], cz[z]);
      } else {
        const double s = fmin(body.a[0], fmin(body.a[1], body.a[2]));
        const double qx = (cx[x] - body.c[0]) / body.a[0], qy = (cy[y] - body.c[1]) / body.a[1],
                     qz = (cz[z] - body.c[2]) / body.a[2];
        v = s * (1.0 - sqrt(qx * qx + qy * qy + qz * qz));
      }
    }
    phi[i] = v;
  }
}

__global__ void k_osc_flags(const double* __restrict__ phi, int N, int M, int K, Box box, unsigned char* __restrict__ flag) {
  const long n = (long)N * M * K, NM = (long)N * M;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    unsigned char f = 0;
    if (in_box(x, y, box) && phi[i] > EQ_DOUBLE) {
      const bool mid = z > 0 && z < K - 1;
      f = (phi[i + 1] < 0) || (phi[i + N] < 0) || (mid && phi[i + NM] < 0);
    }
    flag[i] = f;
  }
}

__device__ inline void check_osc(double& f, double lower, double upper) {
  if ((f > EQ_DOUBLE) && (lower < 0) && (upper < 0)) f = 0.5 * (lower + upper);
}

// The reference visits nodes in index order and updates phi in place, so a flagged node reads
// the final values of its lower neighbours (j-1, j-N, j-NM: visited before it) and the original
// values of its upper ones.  One pass recomputes every flagged node from the current lower
// values (`cur`) and the original upper values (`phi0`) into `nxt`; node j depends only on nodes
// below it, so repeating the pass until nothing changes reproduces the ordered sweep exactly.
__global__ void k_osc_iter(const double* __restrict__ phi0, const unsigned char* __restrict__ flag,
                           const double* __restrict__ cur, double* __restrict__ nxt, int N, int M, int K,
                           int* __restrict__ changed) {
  const long n = (long)N * M * K, NM = (long)N * M;
  for (long j = blockIdx.x * (long)blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x) {
    if (!flag[j]) continue;
    const int z = (int)(j / NM);
    double f = phi0[j];
    check_osc(f, cur[j - 1], phi0[j + 1]);
    check_osc(f, cur[j - N], phi0[j + N]);
    if (z > 0 && z < K - 1) check_osc(f, cur[j - NM], phi0[j + NM]);
    if (f != cur[j]) *changed = 1;
    nxt[j] = f;
  }
}

__device__ inline double edge_fraction(double f, double fo) {
  return (f * fo < 0) ? ((f > 0) ? (fo / (fo - f)) : (f / (f - fo))) : ((f > -fo) ? 0.0 : 1.0);
}

__global__ void k_alpha(const double* __restrict__ phi, int N, int M, int K, Box box, double* __restrict__ ax,
                        double* __restrict__ ay, double* __restrict__ az) {
  const long n = (long)N * M * K, NM = (long)N * M;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    double a0 = 1.0, a1 = 1.0, a2 = 1.0;
    if (in_box(x, y, box)) {
      const double f = phi[i];
      a0 = edge_fraction(f, phi[i - 1]);
      a1 = edge_fraction(f, phi[i - N]);
      a2 = (z == 0 || z == K - 1) ? ((f > 0.0) ? 0.0 : 1.0) : edge_fraction(f, phi[i - NM]);
    }
    ax[i] = a0;
    ay[i] = a1;
    az[i] = a2;
  }
}

// ---- types -----------------------------------------------------------------------------------
__device__ inline int type_of_side(double phi, double phiW, double phiSW, double phiS) {
  int number = (phiS < EQ_DOUBLE) ? 1 : 0;
  if (phiSW < EQ_DOUBLE) number += 10;
  if (phiW < EQ_DOUBLE) number += 100;
  if (phi < EQ_DOUBLE) number += 1000;
  return number;
}

__device__ inline int border_type(int x, int y, int z, int N, int M, int K) {
  const bool w = x == 0, e = x == N - 1, s = y == 0, nn = y == M - 1, b = z == 0, t = z == K - 1;
  if (b) {
    if (s) return w ? BSW_Corner : (e ? BSE_Corner : BS_Edge);
    if (nn) return w ? BNW_Corner : (e ? BNE_Corner : BN_Edge);
    return w ? BW_Edge : (e ? BE_Edge : B_Border);
  }
  if (t) {
    if (s) return w ? TSW_Corner : (e ? TSE_Corner : TS_Edge);
    if (nn) return w ? TNW_Corner : (e ? TNE_Corner : TN_Edge);
    return w ? TW_Edge : (e ? TE_Edge : T_Border);
  }
  if (s) return w ? SW_Edge : (e ? SE_Edge : S_Border);
  if (nn) return w ? NW_Edge : (e ? NE_Edge : N_Border);
  if (w) return W_Border;
  if (e) return E_Border;
  return FluidCell;
}

__global__ void k_types(const double* __restrict__ p, int N, int M, int K, Box box, int* __restrict__ type,
                        int* __restrict__ sx, int* __restrict__ sy, int* __restrict__ sz) {
  const long n = (long)N * M * K, NM = (long)N * M;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    int t = border_type(x, y, z, N, M, K);
    int s0 = FLUID_SIDE, s1 = FLUID_SIDE, s2 = FLUID_SIDE;
    if (in_box(x, y, box)) {
      const double phi = p[i], phiW = p[i - 1], phiS = p[i - N];
      const int tT = type_of_side(phi, phiW, p[i - N - 1], phiS);
      if (z == 0 || z == K - 1) {
        s0 = type_of_side(phi, phiS, phiS, phi);
        s1 = type_of_side(phi, phiW, phiW, phi);
        s2 = tT;
        t = (tT == 0) ? SolidCell : (z == 0 ? B_Border : T_Border);
      } else {
        const double phiSB = p[i - N - NM], phiB = p[i - NM], phiBW = p[i - NM - 1];
        s0 = type_of_side(phi, phiS, phiSB, phiB);
        s1 = type_of_side(phi, phiW, phiBW, phiB);
        s2 = tT;
        t = tT * 10000 + type_of_side(phiB, phiBW, p[i - N - 1 - NM], phiSB);
      }
    }
    type[i] = t;
    sx[i] = s0;
    sy[i] = s1;
    sz[i] = s2;
  }
}

// ---- face fractions ---------------------------------------------------------------------------
__device__ inline double face_theta(int side, double P0, double P1, double Q0, double Q1) {
  double t;
  switch (side) {
    case FLUID_SIDE: return 1.0;
    case 0: return 0.0;
    case 11: case 1100: t = 0.5 * (P0 + P1); break;
    case 1001: case 110: t = 0.5 * (Q0 + Q1); break;
    case 1011: t = 1.0 - 0.5 * (1.0 - P1) * (1.0 - Q0); break;
    case 111: t = 1.0 - 0.5 * (1.0 - P0) * (1.0 - Q0); break;
    case 1101: t = 1.0 - 0.5 * (1.0 - P1) * (1.0 - Q1); break;
    case 1110: t = 1.0 - 0.5 * (1.0 - P0) * (1.0 - Q1); break;
    case 1: t = 0.5 * P0 * Q1; break;
    case 10: t = 0.5 * P1 * Q1; break;
    case 1000: t = 0.5 * P0 * Q0; break;
    case 100: t = 0.5 * P1 * Q0; break;
    default: return 1.0;
  }
  return t < THETA_EQ_ZERO ? 0.0 : t;
}

__global__ void k_theta(const double* __restrict__ ax, const double* __restrict__ ay, const double* __restrict__ az,
                        const int* __restrict__ sx, const int* __restrict__ sy, const int* __restrict__ sz, int N, int M,
                        int K, Box box, double* __restrict__ tx, double* __restrict__ ty, double* __restrict__ tz) {
  const long n = (long)N * M * K, NM = (long)N * M;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    double t0 = 1.0, t1 = 1.0, t2 = 1.0;
    if (in_box(x, y, box)) {
      t2 = face_theta(sz[i], ay[i], ay[i - 1], ax[i], ax[i - N]);
      if (z == 0 || z == K - 1) {
        const double a = ay[i], b = ax[i];
        t0 = a < THETA_EQ_ZERO ? 0.0 : a;
        t1 = b < THETA_EQ_ZERO ? 0.0 : b;
      } else {
        t0 = face_theta(sx[i], az[i], az[i - N], ay[i], ay[i - NM]);
        t1 = face_theta(sy[i], az[i], az[i - 1], ax[i], ax[i - NM]);
      }
    }
    tx[i] = t0;
    ty[i] = t1;
    tz[i] = t2;
  }
}

// ---- cut volumes --------------------------------------------------------------------------------
// Formula families of the reference volume table; each cut code names a family and the
// alpha/theta entries (component, corner offset) it reads.
enum Fam : unsigned char { F_TET, F_HEP3, F_PRISM, F_HEP5, F_HEX5, F_HEP6 };
struct Ref {
  signed char dx, dy, dz, comp;
};
struct Rule {
  int code;
  unsigned char fam, axis, sel, pad;
  Ref a, b, a2, b2;
};

#define R_(dx, dy, dz, c) Ref{dx, dy, dz, c}
#define NONE R_(0, 0, 0, 0)
// sorted by code for binary search
__constant__ Rule c_rules[] = {
    {1, F_TET, 0, 0, 0, R_(0, -1, -1, 0), R_(0, 0, 0, 0), NONE, NONE},
    {10, F_TET, 0, 0, 0, R_(0, -1, -1, 0), R_(-1, 0, 0, 0), NONE, NONE},
    {11, F_PRISM, 0, 0, 0, NONE, NONE, NONE, NONE},
    {100, F_TET, 0, 0, 0, R_(0, 0, -1, 0), R_(-1, 0, 0, 0), NONE, NONE},
    {110, F_PRISM, 1, 0, 0, NONE, NONE, NONE, NONE},
    {111, F_HEX5, 1, 1, 0, R_(0, 0, -1, 0), R_(0, 0, 0, 1), NONE, NONE},
    {1000, F_TET, 0, 0, 0, R_(0, 0, -1, 0), R_(0, 0, 0, 0), NONE, NONE},
    {1001, F_PRISM, 1, 0, 0, NONE, NONE, NONE, NONE},
    {1011, F_HEX5, 1, 1, 0, R_(0, 0, -1, 0), R_(-1, 0, 0, 1), NONE, NONE},
    {1100, F_PRISM, 0, 0, 0, NONE, NONE, NONE, NONE},
    {1101, F_HEX5, 1, 0, 0, R_(0, -1, -1, 0), R_(-1, 0, 0, 1), NONE, NONE},
    {1110, F_HEX5, 1, 0, 0, R_(0, -1, -1, 0), R_(0, 0, 0, 1), NONE, NONE},
    {1111, F_PRISM, 0, 0, 0, NONE, NONE, NONE, NONE},
    {10000, F_TET, 0, 0, 0, R_(0, -1, 0, 0), R_(0, 0, 0, 0), NONE, NONE},
    {10001, F_PRISM, 2, 0, 0, NONE, NONE, NONE, NONE},
    {10011, F_HEX5, 0, 0, 0, R_(-1, -1, 0, 2), R_(0, -1, 0, 0), NONE, NONE},
    {11001, F_HEX5, 2, 1, 0, R_(0, 0, 0, 1), R_(0, 0, 0, 2), NONE, NONE},
    {11011, F_HEP6, 0, 1, 0, R_(0, 0, 0, 2), R_(0, 0, -1, 0), R_(0, 0, 0, 1), R_(0, -1, 0, 0)},
    {11111, F_HEP5, 1, 0, 0, R_(0, -1, 0, 0), R_(-1, 0, 0, 1), NONE, NONE},
    {100000, F_TET, 0, 0, 0, R_(0, -1, 0, 0), R_(-1, 0, 0, 0), NONE, NONE},
    {100010, F_PRISM, 2, 0, 0, NONE, NONE, NONE, NONE},
    {100011, F_HEX5, 0, 1, 0, R_(0, -1, 0, 2), R_(0, -1, 0, 0), NONE, NONE},
    {100110, F_HEX5, 2, 1, 0, R_(-1, 0, 0, 1), R_(-1, 0, 0, 2), NONE, NONE},
    {100111, F_HEP6, 0, 0, 0, R_(-1, 0, 0, 2), R_(0, 0, -1, 0), R_(-1, 0, 0, 1), R_(0, -1, 0, 0)},
    {101111, F_HEP5, 1, 0, 0, R_(0, -1, 0, 0), R_(0, 0, 0, 1), NONE, NONE},
    {110000, F_PRISM, 0, 0, 0, NONE, NONE, NONE, NONE},
    {110001, F_HEX5, 0, 0, 0, R_(-1, -1, 0, 2), R_(0, -1, -1, 0), NONE, NONE},
    {110010, F_HEX5, 0, 1, 0, R_(0, -1, 0, 2), R_(0, -1, -1, 0), NONE, NONE},
    {110011, F_PRISM, 0, 0, 0, NONE, NONE, NONE, NONE},
    {110111, F_HEP5, 0, 0, 0, R_(-1, 0, 0, 2), R_(0, 0, -1, 0), NONE, NONE},
    {111011, F_HEP5, 0, 1, 0, R_(0, 0, 0, 2), R_(0, 0, -1, 2), NONE, NONE},
    {111111, F_PRISM, 0, 0, 0, NONE, NONE, NONE, NONE},
    {1000000, F_TET, 0, 0, 0, R_(0, 0, 0, 0), R_(-1, 0, 0, 0), NONE, NONE},
    {1000100, F_PRISM, 2, 0, 0, NONE, NONE, NONE, NONE},
    {1000110, F_HEX5, 2, 1, 0, R_(-1, 0, 0, 1), R_(-1, -1, 0, 2), NONE, NONE},
    {1001100, F_HEX5, 0, 1, 0, R_(0, 0, 0, 2), R_(0, 0, 0, 0), NONE, NONE},
    {1001110, F_HEP6, 0, 0, 0, R_(-1, -1, 0, 2), R_(0, -1, -1, 0), R_(-1, 0, 0, 1), R_(0, 0, 0, 0)},
    {1001111, F_HEP5, 1, 1, 0, R_(0, 0, 0, 0), R_(-1, 0, 0, 1), NONE, NONE},
    {1100000, F_PRISM, 1, 0, 0, NONE, NONE, NONE, NONE},
    {1100010, F_HEX5, 2, 0, 0, R_(-1, 0, -1, 1), R_(-1, 0, 0, 2), NONE, NONE},
    {1100100, F_HEX5, 2, 0, 0, R_(-1, 0, -1, 1), R_(-1, -1, 0, 2), NONE, NONE},
    {1100110, F_PRISM, 1, 0, 0, NONE, NONE, NONE, NONE},
    {1100111, F_HEP5, 2, 0, 0, R_(0, 0, -1, 1), R_(0, -1, 0, 2), NONE, NONE},
    {1101110, F_HEP5, 2, 0, 0, R_(0, 0, -1, 1), R_(0, 0, 0, 2), NONE, NONE},
    {1101111, F_PRISM, 1, 0, 0, NONE, NONE, NONE, NONE},
    {1110000, F_HEX5, 1, 1, 0, R_(0, 0, 0, 0), R_(0, 0, 0, 1), NONE, NONE},
    {1110010, F_HEP6, 0, 0, 0, R_(-1, 0, 0, 2), R_(0, 0, 0, 0), R_(-1, 0, -1, 1), R_(0, -1, -1, 0)},
    {1110011, F_HEP5, 0, 0, 0, R_(-1, 0, 0, 2), R_(0, 0, 0, 0), NONE, NONE},
    {1110110, F_HEP5, 2, 1, 0, R_(0, 0, 0, 1), R_(0, -1, 0, 2), NONE, NONE},
    {1110111, F_PRISM, 2, 0, 0, NONE, NONE, NONE, NONE},
    {1111111, F_HEP3, 0, 0, 0, R_(0, 0, 0, 0), R_(0, 0, 0, 0), NONE, NONE},
    {10000000, F_TET, 0, 0, 0, R_(0, 0, 0, 0), R_(0, 0, 0, 0), NONE, NONE},
    {10001000, F_PRISM, 2, 0, 0, NONE, NONE, NONE, NONE},
    {10001001, F_HEX5, 2, 1, 0, R_(0, 0, 0, 1), R_(0, -1, 0, 2), NONE, NONE},
    {10001100, F_HEX5, 0, 0, 0, R_(-1, 0, 0, 2), R_(0, 0, 0, 0), NONE, NONE},
    {10001101, F_HEP6, 0, 1, 0, R_(0, -1, 0, 2), R_(0, -1, -1, 0), R_(0, 0, 0, 1), R_(0, 0, 0, 0)},
    {10001111, F_HEP5, 1, 1, 0, R_(0, 0, 0, 0), R_(0, 0, 0, 1), NONE, NONE},
    {10010000, F_PRISM, 1, 0, 0, NONE, NONE, NONE, NONE},
    {10010001, F_HEX5, 2, 0, 0, R_(0, 0, -1, 1), R_(0, 0, 0, 2), NONE, NONE},
    {10011000, F_HEX5, 2, 0, 0, R_(0, 0, -1, 1), R_(0, -1, 0, 2), NONE, NONE},
    {10011001, F_PRISM, 1, 0, 0, NONE, NONE, NONE, NONE},
    {10011011, F_HEP5, 2, 0, 0, R_(-1, 0, -1, 1), R_(-1, -1, 0, 2), NONE, NONE},
    {10011101, F_HEP5, 2, 0, 0, R_(-1, 0, -1, 1), R_(-1, 0, 0, 2), NONE, NONE},
    {10011111, F_PRISM, 1, 0, 0, NONE, NONE, NONE, NONE},
    {10110000, F_HEX5, 1, 1, 0, R_(0, 0, 0, 0), R_(-1, 0, 0, 1), NONE, NONE},
    {10110001, F_HEP6, 0, 1, 0, R_(0, 0, 0, 2), R_(0, 0, 0, 0), R_(0, 0, -1, 1), R_(0, -1, -1, 0)},
    {10110011, F_HEP5, 0, 1, 0, R_(0, 0, 0, 2), R_(0, 0, 0, 0), NONE, NONE},
    {10111001, F_HEP5, 2, 1, 0, R_(-1, 0, 0, 1), R_(-1, -1, 0, 2), NONE, NONE},
    {10111011, F_PRISM, 2, 0, 0, NONE, NONE, NONE, NONE},
    {10111111, F_HEP3, 0, 0, 0, R_(0, 0, 0, 0), R_(-1, 0, 0, 0), NONE, NONE},
    {11000000, F_PRISM, 0, 0, 0, NONE, NONE, NONE, NONE},
    {11000100, F_HEX5, 0, 1, 0, R_(0, 0, 0, 2), R_(0, 0, -1, 0), NONE, NONE},
    {11001000, F_HEX5, 0, 0, 0, R_(-1, 0, 0, 2), R_(0, 0, -1, 0), NONE, NONE},
    {11001100, F_PRISM, 0, 0, 0, NONE, NONE, NONE, NONE},
    {11001101, F_HEP5, 0, 1, 0, R_(0, -1, 0, 2), R_(0, -1, -1, 0), NONE, NONE},
    {11001110, F_HEP5, 0, 0, 0, R_(-1, -1, 0, 2), R_(0, -1, -1, 0), NONE, NONE},
    {11001111, F_PRISM, 0, 0, 0, NONE, NONE, NONE, NONE},
    {11010000, F_HEX5, 1, 0, 0, R_(0, -1, 0, 0), R_(-1, 0, 0, 1), NONE, NONE},
    {11011000, F_HEP6, 0, 1, 0, R_(0, -1, 0, 2), R_(0, 0, -1, 0), R_(0, 0, -1, 1), R_(0, -1, 0, 0)},
    {11011001, F_HEP5, 2, 1, 0, R_(-1, 0, 0, 1), R_(-1, 0, 0, 2), NONE, NONE},
    {11011100, F_HEP5, 0, 1, 0, R_(0, -1, 0, 2), R_(0, -1, 0, 0), NONE, NONE},
    {11011101, F_PRISM, 2, 0, 0, NONE, NONE, NONE, NONE},
    {11011111, F_HEP3, 0, 0, 0, R_(0, -1, 0, 0), R_(-1, 0, 0, 0), NONE, NONE},
    {11100000, F_HEX5, 1, 0, 0, R_(0, -1, 0, 0), R_(0, 0, 0, 1), NONE, NONE},
    {11100100, F_HEP6, 0, 0, 0, R_(-1, -1, 0, 2), R_(0, -1, 0, 0), R_(-1, 0, -1, 1), R_(0, 0, -1, 0)},
    {11100110, F_HEP5, 2, 1, 0, R_(0, 0, 0, 1), R_(0, 0, 0, 2), NONE, NONE},
    {11101100, F_HEP5, 0, 0, 0, R_(-1, -1, 0, 2), R_(0, -1, 0, 0), NONE, NONE},
    {11101110, F_PRISM, 2, 0, 0, NONE, NONE, NONE, NONE},
    {11101111, F_HEP3, 0, 0, 0, R_(0, -1, 0, 0), R_(0, 0, 0, 0), NONE, NONE},
    {11110000, F_PRISM, 0, 0, 0, NONE, NONE, NONE, NONE},
    {11110001, F_HEP5, 1, 0, 0, R_(0, -1, -1, 0), R_(0, 0, -1, 1), NONE, NONE},
    {11110010, F_HEP5, 1, 0, 0, R_(0, -1, -1, 0), R_(-1, 0, -1, 1), NONE, NONE},
    {11110011, F_PRISM, 0, 0, 0, NONE, NONE, NONE, NONE},
    {11110100, F_HEP5, 1, 1, 0, R_(0, 0, -1, 0), R_(-1, 0, -1, 1), NONE, NONE},
    {11110110, F_PRISM, 1, 0, 0, NONE, NONE, NONE, NONE},
    {11110111, F_HEP3, 0, 0, 0, R_(0, 0, -1, 0), R_(0, 0, 0, 0), NONE, NONE},
    {11111000, F_HEP5, 1, 1, 0, R_(0, 0, -1, 0), R_(0, 0, -1, 1), NONE, NONE},
    {11111001, F_PRISM, 1, 0, 0, NONE, NONE, NONE, NONE},
    {11111011, F_HEP3, 0, 0, 0, R_(0, 0, -1, 0), R_(-1, 0, 0, 0), NONE, NONE},
    {11111100, F_PRISM, 0, 0, 0, NONE, NONE, NONE, NONE},
    {11111101, F_HEP3, 0, 0, 0, R_(0, -1, -1, 0), R_(-1, 0, 0, 0), NONE, NONE},
    {11111110, F_HEP3, 0, 0, 0, R_(0, -1, -1, 0), R_(0, 0, 0, 0), NONE, NONE},
};
#undef NONE
#undef R_
constexpr int N_RULES = sizeof(c_rules) / sizeof(Rule);

__device__ inline int find_rule(int code) {
  int lo = 0, hi = N_RULES - 1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int c = c_rules[mid].code;
    if (c == code) return mid;
    if (c < code) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

struct Fields {
  const double* a[3];
  const double* t[3];
  int N, NM;
};

__device__ inline double at(const double* const* arr, const Fields& F, long i, Ref r) {
  return arr[(int)r.comp][i + r.dx + (long)F.N * r.dy + (long)F.NM * r.dz];
}

__global__ void k_vof(const int* __restrict__ type, Fields F, const double* __restrict__ dx,
                      const double* __restrict__ dy, const double* __restrict__ dz, int N, int M, int K,
                      double* __restrict__ vof) {
  const long n = (long)N * M * K;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    const double fullV = dx[x] * dy[y] * dz[z];
    double vol = fullV;
    const int t = type[i];
    if (t == B_Border || t == T_Border) {
      vol *= F.t[2][i];
    } else if (t == SolidCell) {
      vol = 0.0;
    } else if (t > 0 && t != FluidCell) {
      const int k = find_rule(t);
      if (k >= 0) {
        const Rule r = c_rules[k];
        const long lowOff = r.axis == 0 ? 1 : (r.axis == 1 ? F.N : F.NM);
        const double tUp = F.t[r.axis][i], tLo = F.t[r.axis][i - lowOff];
        switch (r.fam) {
          case F_TET:
            vol = at(F.a, F, i, r.a) * at(F.t, F, i, r.b) * fullV * ONE_THIRD;
            break;
          case F_HEP3:
            vol = (3.0 - (1.0 - at(F.a, F, i, r.a)) * (1.0 - at(F.t, F, i, r.b))) * fullV * ONE_THIRD;
            break;
          case F_PRISM:
            vol = (tUp + tLo) * fullV * 0.5;
            break;
          case F_HEP5: {
            const double tsel = r.sel ? tUp : tLo;
            double ap = at(F.a, F, i, r.a);
            const double denom = 1.0 - ap;
            double term = 0.0;
            if (denom > EQ_DOUBLE) {
              ap /= denom;
              term = (1.0 - tsel) * (3.0 - 2.0 * at(F.a, F, i, r.b)) * ap * ap * ONE_THIRD;
            }
            vol = (tLo + tUp + term) * fullV * 0.5;
            break;
          }
          case F_HEX5: {
            const double denom = at(F.a, F, i, r.a);
            double term2 = r.sel ? tUp : tLo;
            const double other = r.sel ? tLo : tUp;
            if (denom > EQ_DOUBLE) {
              const double coef = (1.0 - denom) / denom;
              term2 *= (1.0 - (1.0 + 2.0 * at(F.a, F, i, r.b)) * coef * coef * ONE_THIRD);
            }
            vol = (other + term2) * fullV * 0.5;
            break;
          }
          default: {  // F_HEP6
            const double tsel = r.sel ? tUp : tLo;
            const double other = r.sel ? tLo : tUp;
            double term = 0.0;
            double al = at(F.a, F, i, r.a);
            double denom = 1.0 - al;
            if (denom > EQ_DOUBLE) {
              al /= denom;
              term = (3.0 - 2.0 * at(F.a, F, i, r.b)) * al * al;
            }
            al = at(F.a, F, i, r.a2);
            denom = 1.0 - al;
            if (denom > EQ_DOUBLE) {
              al /= denom;
              term += (3.0 - 2.0 * at(F.a, F, i, r.b2)) * al * al;
            }
            vol = (other + tsel + (1.0 - tsel) * term * ONE_THIRD) * fullV * 0.5;
          }
        }
      }
    }
    vof[i] = vol + PLUS_TINY;
  }
}

// ---- masses and operator coefficients -------------------------------------------------------------
__device__ inline double inv_mass(double v) { return (v < EQ_TINY_DOUBLE) ? 0.0 : (1.0 / v); }

__device__ inline bool no_b(int t) { return t == B_Border || t == BN_Edge || t == BS_Edge || t == BE_Edge || t == BW_Edge || t == BNE_Corner || t == BNW_Corner || t == BSE_Corner || t == BSW_Corner; }
__device__ inline bool no_t(int t) { return t == T_Border || t == TN_Edge || t == TS_Edge || t == TE_Edge || t == TW_Edge || t == TNE_Corner || t == TNW_Corner || t == TSE_Corner || t == TSW_Corner; }
__device__ inline bool no_s(int t) { return t == S_Border || t == TS_Edge || t == BS_Edge || t == SE_Edge || t == SW_Edge || t == TSE_Corner || t == TSW_Corner || t == BSE_Corner || t == BSW_Corner; }
__device__ inline bool no_n(int t) { return t == N_Border || t == TN_Edge || t == BN_Edge || t == NE_Edge || t == NW_Edge || t == TNE_Corner || t == TNW_Corner || t == BNE_Corner || t == BNW_Corner; }
__device__ inline bool no_w(int t) { return t == W_Border || t == TW_Edge || t == BW_Edge || t == NW_Edge || t == SW_Edge || t == TNW_Corner || t == TSW_Corner || t == BNW_Corner || t == BSW_Corner; }
__device__ inline bool no_e(int t) { return t == E_Border || t == TE_Edge || t == BE_Edge || t == NE_Edge || t == SE_Edge || t == TNE_Corner || t == TSE_Corner || t == BNE_Corner || t == BSE_Corner; }

// Pressure operator row of cell i as 7 coefficients c = [b,s,w,p,e,n,t] (absent = 0): the
// values of LaplaceDiscreteOperator::calcLap (laplace_discrete_operator.h:107-686) with the
// staggered masses M = (V_i + V_nb)/2 inlined, and identity rows at the two pinned cells.
__device__ inline void laplace_row(long i, int x, int y, int z, const int* __restrict__ type,
                                   const double* __restrict__ tx, const double* __restrict__ ty,
                                   const double* __restrict__ tz, const double* __restrict__ vof,
                                   const double* __restrict__ dxs, const double* __restrict__ dys,
                                   const double* __restrict__ dzs, int N, int M, int pin, double* c) {
  const long NM = (long)N * M;
  const int t = type[i];
  double a_b = 0, a_s = 0, a_w = 0, a_p = 0, a_e = 0, a_n = 0, a_t = 0;
  if (t == SolidCell) {
    a_p = PLUS_TINY;
  } else {
    const double v = vof[i];
    const double ddx = dxs[x], ddy = dys[y], ddz = dzs[z];
    const bool hb = !no_b(t), hs = !no_s(t), hw = !no_w(t), he = !no_e(t), hn = !no_n(t), ht = !no_t(t);
    // masses of the six faces (only the present ones are read)
    const double imw = hw ? inv_mass(0.5 * (vof[i - 1] + v)) : 0.0;
    const double ime = he ? inv_mass(0.5 * (v + vof[i + 1])) : 0.0;
    const double ims = hs ? inv_mass(0.5 * (vof[i - N] + v)) : 0.0;
    const double imn = hn ? inv_mass(0.5 * (v + vof[i + N])) : 0.0;
    const double imb = hb ? inv_mass(0.5 * (vof[i - NM] + v)) : 0.0;
    const double imt = ht ? inv_mass(0.5 * (v + vof[i + NM])) : 0.0;
    if (t == FluidCell) {
      const double dyz = ddy * ddz, dxz = ddx * ddz, dxy = ddy * ddx;
      double aux = dyz * dyz;
      a_w = aux * imw;
      a_e = aux * ime;
      aux = dxz * dxz;
      a_s = aux * ims;
      a_n = aux * imn;
      aux = dxy * dxy;
      a_b = aux * imb;
      a_t = aux * imt;
      a_p = -a_b - a_s - a_w - a_e - a_n - a_t;
    } else {
      const double dyz = ddy * ddz, dxz = ddx * ddz, dxy = ddx * ddy;
      double aux;
      if (hw) { aux = tx[i - 1] * dyz; a_w = aux * aux * imw; }
      if (he) { aux = tx[i] * dyz; a_e = aux * aux * ime; }
      if (hs) { aux = ty[i - N] * dxz; a_s = aux * aux * ims; }
      if (hn) { aux = ty[i] * dxz; a_n = aux * aux * imn; }
      if (hb) { aux = tz[i - NM] * dxy; a_b = aux * aux * imb; }
      if (ht) { aux = tz[i] * dxy; a_t = aux * aux * imt; }
      bool first = true;
      if (hb) { a_p = -a_b; first = false; }
      if (hs) { a_p = first ? -a_s : a_p - a_s; first = false; }
      if (hw) { a_p = first ? -a_w : a_p - a_w; first = false; }
      if (he) { a_p = first ? -a_e : a_p - a_e; first = false; }
      if (hn) { a_p = first ? -a_n : a_p - a_n; first = false; }
      if (ht) { a_p = first ? -a_t : a_p - a_t; first = false; }
    }
    if (pin && (i == NM - 1 || i == N - 1)) {
      a_b = a_s = a_w = a_e = a_n = a_t = 0.0;
      a_p = 1.0;
    }
  }
  c[0] = a_b;
  c[1] = a_s;
  c[2] = a_w;
  c[3] = a_p;
  c[4] = a_e;
  c[5] = a_n;
  c[6] = a_t;
}

__global__ void k_laplace7(const int* __restrict__ type, const double* __restrict__ tx, const double* __restrict__ ty,
                           const double* __restrict__ tz, const double* __restrict__ vof,
                           const double* __restrict__ dxs, const double* __restrict__ dys,
                           const double* __restrict__ dzs, int N, int M, int K, int pin, double* __restrict__ coef) {
  const long n = (long)N * M * K;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    double c[7];
    laplace_row(i, x, y, z, type, tx, ty, tz, vof, dxs, dys, dzs, N, M, pin, c);
    for (int q = 0; q < 7; ++q) coef[q * n + i] = c[q];
  }
}

// ---- momentum (Helmholtz) operators on the staggered meshes --------------------------------------
// Rows of HelmholtzDiscreteOperatorX / Y / Z::calcHelm* (helmholtz_discrete_operator_x.h:98-689 and
// the _y, _z files) for a box without cut cells, in stencil form [b, s, w, p, e, n, t]: viscous
// couplings across the cross faces ((area nu) (sum of the two widths) / (sum of the two distances)),
// nu (area)^2 / V along the component's own axis, no-slip wall terms (bottom and top ones dropped
// for flow tasks), diagonal -(coef M) - sum in each boundary type's own order, and the outflow term
// added to the last cell of every x row of a flow task.  Face and edge fractions are 1 in such a
// box and the reference's products by them are exact, so they are left out.  One thread per
// staggered cell; the type of a staggered cell is the smaller of its two base cells' types
// (ls_stag_uniform_mesh.h:108-130), the mass half the sum of their volumes (:2804-2826).
__device__ inline bool hb_(int t) { return t == B_Border || t == BN_Edge || t == BS_Edge || t == BE_Edge || t == BW_Edge || t == BNE_Corner || t == BNW_Corner || t == BSE_Corner || t == BSW_Corner; }
__device__ inline bool ht_(int t) { return t == T_Border || t == TN_Edge || t == TS_Edge || t == TE_Edge || t == TW_Edge || t == TNE_Corner || t == TNW_Corner || t == TSE_Corner || t == TSW_Corner; }
__device__ inline bool hs_(int t) { return t == S_Border || t == TS_Edge || t == BS_Edge || t == SE_Edge || t == SW_Edge || t == TSE_Corner || t == TSW_Corner || t == BSE_Corner || t == BSW_Corner; }
__device__ inline bool hn_(int t) { return t == N_Border || t == TN_Edge || t == BN_Edge || t == NE_Edge || t == NW_Edge || t == TNE_Corner || t == TNW_Corner || t == BNE_Corner || t == BNW_Corner; }
__device__ inline bool hw_(int t) { return t == W_Border || t == TW_Edge || t == BW_Edge || t == NW_Edge || t == SW_Edge || t == TNW_Corner || t == TSW_Corner || t == BNW_Corner || t == BSW_Corner; }
__device__ inline bool he_(int t) { return t == E_Border || t == TE_Edge || t == BE_Edge || t == NE_Edge || t == SE_Edge || t == TNE_Corner || t == TSE_Corner || t == BNE_Corner || t == BSE_Corner; }

__global__ void k_helmholtz7(const int* __restrict__ type, const double* __restrict__ thx,
                             const double* __restrict__ thy, const double* __restrict__ thz,
                             const double* __restrict__ vof, const double* __restrict__ dxs,
                             const double* __restrict__ dys, const double* __restrict__ dzs, int N, int M, int K,
                             int dir, double nu, double cf, int flow, double* __restrict__ out) {
  const int Nd = N - (dir == 0), Md = M - (dir == 1), Kd = K - (dir == 2);
  const long nd = (long)Nd * Md * Kd, NM = (long)N * M;
  const long step = dir == 0 ? 1 : (dir == 1 ? N : NM);
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nd; q += (long)gridDim.x * blockDim.x) {
    const int i = (int)(q % Nd), j = (int)((q / Nd) % Md), k = (int)(q / ((long)Nd * Md));
    const long c = i + N * (j + (long)M * k);
    const int tc = type[c], td = type[c + step];
    const double th = dir == 0 ? thx[c] : (dir == 1 ? thy[c] : thz[c]);
    const int t = th < EQ_DOUBLE ? SolidCell : (td < tc ? td : tc);
    auto iv = [&](long cc) { return vof[cc] < EQ_TINY_DOUBLE ? 0.0 : 1.0 / vof[cc]; };
    const double dx = dxs[i], dy = dys[j], dz = dzs[k];
    const double mass = 0.5 * (vof[c] + vof[c + step]);
    double rb = 0, rs = 0, rw = 0, re = 0, rn = 0, rt = 0, rp = 0;
    if (t == FluidCell) {
      if (dir == 0) {
        const double dxx = nu * (dxs[i + 1] + dx);
        double aux = dz * dxx;
        rs = aux / (dy + dys[j - 1]);
        rn = aux / (dy + dys[j + 1]);
        aux = dy * dxx;
        rb = aux / (dz + dzs[k - 1]);
        rt = aux / (dz + dzs[k + 1]);
        aux = dy * dz;
        const double d2 = nu * aux * aux;
        re = d2 * iv(c + 1);
        rw = d2 * iv(c);
      } else if (dir == 1) {
        const double dyy = nu * (dys[j + 1] + dy);
        double aux = dz * dyy;
        rw = aux / (dx + dxs[i - 1]);
        re = aux / (dx + dxs[i + 1]);
        aux = dx * dyy;
        rb = aux / (dz + dzs[k - 1]);
        rt = aux / (dz + dzs[k + 1]);
        aux = dx * dz;
        const double d2 = nu * aux * aux;
        rn = d2 * iv(c + N);
        rs = d2 * iv(c);
      } else {
        const double dzz = nu * (dzs[k + 1] + dz);
        double aux = dzz * dy;
        rw = aux / (dx + dxs[i - 1]);
        re = aux / (dx + dxs[i + 1]);
        aux = dx * dzz;
        rs = aux / (dy + dys[j - 1]);
        rn = aux / (dy + dys[j + 1]);
        aux = dx * dy;
        const double d2 = nu * aux * aux;
        rt = d2 * iv(c + NM);
        rb = d2 * iv(c);
      }
      rp = -(cf * mass) - rb - rs - rw - re - rn - rt;
    } else if (t != SolidCell) {
      const bool hb = !hb_(t), ht = !ht_(t), hs = !hs_(t), hn = !hn_(t), hw = !hw_(t);
      const bool zw = hb_(t) || ht_(t), yw = hn_(t) || hs_(t);
      double p = -(cf * mass);
      if (dir == 0) {
        const double dxx = dxs[i + 1], sx = dx + dxx;
        const double az = dz * nu, ay = dy * nu;
        if (hs) rs = az * sx / (dy + dys[j - 1]);
        if (hn) rn = az * sx / (dy + dys[j + 1]);
        if (hb) rb = ay * sx / (dz + dzs[k - 1]);
        if (ht) rt = ay * sx / (dz + dzs[k + 1]);
        const double a2 = dy * dz, d2 = nu * a2 * a2;
        const double ke = d2 * iv(c + 1), kw = hw ? d2 * iv(c) : 0.0;
        re = ke;
        rw = kw;
        double wz = 0.0, wy = 0.0;
        if (t == TN_Edge) {  // the reference's TN edge: (dy dz) / dz in place of dy
          const double q2 = a2 / dz;
          wz = flow ? 0.0 : q2 * nu * sx / dz;
          wy = nu * dz * sx / q2;
        } else {
          if (zw) wz = flow ? 0.0 : dy * nu * sx / dz;
          if (yw) wy = nu * dz * sx / dy;
        }
        if (hb) p -= rb;
        if (hs) p -= rs;
        p -= hw ? (kw + ke) : ke;
        if (hn) p -= rn;
        if (ht) p -= rt;
        if (zw) p -= wz;
        if (yw) p -= wy;
      } else if (dir == 1) {
        const double dyy = dys[j + 1], sy = dy + dyy;
        const double az = dz * nu, ax = dx * nu;
        if (hw) rw = az * sy / (dx + dxs[i - 1]);
        const double ke = az * sy / (dx + dxs[(c + 1) % N]);  // east boundary: the next row's first cell
        re = ke;
        if (hb) rb = ax * sy / (dz + dzs[k - 1]);
        if (ht) rt = ax * sy / (dz + dzs[k + 1]);
        const double a2 = dx * dz, d2 = nu * a2 * a2;
        const double ks = d2 * iv(c), kn = d2 * iv(c + N);
        rs = ks;
        rn = kn;
        double wz = 0.0;
        if (zw) {
          const double dxw = (t == B_Border || t == T_Border) ? dx : a2;  // edges and corners: dx dz
          wz = flow ? 0.0 : dxw * nu * sy / dz;
        }
        const double wy = hn_(t) ? kn : (hs_(t) ? ks : 0.0);
        if (hb) p -= rb;
        if (hn_(t)) {
          if (t == N_Border) {
            p -= rw;
            p -= ks;
          } else {
            p -= ks;
            if (hw) p -= rw;
          }
          p -= ke;
        } else if (hs_(t)) {
          if (hw) p -= rw;
          p -= ke;
          p -= kn;
        } else {
          p -= ks + kn;
          if (hw) p -= rw;
          p -= ke;
        }
        if (ht) p -= rt;
        if (zw) p -= wz;
        if (yw) p -= wy;
      } else {
        const double dzz = dzs[k + 1], sz = dz + dzz;
        const double ay = dy * nu, ax = dx * nu;
        if (hw) rw = ay * sz / (dx + dxs[i - 1]);
        const double ke = ay * sz / (dx + dxs[(c + 1) % N]);
        re = ke;
        if (hs) rs = ax * sz / (dy + dys[j - 1]);
        if (hn) rn = ax * sz / (dy + dys[j + 1]);
        const double wy = yw ? ax * sz / dy : 0.0;
        const double a2 = dx * dy, d2 = nu * a2 * a2;
        const double kb = d2 * iv(c), kt = d2 * iv(c + NM);
        rb = kb;
        rt = kt;
        if (hb_(t)) {
          const double wb = flow ? 0.0 : kb;
          if (hs) p -= rs;
          if (hw) p -= rw;
          p -= ke;
          if (hn) p -= rn;
          p -= kt + wb;
        } else if (ht_(t)) {
          const double wt = flow ? 0.0 : kt;
          p -= kb + wt;
          if (hs) p -= rs;
          if (hw) p -= rw;
          p -= ke;
          if (hn) p -= rn;
        } else if (t == N_Border) {
          p -= rs;
          p -= rw;
          p -= kb + kt;
          p -= ke;
        } else {
          p -= kb + kt;
          if (hs) p -= rs;
          if (hw) p -= rw;
          p -= ke;
          if (hn) p -= rn;
        }
        if (yw) p -= wy;
      }
      rp = p;
    } else {
      rp = PLUS_TINY;
    }
    if (flow && t != SolidCell && i == Nd - 1) {  // outflow term (addPlusOutFlow)
      if (dir == 0)
        rp += nu * dys[j] * dzs[k] / dxs[N - 1];
      else if (dir == 1)
        rp += nu * (dys[j] + dys[j + 1]) * dzs[k] / dxs[N - 1];
      else
        rp += nu * (dzs[k] + dzs[k + 1]) * dys[j] / dxs[N - 1];
    }
    const bool fl = t == FluidCell, so = t == SolidCell;
    out[q] = (fl || (!so && !hb_(t))) ? rb : 0.0;
    out[nd + q] = (fl || (!so && !hs_(t))) ? rs : 0.0;
    out[2 * nd + q] = (fl || (!so && !hw_(t))) ? rw : 0.0;
    out[3 * nd + q] = rp;
    out[4 * nd + q] = (fl || (!so && !he_(t))) ? re : 0.0;
    out[5 * nd + q] = (fl || (!so && !hn_(t))) ? rn : 0.0;
    out[6 * nd + q] = (fl || (!so && !ht_(t))) ? rt : 0.0;
  }
}

int grid_1d(long n) {
  long g = (n + 255) / 256;
  if (g > 16384) g = 16384;
  return (int)(g < 1 ? 1 : g);
}

}  // namespace

int rule_count() { return N_RULES; }

void level_set(const double* cx, const double* cy, const double* cz, int N, int M, int K, Box box, Body body,
               double* phi, hipStream_t s) {
  k_level_set<<<grid_1d((long)N * M * K), 256, 0, s>>>(cx, cy, cz, N, M, K, box, body, phi);
}

int oscillation_fix(const double* phi0, double* phi, double* tmp, unsigned char* flag, int* changed, int N, int M,
                    int K, Box box, hipStream_t s) {
  const long n = (long)N * M * K;
  k_osc_flags<<<grid_1d(n), 256, 0, s>>>(phi0, N, M, K, box, flag);
  if (hipMemcpyAsync(phi, phi0, n * sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess) return -1;
  if (hipMemcpyAsync(tmp, phi0, n * sizeof(double), hipMemcpyDeviceToDevice, s) != hipSuccess) return -1;
  double *cur = phi, *nxt = tmp;
  for (int pass = 1;; ++pass) {
    int h = 0;
    if (hipMemsetAsync(changed, 0, sizeof(int), s) != hipSuccess) return -1;
    k_osc_iter<<<grid_1d(n), 256, 0, s>>>(phi0, flag, cur, nxt, N, M, K, changed);
    if (hipMemcpyAsync(&h, changed, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    if (hipStreamSynchronize(s) != hipSuccess) return -1;
    double* t = cur;
    cur = nxt;
    nxt = t;
    // an unchanged pass leaves both buffers equal: phi holds the result
    if (!h) return pass;
  }
}

void alpha(const double* phi, int N, int M, int K, Box box, double* ax, double* ay, double* az, hipStream_t s) {
  k_alpha<<<grid_1d((long)N * M * K), 256, 0, s>>>(phi, N, M, K, box, ax, ay, az);
}

void types(const double* phi, int N, int M, int K, Box box, int* type, int* sx, int* sy, int* sz, hipStream_t s) {
  k_types<<<grid_1d((long)N * M * K), 256, 0, s>>>(phi, N, M, K, box, type, sx, sy, sz);
}

void theta(const double* ax, const double* ay, const double* az, const int* sx, const int* sy, const int* sz, int N,
           int M, int K, Box box, double* tx, double* ty, double* tz, hipStream_t s) {
  k_theta<<<grid_1d((long)N * M * K), 256, 0, s>>>(ax, ay, az, sx, sy, sz, N, M, K, box, tx, ty, tz);
}

void vof(const int* type, const double* ax, const double* ay, const double* az, const double* tx, const double* ty,
         const double* tz, const double* dx, const double* dy, const double* dz, int N, int M, int K, double* v,
         hipStream_t s) {
  Fields F;
  F.a[0] = ax;
  F.a[1] = ay;
  F.a[2] = az;
  F.t[0] = tx;
  F.t[1] = ty;
  F.t[2] = tz;
  F.N = N;
  F.NM = N * M;
  k_vof<<<grid_1d((long)N * M * K), 256, 0, s>>>(type, F, dx, dy, dz, N, M, K, v);
}

void helmholtz7(const int* type, const double* tx, const double* ty, const double* tz, const double* v,
                const double* dx, const double* dy, const double* dz, int N, int M, int K, int dir, double nu,
                double coef_time, int flow, double* coef, hipStream_t s) {
  const long nd = (long)(N - (dir == 0)) * (M - (dir == 1)) * (K - (dir == 2));
  k_helmholtz7<<<grid_1d(nd), 256, 0, s>>>(type, tx, ty, tz, v, dx, dy, dz, N, M, K, dir, nu, coef_time, flow, coef);
}

void laplace7(const int* type, const double* tx, const double* ty, const double* tz, const double* v, const double* dx,
              const double* dy, const double* dz, int N, int M, int K, int pin, double* coef, hipStream_t s) {
  k_laplace7<<<grid_1d((long)N * M * K), 256, 0, s>>>(type, tx, ty, tz, v, dx, dy, dz, N, M, K, pin, coef);
}

// ---- pressure right-hand side ------------------------------------------------------------------
// b = Dx u* + Dy v* + Dz w* + flux, b = 0 at the pinned cells: PoissonDiscreteEquation::
// recalcRightMember (poisson_discrete_equation.h:145-159) with the divergence rows of calcDx/Dy/Dz
// (divergence_discrete_operator_{x,y,z}.h) summed in their portrait order, and the static-body
// flux of initFlux / addFluxNSp / addFluxTBp / setBaseFlux (laplace_discrete_operator.h:48-83,
// 744-750) applied in that order.  One thread per cell; no FMA contraction.
namespace {

__device__ inline bool x_east(int t) {
  return t == E_Border || t == NE_Edge || t == SE_Edge || t == TE_Edge || t == BE_Edge || t == TNE_Corner ||
         t == TSE_Corner || t == BNE_Corner || t == BSE_Corner;
}
__device__ inline bool x_west(int t) {
  return t == W_Border || t == NW_Edge || t == SW_Edge || t == TW_Edge || t == BW_Edge || t == TNW_Corner ||
         t == TSW_Corner || t == BNW_Corner || t == BSW_Corner;
}
__device__ inline bool y_north(int t) {
  return t == N_Border || t == NE_Edge || t == NW_Edge || t == TN_Edge || t == BN_Edge || t == TNE_Corner ||
         t == TNW_Corner || t == BNE_Corner || t == BNW_Corner;
}
__device__ inline bool y_south(int t) {
  return t == S_Border || t == SE_Edge || t == SW_Edge || t == TS_Edge || t == BS_Edge || t == TSE_Corner ||
         t == TSW_Corner || t == BSE_Corner || t == BSW_Corner;
}
__device__ inline bool z_bottom(int t) {
  return t == B_Border || t == BN_Edge || t == BS_Edge || t == BE_Edge || t == BW_Edge || t == BNE_Corner ||
         t == BNW_Corner || t == BSE_Corner || t == BSW_Corner;
}
__device__ inline bool z_top(int t) {
  return t == T_Border || t == TN_Edge || t == TS_Edge || t == TE_Edge || t == TW_Edge || t == TNE_Corner ||
         t == TNW_Corner || t == TSE_Corner || t == TSW_Corner;
}

struct RhsIn {
  const double *u, *v, *w, *u_west, *u_out;
  FlowBc bc;
};

__device__ inline double rhs_cell(long i, int x, int y, int z, const int* __restrict__ type,
                                  const double* __restrict__ tx, const double* __restrict__ ty,
                                  const double* __restrict__ tz, const double* __restrict__ dxa,
                                  const double* __restrict__ dya, const double* __restrict__ dza, int N, int M, int K,
                                  const RhsIn& in) {
  const long NM = (long)N * M;
  const double* __restrict__ u = in.u;
  const double* __restrict__ v = in.v;
  const double* __restrict__ w = in.w;
  const FlowBc& bc = in.bc;
  const int t = type[i];
  const double dx = dxa[x], dy = dya[y], dz = dza[z];
  const long row = y + (long)z * M;
  // flux: initFlux, addFluxNSp, addFluxTBp, setBaseFlux (static body: ibFlux = 0)
  double flux = 0.0;
  if (x == 0) flux = -in.u_west[row] * (dy * dz);
  if (x == N - 1) flux = bc.U_E * (dy * dz);
  if (y == 0) flux -= (dx * dz) * bc.V_S;
  if (y == M - 1) flux += (dx * dz) * bc.V_N;
  if (z == 0) flux -= (dy * dx) * tz[i] * bc.W_B;
  if (z == K - 1) flux += (dy * dx) * tz[i] * bc.W_T;
  if (in.u_out && x == N - 1) flux = in.u_out[row * (N - 1) + N - 2] * dy * dz;
  // Dx u
  double aux = 0.0;
  {
    const long p = row * (N - 1) + x;
    if (t == SolidCell) aux += 0.0 * u[x == 0 ? p : p - 1];
    else if (x_east(t)) aux += (-tx[i - 1] * dy * dz) * u[p - 1];
    else if (x_west(t)) aux += (tx[i] * dy * dz) * u[p];
    else {
      const double dyz = dy * dz;
      aux += (-tx[i - 1] * dyz) * u[p - 1];
      aux += (tx[i] * dyz) * u[p];
    }
  }
  // (aux + flux) + Dy v
  double r = aux + flux;
  {
    const long p = (long)z * N * (M - 1) + (long)y * N + x;
    if (t == SolidCell) r += 0.0 * v[y == 0 ? p : p - N];
    else if (y_north(t)) r += (-ty[i - N] * dx * dz) * v[p - N];
    else if (y_south(t)) r += (ty[i] * dx * dz) * v[p];
    else {
      const double dxz = dx * dz;
      r += (-ty[i - N] * dxz) * v[p - N];
      r += (ty[i] * dxz) * v[p];
    }
  }
  // + Dz w
  double auxz = 0.0;
  {
    if (t == SolidCell) auxz += 0.0 * w[z == 0 ? i : i - NM];
    else if (z_bottom(t)) auxz += (tz[i] * dx * dy) * w[i];
    else if (z_top(t)) auxz += (-tz[i - NM] * dx * dy) * w[i - NM];
    else {
      const double dxy = dx * dy;
      auxz += (-tz[i - NM] * dxy) * w[i - NM];
      auxz += (tz[i] * dxy) * w[i];
    }
  }
  r += auxz;
  if (bc.pin && (i == N - 1 || i == NM - 1)) r = 0.0;
  return r;
}

__global__ void k_pressure_rhs(const int* __restrict__ type, const double* __restrict__ tx,
                               const double* __restrict__ ty, const double* __restrict__ tz,
                               const double* __restrict__ dxa, const double* __restrict__ dya,
                               const double* __restrict__ dza, int N, int M, int K, RhsIn in,
                               double* __restrict__ b) {
  const long n = (long)N * M * K, NM = (long)N * M;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int z = (int)(i / NM);
    const long r2 = i - (long)z * NM;
    const int y = (int)(r2 / N), x = (int)(r2 - (long)y * N);
    b[i] = rhs_cell(i, x, y, z, type, tx, ty, tz, dxa, dya, dza, N, M, K, in);
  }
}

// Velocity correction u = (G dp) inv_M + u* (SetSolutions::correction, set_solutions.h:215-237,
// with GradientDiscreteOperatorX/Y/Z::calc and the staggered types of cvu/cvv/cvw,
// ls_stag_uniform_mesh.h:108-130).  One thread per face of the three staggered grids; the
// inverse staggered mass is 1 / (0.5 (V_a + V_b)), 0 below 1e-12 (inverseVOF).
__device__ inline double inv_mass(double va, double vb) {
  const double m = 0.5 * (va + vb);
  return m < EQ_TINY_DOUBLE ? 0.0 : 1.0 / m;
}

__device__ inline bool staggered_solid(double theta, int a, int b) {
  return theta < EQ_DOUBLE || (b < a ? b : a) == SolidCell;
}

__global__ void k_correct(const int* __restrict__ type, const double* __restrict__ tx, const double* __restrict__ ty,
                          const double* __restrict__ tz, const double* __restrict__ vof,
                          const double* __restrict__ dxa, const double* __restrict__ dya,
                          const double* __restrict__ dza, int N, int M, int K, const double* __restrict__ dp,
                          const double* __restrict__ us, const double* __restrict__ vs, const double* __restrict__ ws,
                          double* __restrict__ u, double* __restrict__ v, double* __restrict__ w) {
  const long nu = (long)(N - 1) * M * K, nv = (long)N * (M - 1) * K, nw = (long)N * M * (K - 1);
  const long NM = (long)N * M;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nu + nv + nw; q += (long)gridDim.x * blockDim.x) {
    if (q < nu) {  // x faces: index x + (N-1)(y + M z)
      const long f = q, row = f / (N - 1);
      const int x = (int)(f - row * (N - 1)), y = (int)(row % M), z = (int)(row / M);
      const long a = x + (long)N * row, b = a + 1;
      double aux = 0.0;
      if (staggered_solid(tx[a], type[a], type[b])) {
        aux += 0.0 * dp[a];
      } else {
        const double cell = tx[a] * dya[y] * dza[z];
        aux += cell * dp[a];
        aux += -cell * dp[b];
      }
      aux *= inv_mass(vof[a], vof[b]);
      u[f] = aux + us[f];
    } else if (q < nu + nv) {  // y faces: index x + N (y + (M-1) z)
      const long f = q - nu, xy = f % ((long)N * (M - 1));
      const int z = (int)(f / ((long)N * (M - 1))), y = (int)(xy / N), x = (int)(xy - (long)y * N);
      const long a = x + (long)N * y + NM * z, b = a + N;
      double aux = 0.0;
      if (staggered_solid(ty[a], type[a], type[b])) {
        aux += 0.0 * dp[a];
      } else {
        const double cell = ty[a] * dxa[x] * dza[z];
        aux += cell * dp[a];
        aux += -cell * dp[b];
      }
      aux *= inv_mass(vof[a], vof[b]);
      v[f] = aux + vs[f];
    } else {  // z faces: index of the lower cell
      const long a = q - nu - nv, b = a + NM;
      const int z = (int)(a / NM);
      const long r2 = a - (long)z * NM;
      const int y = (int)(r2 / N), x = (int)(r2 - (long)y * N);
      double aux = 0.0;
      if (staggered_solid(tz[a], type[a], type[b])) {
        aux += 0.0 * dp[a];
      } else {
        const double cell = tz[a] * dxa[x] * dya[y];
        aux += cell * dp[a];
        aux += -cell * dp[b];
      }
      aux *= inv_mass(vof[a], vof[b]);
      w[a] = aux + ws[a];
    }
  }
}

}  // namespace

namespace {
__global__ void k_pressure_update(double* __restrict__ p, const double* __restrict__ dp, double coef, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = p[i] + dp[i] * coef;
}
}  // namespace

void pressure_update(double* p, const double* dp, double coef, long n, hipStream_t s) {
  k_pressure_update<<<grid_1d(n), 256, 0, s>>>(p, dp, coef, n);
}

void correct(const int* type, const double* tx, const double* ty, const double* tz, const double* vof, const double* dx,
             const double* dy, const double* dz, int N, int M, int K, const double* dp, const double* us,
             const double* vs, const double* ws, double* u, double* v, double* w, hipStream_t s) {
  const long total = (long)(N - 1) * M * K + (long)N * (M - 1) * K + (long)N * M * (K - 1);
  k_correct<<<grid_1d(total), 256, 0, s>>>(type, tx, ty, tz, vof, dx, dy, dz, N, M, K, dp, us, vs, ws, u, v, w);
}

void pressure_rhs(const int* type, const double* tx, const double* ty, const double* tz, const double* dx,
                  const double* dy, const double* dz, int N, int M, int K, const double* u, const double* v,
                  const double* w, const double* u_west, const double* u_out, FlowBc bc, double* b, hipStream_t s) {
  const RhsIn in{u, v, w, u_west, u_out, bc};
  k_pressure_rhs<<<grid_1d((long)N * M * K), 256, 0, s>>>(type, tx, ty, tz, dx, dy, dz, N, M, K, in, b);
}

// ---- objective and shape gradient ---------------------------------------------------------------
// Pressure drag J = sum over cut cells of (theta_x west - theta_x east) dy dz p: the pressure term
// of LoadsCalculator::calc (loads_calculator.h:97-116, n_x dS = (theta^u_{i-1} - theta^u_i) dy dz).
//
// Shape gradient.  With A p = b solved and lambda = A^{-T} dJ/dp, the derivative of J along a
// change of the geometry is that of the Lagrangian F = J(theta, p) + lambda . (b - A p) with p and
// lambda held fixed.  F is a sum over cells of terms that read the level set at the corners of
// the cell and of its six neighbours, after an oscillation fix that reads one node further:
// corner c reaches the cells c + [-2, 3]^3.  Corners 6 apart in every direction therefore have
// disjoint reach, so one pair of geometry rebuilds with all corners of one of the 216 classes
// (x mod 6, y mod 6, z mod 6) moved by +h and -h gives the central difference dF/dphi_c of
// every corner of the class at once, summed over its reach.  Differences are taken per
// coefficient (c+ - c-) before they meet p, so the large cancelling products of A p never enter
// them.  A Bezier body then chains dF/dphi to its control values through the closest point:
// phi = sgn |X - S(u*, v*)| gives dphi/dP_ij = -sgn (X - S)/|X - S| B_i(u*) B_j(v*) (the
// minimiser's own change does not enter at a minimum).
namespace {

constexpr int SG_SPACING = 6;

__device__ inline bool cut_cell(int t) { return t > 0 && t != FluidCell; }

__global__ void k_drag(const int* __restrict__ type, const double* __restrict__ tx, const double* __restrict__ dys,
                       const double* __restrict__ dzs, int N, int M, int K, Box box, const double* __restrict__ p,
                       double* __restrict__ dJdp, double* __restrict__ partial) {
  __shared__ double lds[256];
  const long n = (long)N * M * K;
  const long per = (n + gridDim.x - 1) / gridDim.x;
  const long lo = per * blockIdx.x, hi = lo + per < n ? lo + per : n;
  double acc = 0.0;
  for (long i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    double g = 0.0;
    if (in_box(x, y, box) && cut_cell(type[i])) g = (tx[i - 1] - tx[i]) * dys[y] * dzs[z];
    if (dJdp) dJdp[i] = g;
    acc += g * p[i];
  }
  lds[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) lds[threadIdx.x] += lds[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}

__device__ inline int corner_class(int x, int y, int z) {
  return x % SG_SPACING + SG_SPACING * (y % SG_SPACING) + SG_SPACING * SG_SPACING * (z % SG_SPACING);
}

__global__ void k_perturb(const double* __restrict__ phi_raw, int N, int M, int K, Box box, int cls, double h,
                          double* __restrict__ phi_p, double* __restrict__ phi_m) {
  const long n = (long)N * M * K;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    const double v = phi_raw[i];
    const bool on = in_box(x, y, box) && corner_class(x, y, z) == cls;
    phi_p[i] = on ? v + h : v;
    phi_m[i] = on ? v - h : v;
  }
}

__device__ inline double drag_cell(const GeoView& g, long i, int x, int y, int z, const double* dys,
                                   const double* dzs, const Box& box, double p) {
  if (!in_box(x, y, box) || !cut_cell(g.type[i])) return 0.0;
  return (g.tx[i - 1] - g.tx[i]) * dys[y] * dzs[z] * p;
}

__global__ void k_lagr_diff(GeoView gp, GeoView gm, const double* __restrict__ dxs, const double* __restrict__ dys,
                            const double* __restrict__ dzs, int N, int M, int K, Box box, int pin,
                            const double* __restrict__ p, const double* __restrict__ lam, int objective, int has_rhs,
                            RhsIn rin, double inv2h, double* __restrict__ dF) {
  const long n = (long)N * M * K, NM = (long)N * M;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    double dJ = 0.0;
    if (objective == OBJ_DRAG_X && in_box(x, y, box)) {
      const bool cp = cut_cell(gp.type[i]), cm = cut_cell(gm.type[i]);
      if (cp && cm) {
        const double dw = gp.tx[i - 1] - gm.tx[i - 1], de = gp.tx[i] - gm.tx[i];
        dJ = (dw - de) * dys[y] * dzs[z] * p[i];
      } else {
        dJ = drag_cell(gp, i, x, y, z, dys, dzs, box, p[i]) - drag_cell(gm, i, x, y, z, dys, dzs, box, p[i]);
      }
    }
    double ca[7], cb[7];
    laplace_row(i, x, y, z, gp.type, gp.tx, gp.ty, gp.tz, gp.vof, dxs, dys, dzs, N, M, pin, ca);
    laplace_row(i, x, y, z, gm.type, gm.tx, gm.ty, gm.tz, gm.vof, dxs, dys, dzs, N, M, pin, cb);
    const long nb[7] = {i - NM, i - N, i - 1, i, i + 1, i + N, i + NM};
    const bool has[7] = {z > 0, y > 0, x > 0, true, x < N - 1, y < M - 1, z < K - 1};
    double dAp = 0.0;
    for (int q = 0; q < 7; ++q)
      if (has[q]) dAp += (ca[q] - cb[q]) * p[nb[q]];
    double db = 0.0;
    if (has_rhs)
      db = rhs_cell(i, x, y, z, gp.type, gp.tx, gp.ty, gp.tz, dxs, dys, dzs, N, M, K, rin) -
           rhs_cell(i, x, y, z, gm.type, gm.tx, gm.ty, gm.tz, dxs, dys, dzs, N, M, K, rin);
    dF[i] = (dJ + lam[i] * (db - dAp)) * inv2h;
  }
}

// g[c] = sum of dF over the reach of corner c, for the corners of class `cls`
__global__ void k_corner_gather(const double* __restrict__ dF, int N, int M, int K, Box box, int cls,
                                double* __restrict__ g) {
  const long n = (long)N * M * K, NM = (long)N * M;
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(c, N, M, &x, &y, &z);
    if (!in_box(x, y, box) || corner_class(x, y, z) != cls) continue;
    double acc = 0.0;
    for (int zz = z - 2; zz <= z + 3; ++zz) {
      if (zz < 0 || zz >= K) continue;
      for (int yy = y - 2; yy <= y + 3; ++yy) {
        if (yy < 0 || yy >= M) continue;
        for (int xx = x - 2; xx <= x + 3; ++xx) {
          if (xx < 0 || xx >= N) continue;
          acc += dF[xx + (long)N * yy + NM * zz];
        }
      }
    }
    g[c] = acc;
  }
}

// per corner with g != 0: winning patch, its parameters and w = g dphi/dS (3 values)
__global__ void k_bz_hits(const double* __restrict__ g, const double* __restrict__ cx, const double* __restrict__ cy,
                          const double* __restrict__ cz, int N, int M, int K, Box box, Body body,
                          int* __restrict__ hf, double* __restrict__ huv, double* __restrict__ hw) {
  const long n = (long)N * M * K;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    int x, y, z;
    decode(i, N, M, &x, &y, &z);
    int f = -1;
    if (in_box(x, y, box) && g[i] != 0.0) {
      BzHit hit;
      bezier_closest(body.ctrl, body.c, body.a, cx[x], cy[y], cz[z], hit);
      if (hit.dist > 0.0) {
        f = hit.f;
        huv[2 * i] = hit.u;
        huv[2 * i + 1] = hit.v;
        for (int c = 0; c < 3; ++c) hw[3 * i + c] = g[i] * (-hit.sgn * hit.r[c] / hit.dist);
      }
    }
    hf[i] = f;
  }
}

// one workgroup per control point (f, i, j): out[(f*16 + i*4 + j)*3 + c] = sum over its corners
__global__ void k_bz_reduce(const int* __restrict__ hf, const double* __restrict__ huv, const double* __restrict__ hw,
                            long n, double* __restrict__ out) {
  __shared__ double lds[3][256];
  const int cp = blockIdx.x, f = cp / 16, bi = (cp / 4) % 4, bj = cp % 4;
  double acc[3] = {0.0, 0.0, 0.0};
  for (long k = threadIdx.x; k < n; k += blockDim.x) {
    if (hf[k] != f) continue;
    double bu[4], dbu[4], ddbu[4], bv[4], dbv[4], ddbv[4];
    bz_basis(huv[2 * k], bu, dbu, ddbu);
    bz_basis(huv[2 * k + 1], bv, dbv, ddbv);
    const double wgt = bu[bi] * bv[bj];
    for (int c = 0; c < 3; ++c) acc[c] += wgt * hw[3 * k + c];
  }
  for (int c = 0; c < 3; ++c) lds[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int c = 0; c < 3; ++c) lds[c][threadIdx.x] += lds[c][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 3) out[cp * 3 + threadIdx.x] = lds[threadIdx.x][0];
}

}  // namespace

constexpr int DRAG_BLOCKS = 256;

// ---- allocation-free drag -------------------------------------------------------------------
// The partial sums of the drag reduction (one per workgroup of the fixed DRAG_BLOCKS-workgroup
// launch, for any grid size) in a device array of the library instead of caller- or
// library-allocated memory: 2 KiB per GPU, held until the process ends.  Each GPU has its own
// copy at its own address; host code uses the address hipGetSymbolAddress returns on that
// device (kept per device below), never the array's name.  Every call on one GPU shares it.
constexpr int DRAG_SCRATCH = 256;
static_assert(DRAG_BLOCKS <= DRAG_SCRATCH, "drag partials exceed the scratch array");
static __device__ __align__(256) double scratch[DRAG_SCRATCH];

namespace {
constexpr int MAX_DEVICES = 64;
double* scratch_dev[MAX_DEVICES] = {};
std::mutex scratch_mu;

// J = sum of the partials in block order, the order of the host sum in drag() (bit-identical)
__global__ void k_drag_sum(const double* __restrict__ partial, double* __restrict__ J) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double acc = 0.0;
    for (int b = 0; b < DRAG_BLOCKS; ++b) acc += partial[b];
    J[0] = acc;
  }
}
}  // namespace

int drag_scratch(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return -1;
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (!scratch_dev[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return -2;
    scratch_dev[dev] = (double*)ptr;
  }
  *out = scratch_dev[dev];
  return 0;
}

void drag_noalloc(const int* type, const double* tx, const double* dy, const double* dz, int N, int M, int K, Box box,
                  const double* p, double* dJdp, double* partial, double* J, hipStream_t s) {
  k_drag<<<DRAG_BLOCKS, 256, 0, s>>>(type, tx, dy, dz, N, M, K, box, p, dJdp, partial);
  k_drag_sum<<<1, 64, 0, s>>>(partial, J);
}

double drag(const int* type, const double* tx, const double* dy, const double* dz, int N, int M, int K, Box box,
            const double* p, double* dJdp, double* partial, hipStream_t s) {
  k_drag<<<DRAG_BLOCKS, 256, 0, s>>>(type, tx, dy, dz, N, M, K, box, p, dJdp, partial);
  double h[DRAG_BLOCKS];
  if (hipMemcpyAsync(h, partial, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess) return NAN;
  if (hipStreamSynchronize(s) != hipSuccess) return NAN;
  double J = 0.0;
  for (int b = 0; b < DRAG_BLOCKS; ++b) J += h[b];
  return J;
}

void perturb_class(const double* phi_raw, int N, int M, int K, Box box, int cls, double h, double* phi_p,
                   double* phi_m, hipStream_t s) {
  k_perturb<<<grid_1d((long)N * M * K), 256, 0, s>>>(phi_raw, N, M, K, box, cls, h, phi_p, phi_m);
}

void lagrangian_diff(const GeoView& gp, const GeoView& gm, const double* dx, const double* dy, const double* dz, int N,
                     int M, int K, Box box, int pin, const double* p, const double* lam, int objective,
                     const RhsArgs* rhs, double h, double* dF, hipStream_t s) {
  RhsIn in{nullptr, nullptr, nullptr, nullptr, nullptr, FlowBc{0, 0, 0, 0, 0, 0}};
  if (rhs) in = RhsIn{rhs->u, rhs->v, rhs->w, rhs->u_west, rhs->u_out, rhs->bc};
  k_lagr_diff<<<grid_1d((long)N * M * K), 256, 0, s>>>(gp, gm, dx, dy, dz, N, M, K, box, pin, p, lam, objective,
                                                       rhs ? 1 : 0, in, 0.5 / h, dF);
}

void corner_gather(const double* dF, int N, int M, int K, Box box, int cls, double* g, hipStream_t s) {
  k_corner_gather<<<grid_1d((long)N * M * K), 256, 0, s>>>(dF, N, M, K, box, cls, g);
}

void bezier_chain(const double* g, const double* cx, const double* cy, const double* cz, int N, int M, int K, Box box,
                  Body body, int* hf, double* huv, double* hw, double* out, hipStream_t s) {
  const long n = (long)N * M * K;
  k_bz_hits<<<grid_1d(n), 256, 0, s>>>(g, cx, cy, cz, N, M, K, box, body, hf, huv, hw);
  k_bz_reduce<<<96, 256, 0, s>>>(hf, huv, hw, n, out);
}

}  // namespace geo
}  // namespace lsx
