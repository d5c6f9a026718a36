// ORACLE — test infrastructure only (see oracle_solver.cpp header).
//
// Six-patch bicubic Bezier body and its signed distance.  The reference has no 3D Bezier surface
// (src/Bezier holds 2D curves only); this restates the body definition the pressure benchmark
// uses, so the GPU level set of the same control points can be checked against it:
//   * patches: the six faces of a cube, face f spanned by (u, v) in [0,1]^2, oriented so that
//     S_u x S_v points outwards; 16 control points per patch (96 in all), i along u, j along v;
//   * ellipsoid fit: each patch interpolates the cube-sphere map c + a * q/|q| (q on the cube
//     face) at u, v in {0, 1/3, 2/3, 1}; P = Binv S Binv^T with Binv the inverse cubic Bernstein
//     matrix at those nodes.  Control points on a shared cube edge or corner take the value of
//     the first patch that holds them, so the surface is closed;
//   * level set: phi(x) = +distance inside, -distance outside; the closest point of each patch
//     comes from 12 Newton steps (a coordinate at a bound with the gradient pointing out
//     stays there) started at the cube-sphere inverse of x, the patch with
//     the smallest distance wins (lowest index on ties), and the side follows the sign of
//     (x - S) . (S_u x S_v) there.

#include <cmath>

#include "oracle_geometry.hpp"

namespace lsxo {

namespace {

// face f: normal axis, its sign, and the axes of u and v
const int FACE_A[6] = {0, 0, 1, 1, 2, 2};
const double FACE_S[6] = {1.0, -1.0, 1.0, -1.0, 1.0, -1.0};
const int FACE_U[6] = {1, 2, 2, 0, 0, 1};
const int FACE_V[6] = {2, 1, 0, 2, 1, 0};

const double BINV[4][4] = {{1.0, 0.0, 0.0, 0.0},
                           {-5.0 / 6.0, 3.0, -3.0 / 2.0, 1.0 / 3.0},
                           {1.0 / 3.0, -3.0 / 2.0, 3.0, -5.0 / 6.0},
                           {0.0, 0.0, 0.0, 1.0}};

// Bernstein basis of degree 3 and its first and second derivatives at t
void bern(double t, double* b, double* db, double* ddb) {
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
void eval_patch(const double* P, double u, double v, double* S, double* Su, double* Sv) {
  double bu[4], dbu[4], ddbu[4], bv[4], dbv[4], ddbv[4];
  bern(u, bu, dbu, ddbu);
  bern(v, bv, dbv, ddbv);
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
void eval_patch_second(const double* P, double u, double v, const double* r, double* ruu, double* ruv, double* rvv) {
  double bu[4], dbu[4], ddbu[4], bv[4], dbv[4], ddbv[4];
  bern(u, bu, dbu, ddbu);
  bern(v, bv, dbv, ddbv);
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

double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

double clamp01(double t) { return t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t); }

}  // namespace

// 288 doubles: patch f, control point (i, j), coordinate c at ((f * 16 + i * 4 + j) * 3 + c)
void bezier_ellipsoid(const double* center, const double* axes, double* ctrl) {
  double have[4][4][4][3];
  bool set[4][4][4] = {};
  for (int f = 0; f < 6; ++f) {
    const int a = FACE_A[f], bu = FACE_U[f], bv = FACE_V[f];
    double S[4][4][3];
    for (int iu = 0; iu < 4; ++iu)
      for (int iv = 0; iv < 4; ++iv) {
        double q[3];
        q[a] = FACE_S[f];
        q[bu] = (2.0 * iu - 3.0) / 3.0;
        q[bv] = (2.0 * iv - 3.0) / 3.0;
        const double r = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        for (int c = 0; c < 3; ++c) S[iu][iv][c] = center[c] + axes[c] * (q[c] / r);
      }
    double T[4][4][3], P[4][4][3];
    for (int i = 0; i < 4; ++i)
      for (int iv = 0; iv < 4; ++iv)
        for (int c = 0; c < 3; ++c) {
          double t = 0.0;
          for (int k = 0; k < 4; ++k) t += BINV[i][k] * S[k][iv][c];
          T[i][iv][c] = t;
        }
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j)
        for (int c = 0; c < 3; ++c) {
          double p = 0.0;
          for (int k = 0; k < 4; ++k) p += T[i][k][c] * BINV[j][k];
          P[i][j][c] = p;
        }
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        int l[3];
        l[a] = FACE_S[f] > 0 ? 3 : 0;
        l[bu] = i;
        l[bv] = j;
        double* dst = have[l[0]][l[1]][l[2]];
        if (!set[l[0]][l[1]][l[2]]) {
          for (int c = 0; c < 3; ++c) dst[c] = P[i][j][c];
          set[l[0]][l[1]][l[2]] = true;
        }
        for (int c = 0; c < 3; ++c) ctrl[(f * 16 + i * 4 + j) * 3 + c] = dst[c];
      }
  }
}

double bezier_level_set(const double* ctrl, const double* center, const double* axes, double x, double y, double z) {
  const double X[3] = {x, y, z};
  double d[3];
  for (int c = 0; c < 3; ++c) d[c] = (X[c] - center[c]) / axes[c];
  double best = -1.0, side = 0.0;
  for (int f = 0; f < 6; ++f) {
    const double* P = ctrl + f * 48;
    const int a = FACE_A[f], bu = FACE_U[f], bv = FACE_V[f];
    const double da = std::fabs(d[a]);
    double u = 0.5, v = 0.5;
    if (da > 0.0) {
      u = clamp01(0.5 * (d[bu] / da + 1.0));
      v = clamp01(0.5 * (d[bv] / da + 1.0));
    }
    double S[3], Su[3], Sv[3];
    for (int it = 0; it < 12; ++it) {
      eval_patch(P, u, v, S, Su, Sv);
      double r[3];
      for (int c = 0; c < 3; ++c) r[c] = S[c] - X[c];
      const double gu = dot3(Su, r), gv = dot3(Sv, r);
      const double juu = dot3(Su, Su), juv = dot3(Su, Sv), jvv = dot3(Sv, Sv);
      double ruu, ruv, rvv;
      eval_patch_second(P, u, v, r, &ruu, &ruv, &rvv);
      double huu = juu + ruu, huv = juv + ruv, hvv = jvv + rvv;
      double det = huu * hvv - huv * huv;
      if (!(det > 0.0 && huu > 0.0)) {  // Gauss-Newton where the Hessian is not positive definite
        huu = juu;
        huv = juv;
        hvv = jvv;
        det = huu * hvv - huv * huv;
      }
      // a coordinate held at its bound while the gradient points out of the patch stays there,
      // and the other one takes a one-dimensional Newton step
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
      u = clamp01(u + du);
      v = clamp01(v + dv);
    }
    eval_patch(P, u, v, S, Su, Sv);
    double r[3];
    for (int c = 0; c < 3; ++c) r[c] = X[c] - S[c];
    const double dist2 = dot3(r, r);
    if (best < 0.0 || dist2 < best) {
      best = dist2;
      const double n[3] = {Su[1] * Sv[2] - Su[2] * Sv[1], Su[2] * Sv[0] - Su[0] * Sv[2],
                           Su[0] * Sv[1] - Su[1] * Sv[0]};
      side = dot3(r, n);
    }
  }
  const double dist = std::sqrt(best);
  return side > 0.0 ? -dist : dist;
}

}  // namespace lsxo
