// Six-patch bicubic Bezier fit of an ellipsoid (host).  The control net is a 4x4x4 lattice on
// the cube surface (56 distinct points); patch f is the lattice face with normal axis A[f], u
// along U[f] and v along V[f].  A patch parameter t in [0,1] maps to the cube coordinate
// tan(pi/4 (2t - 1)) (equal angles, which a cubic follows ten times more closely than the
// plain cube coordinate) and the cube to the ellipsoid by c + a * q / |q|.
//   corners: the image of the cube corner;
//   edge interiors: least-squares cubic fit of the edge image (65 samples), corners fixed;
//   face interiors: least-squares bicubic fit of the face image (65 x 65 samples), boundary fixed.

#include <cmath>
#include <vector>

#include "lsx_solver.hpp"

namespace lsx {

namespace {

constexpr double QUARTER_PI = 0.78539816339744830962;
double cube_of(double t) { return std::tan(QUARTER_PI * (2.0 * t - 1.0)); }

const int FA[6] = {0, 0, 1, 1, 2, 2};
const int FS[6] = {1, -1, 1, -1, 1, -1};
const int FU[6] = {1, 2, 2, 0, 0, 1};
const int FV[6] = {2, 1, 0, 2, 1, 0};
constexpr int SAMPLES = 65;

void bernstein3(double t, double* b) {
  const double s = 1.0 - t;
  b[0] = s * s * s;
  b[1] = 3.0 * t * s * s;
  b[2] = 3.0 * t * t * s;
  b[3] = t * t * t;
}

struct Fit {
  double c[3], a[3];
  double P[4][4][4][3];  // lattice control points (only surface entries used)

  void image(const double* q, double* out) const {
    const double r = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    for (int k = 0; k < 3; ++k) out[k] = c[k] + a[k] * (q[k] / r);
  }
  double* at(const int* l) { return P[l[0]][l[1]][l[2]]; }
};

// Solve the n x n system G x = r (n <= 4) by Gaussian elimination with partial pivoting.
void solve_small(int n, double G[4][4], double* r) {
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int i = c + 1; i < n; ++i)
      if (std::fabs(G[i][c]) > std::fabs(G[p][c])) p = i;
    if (p != c) {
      for (int j = 0; j < n; ++j) std::swap(G[c][j], G[p][j]);
      std::swap(r[c], r[p]);
    }
    if (G[c][c] == 0.0) throw Error(LSX_ERR_ARG, "singular Bezier fit");
    for (int i = c + 1; i < n; ++i) {
      const double f = G[i][c] / G[c][c];
      for (int j = c; j < n; ++j) G[i][j] -= f * G[c][j];
      r[i] -= f * r[c];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = r[i];
    for (int j = i + 1; j < n; ++j) s -= G[i][j] * r[j];
    r[i] = s / G[i][i];
  }
}

}  // namespace

void bezier_ellipsoid(const double* center, const double* axes, double* ctrl) {
  for (int k = 0; k < 3; ++k)
    if (!(axes[k] > 0.0) || !std::isfinite(center[k])) throw Error(LSX_ERR_ARG, "bad ellipsoid");
  Fit F;
  for (int k = 0; k < 3; ++k) {
    F.c[k] = center[k];
    F.a[k] = axes[k];
  }
  auto cube = [](int l) { return (2.0 * l - 3.0) / 3.0; };
  // corners
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) {
        const int l[3] = {3 * i, 3 * j, 3 * k};
        const double q[3] = {cube(l[0]), cube(l[1]), cube(l[2])};
        F.image(q, F.at(l));
      }
  // edges: axis e varies, the other two lattice coordinates are 0 or 3
  for (int e = 0; e < 3; ++e)
    for (int s1 = 0; s1 < 2; ++s1)
      for (int s2 = 0; s2 < 2; ++s2) {
        const int o1 = (e + 1) % 3, o2 = (e + 2) % 3;
        int l[3];
        l[o1] = 3 * s1;
        l[o2] = 3 * s2;
        l[e] = 0;
        double p0[3], p3[3];
        for (int k = 0; k < 3; ++k) p0[k] = F.at(l)[k];
        l[e] = 3;
        for (int k = 0; k < 3; ++k) p3[k] = F.at(l)[k];
        double G[4][4] = {}, R[3][2] = {};
        for (int m = 0; m < SAMPLES; ++m) {
          const double t = (double)m / (SAMPLES - 1);
          double b[4];
          bernstein3(t, b);
          double q[3];
          q[o1] = cube(3 * s1);
          q[o2] = cube(3 * s2);
          q[e] = cube_of(t);
          double C[3];
          F.image(q, C);
          G[0][0] += b[1] * b[1];
          G[0][1] += b[1] * b[2];
          G[1][1] += b[2] * b[2];
          for (int k = 0; k < 3; ++k) {
            const double rest = C[k] - b[0] * p0[k] - b[3] * p3[k];
            R[k][0] += b[1] * rest;
            R[k][1] += b[2] * rest;
          }
        }
        G[1][0] = G[0][1];
        for (int k = 0; k < 3; ++k) {
          double g[4][4] = {{G[0][0], G[0][1]}, {G[1][0], G[1][1]}};
          double r[2] = {R[k][0], R[k][1]};
          solve_small(2, g, r);
          l[e] = 1;
          F.at(l)[k] = r[0];
          l[e] = 2;
          F.at(l)[k] = r[1];
        }
      }
  // face interiors
  for (int f = 0; f < 6; ++f) {
    const int A = FA[f], U = FU[f], V = FV[f];
    const int la = FS[f] > 0 ? 3 : 0;
    double G[4][4] = {}, R[3][4] = {};
    for (int mu = 0; mu < SAMPLES; ++mu)
      for (int mv = 0; mv < SAMPLES; ++mv) {
        const double u = (double)mu / (SAMPLES - 1), v = (double)mv / (SAMPLES - 1);
        double bu[4], bv[4];
        bernstein3(u, bu);
        bernstein3(v, bv);
        double q[3];
        q[A] = cube(la);
        q[U] = cube_of(u);
        q[V] = cube_of(v);
        double C[3];
        F.image(q, C);
        double rest[3] = {C[0], C[1], C[2]}, phi[4];
        int idx = 0;
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 4; ++j) {
            const double w = bu[i] * bv[j];
            if (i >= 1 && i <= 2 && j >= 1 && j <= 2) {
              phi[idx++] = w;
            } else {
              int l[3];
              l[A] = la;
              l[U] = i;
              l[V] = j;
              for (int k = 0; k < 3; ++k) rest[k] -= w * F.at(l)[k];
            }
          }
        for (int p = 0; p < 4; ++p) {
          for (int r = 0; r < 4; ++r) G[p][r] += phi[p] * phi[r];
          for (int k = 0; k < 3; ++k) R[k][p] += phi[p] * rest[k];
        }
      }
    for (int k = 0; k < 3; ++k) {
      double g[4][4];
      for (int p = 0; p < 4; ++p)
        for (int r = 0; r < 4; ++r) g[p][r] = G[p][r];
      double r[4] = {R[k][0], R[k][1], R[k][2], R[k][3]};
      solve_small(4, g, r);
      int idx = 0;
      for (int i = 1; i <= 2; ++i)
        for (int j = 1; j <= 2; ++j) {
          int l[3];
          l[A] = la;
          l[U] = i;
          l[V] = j;
          F.at(l)[k] = r[idx++];
        }
    }
  }
  for (int f = 0; f < 6; ++f)
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        int l[3];
        l[FA[f]] = FS[f] > 0 ? 3 : 0;
        l[FU[f]] = i;
        l[FV[f]] = j;
        for (int k = 0; k < 3; ++k) ctrl[(f * 16 + i * 4 + j) * 3 + k] = F.at(l)[k];
      }
}

}  // namespace lsx
