// ORACLE — test infrastructure only (see oracle_solver.cpp header).
//
// Restatement of the reference's cut-cell geometry on the base (pressure) grid:
//   level set at each cell's upper corner      flow_sphere_task_description.h:57-74
//   edge fractions alpha (+ oscillation fix)   ls_stag_uniform_mesh.h:2906-2963, driver
//                                              flow_task_description.h:102-137 (serial sweep)
//   face types / cell types                    ls_stag_uniform_mesh.h:2969-2977, 3624-3675,
//                                              driver flow_task_description.h:175-241
//   face fractions theta                       ls_stag_uniform_mesh.h:1652-1837, driver
//                                              flow_task_description.h:138-174
//   cut volumes VOF (+ PLUS_TINY)              ls_stag_uniform_mesh.h:1838-2713
//
// The geometry region ("box") where cut cells may appear is x in [hs,hf), y in [vs,vf), every
// z plane; the bottom and top planes use the reference's plane variants.  Outside the box
// edges and faces are fully open.  The volume table is expressed as formula families with
// per-type parameters (which alpha / theta component at which corner offset), keyed by the
// reference's decimal corner codes (cut_cells.h:26-168).

#include <cmath>
#include <functional>
#include <map>
#include <stdexcept>

#include "oracle_geometry.hpp"

namespace lsxo {

namespace {

constexpr int FluidSide = 1111, SolidSide = 0;

// Reference offsets: a corner reference (dx, dy, dz in {0,-1}) and a component 0=x 1=y 2=z
struct Ref {
  int dx, dy, dz, comp;
};

inline double get(const std::vector<double>* arr3, const Geometry& g, long num, Ref r) {
  const long off = r.dx + (long)g.N * r.dy + (long)g.N * g.M * r.dz;
  return arr3[r.comp][num + off];
}

int type_of_side(double phi, double phiW, double phiSW, double phiS) {
  // ls_stag_uniform_mesh.h:2969-2977: a corner counts as fluid when phi < EQ_DOUBLE
  int number = (phiS < EQ_DOUBLE) ? 1 : 0;
  if (phiSW < EQ_DOUBLE) number += 10;
  if (phiW < EQ_DOUBLE) number += 100;
  if (phi < EQ_DOUBLE) number += 1000;
  return number;
}

// Face fraction from the side type: getThetaZ and BuildTheta (ls_stag_uniform_mesh.h:1652-1837).
// P0/P1 and Q0/Q1 are the in-plane edge fractions of the face at its two corner offsets.
double face_theta(int side, double P0, double P1, double Q0, double Q1) {
  double t = 1.0;
  switch (side) {
    case FluidSide: return 1.0;
    case SolidSide: return 0.0;
    case 11: case 1100: t = 0.5 * (P0 + P1); break;                          // N/S trapezoid
    case 1001: case 110: t = 0.5 * (Q0 + Q1); break;                         // W/E trapezoid
    case 1011: t = 1.0 - 0.5 * (1.0 - P1) * (1.0 - Q0); break;               // NW pentagon
    case 111: t = 1.0 - 0.5 * (1.0 - P0) * (1.0 - Q0); break;                // NE pentagon
    case 1101: t = 1.0 - 0.5 * (1.0 - P1) * (1.0 - Q1); break;               // SW pentagon
    case 1110: t = 1.0 - 0.5 * (1.0 - P0) * (1.0 - Q1); break;               // SE pentagon
    case 1: t = 0.5 * P0 * Q1; break;                                        // NW triangle
    case 10: t = 0.5 * P1 * Q1; break;                                       // NE triangle
    case 1000: t = 0.5 * P0 * Q0; break;                                     // SW triangle
    case 100: t = 0.5 * P1 * Q0; break;                                      // SE triangle
    default: return 1.0;                                                     // other codes: open
  }
  if (t < THETA_EQ_ZERO) t = 0.0;
  return t;
}

// ---- cut-volume formula families -----------------------------------------------------------
enum Family { F_TET, F_HEP3, F_PRISM, F_HEP5, F_HEX5, F_HEP6 };
struct VofRule {
  Family fam;
  int axis;     // theta component used by the prism-like families (0 x, 1 y, 2 z)
  int sel;      // F_HEP5 / F_HEX5 / F_HEP6: 0 -> the lower face theta, 1 -> the upper face
  Ref a, b;     // F_TET/F_HEP3: a = alpha ref, b = theta ref ; F_HEP5/F_HEX5: a = ratio alpha, b = weight alpha
  Ref a2, b2;   // F_HEP6 second term
};

const Ref R0{0, 0, 0, 0};
Ref A(int dx, int dy, int dz, int comp) { return Ref{dx, dy, dz, comp}; }

std::map<int, VofRule> make_vof_table() {
  std::map<int, VofRule> t;
  const int X = 0, Y = 1, Z = 2;
  auto tet = [&](int code, Ref a, Ref th) { t[code] = VofRule{F_TET, 0, 0, a, th, R0, R0}; };
  auto hep3 = [&](int code, Ref a, Ref th) { t[code] = VofRule{F_HEP3, 0, 0, a, th, R0, R0}; };
  auto prism = [&](int code, int axis) { t[code] = VofRule{F_PRISM, axis, 0, R0, R0, R0, R0}; };
  auto hep5 = [&](int code, int axis, int sel, Ref ap, Ref aq) { t[code] = VofRule{F_HEP5, axis, sel, ap, aq, R0, R0}; };
  auto hex5 = [&](int code, int axis, int sel, Ref ap, Ref aq) { t[code] = VofRule{F_HEX5, axis, sel, ap, aq, R0, R0}; };
  auto hep6 = [&](int code, int sel, Ref ap1, Ref aq1, Ref ap2, Ref aq2) {
    t[code] = VofRule{F_HEP6, 0, sel, ap1, aq1, ap2, aq2};
  };
  // tetrahedra: alpha_x * theta_x * V / 3
  tet(10000000, A(0, 0, 0, X), A(0, 0, 0, X));
  tet(1000000, A(0, 0, 0, X), A(-1, 0, 0, X));
  tet(100000, A(0, -1, 0, X), A(-1, 0, 0, X));
  tet(10000, A(0, -1, 0, X), A(0, 0, 0, X));
  tet(1000, A(0, 0, -1, X), A(0, 0, 0, X));
  tet(100, A(0, 0, -1, X), A(-1, 0, 0, X));
  tet(10, A(0, -1, -1, X), A(-1, 0, 0, X));
  tet(1, A(0, -1, -1, X), A(0, 0, 0, X));
  // 3-gon heptahedra: (3 - (1-alpha_x)(1-theta_x)) V / 3
  hep3(1111111, A(0, 0, 0, X), A(0, 0, 0, X));
  hep3(10111111, A(0, 0, 0, X), A(-1, 0, 0, X));
  hep3(11011111, A(0, -1, 0, X), A(-1, 0, 0, X));
  hep3(11101111, A(0, -1, 0, X), A(0, 0, 0, X));
  hep3(11110111, A(0, 0, -1, X), A(0, 0, 0, X));
  hep3(11111011, A(0, 0, -1, X), A(-1, 0, 0, X));
  hep3(11111101, A(0, -1, -1, X), A(-1, 0, 0, X));
  hep3(11111110, A(0, -1, -1, X), A(0, 0, 0, X));
  // pentahedra, 4-gon heptahedra and 4-gon hexahedra: mean of two opposite face fractions
  for (int c : {10001000, 1000100, 100010, 10001, 11101110, 11011101, 1110111, 10111011}) prism(c, Z);
  for (int c : {11000000, 110000, 1100, 11, 11111100, 11110011, 11001111, 111111, 1111, 11110000, 110011, 11001100})
    prism(c, X);
  for (int c : {10010000, 1100000, 1001, 110, 11111001, 11110110, 10011111, 1101111, 1100110, 10011001}) prism(c, Y);
  // 5-gon heptahedra
  hep5(1100111, Z, 0, A(0, 0, -1, Y), A(0, -1, 0, Z));    // TENE
  hep5(11100110, Z, 1, A(0, 0, 0, Y), A(0, 0, 0, Z));     // BESE
  hep5(1110110, Z, 1, A(0, 0, 0, Y), A(0, -1, 0, Z));     // BENE
  hep5(1101110, Z, 0, A(0, 0, -1, Y), A(0, 0, 0, Z));     // TESE
  hep5(11011001, Z, 1, A(-1, 0, 0, Y), A(-1, 0, 0, Z));   // BWSW
  hep5(10111001, Z, 1, A(-1, 0, 0, Y), A(-1, -1, 0, Z));  // BWNW
  hep5(10011101, Z, 0, A(-1, 0, -1, Y), A(-1, 0, 0, Z));  // TWSW
  hep5(10011011, Z, 0, A(-1, 0, -1, Y), A(-1, -1, 0, Z)); // TWNW
  hep5(10110011, X, 1, A(0, 0, 0, Z), A(0, 0, 0, X));     // BNNW
  hep5(1110011, X, 0, A(-1, 0, 0, Z), A(0, 0, 0, X));     // BNNE
  hep5(111011, X, 1, A(0, 0, 0, Z), A(0, 0, -1, Z));      // TNNW (the reference weights with alpha_z here)
  hep5(110111, X, 0, A(-1, 0, 0, Z), A(0, 0, -1, X));     // TNNE
  hep5(11101100, X, 0, A(-1, -1, 0, Z), A(0, -1, 0, X));  // BSSE
  hep5(11011100, X, 1, A(0, -1, 0, Z), A(0, -1, 0, X));   // BSSW
  hep5(11001110, X, 0, A(-1, -1, 0, Z), A(0, -1, -1, X)); // TSSE
  hep5(11001101, X, 1, A(0, -1, 0, Z), A(0, -1, -1, X));  // TSSW
  hep5(10001111, Y, 1, A(0, 0, 0, X), A(0, 0, 0, Y));     // TSTW
  hep5(1001111, Y, 1, A(0, 0, 0, X), A(-1, 0, 0, Y));     // TSTE
  hep5(11111, Y, 0, A(0, -1, 0, X), A(-1, 0, 0, Y));      // TNTW
  hep5(101111, Y, 0, A(0, -1, 0, X), A(0, 0, 0, Y));      // TNTE
  hep5(11111000, Y, 1, A(0, 0, -1, X), A(0, 0, -1, Y));   // BSBW
  hep5(11110001, Y, 0, A(0, -1, -1, X), A(0, 0, -1, Y));  // BNBW
  hep5(11110010, Y, 0, A(0, -1, -1, X), A(-1, 0, -1, Y)); // BNBE
  hep5(11110100, Y, 1, A(0, 0, -1, X), A(-1, 0, -1, Y));  // BSBE
  // 5-gon hexahedra
  hex5(10011000, Z, 0, A(0, 0, -1, Y), A(0, -1, 0, Z));   // BWSW
  hex5(11001, Z, 1, A(0, 0, 0, Y), A(0, 0, 0, Z));        // TWNW
  hex5(10001001, Z, 1, A(0, 0, 0, Y), A(0, -1, 0, Z));    // TWSW
  hex5(10010001, Z, 0, A(0, 0, -1, Y), A(0, 0, 0, Z));    // BWNW
  hex5(100110, Z, 1, A(-1, 0, 0, Y), A(-1, 0, 0, Z));     // TENE
  hex5(1000110, Z, 1, A(-1, 0, 0, Y), A(-1, -1, 0, Z));   // TESE
  hex5(1100010, Z, 0, A(-1, 0, -1, Y), A(-1, 0, 0, Z));   // BENE
  hex5(1100100, Z, 0, A(-1, 0, -1, Y), A(-1, -1, 0, Z));  // BESE
  hex5(10001100, X, 0, A(-1, 0, 0, Z), A(0, 0, 0, X));    // TSSW
  hex5(1001100, X, 1, A(0, 0, 0, Z), A(0, 0, 0, X));      // TSSE
  hex5(11001000, X, 0, A(-1, 0, 0, Z), A(0, 0, -1, X));   // BSSW
  hex5(11000100, X, 1, A(0, 0, 0, Z), A(0, 0, -1, X));    // BSSE
  hex5(10011, X, 0, A(-1, -1, 0, Z), A(0, -1, 0, X));     // TNNW
  hex5(100011, X, 1, A(0, -1, 0, Z), A(0, -1, 0, X));     // TNNE
  hex5(110001, X, 0, A(-1, -1, 0, Z), A(0, -1, -1, X));   // BNNW
  hex5(110010, X, 1, A(0, -1, 0, Z), A(0, -1, -1, X));    // BNNE
  hex5(1110000, Y, 1, A(0, 0, 0, X), A(0, 0, 0, Y));      // BNBE
  hex5(10110000, Y, 1, A(0, 0, 0, X), A(-1, 0, 0, Y));    // BNBW
  hex5(11010000, Y, 0, A(0, -1, 0, X), A(-1, 0, 0, Y));   // BSBW
  hex5(11100000, Y, 0, A(0, -1, 0, X), A(0, 0, 0, Y));    // BSBE
  hex5(111, Y, 1, A(0, 0, -1, X), A(0, 0, 0, Y));         // TNTE
  hex5(1011, Y, 1, A(0, 0, -1, X), A(-1, 0, 0, Y));       // TNTW
  hex5(1101, Y, 0, A(0, -1, -1, X), A(-1, 0, 0, Y));      // TSTW
  hex5(1110, Y, 0, A(0, -1, -1, X), A(0, 0, 0, Y));       // TSTE
  // 6-gon heptahedra (through theta_x)
  hep6(11011000, 1, A(0, -1, 0, Z), A(0, 0, -1, X), A(0, 0, -1, Y), A(0, -1, 0, X));     // BSW
  hep6(11100100, 0, A(-1, -1, 0, Z), A(0, -1, 0, X), A(-1, 0, -1, Y), A(0, 0, -1, X));   // BSE
  hep6(1110010, 0, A(-1, 0, 0, Z), A(0, 0, 0, X), A(-1, 0, -1, Y), A(0, -1, -1, X));    // BNE
  hep6(10110001, 1, A(0, 0, 0, Z), A(0, 0, 0, X), A(0, 0, -1, Y), A(0, -1, -1, X));     // BNW
  hep6(10001101, 1, A(0, -1, 0, Z), A(0, -1, -1, X), A(0, 0, 0, Y), A(0, 0, 0, X));     // TSW
  hep6(1001110, 0, A(-1, -1, 0, Z), A(0, -1, -1, X), A(-1, 0, 0, Y), A(0, 0, 0, X));    // TSE
  hep6(100111, 0, A(-1, 0, 0, Z), A(0, 0, -1, X), A(-1, 0, 0, Y), A(0, -1, 0, X));      // TNE
  hep6(11011, 1, A(0, 0, 0, Z), A(0, 0, -1, X), A(0, 0, 0, Y), A(0, -1, 0, X));         // TNW
  return t;
}

const std::map<int, VofRule>& vof_table() {
  static const std::map<int, VofRule> t = make_vof_table();
  return t;
}

}  // namespace

int vof_table_size() { return (int)vof_table().size(); }

// Cell centres along one axis.  A uniform axis uses a + (k + 1/2) h as the reference's
// buildDeltaNodes does (ls_stag_uniform_mesh.h:174-185); a stretched axis accumulates the
// widths left to right and takes the midpoint of each cell.
std::vector<double> cell_centres(const std::vector<double>& w, double a) {
  const int n = (int)w.size();
  std::vector<double> c(n);
  bool uniform = true;
  for (int i = 1; i < n; ++i) uniform = uniform && (w[i] == w[0]);
  if (uniform) {
    for (int i = 0; i < n; ++i) c[i] = a + (i + 0.5) * w[0];
  } else {
    double e = a;
    for (int i = 0; i < n; ++i) {
      c[i] = e + 0.5 * w[i];
      e += w[i];
    }
  }
  return c;
}

// ---- level set at the upper corner of each cell (calcLevelSetSphere order of operations) -----
void corner_level_set(Geometry& g, const Box& box, const std::function<double(double, double, double)>& phi_fn) {
  const long n = g.ncells();
  g.phi.assign(n, 0.0);
  const std::vector<double> xs = cell_centres(g.dx, g.x0), ys = cell_centres(g.dy, g.y0), zs = cell_centres(g.dz, g.z0);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int z = 0; z < g.K; ++z)
    for (int y = box.vs; y < box.vf; ++y) {
      // corner of the first box cell, then step by the next cell's width
      double px = xs[box.hs] + 0.5 * g.dx[box.hs];
      const double py = ys[y] + 0.5 * g.dy[y];
      const double pz = zs[z] + 0.5 * g.dz[z];
      for (int x = box.hs; x < box.hf; ++x) {
        g.phi[g.idx(x, y, z)] = phi_fn(px, py, pz);
        if (x + 1 < g.N) px += g.dx[x + 1];
      }
    }
}

// ---- edge fractions ------------------------------------------------------------------------
static void check_osc(std::vector<double>& phi, long i, double& f, double fwCur, double fw2) {
  if ((f > EQ_DOUBLE) && (fwCur < 0) && (fw2 < 0)) {
    f = 0.5 * (fwCur + fw2);
    phi[i] = f;
  }
}

static inline double edge_fraction(double f, double fo) {
  return (f * fo < 0) ? ((f > 0) ? (fo / (fo - f)) : (f / (f - fo))) : ((f > -fo) ? 0.0 : 1.0);
}

void build_alpha(Geometry& g, const Box& box) {
  const long n = g.ncells();
  const long N = g.N, NM = (long)g.N * g.M;
  for (int c = 0; c < 3; ++c) g.alpha[c].assign(n, 1.0);
  for (int z = 0; z < g.K; ++z)
    for (int y = box.vs; y < box.vf; ++y)
      for (int x = box.hs; x < box.hf; ++x) {
        const long num = g.idx(x, y, z);
        double f = g.phi[num];
        const double fw = g.phi[num - 1], fs = g.phi[num - N];
        if (z == 0 || z == g.K - 1) {  // bottomBuildAlpha / topBuildAlpha
          check_osc(g.phi, num, f, fw, g.phi[num + 1]);
          check_osc(g.phi, num, f, fs, g.phi[num + N]);
          g.alpha[0][num] = edge_fraction(f, fw);
          g.alpha[1][num] = edge_fraction(f, fs);
          g.alpha[2][num] = (f > 0.0) ? 0.0 : 1.0;
        } else {  // BuildAlpha
          const double fb = g.phi[num - NM];
          check_osc(g.phi, num, f, fw, g.phi[num + 1]);
          check_osc(g.phi, num, f, fs, g.phi[num + N]);
          check_osc(g.phi, num, f, fb, g.phi[num + NM]);
          g.alpha[0][num] = edge_fraction(f, fw);
          g.alpha[1][num] = edge_fraction(f, fs);
          g.alpha[2][num] = edge_fraction(f, fb);
        }
      }
}

// ---- face and cell types ---------------------------------------------------------------------
void build_types(Geometry& g, const Box& box) {
  const long n = g.ncells();
  const long N = g.N, NM = (long)g.N * g.M;
  for (int c = 0; c < 3; ++c) g.side[c].assign(n, FluidSide);
  for (int z = 0; z < g.K; ++z)
    for (int y = 0; y < g.M; ++y)
      for (int x = 0; x < g.N; ++x) g.type[g.idx(x, y, z)] = boundary_type(x, y, z, g.N, g.M, g.K);
  const auto& p = g.phi;
  for (int z = 0; z < g.K; ++z)
    for (int y = box.vs; y < box.vf; ++y)
      for (int x = box.hs; x < box.hf; ++x) {
        const long num = g.idx(x, y, z);
        const double phi = p[num], phiW = p[num - 1], phiS = p[num - N];
        if (z == 0 || z == g.K - 1) {  // bottomTypeOfCell / topTypeOfCell
          const int tT = type_of_side(phi, phiW, p[num - N - 1], phiS);
          g.side[0][num] = type_of_side(phi, phiS, phiS, phi);
          g.side[1][num] = type_of_side(phi, phiW, phiW, phi);
          g.side[2][num] = tT;
          g.type[num] = (tT == SolidSide) ? SolidCell : (z == 0 ? B_Border : T_Border);
        } else {  // TypeOfCell
          const double phiSB = p[num - N - NM], phiB = p[num - NM], phiBW = p[num - NM - 1];
          const int tT = type_of_side(phi, phiW, p[num - N - 1], phiS);
          g.side[0][num] = type_of_side(phi, phiS, phiSB, phiB);
          g.side[1][num] = type_of_side(phi, phiW, phiBW, phiB);
          g.side[2][num] = tT;
          g.type[num] = tT * 10000 + type_of_side(phiB, phiBW, p[num - N - 1 - NM], phiSB);
        }
      }
}

// ---- face fractions ----------------------------------------------------------------------------
void build_theta(Geometry& g, const Box& box) {
  const long n = g.ncells();
  const long N = g.N, NM = (long)g.N * g.M;
  g.thx.assign(n, 1.0);
  g.thy.assign(n, 1.0);
  g.thz.assign(n, 1.0);
  const auto* a = g.alpha;
  for (int z = 0; z < g.K; ++z)
    for (int y = box.vs; y < box.vf; ++y)
      for (int x = box.hs; x < box.hf; ++x) {
        const long num = g.idx(x, y, z);
        const double tz = face_theta(g.side[2][num], a[1][num], a[1][num - 1], a[0][num], a[0][num - N]);
        if (z == 0 || z == g.K - 1) {  // tbBuildTheta
          const double tx = a[1][num], ty = a[0][num];
          g.thx[num] = (tx < THETA_EQ_ZERO) ? 0.0 : tx;
          g.thy[num] = (ty < THETA_EQ_ZERO) ? 0.0 : ty;
          g.thz[num] = tz;
        } else {  // BuildTheta
          g.thx[num] = face_theta(g.side[0][num], a[2][num], a[2][num - N], a[1][num], a[1][num - NM]);
          g.thy[num] = face_theta(g.side[1][num], a[2][num], a[2][num - 1], a[0][num], a[0][num - NM]);
          g.thz[num] = tz;
        }
      }
}

// ---- cut volumes ------------------------------------------------------------------------------
void build_vof(Geometry& g) {
  const std::vector<double>* th3[3] = {&g.thx, &g.thy, &g.thz};
  std::vector<double> T[3] = {g.thx, g.thy, g.thz};
  (void)th3;
  const auto& table = vof_table();
  const long N = g.N, NM = (long)g.N * g.M;
  for (int z = 0; z < g.K; ++z)
    for (int y = 0; y < g.M; ++y)
      for (int x = 0; x < g.N; ++x) {
        const long i = g.idx(x, y, z);
        const double fullV = g.dx[x] * g.dy[y] * g.dz[z];
        double vol = fullV;
        const int t = g.type[i];
        if (t == B_Border || t == T_Border) {
          vol *= g.thz[i];
        } else if (t == SolidCell) {
          vol = 0.0;
        } else {
          auto it = table.find(t);
          if (it != table.end()) {
            const VofRule& r = it->second;
            const long lowOff = r.axis == 0 ? 1 : (r.axis == 1 ? N : NM);
            const std::vector<double>& Ta = T[r.axis];
            const double tUp = Ta[i], tLo = Ta[i - lowOff];
            switch (r.fam) {
              case F_TET:
                vol = get(g.alpha, g, i, r.a) * get(T, g, i, r.b) * fullV * ONE_THIRD;
                break;
              case F_HEP3:
                vol = (3.0 - (1.0 - get(g.alpha, g, i, r.a)) * (1.0 - get(T, g, i, r.b))) * fullV * ONE_THIRD;
                break;
              case F_PRISM:
                vol = (tUp + tLo) * fullV * 0.5;
                break;
              case F_HEP5: {
                const double tsel = r.sel ? tUp : tLo;
                double ap = get(g.alpha, g, i, r.a);
                const double denom = 1.0 - ap;
                double term = 0.0;
                if (denom > EQ_DOUBLE) {
                  ap /= denom;
                  term = (1.0 - tsel) * (3.0 - 2.0 * get(g.alpha, g, i, r.b)) * ap * ap * ONE_THIRD;
                }
                vol = (tLo + tUp + term) * fullV * 0.5;
                break;
              }
              case F_HEX5: {
                const double denom = get(g.alpha, g, i, r.a);
                double term2 = r.sel ? tUp : tLo;
                const double other = r.sel ? tLo : tUp;
                if (denom > EQ_DOUBLE) {
                  const double coef = (1.0 - denom) / denom;
                  term2 *= (1.0 - (1.0 + 2.0 * get(g.alpha, g, i, r.b)) * coef * coef * ONE_THIRD);
                }
                vol = (other + term2) * fullV * 0.5;
                break;
              }
              case F_HEP6: {
                const double tsel = r.sel ? tUp : tLo;
                const double other = r.sel ? tLo : tUp;
                double term = 0.0;
                double al = get(g.alpha, g, i, r.a);
                double denom = 1.0 - al;
                if (denom > EQ_DOUBLE) {
                  al /= denom;
                  term = (3.0 - 2.0 * get(g.alpha, g, i, r.b)) * al * al;
                }
                al = get(g.alpha, g, i, r.a2);
                denom = 1.0 - al;
                if (denom > EQ_DOUBLE) {
                  al /= denom;
                  term += (3.0 - 2.0 * get(g.alpha, g, i, r.b2)) * al * al;
                }
                vol = (other + tsel + (1.0 - tsel) * term * ONE_THIRD) * fullV * 0.5;
                break;
              }
            }
          }
        }
        g.vof[i] = vol + PLUS_TINY;
      }
}

void build_cut_geometry(Geometry& g, const Box& box, const std::function<double(double, double, double)>& phi_fn) {
  if (box.hs < 1 || box.vs < 1 || box.hf > g.N - 1 || box.vf > g.M - 1 || box.hs >= box.hf || box.vs >= box.vf)
    throw std::runtime_error("geometry box must lie inside the domain's interior");
  corner_level_set(g, box, phi_fn);
  build_alpha(g, box);
  build_types(g, box);
  build_theta(g, box);
  build_vof(g);
  g.build_masses();
}

}  // namespace lsxo
