// ORACLE — test infrastructure only (see oracle_solver.cpp header).
//
// Right-hand side of the pressure-correction equation, b = Dx u* + Dy v* + Dz w* + flux, with b
// set to 0 at the pinned cells.  Restates, term by term and in the reference's summation order:
//   PoissonDiscreteEquation::recalcRightMember  src/DiscreteEquations/poisson_discrete_equation.h:145-159
//     aux = Dx u*;  b = (aux + flux) + Dy v*;  b += Dz w*;  fixPoint (:87-91)
//   DivergenceDiscreteOperatorX::calcDx  src/DiscreteOperators/Divergence/divergence_discrete_operator_x.h:119-159
//   DivergenceDiscreteOperatorY::calcDy  src/DiscreteOperators/Divergence/divergence_discrete_operator_y.h:91-117
//   DivergenceDiscreteOperatorZ::calcDz  src/DiscreteOperators/Divergence/divergence_discrete_operator_z.h:86-112
//   CSR row sums: RealValuedMeshFunction::mvProd  src/Base/MeshFunctions/real_valued_mesh_function.h:81-106
//   boundary flux of a static body: LaplaceDiscreteOperator::initFlux / addFluxNSp / addFluxTBp /
//     setBaseFlux / ibFlux  src/DiscreteOperators/WithFlux/Sym/laplace_discrete_operator.h:48-83, 687-750, 799-808
//     (ibFlux is zero when the body does not move)
//
// Face arrays: u* on x faces, (N-1) per x-row, index (y + M z)(N-1) + x for the face between cells
// x and x+1; v* on y faces, N(M-1) per z-plane, index z N(M-1) + y N + x for the face between
// rows y and y+1; w* on z faces, index z NM + y N + x for the face between planes z and z+1.

#include <cmath>

#include "oracle_geometry.hpp"

namespace lsxo {

namespace {

bool is_x_east(int t) {
  switch (t) {
    case E_Border: case NE_Edge: case SE_Edge: case TE_Edge: case BE_Edge:
    case TNE_Corner: case TSE_Corner: case BNE_Corner: case BSE_Corner: return true;
    default: return false;
  }
}
bool is_x_west(int t) {
  switch (t) {
    case W_Border: case NW_Edge: case SW_Edge: case TW_Edge: case BW_Edge:
    case TNW_Corner: case TSW_Corner: case BNW_Corner: case BSW_Corner: return true;
    default: return false;
  }
}
bool is_y_north(int t) {
  switch (t) {
    case N_Border: case NE_Edge: case NW_Edge: case TN_Edge: case BN_Edge:
    case TNE_Corner: case TNW_Corner: case BNE_Corner: case BNW_Corner: return true;
    default: return false;
  }
}
bool is_y_south(int t) {
  switch (t) {
    case S_Border: case SE_Edge: case SW_Edge: case TS_Edge: case BS_Edge:
    case TSE_Corner: case TSW_Corner: case BSE_Corner: case BSW_Corner: return true;
    default: return false;
  }
}
bool is_z_bottom(int t) {
  switch (t) {
    case B_Border: case BN_Edge: case BS_Edge: case BE_Edge: case BW_Edge:
    case BNE_Corner: case BNW_Corner: case BSE_Corner: case BSW_Corner: return true;
    default: return false;
  }
}
bool is_z_top(int t) {
  switch (t) {
    case T_Border: case TN_Edge: case TS_Edge: case TE_Edge: case TW_Edge:
    case TNE_Corner: case TNW_Corner: case TSE_Corner: case TSW_Corner: return true;
    default: return false;
  }
}

// One CSR row of a divergence operator: up to two (value, column) entries in portrait order.
struct Row {
  int cnt = 0;
  double val[2];
  long col[2];
  void add(double v, long c) {
    val[cnt] = v;
    col[cnt] = c;
    ++cnt;
  }
  double apply(const double* f, double init) const {
    double aux = init;
    for (int q = 0; q < cnt; ++q) aux += val[q] * f[col[q]];
    return aux;
  }
};

}  // namespace

// Flux vector of the static-body pressure equation (laplace_discrete_operator.h:799-808 then
// setBaseFlux :744-750 for an outflow east boundary).
std::vector<double> pressure_flux(const Geometry& g, const double* u_west, double U_E, double V_S, double V_N,
                                  double W_B, double W_T, const double* u_out) {
  const int N = g.N, M = g.M, K = g.K;
  const long NM = (long)N * M, n = g.ncells();
  std::vector<double> flux(n, 0.0);  // ibFlux of a static body: 0
  // initFlux (:48-56): west and east cells of every x-row, dyz of the row's first cell
  for (int z = 0; z < K; ++z)
    for (int y = 0; y < M; ++y) {
      const long numF = (long)N * y + NM * z;
      const double dyz = g.dy[y] * g.dz[z];
      flux[numF] = -u_west[y + (long)z * M] * dyz;
      flux[numF + N - 1] = U_E * dyz;
    }
  // addFluxNSp (:57-70): south and north rows of every z-plane
  for (int z = 0; z < K; ++z) {
    const double dz = g.dz[z];
    for (int x = 0; x < N; ++x) {
      const long numS = NM * z + x, numN = numS + NM - N;
      const double coef = g.dx[x] * dz;
      flux[numS] -= coef * V_S;
      flux[numN] += coef * V_N;
    }
  }
  // addFluxTBp (:71-83): bottom and top planes, weighted by the cells' theta_z
  for (int y = 0; y < M; ++y) {
    const double dy = g.dy[y];
    for (int x = 0; x < N; ++x) {
      const long numB = (long)y * N + x, numT = n - NM + numB;
      const double coef = dy * g.dx[x];
      flux[numB] -= coef * g.thz[numB] * W_B;
      flux[numT] += coef * g.thz[numT] * W_T;
    }
  }
  // setBaseFlux (:744-750): outflow east boundary carries the last interior u face
  if (u_out)
    for (int z = 0; z < K; ++z)
      for (int y = 0; y < M; ++y) {
        const long num = (long)(y + 1) * N - 1 + NM * z;
        const long row = y + (long)z * M;
        flux[num] = u_out[row * (N - 1) + N - 2] * g.dy[y] * g.dz[z];
      }
  return flux;
}

void pressure_rhs(const Geometry& g, const double* u, const double* v, const double* w, const std::vector<double>& flux,
                  bool pin_outlet, double* b) {
  const int N = g.N, M = g.M, K = g.K;
  const long NM = (long)N * M, n = g.ncells();
  for (long i = 0; i < n; ++i) {
    const int x = (int)(i % N), y = (int)((i / N) % M), z = (int)(i / NM);
    const int t = g.type[i];
    const double dx = g.dx[x], dy = g.dy[y], dz = g.dz[z];
    // Dx row (calcDx): u faces west (w) and east (p) of the cell
    Row rx;
    {
      const long p = (y + (long)z * M) * (N - 1) + x, wcol = p - 1;
      if (t == SolidCell) rx.add(0.0, x == 0 ? p : wcol);
      else if (is_x_east(t)) rx.add(-g.thx[i - 1] * dy * dz, wcol);
      else if (is_x_west(t)) rx.add(g.thx[i] * dy * dz, p);
      else {
        const double dyz = dy * dz;
        rx.add(-g.thx[i - 1] * dyz, wcol);
        rx.add(g.thx[i] * dyz, p);
      }
    }
    // Dy row (calcDy): v faces south (s) and north (p)
    Row ry;
    {
      const long p = (long)z * N * (M - 1) + (long)y * N + x, s = p - N;
      if (t == SolidCell) ry.add(0.0, y == 0 ? p : s);
      else if (is_y_north(t)) ry.add(-g.thy[i - N] * dx * dz, s);
      else if (is_y_south(t)) ry.add(g.thy[i] * dx * dz, p);
      else {
        const double dxz = dx * dz;
        ry.add(-g.thy[i - N] * dxz, s);
        ry.add(g.thy[i] * dxz, p);
      }
    }
    // Dz row (calcDz): w faces below (b) and above (p)
    Row rz;
    {
      const long p = i, bcol = i - NM;
      if (t == SolidCell) rz.add(0.0, z == 0 ? p : bcol);
      else if (is_z_bottom(t)) rz.add(g.thz[i] * dx * dy, p);
      else if (is_z_top(t)) rz.add(-g.thz[i - NM] * dx * dy, bcol);
      else {
        const double dxy = dx * dy;
        rz.add(-g.thz[i - NM] * dxy, bcol);
        rz.add(g.thz[i] * dxy, p);
      }
    }
    const double aux = rx.apply(u, 0.0);
    double r = ry.apply(v, aux + flux[i]);
    const double auxz = rz.apply(w, 0.0);
    r += auxz;
    b[i] = r;
  }
  // fixPoint (poisson_discrete_equation.h:87-91): the two pinned cells of a flow task (pin_outlet
  // false leaves b unpinned, for an operator built without pins)
  if (pin_outlet) {
    b[N - 1] = 0.0;
    b[NM - 1] = 0.0;
  }
}

// Velocity and pressure correction after the pressure solve: SetSolutions::correction
// (src/Base/set_solutions.h:215-237, without the optional outflow clamps) with the gradient rows
// of GradientDiscreteOperatorX/Y/Z::calc (src/DiscreteOperators/Gradient/gradient_discrete_
// operator_{x,y,z}.h) and the staggered cell types cvu/cvv/cvw (src/Base/LS_STAG_Mesh/
// ls_stag_uniform_mesh.h:108-130):
//   u = (G_x dp) * inv_Mx + u*   (G row: [theta_x dy dz, -theta_x dy dz] on [west, east] cells,
//                                  a single 0 entry when the x control volume is solid)
//   p = p + dp * coef
namespace {
int staggered_type(double theta, int a, int b) { return theta < EQ_DOUBLE ? (int)SolidCell : (b < a ? b : a); }
}  // namespace

void velocity_correction(const Geometry& g, const double* dp, const double* us, const double* vs, const double* ws,
                         double* u, double* v, double* w) {
  const int N = g.N, M = g.M, K = g.K;
  const long NM = (long)N * M;
  for (int z = 0; z < K; ++z)
    for (int y = 0; y < M; ++y)
      for (int x = 0; x + 1 < N; ++x) {
        const long W = g.idx(x, y, z), f = x + (long)(N - 1) * (y + (long)M * z);
        double aux = 0.0;
        if (staggered_type(g.thx[W], g.type[W], g.type[W + 1]) == SolidCell) {
          aux += 0.0 * dp[W];
        } else {
          const double cell = g.thx[W] * g.dy[y] * g.dz[z];
          aux += cell * dp[W];
          aux += -cell * dp[W + 1];
        }
        aux *= g.inv_mx[f];
        u[f] = aux + us[f];
      }
  for (int z = 0; z < K; ++z)
    for (int y = 0; y + 1 < M; ++y)
      for (int x = 0; x < N; ++x) {
        const long S = g.idx(x, y, z), f = x + (long)N * (y + (long)(M - 1) * z);
        double aux = 0.0;
        if (staggered_type(g.thy[S], g.type[S], g.type[S + N]) == SolidCell) {
          aux += 0.0 * dp[S];
        } else {
          const double cell = g.thy[S] * g.dx[x] * g.dz[z];
          aux += cell * dp[S];
          aux += -cell * dp[S + N];
        }
        aux *= g.inv_my[f];
        v[f] = aux + vs[f];
      }
  for (int z = 0; z + 1 < K; ++z)
    for (int y = 0; y < M; ++y)
      for (int x = 0; x < N; ++x) {
        const long B = g.idx(x, y, z);
        double aux = 0.0;
        if (staggered_type(g.thz[B], g.type[B], g.type[B + NM]) == SolidCell) {
          aux += 0.0 * dp[B];
        } else {
          const double cell = g.thz[B] * g.dx[x] * g.dy[y];
          aux += cell * dp[B];
          aux += -cell * dp[B + NM];
        }
        aux *= g.inv_mz[B];
        w[B] = aux + ws[B];
      }
}

}  // namespace lsxo
