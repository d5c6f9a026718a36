// ORACLE — test infrastructure only (see oracle_solver.cpp header).
//
// Restatement of the cut-cell pressure operator assembly of the reference:
//   * CSR row shapes and column order [b, s, w, p, e, n, t] minus the neighbours missing at the
//     domain boundary: SymDiscreteOperator::numBuild / portraitBuild
//     (src/DiscreteOperators/WithFlux/Sym/sym_discrete_operator.h:36-269);
//   * values: LaplaceDiscreteOperator::calcLap (laplace_discrete_operator.h:107-686):
//     a_nb = (theta_face * face_area)^2 * inv_M[face], a_p = -sum in the order the reference
//     adds them, solid cells a single diagonal PLUS_TINY;
//   * the two pinned outlet-corner rows, fixPointFlow (laplace_discrete_operator.h:84-89,
//     760-763, 809-821): rows (N-1,0,0) and (N-1,M-1,0) become identity rows;
//   * staggered masses M = (V_i + V_nb)/2 and their inverses (ls_stag_uniform_mesh.h:2804-2826,
//     ls_stag_base_mesh_params.h:77-85).

#include "oracle_geometry.hpp"

#include <cmath>
#include <stdexcept>

namespace lsxo {

static bool has_b(int t) { return !(t == B_Border || t == BN_Edge || t == BS_Edge || t == BE_Edge || t == BW_Edge || t == BNE_Corner || t == BNW_Corner || t == BSE_Corner || t == BSW_Corner); }
static bool has_t(int t) { return !(t == T_Border || t == TN_Edge || t == TS_Edge || t == TE_Edge || t == TW_Edge || t == TNE_Corner || t == TNW_Corner || t == TSE_Corner || t == TSW_Corner); }
static bool has_s(int t) { return !(t == S_Border || t == TS_Edge || t == BS_Edge || t == SE_Edge || t == SW_Edge || t == TSE_Corner || t == TSW_Corner || t == BSE_Corner || t == BSW_Corner); }
static bool has_n(int t) { return !(t == N_Border || t == TN_Edge || t == BN_Edge || t == NE_Edge || t == NW_Edge || t == TNE_Corner || t == TNW_Corner || t == BNE_Corner || t == BNW_Corner); }
static bool has_w(int t) { return !(t == W_Border || t == TW_Edge || t == BW_Edge || t == NW_Edge || t == SW_Edge || t == TNW_Corner || t == TSW_Corner || t == BNW_Corner || t == BSW_Corner); }
static bool has_e(int t) { return !(t == E_Border || t == TE_Edge || t == BE_Edge || t == NE_Edge || t == SE_Edge || t == TNE_Corner || t == TSE_Corner || t == BNE_Corner || t == BSE_Corner); }

int boundary_type(int x, int y, int z, int N, int M, int K) {
  // Position-based typing of the domain's outer cells (flow_task_description.h:175-241).
  const bool w = x == 0, e = x == N - 1, s = y == 0, n = y == M - 1, b = z == 0, t = z == K - 1;
  if (b) {
    if (s) return w ? BSW_Corner : (e ? BSE_Corner : BS_Edge);
    if (n) return w ? BNW_Corner : (e ? BNE_Corner : BN_Edge);
    return w ? BW_Edge : (e ? BE_Edge : B_Border);
  }
  if (t) {
    if (s) return w ? TSW_Corner : (e ? TSE_Corner : TS_Edge);
    if (n) return w ? TNW_Corner : (e ? TNE_Corner : TN_Edge);
    return w ? TW_Edge : (e ? TE_Edge : T_Border);
  }
  if (s) return w ? SW_Edge : (e ? SE_Edge : S_Border);
  if (n) return w ? NW_Edge : (e ? NE_Edge : N_Border);
  if (w) return W_Border;
  if (e) return E_Border;
  return FluidCell;
}

void Geometry::init_grid(const std::vector<double>& dxs, const std::vector<double>& dys, const std::vector<double>& dzs,
                         double a, double b, double c) {
  N = (int)dxs.size();
  M = (int)dys.size();
  K = (int)dzs.size();
  dx = dxs;
  dy = dys;
  dz = dzs;
  x0 = a;
  y0 = b;
  z0 = c;
  const long n = ncells();
  type.assign(n, FluidCell);
  thx.assign(n, 1.0);
  thy.assign(n, 1.0);
  thz.assign(n, 1.0);
  vof.assign(n, 0.0);
  for (int z = 0; z < K; ++z)
    for (int y = 0; y < M; ++y)
      for (int x = 0; x < N; ++x) type[idx(x, y, z)] = boundary_type(x, y, z, N, M, K);
}

void Geometry::build_masses() {
  // ls_stag_uniform_mesh.h:2804-2826 (mx, my, mz) + inverseVOF (ls_stag_base_mesh_params.h:77-85)
  const long NM = (long)N * M;
  inv_mx.assign((long)(N - 1) * M * K, 0.0);
  inv_my.assign((long)N * (M - 1) * K, 0.0);
  inv_mz.assign((long)N * M * (K - 1), 0.0);
  auto inv = [](double v) { return (v < EQ_TINY_DOUBLE) ? 0.0 : (1.0 / v); };
  for (int z = 0; z < K; ++z)
    for (int y = 0; y < M; ++y) {
      for (int x = 0; x + 1 < N; ++x) {
        const long num = idx(x, y, z);
        inv_mx[x + (long)(N - 1) * (y + (long)M * z)] = inv(0.5 * (vof[num] + vof[num + 1]));
      }
      if (y + 1 < M)
        for (int x = 0; x < N; ++x) {
          const long num = idx(x, y, z);
          inv_my[x + (long)N * (y + (long)(M - 1) * z)] = inv(0.5 * (vof[num] + vof[num + N]));
        }
      if (z + 1 < K)
        for (int x = 0; x < N; ++x) {
          const long num = idx(x, y, z);
          inv_mz[num] = inv(0.5 * (vof[num] + vof[num + NM]));
        }
    }
}

Csr build_laplace(const Geometry& g, bool pin_outlet) {
  const int N = g.N, M = g.M;
  const long NM = (long)N * M, n = g.ncells();
  Csr A;
  A.rows = A.cols = n;
  A.rowptr.assign(n + 1, 0);
  for (long i = 0; i < n; ++i) {
    const int t = g.type[i];
    int cnt = 1;
    if (t != SolidCell) cnt = 1 + has_b(t) + has_s(t) + has_w(t) + has_e(t) + has_n(t) + has_t(t);
    A.rowptr[i + 1] = A.rowptr[i] + cnt;
  }
  A.col.resize(A.rowptr[n]);
  A.val.resize(A.rowptr[n]);
  for (int z = 0; z < g.K; ++z)
    for (int y = 0; y < M; ++y)
      for (int x = 0; x < N; ++x) {
        const long num = g.idx(x, y, z);
        const long numX = x + (long)(N - 1) * (y + (long)M * z);
        const long numY = x + (long)N * (y + (long)(M - 1) * z);
        const int t = g.type[num];
        int p = A.rowptr[num];
        if (t == SolidCell) {
          A.col[p] = (int)num;
          A.val[p] = PLUS_TINY;
          continue;
        }
        const double ddx = g.dx[x], ddy = g.dy[y], ddz = g.dz[z];
        double a_b = 0, a_s = 0, a_w = 0, a_e = 0, a_n = 0, a_t = 0, a_p;
        if (t == FluidCell) {
          // laplace_discrete_operator.h:117-139: dxy is formed as dy*dx there
          const double dyz = ddy * ddz, dxz = ddx * ddz, dxy = ddy * ddx;
          double aux = dyz * dyz;
          a_w = aux * g.inv_mx[numX - 1];
          a_e = aux * g.inv_mx[numX];
          aux = dxz * dxz;
          a_s = aux * g.inv_my[numY - N];
          a_n = aux * g.inv_my[numY];
          aux = dxy * dxy;
          a_b = aux * g.inv_mz[num - NM];
          a_t = aux * g.inv_mz[num];
          a_p = -a_b - a_s - a_w - a_e - a_n - a_t;
        } else {
          // borders, edges, corners and cut cells: each face area scaled by its theta
          const double dyz = ddy * ddz, dxz = ddx * ddz, dxy = ddx * ddy;
          double aux;
          if (has_w(t)) { aux = g.thx[num - 1] * dyz; a_w = aux * aux * g.inv_mx[numX - 1]; }
          if (has_e(t)) { aux = g.thx[num] * dyz; a_e = aux * aux * g.inv_mx[numX]; }
          if (has_s(t)) { aux = g.thy[num - N] * dxz; a_s = aux * aux * g.inv_my[numY - N]; }
          if (has_n(t)) { aux = g.thy[num] * dxz; a_n = aux * aux * g.inv_my[numY]; }
          if (has_b(t)) { aux = g.thz[num - NM] * dxy; a_b = aux * aux * g.inv_mz[num - NM]; }
          if (has_t(t)) { aux = g.thz[num] * dxy; a_t = aux * aux * g.inv_mz[num]; }
          // a_p = -(present terms) in the order b, s, w, e, n, t (the reference's per-type
          // expressions list exactly the present ones in this order)
          a_p = 0.0;
          bool first = true;
          auto sub = [&](bool present, double v) {
            if (!present) return;
            if (first) { a_p = -v; first = false; }
            else a_p = a_p - v;
          };
          sub(has_b(t), a_b);
          sub(has_s(t), a_s);
          sub(has_w(t), a_w);
          sub(has_e(t), a_e);
          sub(has_n(t), a_n);
          sub(has_t(t), a_t);
        }
        if (has_b(t)) { A.col[p] = (int)(num - NM); A.val[p++] = a_b; }
        if (has_s(t)) { A.col[p] = (int)(num - N); A.val[p++] = a_s; }
        if (has_w(t)) { A.col[p] = (int)(num - 1); A.val[p++] = a_w; }
        A.col[p] = (int)num; A.val[p++] = a_p;
        if (has_e(t)) { A.col[p] = (int)(num + 1); A.val[p++] = a_e; }
        if (has_n(t)) { A.col[p] = (int)(num + N); A.val[p++] = a_n; }
        if (has_t(t)) { A.col[p] = (int)(num + NM); A.val[p++] = a_t; }
      }
  if (pin_outlet) {
    const long pins[2] = {NM - 1, (long)N - 1};  // fixPointFlowNum, fixPointNum
    for (long pr : pins)
      for (int c = A.rowptr[pr]; c < A.rowptr[pr + 1]; ++c) A.val[c] = (A.col[c] == pr) ? 1.0 : 0.0;
  }
  return A;
}

void body_free_vof(Geometry& g) {
  // ls_stag_uniform_mesh.h:1838-1853, 2713: fluid and border cells keep the full volume
  // (B/T borders scale by theta_z, which is 1 without a body), plus PLUS_TINY.
  for (int z = 0; z < g.K; ++z)
    for (int y = 0; y < g.M; ++y)
      for (int x = 0; x < g.N; ++x) {
        const long i = g.idx(x, y, z);
        double vol = g.dx[x] * g.dy[y] * g.dz[z];
        if (g.type[i] == B_Border || g.type[i] == T_Border) vol *= g.thz[i];
        if (g.type[i] == SolidCell) vol = 0.0;
        g.vof[i] = vol + PLUS_TINY;  // ls_stag_uniform_mesh.h:2713
      }
}

}  // namespace lsxo
