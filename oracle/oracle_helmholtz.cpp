// ORACLE — test infrastructure only (see oracle_solver.cpp header).
//
// Restatement of the momentum (Helmholtz) operators of the reference on the staggered meshes,
// for a box without an immersed body (every pressure cell fluid or a domain-boundary type):
//   * staggered control-volume types, the smaller of the two base-cell types a face separates
//     (LSSTAGUniformMesh::cvu / cvv / cvw, ls_stag_uniform_mesh.h:108-130);
//   * staggered masses M = (V_i + V_nb) / 2 (mx / my / mz, ls_stag_uniform_mesh.h:2804-2826) and the
//     base cells' inverse volumes (ls_stag_base_mesh_params.h:77-85);
//   * CSR rows [b, s, w, p, e, n, t] minus the neighbours a boundary type lacks, as for the
//     pressure operator (sym_discrete_operator.h:36-269);
//   * values: HelmholtzDiscreteOperatorX / Y / Z::calcHelm* (helmholtz_discrete_operator_x.h:
//     98-689, _y.h, _z.h) — viscous couplings across the cross faces, nu (Delta area)^2 / V along
//     the component's own axis, wall terms on no-slip walls (those on the bottom and top walls
//     dropped for flow tasks), the diagonal -(COEF_INTEGR_TIME M) - sum, and for flow tasks the
//     outflow term added to the diagonal of the last cell of every x row (addPlusOutFlow /
//     setPlusOutFlow*, helmholtz_discrete_operator_base.h:42-62 and the three operator files).
// Every face fraction and edge fraction is 1 in a body-free box, so the reference's products by
// them are exact and are left out; the remaining operations are evaluated in the reference's order
// for each cell type (the order of the diagonal's terms differs between types and is kept, per
// type, in the tables below).  Behaviours of the reference kept on purpose:
//   * an x-mesh cell of the west-boundary types has no west coupling in its diagonal (the inflow
//     enters through the flux); the y and z operators of the east-boundary types include the east
//     coupling in their diagonal, computed with the x width of base cell num + 1 (the first cell of
//     the next row);
//   * the x operator's TN edge evaluates its wall terms with (dy dz) / dz in place of dy;
//   * the y operator's edges and corners on the bottom and top walls evaluate the wall term with
//     dx dz in place of dx (the y borders B and T use dx).
// Cut cells need the immersed-boundary fractions and surfaces (ls_stag_uniform_mesh.h:186-) that
// this restatement does not build: a geometry with cut cells is rejected.

#include <stdexcept>
#include <string>

#include "oracle_geometry.hpp"

namespace lsxo {

namespace {

bool is_b(int t) { return t == B_Border || t == BN_Edge || t == BS_Edge || t == BE_Edge || t == BW_Edge || t == BNE_Corner || t == BNW_Corner || t == BSE_Corner || t == BSW_Corner; }
bool is_t(int t) { return t == T_Border || t == TN_Edge || t == TS_Edge || t == TE_Edge || t == TW_Edge || t == TNE_Corner || t == TNW_Corner || t == TSE_Corner || t == TSW_Corner; }
bool is_s(int t) { return t == S_Border || t == TS_Edge || t == BS_Edge || t == SE_Edge || t == SW_Edge || t == TSE_Corner || t == TSW_Corner || t == BSE_Corner || t == BSW_Corner; }
bool is_n(int t) { return t == N_Border || t == TN_Edge || t == BN_Edge || t == NE_Edge || t == NW_Edge || t == TNE_Corner || t == TNW_Corner || t == BNE_Corner || t == BNW_Corner; }
bool is_w(int t) { return t == W_Border || t == TW_Edge || t == BW_Edge || t == NW_Edge || t == SW_Edge || t == TNW_Corner || t == TSW_Corner || t == BNW_Corner || t == BSW_Corner; }
bool is_e(int t) { return t == E_Border || t == TE_Edge || t == BE_Edge || t == NE_Edge || t == SE_Edge || t == TNE_Corner || t == TSE_Corner || t == BNE_Corner || t == BSE_Corner; }

struct Row {
  // off-diagonal values of [b, s, w, e, n, t] (present by type) and the diagonal
  double b = 0, s = 0, w = 0, e = 0, n = 0, t = 0, p = 0;
};

}  // namespace

// dir 0: u on the (N-1) x M x K x-face mesh; 1: v on N x (M-1) x K; 2: w on N x M x (K-1)
Csr build_helmholtz(const Geometry& g, int dir, double nu, double coef, bool flow, int* dims_out) {
  const int N = g.N, M = g.M, K = g.K;
  const long NM = (long)N * M;
  if (dir < 0 || dir > 2) throw std::runtime_error("direction must be 0, 1 or 2");
  for (long i = 0; i < g.ncells(); ++i) {
    const int t = g.type[i];
    if (t != FluidCell && t >= 0) throw std::runtime_error("the Helmholtz restatement covers boxes without cut cells");
  }
  const int Nd = N - (dir == 0), Md = M - (dir == 1), Kd = K - (dir == 2);
  if (dims_out) {
    dims_out[0] = Nd;
    dims_out[1] = Md;
    dims_out[2] = Kd;
  }
  const long nd = (long)Nd * Md * Kd;
  const long off = dir == 0 ? 1 : (dir == 1 ? N : NM);  // base-cell step across the face
  auto V = [&](long c) { return g.vof[c]; };
  auto iv = [&](long c) { return g.vof[c] < EQ_TINY_DOUBLE ? 0.0 : 1.0 / g.vof[c]; };

  Csr A;
  A.rows = A.cols = nd;
  A.rowptr.assign(nd + 1, 0);
  for (int k = 0; k < Kd; ++k)
    for (int j = 0; j < Md; ++j)
      for (int i = 0; i < Nd; ++i) {
        const long q = i + (long)Nd * (j + (long)Md * k);
        const long c = i + N * (j + (long)M * k);  // base cell on the lower side of the face
        const int tc = g.type[c], td = g.type[c + off];
        const double th = dir == 0 ? g.thx[c] : (dir == 1 ? g.thy[c] : g.thz[c]);
        const int t = th < EQ_DOUBLE ? SolidCell : (td < tc ? td : tc);
        const double dx = g.dx[i], dy = g.dy[j], dz = g.dz[k];
        const double mass = 0.5 * (V(c) + V(c + off));
        Row r;
        if (t == FluidCell) {
          if (dir == 0) {
            const double dxx = nu * (g.dx[i + 1] + dx);
            double aux = dz * dxx;
            r.s = aux / (dy + g.dy[j - 1]);
            r.n = aux / (dy + g.dy[j + 1]);
            aux = dy * dxx;
            r.b = aux / (dz + g.dz[k - 1]);
            r.t = aux / (dz + g.dz[k + 1]);
            aux = dy * dz;
            const double d2 = nu * aux * aux;
            r.e = d2 * iv(c + 1);
            r.w = d2 * iv(c);
          } else if (dir == 1) {
            const double dyy = nu * (g.dy[j + 1] + dy);
            double aux = dz * dyy;
            r.w = aux / (dx + g.dx[i - 1]);
            r.e = aux / (dx + g.dx[i + 1]);
            aux = dx * dyy;
            r.b = aux / (dz + g.dz[k - 1]);
            r.t = aux / (dz + g.dz[k + 1]);
            aux = dx * dz;
            const double d2 = nu * aux * aux;
            r.n = d2 * iv(c + N);
            r.s = d2 * iv(c);
          } else {
            const double dzz = nu * (g.dz[k + 1] + dz);
            double aux = dzz * dy;
            r.w = aux / (dx + g.dx[i - 1]);
            r.e = aux / (dx + g.dx[i + 1]);
            aux = dx * dzz;
            r.s = aux / (dy + g.dy[j - 1]);
            r.n = aux / (dy + g.dy[j + 1]);
            aux = dx * dy;
            const double d2 = nu * aux * aux;
            r.t = d2 * iv(c + NM);
            r.b = d2 * iv(c);
          }
          r.p = -(coef * mass) - r.b - r.s - r.w - r.e - r.n - r.t;
        } else if (t != SolidCell) {
          const bool hb = !is_b(t), ht = !is_t(t), hs = !is_s(t), hn = !is_n(t), hw = !is_w(t);
          double p = -(coef * mass);
          if (dir == 0) {
            const double dxx = g.dx[i + 1], sx = dx + dxx;
            const double az = dz * nu, ay = dy * nu;
            if (hs) r.s = az * sx / (dy + g.dy[j - 1]);
            if (hn) r.n = az * sx / (dy + g.dy[j + 1]);
            if (hb) r.b = ay * sx / (dz + g.dz[k - 1]);
            if (ht) r.t = ay * sx / (dz + g.dz[k + 1]);
            const double A2 = dy * dz, d2 = nu * A2 * A2;
            const double ke = d2 * iv(c + 1), kw = hw ? d2 * iv(c) : 0.0;
            r.e = ke;
            r.w = kw;
            double wallz = 0.0, wally = 0.0;
            if (t == TN_Edge) {  // (dy dz) / dz in place of dy
              const double q2 = A2 / dz;
              wallz = flow ? 0.0 : q2 * nu * sx / dz;
              wally = nu * dz * sx / q2;
            } else {
              if (is_b(t) || is_t(t)) wallz = flow ? 0.0 : dy * nu * sx / dz;
              if (is_n(t) || is_s(t)) wally = nu * dz * sx / dy;
            }
            if (hb) p -= r.b;
            if (hs) p -= r.s;
            p -= hw ? (kw + ke) : ke;  // (k_w + k_e) tx, or k_e tx on the west boundary (tx = 1)
            if (hn) p -= r.n;
            if (ht) p -= r.t;
            if (is_b(t) || is_t(t)) p -= wallz;
            if (is_n(t) || is_s(t)) p -= wally;
          } else if (dir == 1) {
            const double dyy = g.dy[j + 1], sy = dy + dyy;
            const double az = dz * nu, ax = dx * nu;
            if (hw) r.w = az * sy / (dx + g.dx[i - 1]);
            const long ce = c + 1;  // the reference's east neighbour: on the east boundary the next row's first cell
            const double ke = az * sy / (dx + g.dx[ce % N]);
            r.e = ke;
            if (hb) r.b = ax * sy / (dz + g.dz[k - 1]);
            if (ht) r.t = ax * sy / (dz + g.dz[k + 1]);
            const double A2 = dx * dz, d2 = nu * A2 * A2;
            const double ks = d2 * iv(c), kn = d2 * iv(c + N);
            r.s = ks;
            r.n = kn;
            double wallz = 0.0;
            if (is_b(t) || is_t(t)) {
              const double dxw = (t == B_Border || t == T_Border) ? dx : A2;  // edges and corners: dx dz
              wallz = flow ? 0.0 : dxw * nu * sy / dz;
            }
            const double wally = is_n(t) ? kn : (is_s(t) ? ks : 0.0);  // own-axis walls
            if (hb) p -= r.b;
            if (is_n(t)) {
              if (t == N_Border) {
                p -= r.w;
                p -= ks;
              } else {
                p -= ks;
                if (hw) p -= r.w;
              }
              p -= ke;
            } else if (is_s(t)) {
              if (hw) p -= r.w;
              p -= ke;
              p -= kn;
            } else {
              p -= ks + kn;
              if (hw) p -= r.w;
              p -= ke;
            }
            if (ht) p -= r.t;
            if (is_b(t) || is_t(t)) p -= wallz;
            if (is_n(t) || is_s(t)) p -= wally;
          } else {
            const double dzz = g.dz[k + 1], sz = dz + dzz;
            const double ay = dy * nu, ax = dx * nu;
            if (hw) r.w = ay * sz / (dx + g.dx[i - 1]);
            const long ce = c + 1;
            const double ke = ay * sz / (dx + g.dx[ce % N]);
            r.e = ke;
            if (hs) r.s = ax * sz / (dy + g.dy[j - 1]);
            if (hn) r.n = ax * sz / (dy + g.dy[j + 1]);
            double wally = 0.0;
            if (is_n(t) || is_s(t)) wally = ax * sz / dy;  // cross walls, before dx *= dy
            const double A2 = dx * dy, d2 = nu * A2 * A2;
            const double kb = d2 * iv(c), kt = d2 * iv(c + NM);
            r.b = kb;
            r.t = kt;
            if (is_b(t)) {
              const double wb = flow ? 0.0 : kb;
              if (hs) p -= r.s;
              if (hw) p -= r.w;
              p -= ke;
              if (hn) p -= r.n;
              p -= kt + wb;  // (k_t + k_wallB) tz
            } else if (is_t(t)) {
              const double wt = flow ? 0.0 : kt;
              p -= kb + wt;  // (k_b + k_wallT) tz
              if (hs) p -= r.s;
              if (hw) p -= r.w;
              p -= ke;
              if (hn) p -= r.n;
            } else if (t == N_Border) {
              p -= r.s;
              p -= r.w;
              p -= kb + kt;
              p -= ke;
            } else {
              p -= kb + kt;
              if (hs) p -= r.s;
              if (hw) p -= r.w;
              p -= ke;
              if (hn) p -= r.n;
            }
            if (is_n(t) || is_s(t)) p -= wally;
          }
          r.p = p;
        }
        // outflow: the last cell of every x row of a flow task (addPlusOutFlow)
        if (flow && t != SolidCell && i == Nd - 1) {
          const long cb = (N - 1) + N * (j + (long)M * k);  // base cell (N-1, j, k)
          if (dir == 0)
            r.p += nu * g.dy[j] * g.dz[k] / g.dx[N - 1];
          else if (dir == 1)
            r.p += nu * (g.dy[j] + g.dy[j + 1]) * g.dz[k] / g.dx[N - 1];
          else
            r.p += nu * (g.dz[k] + g.dz[k + 1]) * g.dy[j] / g.dx[N - 1];
          (void)cb;
        }
        // the row, in the reference's column order
        auto put = [&](long col, double v) {
          A.col.push_back((int)col);
          A.val.push_back(v);
        };
        if (t == SolidCell) {
          put(q, PLUS_TINY);
        } else {
          const bool hb = t == FluidCell || !is_b(t), ht = t == FluidCell || !is_t(t);
          const bool hs = t == FluidCell || !is_s(t), hn = t == FluidCell || !is_n(t);
          const bool hw = t == FluidCell || !is_w(t), he = t == FluidCell || !is_e(t);
          const long ND = Nd, NMD = (long)Nd * Md;
          if (hb) put(q - NMD, r.b);
          if (hs) put(q - ND, r.s);
          if (hw) put(q - 1, r.w);
          put(q, r.p);
          if (he) put(q + 1, r.e);
          if (hn) put(q + ND, r.n);
          if (ht) put(q + NMD, r.t);
        }
        A.rowptr[q + 1] = (int)A.col.size();
      }
  return A;
}

}  // namespace lsxo
