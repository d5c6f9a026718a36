// ORACLE — test infrastructure only (see oracle_solver.cpp header).
#pragma once

#include <functional>
#include <vector>

#include "oracle_internal.hpp"

namespace lsxo {

// Cell type codes of the reference, src/Base/LS_STAG_Mesh/cut_cells.h:26-168.  Interior cut
// cells carry an 8-digit decimal code of their fluid corners (BSW BSE BNE BNW TSW TSE TNE TNW,
// most significant first); domain border cells carry negative codes.
enum CellType : int {
  SolidCell = 0,
  FluidCell = 11111111,
  T_Border = -2, B_Border = -3, N_Border = -4, S_Border = -5, E_Border = -6, W_Border = -7,
  TN_Edge = -8, TS_Edge = -9, TE_Edge = -10, TW_Edge = -11, BN_Edge = -12, BS_Edge = -13,
  BE_Edge = -14, BW_Edge = -15, NE_Edge = -16, NW_Edge = -17, SE_Edge = -18, SW_Edge = -19,
  TNE_Corner = -20, TNW_Corner = -21, TSE_Corner = -22, TSW_Corner = -23,
  BNE_Corner = -24, BNW_Corner = -25, BSE_Corner = -26, BSW_Corner = -27
};

// Cut-cell geometry of the base (pressure) grid.  Index num = x + N*y + N*M*z.
// thx/thy/thz: fluid fraction of the east/north/top face of each cell (reference theta).
struct Geometry {
  int N = 0, M = 0, K = 0;
  double x0 = 0, y0 = 0, z0 = 0;
  std::vector<double> dx, dy, dz;  // per-axis cell widths (tensor grid)
  std::vector<int> type;
  std::vector<double> thx, thy, thz, vof;
  std::vector<double> inv_mx, inv_my, inv_mz;  // inverse staggered masses on x/y/z faces
  std::vector<double> phi;                     // level set at each cell's upper corner (> 0 in the body)
  std::vector<double> alpha[3];                // fluid fraction of the x/y/z edges ending at that corner
  std::vector<int> side[3];                    // type of the east/north/top face (corner code)
  long ncells() const { return (long)N * M * K; }
  long idx(int x, int y, int z) const { return x + (long)N * (y + (long)M * z); }
  void init_grid(const std::vector<double>& dxs, const std::vector<double>& dys, const std::vector<double>& dzs,
                 double a, double b, double c);
  void build_masses();
};

// Region where cut cells may appear: x in [hs,hf), y in [vs,vf), all z planes.
struct Box {
  int hs, hf, vs, vf;
};

int boundary_type(int x, int y, int z, int N, int M, int K);
void corner_level_set(Geometry& g, const Box& box, const std::function<double(double, double, double)>& phi_fn);
void build_alpha(Geometry& g, const Box& box);
void build_types(Geometry& g, const Box& box);
void build_theta(Geometry& g, const Box& box);
void build_vof(Geometry& g);
void build_cut_geometry(Geometry& g, const Box& box, const std::function<double(double, double, double)>& phi_fn);
int vof_table_size();
std::vector<double> cell_centres(const std::vector<double>& widths, double origin);
Csr build_laplace(const Geometry& g, bool pin_outlet);
void body_free_vof(Geometry& g);
// Six-patch bicubic Bezier body (oracle_bezier.cpp): ellipsoid fit into 288 control values, and
// the signed distance (positive inside) at a point.
void bezier_ellipsoid(const double* center, const double* axes, double* ctrl);
double bezier_level_set(const double* ctrl, const double* center, const double* axes, double x, double y, double z);
// Pressure-equation right-hand side (oracle_rhs.cpp).  u_west: M*K inflow velocities (one per
// x-row); u_out: the x-face velocity for an outflow east boundary, or null for a fixed U_E.
std::vector<double> pressure_flux(const Geometry& g, const double* u_west, double U_E, double V_S, double V_N,
                                  double W_B, double W_T, const double* u_out);
void pressure_rhs(const Geometry& g, const double* u, const double* v, const double* w, const std::vector<double>& flux,
                  bool pin_outlet, double* b);
// Momentum (Helmholtz) operator of velocity component dir (0 u, 1 v, 2 w) on its staggered mesh,
// for a box without cut cells (oracle_helmholtz.cpp); dims_out receives the mesh dimensions.
Csr build_helmholtz(const Geometry& g, int dir, double nu, double coef, bool flow, int* dims_out);
// u = (G_x dp) inv_Mx + u* (and v, w), the velocity part of SetSolutions::correction
void velocity_correction(const Geometry& g, const double* dp, const double* us, const double* vs, const double* ws,
                         double* u, double* v, double* w);

}  // namespace lsxo
