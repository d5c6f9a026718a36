// State formats of the projection step: the reference's binary checkpoint and its Tecplot
// point-format solution file.
//
// Checkpoint (SetSolutions::saveState / load, set_solutions.h:317-339 / 247-269): int32 it,
// double time, per immersed body its centre and centre velocity (3 + 3 doubles), per body the
// task's loads F_aer_x, F_aer_y, F_aer_z, moment_aer (BaseTaskDescription::saveState,
// base_task_description.h:151-160), then the current solution u, v, w, p and the previous one,
// each array written raw (Solution::saveState, solution.h:76-81; ArrayCPU::saveState writes
// size * sizeof(double) bytes, array.h:60).  Native byte order, no padding.
//
// Tecplot file (SolTecplotPrinter::printField, sol_tecplot_printer.h:254-347, header from
// BaseTecplotPrinter::init, base_tecplot_printer.h:51-53): an (N+2) x (M+2) x (K+2) point zone,
// x fastest, variables x y z P SPD_x SPD_y SPD_z SPD.  Every row is a wall point, the N cell
// centres and a wall point; every plane a wall row, the M cell rows and a wall row; the planes a
// wall plane, the K cell planes and a wall plane.  Face velocities are averaged to the cell
// centres (cut cells: weighted by their face fractions); wall points take the wall velocities.
// The point order and the values follow the reference's traversal, including the values it
// carries along a row: an edge row (first and last row of a plane) uses the z velocity of the
// row's first cell for every cell of the row, and the pressure of a wall point is that of the
// nearest cell printed before it.  Numbers are written with the default stream format (six
// significant digits), as the reference's std::ofstream does.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/lsx.h"
#include "lsx_geometry_host.hpp"
#include "lsx_solver.hpp"

namespace lsx {

namespace {

constexpr double EQ_DOUBLE = 1e-5;  // ls_stag_real_constants.h

// host copy of n doubles from `mem` space (null stays null)
std::vector<double> fetch(const double* p, long n, int mem) {
  std::vector<double> h;
  if (!p) return h;
  h.resize(n);
  if (mem == LSX_MEM_DEVICE) {
    LSX_HIP(hipMemcpy(h.data(), p, n * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::copy(p, p + n, h.begin());
  }
  return h;
}

void store(double* dst, const std::vector<double>& h, int mem) {
  if (mem == LSX_MEM_DEVICE) {
    LSX_HIP(hipMemcpy(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::copy(h.begin(), h.end(), dst);
  }
}

struct Sizes {
  long nu, nv, nw, np;
};

Sizes face_sizes(int N, int M, int K) {
  if (N < 2 || M < 2 || K < 2) throw Error(LSX_ERR_ARG, "grid dimensions must be >= 2");
  return Sizes{(long)(N - 1) * M * K, (long)N * (M - 1) * K, (long)N * M * (K - 1), (long)N * M * K};
}

}  // namespace

long state_bytes(int N, int M, int K, int n_bodies) {
  const Sizes s = face_sizes(N, M, K);
  return 4 + 8 + (long)n_bodies * (6 + 4) * 8 + 2L * (s.nu + s.nv + s.nw + s.np) * 8;
}

void state_save(const char* path, int it, double time, int n_bodies, const double* body_pos, const double* body_vel,
                const double* loads, int N, int M, int K, const double* const cur[4], const double* const prev[4],
                int mem) {
  if (!path) throw Error(LSX_ERR_ARG, "null path");
  if (n_bodies < 0 || (n_bodies > 0 && (!body_pos || !body_vel || !loads)))
    throw Error(LSX_ERR_ARG, "body arrays missing");
  const Sizes s = face_sizes(N, M, K);
  const long len[4] = {s.nu, s.nv, s.nw, s.np};
  for (int q = 0; q < 4; ++q)
    if (!cur[q] || !prev[q]) throw Error(LSX_ERR_ARG, "null solution array");
  std::ofstream f(path, std::ios::binary);
  if (!f) throw Error(LSX_ERR_ARG, std::string("cannot open ") + path);
  const int32_t it32 = it;
  f.write(reinterpret_cast<const char*>(&it32), sizeof(it32));
  f.write(reinterpret_cast<const char*>(&time), sizeof(time));
  for (int b = 0; b < n_bodies; ++b) {
    f.write(reinterpret_cast<const char*>(body_pos + 3 * b), 3 * sizeof(double));
    f.write(reinterpret_cast<const char*>(body_vel + 3 * b), 3 * sizeof(double));
  }
  f.write(reinterpret_cast<const char*>(loads), (long)n_bodies * 4 * sizeof(double));
  for (const double* const* set : {cur, prev})
    for (int q = 0; q < 4; ++q) {
      const std::vector<double> h = fetch(set[q], len[q], mem);
      f.write(reinterpret_cast<const char*>(h.data()), len[q] * sizeof(double));
    }
  if (!f) throw Error(LSX_ERR_ARG, std::string("write failed: ") + path);
}

void state_load(const char* path, int* it, double* time, int n_bodies, double* body_pos, double* body_vel,
                double* loads, int N, int M, int K, double* const cur[4], double* const prev[4], int mem) {
  if (!path || !it || !time) throw Error(LSX_ERR_ARG, "null argument");
  if (n_bodies < 0 || (n_bodies > 0 && (!body_pos || !body_vel || !loads)))
    throw Error(LSX_ERR_ARG, "body arrays missing");
  const Sizes s = face_sizes(N, M, K);
  const long len[4] = {s.nu, s.nv, s.nw, s.np};
  for (int q = 0; q < 4; ++q)
    if (!cur[q] || !prev[q]) throw Error(LSX_ERR_ARG, "null solution array");
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw Error(LSX_ERR_ARG, std::string("cannot open ") + path);
  const long size = (long)f.tellg();
  if (size != state_bytes(N, M, K, n_bodies))
    throw Error(LSX_ERR_ARG, "checkpoint size does not match the grid and body count (" + std::to_string(size) +
                                 " bytes, expected " + std::to_string(state_bytes(N, M, K, n_bodies)) + ")");
  f.seekg(0);
  int32_t it32 = 0;
  f.read(reinterpret_cast<char*>(&it32), sizeof(it32));
  f.read(reinterpret_cast<char*>(time), sizeof(double));
  *it = it32;
  for (int b = 0; b < n_bodies; ++b) {
    f.read(reinterpret_cast<char*>(body_pos + 3 * b), 3 * sizeof(double));
    f.read(reinterpret_cast<char*>(body_vel + 3 * b), 3 * sizeof(double));
  }
  f.read(reinterpret_cast<char*>(loads), (long)n_bodies * 4 * sizeof(double));
  for (double* const* set : {cur, prev})
    for (int q = 0; q < 4; ++q) {
      std::vector<double> h(len[q]);
      f.read(reinterpret_cast<char*>(h.data()), len[q] * sizeof(double));
      if (!f) throw Error(LSX_ERR_ARG, std::string("read failed: ") + path);
      store(set[q], h, mem);
    }
}

// ---- Tecplot point file ------------------------------------------------------------------------

namespace {

// The traversal state of the writer: the base-cell counter c (with the centre and width arrays
// addressed through it), the x-face counter cu, the y-face counter cv and the inflow-profile
// counter cw, advanced exactly as the reference printer advances num, num2, num3 and numW.
struct TecWriter {
  std::ofstream& out;
  int N, M, K;
  long NM;
  const std::vector<double>&xc, &yc, &zc, &wx, &wy;
  const std::vector<int>& type;
  const std::vector<double>&thx, &thy, &thz;
  const std::vector<double>&u, &v, &w, &p;
  const std::vector<double>& uw;  // inflow profile (empty: none)
  lsx_walls bc;
  bool outflow;

  long c = 0, cu = 0, cv = 0, cw = 0;
  double X = 0, Y = 0, Z = 0, P = 0, Ux = 0, Uy = 0, Uz = 0;

  double cx() const { return xc[c % N]; }
  double cy() const { return yc[(c / N) % M]; }
  double cz() const { return zc[c / NM]; }
  double dxw(long i) const { return wx[i % N]; }
  double dyw(long i) const { return wy[(i / N) % M]; }

  void node() {
    const double s = std::sqrt(Ux * Ux + Uy * Uy + Uz * Uz);
    out << X << " " << Y << " " << Z << " " << P << " " << Ux << " " << Uy << " " << Uz << " " << s << std::endl;
  }
  static double weighted(double ta, double va, double tb, double vb) {
    const double sum = ta + tb;
    return (sum > EQ_DOUBLE) ? ((ta * va + tb * vb) / sum) : 0.0;
  }
  // a row of wall points: the wall corner, the N cell centres, the far corner
  void wall_row(double V, double U, double W) {
    X = 0.0;
    P = p[c];
    Ux = U;
    Uy = V;
    Uz = W;
    node();
    for (int i = 0; i < N; ++i, ++c) {
      X = cx();
      P = p[c];
      node();
    }
    X += 0.5 * dxw(c - 1);
    node();
  }
  // a plane of wall points (reference PrintWallPlane)
  void wall_plane(double V, double U, double W) {
    Y = 0.0;
    wall_row(V, U, W);
    c -= N;
    for (int j = 0; j < M; ++j) {
      X = 0.0;
      Y = cy();
      wall_row(V, U, W);
    }
    c -= N;
    Y += 0.5 * dyw(c);
    wall_row(V, U, W);
  }
  // the west wall point and the first cell of a row
  void west(double V, double W, double W2) {
    Y = cy();
    X = 0.0;
    P = p[c];
    Ux = uw.empty() ? 0.0 : uw[cw];
    Uy = bc.V_W;
    Uz = bc.W_W;
    node();
    X = cx();
    Ux = uw.empty() ? 0.0 : 0.5 * (u[cu] + uw[cw]);
    Uy = 0.5 * (v[cv] + V);
    Uz = 0.5 * (W2 + W);
    node();
    ++c, ++cw, ++cv, ++cu;
  }
  // the last cell of a row and the east wall point
  void east(double V, double W, double W2) {
    X = cx();
    P = p[c];
    Ux = outflow ? u[cu - 1] : 0.5 * (u[cu - 1] + bc.U_E);
    Uy = 0.5 * (v[cv] + V);
    Uz = 0.5 * (W2 + W);
    node();
    X += 0.5 * dxw(c);
    if (!outflow) {
      Ux = bc.U_E;
      Uy = bc.V_E;
      Uz = bc.W_E;
    }
    node();
    ++c, ++cv;
  }
  // first or last row of a plane: every point takes V (wall) and the pair W, W2 of the row
  void edge_row(double V, double W, double W2) {
    west(V, W, W2);
    for (int i = 1; i < N - 1; ++i, ++c, ++cu, ++cv) {
      X = cx();
      P = p[c];
      Ux = 0.5 * (u[cu] + u[cu - 1]);
      Uy = 0.5 * (v[cv] + V);
      Uz = 0.5 * (W2 + W);
      node();
    }
    east(V, W, W2);
  }
  // interior cell of a row; wlo / whi: the z velocities below and above the cell
  void cell(double wlo, double whi) {
    X = cx();
    P = p[c];
    const int t = type[c];
    if (t == 0) {  // SolidCell
      Ux = Uy = Uz = 0.0;
    } else if (t == 11111111) {  // FluidCell
      Ux = 0.5 * (u[cu] + u[cu - 1]);
      Uy = 0.5 * (v[cv] + v[cv - N]);
      Uz = 0.5 * (whi + wlo);
    } else {
      Ux = weighted(thx[c], u[cu], thx[c - 1], u[cu - 1]);
      Uy = weighted(thy[c], v[cv], thy[c - N], v[cv - N]);
      Uz = 0.0;  // set by the caller (the z weights differ between the layers)
    }
    node_z_pending = (t != 0 && t != 11111111);
  }
  bool node_z_pending = false;

  // plane k of cells: kind 0 the first plane (bottom wall velocity W_B below), 1 an interior
  // plane, 2 the last plane (W_T above); as_wall: the plane printed at the wall position
  void cell_plane(int kind, int k, bool as_wall) {
    if (kind == 0) {
      c = 0, cu = 0, cv = 0, cw = 0;
      Y = 0.0;
      Z = as_wall ? 0.0 : cz();
    } else if (kind == 2 && as_wall) {
      const long nt = NM * (K - 1);
      Y = 0.0;
      Z = zc[nt / NM] + 0.5 * wz_last;
    } else {
      Y = 0.0;
      Z = cz();
    }
    wall_row(bc.V_S, bc.U_S, bc.W_S);
    c -= N;
    const double WB = bc.W_B, WT = bc.W_T;
    auto wbelow = [&]() { return kind == 0 ? WB : w[c - NM]; };
    auto wabove = [&]() { return kind == 2 ? WT : w[c]; };
    {
      const double W = wbelow(), W2 = wabove();
      edge_row(bc.V_S, W, W2);
    }
    for (int j = 1; j < M - 1; ++j) {
      {
        const double V = v[cv - N], W = wbelow(), W2 = wabove();
        west(V, W, W2);
      }
      cu = c - (j + (long)k * M);
      cv = c - (long)k * N;
      for (int i = 1; i < N - 1; ++i, ++c, ++cu, ++cv) {
        const double wlo = wbelow(), whi = wabove();
        cell(wlo, whi);
        if (node_z_pending) {
          if (kind == 0)
            Uz = weighted(thz[c], w[c], thz[c], WB);
          else if (kind == 2)
            Uz = weighted(thz[c], w[c - NM], thz[c], WT);
          else
            Uz = weighted(thz[c], w[c], thz[c - NM], w[c - NM]);
        }
        node();
      }
      {
        const double V = v[cv - N], W = wbelow(), W2 = wabove();
        east(V, W, W2);
      }
    }
    cv -= N;
    {
      const double W = wbelow(), W2 = wabove();
      edge_row(bc.V_N, W, W2);
    }
    c -= N;
    X = 0.0;
    Y += 0.5 * dyw(c);
    wall_row(bc.V_N, bc.U_N, bc.W_N);
  }
  double wz_last = 0.0;
};

}  // namespace

void write_tecplot(const char* path, int N, int M, int K, const std::vector<double>& xc, const std::vector<double>& yc,
                   const std::vector<double>& zc, const std::vector<double>& wx, const std::vector<double>& wy,
                   const std::vector<double>& wz, const std::vector<int>& type, const std::vector<double>& thx,
                   const std::vector<double>& thy, const std::vector<double>& thz, const std::vector<double>& u,
                   const std::vector<double>& v, const std::vector<double>& w, const std::vector<double>& p,
                   const std::vector<double>& uw, const lsx_walls& bc, bool outflow) {
  if (N < 3 || M < 3 || K < 3) throw Error(LSX_ERR_ARG, "the Tecplot writer needs at least 3 cells per axis");
  std::ofstream f(path);
  if (!f) throw Error(LSX_ERR_ARG, std::string("cannot open ") + path);
  f << "TITLE     = \"Velocity Vectors\" " << std::endl;
  f << "VARIABLES = \"x\" " << std::endl;
  for (const char* name : {"\"y\"", "\"z\"", "\"P\"", "\"SPD_x\"", "\"SPD_y\"", "\"SPD_z\"", "\"SPD\""})
    f << name << std::endl;
  f << "ZONE T=\"Data\" " << std::endl;
  f << "I=" << N + 2 << ", J=" << M + 2 << ", K=" << K + 2 << ", F=POINT" << std::endl;
  TecWriter tw{f, N, M, K, (long)N * M, xc, yc, zc, wx, wy, type, thx, thy, thz, u, v, w, p, uw, bc, outflow};
  tw.wz_last = wz[K - 1];
  // bottom: the wall plane (flow tasks: a copy of the first cell plane at z = 0)
  tw.X = tw.Y = tw.Z = 0.0;
  tw.c = 0;
  if (outflow) {
    tw.cell_plane(0, 0, true);
  } else {
    tw.wall_plane(bc.V_B, bc.U_B, bc.W_B);
  }
  tw.cell_plane(0, 0, false);
  for (int k = 1; k < K - 1; ++k) tw.cell_plane(1, k, false);
  const long s_c = tw.c, s_cu = tw.cu, s_cv = tw.cv, s_cw = tw.cw;
  tw.cell_plane(2, K - 1, false);
  tw.c = s_c, tw.cu = s_cu, tw.cv = s_cv, tw.cw = s_cw;
  if (outflow) {
    tw.cell_plane(2, K - 1, true);
  } else {
    const long nt = (long)N * M * (K - 1);
    tw.Z = zc[nt / ((long)N * M)] + 0.5 * wz[K - 1];
    tw.wall_plane(bc.V_T, bc.U_T, bc.W_T);
  }
  if (!f) throw Error(LSX_ERR_ARG, std::string("write failed: ") + path);
}

}  // namespace lsx
