// ORACLE — test infrastructure only (see oracle_solver.cpp header).
// Plain C entry points so tests/ can drive the restatement through ctypes.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <string>

#include "oracle_geometry.hpp"

using namespace lsxo;

namespace {
thread_local std::string g_err;

struct SolverH {
  Csr A;
  int N, M, K;
  int kind;  // 0 FGMRES, 1 BiCGStab
  int restart, max_it;
  double eps;
  int precond;  // 0 none, 1 multigrid cascade
  Multigrid mg;
  Identity id;
};

struct GeomH {
  Geometry g;
  Csr A;
};

Csr make_csr(long n, long ncols, const int* rowptr, const int* col, const double* val) {
  Csr A;
  A.rows = n;
  A.cols = ncols;
  A.rowptr.assign(rowptr, rowptr + n + 1);
  const long nnz = rowptr[n];
  A.col.assign(col, col + nnz);
  A.val.assign(val, val + nnz);
  return A;
}
}  // namespace

extern "C" {

const char* lsxo_last_error() { return g_err.c_str(); }

// ---- geometry ------------------------------------------------------------------------
void* lsxo_geom_create(int N, int M, int K, const double* dx, const double* dy, const double* dz, double x0,
                       double y0, double z0) {
  try {
    auto* h = new GeomH();
    h->g.init_grid(std::vector<double>(dx, dx + N), std::vector<double>(dy, dy + M), std::vector<double>(dz, dz + K), x0,
                   y0, z0);
    body_free_vof(h->g);
    h->g.build_masses();
    return h;
  } catch (std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void lsxo_geom_destroy(void* p) { delete static_cast<GeomH*>(p); }

// Overwrite types/theta/vof (used by tests to build synthetic cut geometries), then
// recompute masses.
int lsxo_geom_set_fields(void* p, const int* type, const double* thx, const double* thy, const double* thz,
                         const double* vof) {
  auto* h = static_cast<GeomH*>(p);
  const long n = h->g.ncells();
  if (type) std::memcpy(h->g.type.data(), type, n * sizeof(int));
  if (thx) std::memcpy(h->g.thx.data(), thx, n * sizeof(double));
  if (thy) std::memcpy(h->g.thy.data(), thy, n * sizeof(double));
  if (thz) std::memcpy(h->g.thz.data(), thz, n * sizeof(double));
  if (vof) std::memcpy(h->g.vof.data(), vof, n * sizeof(double));
  h->g.build_masses();
  return 0;
}

int lsxo_geom_get_fields(void* p, int* type, double* thx, double* thy, double* thz, double* vof) {
  auto* h = static_cast<GeomH*>(p);
  const long n = h->g.ncells();
  if (type) std::memcpy(type, h->g.type.data(), n * sizeof(int));
  if (thx) std::memcpy(thx, h->g.thx.data(), n * sizeof(double));
  if (thy) std::memcpy(thy, h->g.thy.data(), n * sizeof(double));
  if (thz) std::memcpy(thz, h->g.thz.data(), n * sizeof(double));
  if (vof) std::memcpy(vof, h->g.vof.data(), n * sizeof(double));
  return 0;
}

// Cut-cell geometry of a sphere (phi = r - |x - c|, flow_sphere_task_description.h:57-74)
// or an axis-aligned ellipsoid (phi = 1 - |(x - c)/a|, scaled by min(a) so it is a distance
// near the surface) inside the box x in [hs,hf), y in [vs,vf).
int lsxo_geom_build_body(void* p, int kind, double cx, double cy, double cz, double a, double b, double c, int hs,
                         int hf, int vs, int vf) {
  auto* h = static_cast<GeomH*>(p);
  try {
    Box box{hs, hf, vs, vf};
    if (kind == 0) {
      build_cut_geometry(h->g, box, [=](double x, double y, double z) {
        const double qx = x - cx, qy = y - cy, qz = z - cz;
        return a - std::sqrt(qx * qx + qy * qy + qz * qz);
      });
    } else {
      const double s = std::min(a, std::min(b, c));
      build_cut_geometry(h->g, box, [=](double x, double y, double z) {
        const double qx = (x - cx) / a, qy = (y - cy) / b, qz = (z - cz) / c;
        return s * (1.0 - std::sqrt(qx * qx + qy * qy + qz * qz));
      });
    }
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return 0;
}

// Six-patch Bezier body: fit, point distances and the cut-cell geometry of a control net.
void lsxo_bezier_ellipsoid(const double* center, const double* axes, double* ctrl) {
  bezier_ellipsoid(center, axes, ctrl);
}

void lsxo_bezier_phi(const double* ctrl, const double* center, const double* axes, const double* pts, long n,
                     double* out) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) out[i] = bezier_level_set(ctrl, center, axes, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
}

int lsxo_geom_build_bezier(void* p, const double* ctrl, const double* center, const double* axes, int hs, int hf,
                           int vs, int vf) {
  auto* h = static_cast<GeomH*>(p);
  try {
    std::vector<double> cp(ctrl, ctrl + 288);
    const double c0[3] = {center[0], center[1], center[2]}, a0[3] = {axes[0], axes[1], axes[2]};
    build_cut_geometry(h->g, Box{hs, hf, vs, vf}, [cp, c0, a0](double x, double y, double z) {
      return bezier_level_set(cp.data(), c0, a0, x, y, z);
    });
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return 0;
}

int lsxo_geom_get_level_set(void* p, double* phi, double* ax, double* ay, double* az, int* sx, int* sy, int* sz) {
  auto* h = static_cast<GeomH*>(p);
  const long n = h->g.ncells();
  if (h->g.phi.size() != (size_t)n) return -1;
  std::memcpy(phi, h->g.phi.data(), n * sizeof(double));
  std::memcpy(ax, h->g.alpha[0].data(), n * sizeof(double));
  std::memcpy(ay, h->g.alpha[1].data(), n * sizeof(double));
  std::memcpy(az, h->g.alpha[2].data(), n * sizeof(double));
  std::memcpy(sx, h->g.side[0].data(), n * sizeof(int));
  std::memcpy(sy, h->g.side[1].data(), n * sizeof(int));
  std::memcpy(sz, h->g.side[2].data(), n * sizeof(int));
  return 0;
}

int lsxo_vof_table_size() { return vof_table_size(); }

// Velocity correction u = (G dp) inv_M + u* (oracle_rhs.cpp).
int lsxo_geom_correct(void* p, const double* dp, const double* us, const double* vs, const double* ws, double* u,
                      double* v, double* w) {
  auto* h = static_cast<GeomH*>(p);
  velocity_correction(h->g, dp, us, vs, ws, u, v, w);
  return 0;
}

// Pressure right-hand side from predicted face velocities (oracle_rhs.cpp).
int lsxo_geom_rhs(void* p, const double* u, const double* v, const double* w, const double* u_west, double U_E,
                  double V_S, double V_N, double W_B, double W_T, const double* u_out, int pin, double* b,
                  double* flux_out) {
  auto* h = static_cast<GeomH*>(p);
  try {
    std::vector<double> flux = pressure_flux(h->g, u_west, U_E, V_S, V_N, W_B, W_T, u_out);
    pressure_rhs(h->g, u, v, w, flux, pin != 0, b);
    if (flux_out) std::memcpy(flux_out, flux.data(), flux.size() * sizeof(double));
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return 0;
}

long lsxo_geom_build_laplace(void* p, int pin) {
  auto* h = static_cast<GeomH*>(p);
  try {
    h->A = build_laplace(h->g, pin != 0);
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return (long)h->A.rowptr[h->A.rows];
}

// Helmholtz operator of component dir into the handle's CSR; returns nnz (dims: 3 ints)
long lsxo_geom_build_helmholtz(void* p, int dir, double nu, double coef, int flow, int* dims) {
  auto* h = static_cast<GeomH*>(p);
  try {
    h->A = build_helmholtz(h->g, dir, nu, coef, flow != 0, dims);
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return (long)h->A.rowptr[h->A.rows];
}

int lsxo_geom_get_csr(void* p, int* rowptr, int* col, double* val) {
  auto* h = static_cast<GeomH*>(p);
  std::memcpy(rowptr, h->A.rowptr.data(), (h->A.rows + 1) * sizeof(int));
  std::memcpy(col, h->A.col.data(), h->A.col.size() * sizeof(int));
  std::memcpy(val, h->A.val.data(), h->A.val.size() * sizeof(double));
  return 0;
}

// ---- solver ----------------------------------------------------------------------------
void* lsxo_solver_create(int N, int M, int K, const int* rowptr, const int* col, const double* val, int kind,
                         int restart, int max_it, double eps, int precond, int n_post, double omega) {
  try {
    auto h = std::make_unique<SolverH>();
    const long n = (long)N * M * K;
    h->A = make_csr(n, n, rowptr, col, val);
    h->N = N;
    h->M = M;
    h->K = K;
    h->kind = kind;
    h->restart = restart;
    h->max_it = max_it;
    h->eps = eps;
    h->precond = precond;
    h->id.n = n;
    if (precond == 1 || precond == 2) h->mg.set(h->A, N, M, K, n_post, omega, precond == 2);
    return h.release();
  } catch (std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void lsxo_solver_destroy(void* p) { delete static_cast<SolverH*>(p); }

void lsxo_solver_set_eps(void* p, double eps) { static_cast<SolverH*>(p)->eps = eps; }

int lsxo_solver_solve(void* p, const double* b, double* x, double* norm_r) {
  auto* h = static_cast<SolverH*>(p);
  Preconditioner& pr = h->precond != 0 ? (Preconditioner&)h->mg : (Preconditioner&)h->id;
  if (h->kind == 0) return fgmres_solve(h->A, pr, b, x, h->eps, h->max_it, h->restart, norm_r);
  return bicgstab_solve(h->A, pr, b, x, h->eps, h->max_it, norm_r);
}

void lsxo_mg_apply(void* p, const double* b, double* x) { static_cast<SolverH*>(p)->mg.apply(b, x); }

int lsxo_mg_levels(void* p) { return (int)static_cast<SolverH*>(p)->mg.levels.size(); }

long lsxo_mg_level_info(void* p, int l, int* dims) {
  auto& L = static_cast<SolverH*>(p)->mg.levels[l];
  dims[0] = L.N;
  dims[1] = L.M;
  dims[2] = L.K;
  return (long)L.A.rowptr[L.A.rows];
}

int lsxo_mg_level_csr(void* p, int l, int* rowptr, int* col, double* val) {
  auto& A = static_cast<SolverH*>(p)->mg.levels[l].A;
  std::memcpy(rowptr, A.rowptr.data(), (A.rows + 1) * sizeof(int));
  std::memcpy(col, A.col.data(), A.col.size() * sizeof(int));
  std::memcpy(val, A.val.data(), A.val.size() * sizeof(double));
  return 0;
}

// ADLJ factors of level l, direction dir (0 x, 1 y, 2 z): c (n-1), d (n), e (n-1)
int lsxo_mg_level_adlj(void* p, int l, int dir, double* c, double* d, double* e) {
  auto& L = static_cast<SolverH*>(p)->mg.levels[l];
  std::memcpy(c, L.adlj.c[dir].data(), L.adlj.c[dir].size() * sizeof(double));
  std::memcpy(d, L.adlj.dg[dir].data(), L.adlj.dg[dir].size() * sizeof(double));
  std::memcpy(e, L.adlj.e[dir].data(), L.adlj.e[dir].size() * sizeof(double));
  return 0;
}

// Free functions on a caller CSR
void lsxo_spmv(long n, const int* rowptr, const int* col, const double* val, const double* x, double* y) {
  Csr A = make_csr(n, n, rowptr, col, val);
  spmv(A, x, y);
}

}  // extern "C"
