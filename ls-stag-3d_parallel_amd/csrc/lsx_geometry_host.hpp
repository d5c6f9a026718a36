// Host side of the GPU cut-cell geometry: grid description, device fields, build pipeline.
#pragma once

#include <vector>

#include "lsx_geometry.hpp"
#include "lsx_solver.hpp"

namespace lsx {

// six-patch bicubic Bezier fit of an ellipsoid into ctrl[288] (lsx_bezier.cpp)
void bezier_ellipsoid(const double* center, const double* axes, double* ctrl);

// Flow fields of a pressure right-hand side staged in device memory (copies when given on the host).
struct FlowStage {
  DevBuf<double> u, v, w, west, out;
  geo::RhsArgs args{};
};

// Checkpoint in the reference's binary layout (lsx_state.cpp); cur / prev: u, v, w, p in `mem`
// space.  state_bytes: the file size for a grid and body count.
long state_bytes(int N, int M, int K, int n_bodies);
void state_save(const char* path, int it, double time, int n_bodies, const double* body_pos, const double* body_vel,
                const double* loads, int N, int M, int K, const double* const cur[4], const double* const prev[4],
                int mem);
void state_load(const char* path, int* it, double* time, int n_bodies, double* body_pos, double* body_vel,
                double* loads, int N, int M, int K, double* const cur[4], double* const prev[4], int mem);
// Tecplot point file of the solution (lsx_state.cpp), host arrays
void write_tecplot(const char* path, int N, int M, int K, const std::vector<double>& xc, const std::vector<double>& yc,
                   const std::vector<double>& zc, const std::vector<double>& wx, const std::vector<double>& wy,
                   const std::vector<double>& wz, const std::vector<int>& type, const std::vector<double>& thx,
                   const std::vector<double>& thy, const std::vector<double>& thz, const std::vector<double>& u,
                   const std::vector<double>& v, const std::vector<double>& w, const std::vector<double>& p,
                   const std::vector<double>& uw, const lsx_walls& bc, bool outflow);

class Geometry {
 public:
  // stage u*, v*, w*, u_west, u_out (`mem` space) for rhs-dependent gradients
  void stage_flow(const double* u, const double* v, const double* w, const double* u_west, const double* u_out,
                  const lsx_flow_bc& bc, int mem, FlowStage& st) const;
  Geometry(int N, int M, int K, const double* dx, const double* dy, const double* dz, double x0, double y0, double z0,
           hipStream_t stream);
  ~Geometry();
  // Build every field for `body` with cut cells allowed in box [hs,hf) x [vs,vf) x all z.
  void build(const geo::Body& body, geo::Box box, const double* host_ctrl = nullptr);
  // Copy fields to host (any pointer may be null).
  void fields(int* type, double* thx, double* thy, double* thz, double* vof, double* phi) const;
  // 7 coefficient planes [b,s,w,p,e,n,t] of the pressure operator (device pointer, 7*n doubles).
  void laplace7(int pin, double* coef_dev, hipStream_t s) const;
  // the same for the momentum operator of component dir on its staggered mesh (7 * that mesh's
  // cell count doubles); the geometry must have no cut cells
  void helmholtz7(int dir, double nu, double coef_time, int flow, double* coef_dev, hipStream_t s) const;
  // b = D u + flux (pressure right-hand side); arrays in `mem` space (LSX_MEM_HOST / DEVICE)
  void pressure_rhs(const double* u, const double* v, const double* w, const double* u_west, const double* u_out,
                    const lsx_flow_bc& bc, double* b, int mem);
  // u = (G dp) inv_M + u*, v, w likewise; p += dp * coef when p is given; `mem` space arrays
  void correct(const double* dp, const double* us, const double* vs, const double* ws, double* u, double* v, double* w,
               double* p, double coef, int mem);

  // pressure drag J and (dJdp non-null) its gradient; p and dJdp in `mem` space
  double drag(const double* p, double* dJdp, int mem);
  // the same J (written to the device pointer J) and dJdp with no device allocation and no host
  // synchronisation; p, J, dJdp device pointers; partial sums in the library's scratch array
  static constexpr long DRAG_NOALLOC_MAX_CELLS = (1L << 31) - 1;
  void drag_noalloc(const double* p, double* J, double* dJdp);
  // Shape gradient of J at the solution p with the adjoint lambda = A^{-T} dJ/dp: dF/dphi per
  // level-set corner (dJdphi, `mem` space, may be null) and, for a Bezier body, dJ/dctrl
  // (288 host doubles, may be null).  `rhs` null: the right-hand side does not depend on the
  // geometry; otherwise b = D u* + flux from these fields (device pointers).  h_rel: level-set
  // step as a fraction of the smallest cell width.
  void shape_gradient(const double* p, const double* lam, int objective, const geo::RhsArgs* rhs, int pin, double h_rel,
                      double* dJdphi, double* dJdctrl, int mem);
  // Tecplot point file of u, v, w, p (`mem` space) on this grid; u_west: M*K inflow velocities
  // or null
  void write_tecplot(const char* path, const double* u, const double* v, const double* w, const double* p,
                     const double* u_west, const lsx_walls& walls, bool outflow, int mem) const;
  hipStream_t get_stream() const { return stream; }
  long version() const { return build_count; }

  int N, M, K;
  long n() const { return (long)N * M * K; }
  bool built = false;
  long build_count = 0;

 private:
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::vector<double> hdx, hdy, hdz;
  double x0, y0, z0;
  DevBuf<double> dx, dy, dz, cx, cy, cz;
  DevBuf<double> phi, ax, ay, az, tx, ty, tz, vof;
  DevBuf<double> phi_raw, osc_tmp;  // level set before the oscillation fix; scratch of the fix
  DevBuf<int> dchanged;
  int osc_passes = 0;
  DevBuf<int> type, sx, sy, sz;
  DevBuf<unsigned char> flag;
  DevBuf<double> bez;  // Bezier control values (device)
  geo::Body body_{};
  geo::Box box_{};
  // shape-gradient work: two perturbed geometries, dF per cell, g per corner, closest points
  struct Work {
    DevBuf<double> raw, phi, tmp, ax, ay, az, tx, ty, tz, vof;
    DevBuf<int> type, sx, sy, sz;
    DevBuf<unsigned char> flag;
    void alloc(long n);
    geo::GeoView view() const { return geo::GeoView{type.p, tx.p, ty.p, tz.p, vof.p}; }
  };
  Work wp, wm;
  DevBuf<double> sg_dF, sg_g, sg_uv, sg_w, sg_partial, sg_out;
  DevBuf<int> sg_f;
  void rebuild_from(Work& w);  // fields of w from w.raw (oscillation fix onwards)
};

}  // namespace lsx
