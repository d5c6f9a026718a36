// Host side of the GPU cut-cell geometry: grid description, device fields, build pipeline.
#pragma once

#include <vector>

#include "lsx_geometry.hpp"
#include "lsx_solver.hpp"

namespace lsx {

// six-patch bicubic Bezier fit of an ellipsoid into ctrl[288] (lsx_bezier.cpp)
void bezier_ellipsoid(const double* center, const double* axes, double* ctrl);

class Geometry {
 public:
  Geometry(int N, int M, int K, const double* dx, const double* dy, const double* dz, double x0, double y0, double z0,
           hipStream_t stream);
  ~Geometry();
  // Build every field for `body` with cut cells allowed in box [hs,hf) x [vs,vf) x all z.
  void build(const geo::Body& body, geo::Box box, const double* host_ctrl = nullptr);
  // Copy fields to host (any pointer may be null).
  void fields(int* type, double* thx, double* thy, double* thz, double* vof, double* phi) const;
  // 7 coefficient planes [b,s,w,p,e,n,t] of the pressure operator (device pointer, 7*n doubles).
  void laplace7(int pin, double* coef_dev, hipStream_t s) const;
  // b = D u + flux (pressure right-hand side); arrays in `mem` space (LSX_MEM_HOST / DEVICE)
  void pressure_rhs(const double* u, const double* v, const double* w, const double* u_west, const double* u_out,
                    const lsx_flow_bc& bc, double* b, int mem);
  // u = (G dp) inv_M + u*, v, w likewise; p += dp * coef when p is given; `mem` space arrays
  void correct(const double* dp, const double* us, const double* vs, const double* ws, double* u, double* v, double* w,
               double* p, double coef, int mem);

  int N, M, K;
  long n() const { return (long)N * M * K; }
  bool built = false;

 private:
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::vector<double> hdx, hdy, hdz;
  double x0, y0, z0;
  DevBuf<double> dx, dy, dz, cx, cy, cz;
  DevBuf<double> phi, ax, ay, az, tx, ty, tz, vof;
  DevBuf<int> type, sx, sy, sz;
  DevBuf<unsigned char> flag;
  DevBuf<double> bez;  // Bezier control values (device)
};

}  // namespace lsx
