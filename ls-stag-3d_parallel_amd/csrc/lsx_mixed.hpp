// Mixed-precision multigrid preconditioner (FP32 V-cycle, FP16 coarse faces): device kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "lsx_kernels.hpp"

namespace lsx {
namespace mix {

// One level of the V-cycle.  Faces are stored divided by fscale (FP16 levels) or as is (FP32).
struct Level {
  Dims d{0, 0, 0};
  bool f16 = false;
  const float *f_e = nullptr, *f_n = nullptr, *f_t = nullptr;
  const _Float16 *h_e = nullptr, *h_n = nullptr, *h_t = nullptr;
  const float* s = nullptr;  // 1 / diagonal (0 for solid / closed cells, 1 for pinned cells)
  float fscale = 1.0f;
  int pin0 = -1, pin1 = -1;
};

// setup (FP64 temporaries)
void faces_from7(const double* c7, Dims d, int pin0, int pin1, double* fe, double* fn, double* ft, double* dg,
                 hipStream_t s);
void aggregate(const double* fe, const double* fn, const double* ft, Dims df, double* ce, double* cn, double* ct,
               Dims dc, hipStream_t s);
void diag_from_faces(const double* fe, const double* fn, const double* ft, Dims d, double* dg, hipStream_t s);
void face_absmax(const double* fe, const double* fn, const double* ft, long n, double* partial, double* out,
                 hipStream_t s);
void inv_diag(const double* dg, long n, const double* absmax, int pin0, int pin1, float* sv, hipStream_t s);
void store_faces_f32(const double* fe, const double* fn, const double* ft, long n, double inv_scale, float* ae,
                     float* an, float* at, hipStream_t s);
void store_faces_f16(const double* fe, const double* fn, const double* ft, long n, double inv_scale, _Float16* ae,
                     _Float16* an, _Float16* at, hipStream_t s);
void to27(const double* fe, const double* fn, const double* ft, const double* dg, Dims d, int pin0, int pin1,
          double* c27, hipStream_t s);

// V-cycle pieces
void rbgs(const Level& L, float* x, const float* b, int color, int zero_guess, hipStream_t s);
void resid_restrict(const Level& L, const float* x, const float* b, float* bc, hipStream_t s);
void prolong_add(const float* xc, Dims dc, float* xf, Dims df, hipStream_t s);
void f64_to_f32(const double* a, float* b, long n, hipStream_t s);
void f32_to_f64(const float* a, double* b, long n, hipStream_t s);

}  // namespace mix
}  // namespace lsx
