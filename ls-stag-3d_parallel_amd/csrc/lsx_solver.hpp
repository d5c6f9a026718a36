// Host-side solver: owns the device multigrid hierarchy and Krylov workspace of one handle.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lsx.h"
#include "lsx_kernels.hpp"
#include "lsx_vcycle.hpp"

namespace lsx {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define LSX_HIP(expr)                                                                                   \
  do {                                                                                                  \
    hipError_t e_ = (expr);                                                                             \
    if (e_ != hipSuccess)                                                                               \
      throw ::lsx::Error(LSX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_) + " at " +      \
                                          __FILE__ + ":" + std::to_string(__LINE__));                   \
  } while (0)

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count) LSX_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

struct MGLevel {
  Dims d{0, 0, 0};
  int npts = 7;
  DevBuf<double> coef;          // npts planes
  DevBuf<double> dinv[3], eu[3];  // line LU factors per direction
  DevBuf<double> r, ch;         // level right-hand side / correction
};

class Geometry;

class Solver {
 public:
  Solver(const lsx_options& opt, hipStream_t stream);
  ~Solver();

  void set_csr(int N, int M, int K, const int* rowptr, const int* col, const double* val, long nnz, int mem);
  void set_geometry(const Geometry& g, int pin);
  // the cascade with opt.levels == 1: no coarse levels, the preconditioner is the line sweeps of
  // the fine level from a zero correction (grids of any dimensions)
  bool single_level() const { return opt.mg != LSX_MG_VCYCLE && opt.levels == 1; }
  // the pressure operator assembled on the device from the caller's mesh fields (widths, cell
  // types, face fractions, cut volumes; `mem` space), then the hierarchy
  // the momentum operator of velocity component dir of `g` (a box without cut cells) on its
  // staggered mesh, assembled on the device, then the hierarchy (the one-level cascade for its odd
  // dimension)
  void set_helmholtz(const Geometry& g, int dir, double nu, double coef_time, int flow);
  void set_fields(int N, int M, int K, const double* dx, const double* dy, const double* dz, const int* type,
                  const double* thx, const double* thy, const double* thz, const double* vof, int pin, int mem);
  // the geometry (and its build and pin choice) the current operator was assembled from
  bool operator_from(const Geometry* g, long version, int pin) const {
    return ready && src_geo == g && src_version == version && src_pin == pin;
  }
  void update_values(const double* val, int mem);
  void rebuild();
  void solve(const double* b, double* x, int mem, int* it, double* norm_r);
  void precondition(const double* v, double* z, int mem);
  // adjoint: lambda = A^{-T} g by the configured Krylov method on A^T, right-preconditioned by
  // the transposed V-cycle; transposed operator and preconditioner applications
  void solve_adjoint(const double* g, double* lambda, int mem, int* it, double* norm_r);
  void precondition_transpose(const double* v, double* z, int mem);
  // `sweeps` sweeps of the V-cycle's smoother on the finest level, x updated in place
  void smooth_fine(const double* b, double* x, int sweeps, int zero_guess, int mem);
  void apply_operator_transpose(const double* x, double* y, int mem);
  void apply_operator(int level, const double* x, double* y, int mem);
  long level_dims(int level, int* N, int* M, int* K) const;
  int level_stencil(int level, double* host, long capacity) const;
  void operator_diagonal(double* host) const;
  int fine_storage() const {
    if (!ready) throw Error(LSX_ERR_STATE, "fine_storage before the matrix is set");
    return sym_ok ? 4 : 7;
  }
  void synchronize();

  void set_timing(bool on);
  void kernel_stats(int k, long* launches, double* total_ms, double* bytes);

  lsx_options opt;
  int it = 0;
  double norm_r = 0.0;
  int num_levels() const { return (int)levels.size(); }

 private:
  void setup_levels(int N, int M, int K);
  int coarsenings(int N, int M, int K) const;
  void check_dims(int N, int M, int K) const;
  void build_hierarchy();
  bool from_csr = false;
  const Geometry* src_geo = nullptr;
  long src_version = -1;
  int src_pin = -1;
  void alloc_krylov();
  void mg_apply(const double* v_dev, double* z_dev);
  int fgmres(const double* b, double* x);
  int bicgstab(const double* b, double* x);
  double reduce_to_host(const double* dev_scalar);
  double dot(const double* a, const double* b, long n, double* out_dev);

  hipStream_t stream = nullptr;
  bool own_stream = false;
  Dims d0{0, 0, 0};
  long nnz0 = 0;
  DevBuf<int> rowptr, col;
  DevBuf<double> val;
  DevBuf<int> err;
  std::vector<MGLevel> levels;  // levels.back() is solved by the dense LU
  DevBuf<double> lu_scale, lu_L, lu_U, lu_col;
  DevBuf<int> lu_perm;
  // dense inverse of the coarsest level of a capped V-cycle hierarchy (column- and row-major)
  DevBuf<double> dense_c, dense_r, dense_part;
  template <class T>
  void coarse_solve(bool transposed);
  DevBuf<double> cwork;   // correction increment c (fine size, reused on every level)
  DevBuf<double> partial, scalars;
  DevBuf<double> mpartial, cgs_h;  // block Gram-Schmidt partial sums and coefficients
  // Krylov workspace
  DevBuf<double> V, Z, w;                             // FGMRES
  DevBuf<double> bnu, bt, by, bz, br, bp, bs, brr;   // BiCGStab
  DevBuf<double> hb, hx;                              // staging for host inputs
  std::vector<double> H, cs, sn, y;
  bool ready = false;
  bool values_dirty = false;  // update_values() called, rebuild() not yet
  void require_ready(const char* what) const;

  // fine operator in symmetric four-plane storage (when the operator allows it)
  DevBuf<double> s4;
  SymExc sym_ex;
  bool sym_ok = false;
  std::vector<int> pins;  // identity-like rows of the fine operator (pinned cells)
  void sym_setup();
  void fine_apply(const double* x, const double* b, double* y, int mode);  // StencilMode on level 0
  bool trans = false;  // Krylov loops and preconditioner work on A^T (adjoint solves)
  void require_transpose() const;
  struct TransScope {  // sets `trans` for one adjoint operation
    Solver& s;
    explicit TransScope(Solver& sv) : s(sv) { s.trans = true; }
    ~TransScope() { s.trans = false; }
  };

  // V-cycle hierarchy (FP64, or FP32 values with FP32/FP16 faces in mixed precision)
  struct VStore {
    vc::Level L;
    DevBuf<double> de, dn, dt, ds, dx, db;
    DevBuf<float> fe, fn, ft, fs, fx, fb;
    DevBuf<_Float16> he, hn, ht;
    DevBuf<double> dr;  // level residual of the transposed cycle
    DevBuf<float> fr;
    DevBuf<double> dxt;  // second iterate buffer of the line relaxation
    DevBuf<float> fxt;
  };
  template <class T>
  T* vxt(VStore& S);
  template <class T>
  void smooth(int l, T* x, const T* b, int sweeps, bool zero_guess, bool rev_lines = false);
  // levels whose red-black sweeps run as one fused pass: the finest level when it has at least
  // 2^23 cells, enough 64 x 16 x 32 tiles to fill the device (opt.fused_sweep 1), or every level (2)
  bool fused_level(int l) const {
    return opt.fused_sweep == 2 || (opt.fused_sweep == 1 && l == 0 && levels[l].d.n() >= (1L << 23));
  }
  // levels smoothed by line relaxation: every level, or the finest opt.line_levels of them
  bool line_level(int l) const {
    return opt.smoother == LSX_SMOOTH_LINE && (opt.line_levels <= 0 || l < opt.line_levels);
  }
  // levels smoothed by red-black Gauss-Seidel (the others: lines, or damped Jacobi on every level)
  bool rb_level(int l) const { return opt.smoother != LSX_SMOOTH_JACOBI && !line_level(l); }
  std::vector<VStore> vlev;
  DevBuf<float> Z32, ztmp32;
  DevBuf<float> V32;  // FP32 copies of the FGMRES basis (mixed precision with lsx_options.dot32)
  bool mixed() const { return opt.precision == LSX_PREC_MIXED || opt.precision == LSX_PREC_MIXED16; }
  bool vcyc() const { return opt.mg == LSX_MG_VCYCLE; }
  void build_vcycle();
  template <class T>
  // pre: (mixed) the FP32 copy of v is already in level 0's right-hand side buffer
  void vcycle_apply(const double* v, T* z, const float* b32 = nullptr);
  template <class T>
  void vcycle_apply_tr(const double* v, T* z);
  template <class T>
  void vcycle_run(const double* v, T* z, const float* b32 = nullptr);
  // captured V-cycle applications, one per (input, output) pair (opt.graphs)
  struct CycleGraph {
    const double* v;
    const void* z;
    const float* b32;
    hipGraphExec_t exec;
  };
  std::vector<CycleGraph> cgraphs;
  bool capturing = false;
  void clear_graphs();
  template <class T>
  T* vr(VStore& S);
  template <class T>
  T* vx(VStore& S);
  template <class T>
  T* vb(VStore& S);
  void precond(const double* v, double* z64);  // the configured preconditioner, FP64 in and out

  // event timing of fine-level kernel launches
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_rec[LSX_K_COUNT];
  double ev_ms[LSX_K_COUNT] = {0};
  double ev_bytes[LSX_K_COUNT] = {0};
  long ev_count[LSX_K_COUNT] = {0};
  hipEvent_t ev_take();
  void tbeg(int k, hipEvent_t* e);
  void tend(int k, hipEvent_t b, double bytes = -1.0);  // bytes < 0: the kernel's fixed figure
  void drain_events();
  double kernel_bytes(int k) const;
};

}  // namespace lsx
