// extern "C" entry points of include/lsx.h.  Exceptions never cross the boundary: each call
// returns a status and records the message for lsx_last_error().

#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/lsx.h"
#include "lsx_geometry_host.hpp"
#include "lsx_solver.hpp"

struct lsx_solver {
  lsx::Solver impl;
  lsx_solver(const lsx_options& o, hipStream_t s) : impl(o, s) {}
};

struct lsx_geometry {
  lsx::Geometry impl;
  lsx_geometry(int N, int M, int K, const double* dx, const double* dy, const double* dz, double x0, double y0,
               double z0, hipStream_t s)
      : impl(N, M, K, dx, dy, dz, x0, y0, z0, s) {}
};

struct lsx_tape {
  lsx_solver* s;
  lsx_geometry* g;
  long version;
  int pin;
  lsx::DevBuf<double> p;  // the forward solution
};

namespace {
thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
  try {
    g_last_error.clear();
    f();
    return LSX_OK;
  } catch (const lsx::Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return LSX_ERR_HIP;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return LSX_ERR_ARG;
  }
}
}  // namespace

extern "C" {

void lsx_options_default(lsx_options* o) {
  if (!o) return;
  o->krylov = LSX_KRYLOV_FGMRES;
  o->restart = 13;
  o->max_it = 200;
  o->eps = 1e-5;
  o->mg = LSX_MG_CASCADE;
  o->n_post = 3;
  o->n_pre = 0;
  o->omega = 0.805;
  o->precision = LSX_PREC_FP64;
  o->levels = 0;
  o->z_head_fix = 0;
  o->ortho = LSX_ORTHO_AUTO;
  o->coarse_fuse = 1;
  o->graphs = 0;
  o->smoother = LSX_SMOOTH_RBGS;
  o->fused_sweep = 1;
  o->line_levels = 0;
  o->diag_faces = 1;
  o->dot32 = 1;
}

int lsx_create(lsx_solver** out, const lsx_options* opt, void* stream) {
  if (!out || !opt) {
    g_last_error = "null argument";
    return LSX_ERR_ARG;
  }
  *out = nullptr;
  return guarded([&] { *out = new lsx_solver(*opt, (hipStream_t)stream); });
}

int lsx_destroy(lsx_solver* s) {
  return guarded([&] { delete s; });
}

int lsx_set_csr(lsx_solver* s, int N, int M, int K, const int* rowptr, const int* col, const double* val, long nnz,
                int mem) {
  if (!s) return LSX_ERR_ARG;
  return guarded([&] { s->impl.set_csr(N, M, K, rowptr, col, val, nnz, mem); });
}

int lsx_update_values(lsx_solver* s, const double* val, int mem) {
  if (!s) return LSX_ERR_ARG;
  return guarded([&] { s->impl.update_values(val, mem); });
}

int lsx_rebuild(lsx_solver* s) {
  if (!s) return LSX_ERR_ARG;
  return guarded([&] { s->impl.rebuild(); });
}

int lsx_solve(lsx_solver* s, const double* b, double* x, int mem, int* it, double* norm_r) {
  if (!s) return LSX_ERR_ARG;
  return guarded([&] { s->impl.solve(b, x, mem, it, norm_r); });
}

int lsx_set_eps(lsx_solver* s, double eps) {
  if (!s) return LSX_ERR_ARG;
  if (!(eps >= 0.0)) {
    g_last_error = "eps must be >= 0";
    return LSX_ERR_ARG;
  }
  s->impl.opt.eps = eps;
  return LSX_OK;
}

int lsx_get_stats(const lsx_solver* s, int* it, double* norm_r, int* levels) {
  if (!s) return LSX_ERR_ARG;
  if (it) *it = s->impl.it;
  if (norm_r) *norm_r = s->impl.norm_r;
  if (levels) *levels = s->impl.num_levels();
  return LSX_OK;
}

int lsx_precondition(lsx_solver* s, const double* v, double* z, int mem) {
  if (!s) return LSX_ERR_ARG;
  return guarded([&] { s->impl.precondition(v, z, mem); });
}

int lsx_apply_operator(lsx_solver* s, int level, const double* x, double* y, int mem) {
  if (!s) return LSX_ERR_ARG;
  return guarded([&] { s->impl.apply_operator(level, x, y, mem); });
}

int lsx_solve_adjoint(lsx_solver* s, const double* g, double* lambda, int mem, int* it, double* norm_r) {
  if (!s) return LSX_ERR_ARG;
  return guarded([&] { s->impl.solve_adjoint(g, lambda, mem, it, norm_r); });
}

int lsx_precondition_transpose(lsx_solver* s, const double* v, double* z, int mem) {
  if (!s) return LSX_ERR_ARG;
  return guarded([&] { s->impl.precondition_transpose(v, z, mem); });
}

int lsx_apply_operator_transpose(lsx_solver* s, const double* x, double* y, int mem) {
  if (!s) return LSX_ERR_ARG;
  return guarded([&] { s->impl.apply_operator_transpose(x, y, mem); });
}

int lsx_smooth(lsx_solver* s, const double* b, double* x, int sweeps, int zero_guess, int mem) {
  if (!s) return LSX_ERR_ARG;
  if (mem != LSX_MEM_HOST && mem != LSX_MEM_DEVICE) return LSX_ERR_ARG;
  return guarded([&] { s->impl.smooth_fine(b, x, sweeps, zero_guess, mem); });
}

long lsx_level_dims(const lsx_solver* s, int level, int* N, int* M, int* K) {
  if (!s) return LSX_ERR_ARG;
  long r = 0;
  const int st = guarded([&] { r = s->impl.level_dims(level, N, M, K); });
  return st == LSX_OK ? r : st;
}

int lsx_level_stencil(const lsx_solver* s, int level, double* coef_host, long capacity) {
  if (!s) return LSX_ERR_ARG;
  int r = 0;
  const int st = guarded([&] { r = s->impl.level_stencil(level, coef_host, capacity); });
  return st == LSX_OK ? r : st;
}

int lsx_operator_diagonal(const lsx_solver* s, double* diag_host) {
  if (!s || !diag_host) return LSX_ERR_ARG;
  return guarded([&] { s->impl.operator_diagonal(diag_host); });
}

int lsx_fine_storage(const lsx_solver* s) {
  if (!s) return LSX_ERR_ARG;
  int planes = 0;
  const int st = guarded([&] { planes = s->impl.fine_storage(); });
  return st == LSX_OK ? planes : st;
}

int lsx_synchronize(lsx_solver* s) {
  if (!s) return LSX_ERR_ARG;
  return guarded([&] { s->impl.synchronize(); });
}

const char* lsx_last_error(void) { return g_last_error.c_str(); }

int lsx_geometry_create(lsx_geometry** out, int N, int M, int K, const double* dx, const double* dy,
                        const double* dz, double x0, double y0, double z0, void* stream) {
  if (!out) {
    g_last_error = "null argument";
    return LSX_ERR_ARG;
  }
  *out = nullptr;
  return guarded([&] { *out = new lsx_geometry(N, M, K, dx, dy, dz, x0, y0, z0, (hipStream_t)stream); });
}

int lsx_geometry_destroy(lsx_geometry* g) {
  return guarded([&] { delete g; });
}

int lsx_bezier_ellipsoid(const double center[3], const double axes[3], double ctrl[288]) {
  if (!center || !axes || !ctrl) return LSX_ERR_ARG;
  return guarded([&] { lsx::bezier_ellipsoid(center, axes, ctrl); });
}

int lsx_geometry_build(lsx_geometry* g, const lsx_body* body, int hs, int hf, int vs, int vf) {
  if (!g || !body) return LSX_ERR_ARG;
  return guarded([&] {
    lsx::geo::Body b;
    b.kind = body->kind;
    for (int q = 0; q < 3; ++q) {
      b.c[q] = body->center[q];
      b.a[q] = body->axes[q];
    }
    g->impl.build(b, lsx::geo::Box{hs, hf, vs, vf}, body->kind == LSX_BODY_BEZIER6 ? body->ctrl : nullptr);
  });
}

int lsx_geometry_fields(const lsx_geometry* g, int* type, double* thx, double* thy, double* thz, double* vof,
                        double* phi) {
  if (!g) return LSX_ERR_ARG;
  return guarded([&] { g->impl.fields(type, thx, thy, thz, vof, phi); });
}

int lsx_geometry_rhs(lsx_geometry* g, const double* u, const double* v, const double* w, const double* u_west,
                     const double* u_out, const lsx_flow_bc* bc, double* b, int mem) {
  if (!g || !u || !v || !w || !u_west || !bc || !b) return LSX_ERR_ARG;
  if (mem != LSX_MEM_HOST && mem != LSX_MEM_DEVICE) return LSX_ERR_ARG;
  return guarded([&] { g->impl.pressure_rhs(u, v, w, u_west, u_out, *bc, b, mem); });
}

int lsx_geometry_correct(lsx_geometry* g, const double* dp, const double* us, const double* vs, const double* ws,
                         double* u, double* v, double* w, double* p, double coef, int mem) {
  if (!g || !dp || !us || !vs || !ws || !u || !v || !w) return LSX_ERR_ARG;
  if (mem != LSX_MEM_HOST && mem != LSX_MEM_DEVICE) return LSX_ERR_ARG;
  return guarded([&] { g->impl.correct(dp, us, vs, ws, u, v, w, p, coef, mem); });
}

int lsx_kernel_timing(lsx_solver* s, int enable) {
  if (!s) return LSX_ERR_ARG;
  return guarded([&] { s->impl.set_timing(enable != 0); });
}

int lsx_kernel_stats(lsx_solver* s, int kernel, long* launches, double* total_ms, double* bytes_per_launch) {
  if (!s) return LSX_ERR_ARG;
  return guarded([&] { s->impl.kernel_stats(kernel, launches, total_ms, bytes_per_launch); });
}

int lsx_geometry_drag(lsx_geometry* g, const double* p, int mem, double* J, double* dJdp) {
  if (!g || !p || !J) return LSX_ERR_ARG;
  if (mem != LSX_MEM_HOST && mem != LSX_MEM_DEVICE) return LSX_ERR_ARG;
  return guarded([&] { *J = g->impl.drag(p, dJdp, mem); });
}

int lsx_geometry_drag_noalloc(lsx_geometry* g, const double* p, double* J, double* dJdp) {
  if (!g || !p || !J) return LSX_ERR_ARG;
  return guarded([&] { g->impl.drag_noalloc(p, J, dJdp); });
}

int lsx_state_save(const char* path, int it, double time, int n_bodies, const double* body_pos, const double* body_vel,
                   const double* loads, int N, int M, int K, const double* u, const double* v, const double* w,
                   const double* p, const double* u_prev, const double* v_prev, const double* w_prev,
                   const double* p_prev, int mem) {
  if (mem != LSX_MEM_HOST && mem != LSX_MEM_DEVICE) return LSX_ERR_ARG;
  const double* cur[4] = {u, v, w, p};
  const double* prev[4] = {u_prev, v_prev, w_prev, p_prev};
  return guarded([&] { lsx::state_save(path, it, time, n_bodies, body_pos, body_vel, loads, N, M, K, cur, prev, mem); });
}

int lsx_state_load(const char* path, int* it, double* time, int n_bodies, double* body_pos, double* body_vel,
                   double* loads, int N, int M, int K, double* u, double* v, double* w, double* p, double* u_prev,
                   double* v_prev, double* w_prev, double* p_prev, int mem) {
  if (mem != LSX_MEM_HOST && mem != LSX_MEM_DEVICE) return LSX_ERR_ARG;
  double* cur[4] = {u, v, w, p};
  double* prev[4] = {u_prev, v_prev, w_prev, p_prev};
  return guarded([&] { lsx::state_load(path, it, time, n_bodies, body_pos, body_vel, loads, N, M, K, cur, prev, mem); });
}

long lsx_state_bytes(int N, int M, int K, int n_bodies) {
  long r = LSX_ERR_ARG;
  if (n_bodies < 0) return LSX_ERR_ARG;
  const int st = guarded([&] { r = lsx::state_bytes(N, M, K, n_bodies); });
  return st == LSX_OK ? r : st;
}

int lsx_geometry_write_tecplot(const lsx_geometry* g, const char* path, const double* u, const double* v,
                               const double* w, const double* p, const double* u_west, const lsx_walls* walls,
                               int outflow, int mem) {
  if (!g || !path || !walls) return LSX_ERR_ARG;
  if (mem != LSX_MEM_HOST && mem != LSX_MEM_DEVICE) return LSX_ERR_ARG;
  return guarded([&] { g->impl.write_tecplot(path, u, v, w, p, u_west, *walls, outflow != 0, mem); });
}

int lsx_write_tecplot(const char* path, int N, int M, int K, const double* xc, const double* yc, const double* zc,
                      const double* dx, const double* dy, const double* dz, const int* type, const double* thx,
                      const double* thy, const double* thz, const double* u, const double* v, const double* w,
                      const double* p, const double* u_west, const lsx_walls* walls, int outflow) {
  if (!path || !xc || !yc || !zc || !dx || !dy || !dz || !type || !thx || !thy || !thz || !u || !v || !w || !p ||
      !walls)
    return LSX_ERR_ARG;
  if (N < 3 || M < 3 || K < 3 || (long)N * M * K >= (1L << 31)) return LSX_ERR_ARG;
  return guarded([&] {
    const long n = (long)N * M * K;
    auto vec = [](const double* a, long len) { return std::vector<double>(a, a + len); };
    lsx::write_tecplot(path, N, M, K, vec(xc, N), vec(yc, M), vec(zc, K), vec(dx, N), vec(dy, M), vec(dz, K),
                       std::vector<int>(type, type + n), vec(thx, n), vec(thy, n), vec(thz, n),
                       vec(u, (long)(N - 1) * M * K), vec(v, (long)N * (M - 1) * K), vec(w, (long)N * M * (K - 1)),
                       vec(p, n), u_west ? vec(u_west, (long)M * K) : std::vector<double>(), *walls, outflow != 0);
  });
}

int lsx_geometry_shape_gradient(lsx_geometry* g, const double* p, const double* lambda, int objective,
                                const double* u, const double* v, const double* w, const double* u_west,
                                const double* u_out, const lsx_flow_bc* bc, int pin, double h_rel, int mem,
                                double* dJdphi, double* dJdctrl) {
  if (!g || !p || !lambda) return LSX_ERR_ARG;
  if (mem != LSX_MEM_HOST && mem != LSX_MEM_DEVICE) return LSX_ERR_ARG;
  if (u && !bc) return LSX_ERR_ARG;
  return guarded([&] {
    lsx::FlowStage st;
    if (u) g->impl.stage_flow(u, v, w, u_west, u_out, *bc, mem, st);
    g->impl.shape_gradient(p, lambda, objective, u ? &st.args : nullptr, pin, h_rel, dJdphi, dJdctrl, mem);
  });
}

int lsx_adjoint_forward(lsx_solver* s, lsx_geometry* g, int pin, const double* b, double* p, int mem, lsx_tape** tape,
                        int* it, double* norm_r) {
  if (!s || !g || !b || !p || !tape) return LSX_ERR_ARG;
  if (mem != LSX_MEM_HOST && mem != LSX_MEM_DEVICE) return LSX_ERR_ARG;
  *tape = nullptr;
  return guarded([&] {
    if (!s->impl.operator_from(&g->impl, g->impl.version(), pin)) s->impl.set_geometry(g->impl, pin);
    s->impl.solve(b, p, mem, it, norm_r);
    lsx_tape* t = new lsx_tape{s, g, g->impl.version(), pin, {}};
    const long n = g->impl.n();
    try {
      t->p.alloc(n);
      LSX_HIP(hipMemcpy(t->p.p, p, n * sizeof(double),
                        mem == LSX_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice));
    } catch (...) {
      delete t;
      throw;
    }
    *tape = t;
  });
}

int lsx_adjoint_backward(lsx_tape* t, const double* dJdp, int objective, const double* u, const double* v,
                         const double* w, const double* u_west, const double* u_out, const lsx_flow_bc* bc,
                         double h_rel, int mem, double* dJdb, double* dJdctrl, int* it, double* norm_r) {
  if (!t || !dJdp || !dJdb) return LSX_ERR_ARG;
  if (mem != LSX_MEM_HOST && mem != LSX_MEM_DEVICE) return LSX_ERR_ARG;
  if (u && !bc) return LSX_ERR_ARG;
  return guarded([&] {
    lsx::Geometry& G = t->g->impl;
    lsx::Solver& S = t->s->impl;
    if (G.version() != t->version)
      throw lsx::Error(LSX_ERR_STATE, "the geometry was rebuilt after the forward call");
    // recompute the operator and its multigrid hierarchy when they were replaced meanwhile
    if (!S.operator_from(&G, t->version, t->pin)) S.set_geometry(G, t->pin);
    const long n = G.n();
    if (mem == LSX_MEM_HOST)
      for (long i = 0; i < n; ++i) dJdb[i] = 0.0;
    else
      LSX_HIP(hipMemset(dJdb, 0, n * sizeof(double)));
    S.solve_adjoint(dJdp, dJdb, mem, it, norm_r);
    if (dJdctrl) {
      lsx::FlowStage st;
      if (u) G.stage_flow(u, v, w, u_west, u_out, *bc, mem, st);
      lsx::DevBuf<double> lam;
      const double* dl = dJdb;
      if (mem == LSX_MEM_HOST) {
        lam.alloc(n);
        LSX_HIP(hipMemcpy(lam.p, dJdb, n * sizeof(double), hipMemcpyHostToDevice));
        dl = lam.p;
      }
      G.shape_gradient(t->p.p, dl, objective, u ? &st.args : nullptr, t->pin, h_rel, nullptr, dJdctrl,
                       LSX_MEM_DEVICE);
    }
  });
}

int lsx_tape_destroy(lsx_tape* t) {
  return guarded([&] { delete t; });
}

int lsx_set_helmholtz(lsx_solver* s, const lsx_geometry* g, int component, double nu, double coef_time, int flow) {
  if (!s || !g) return LSX_ERR_ARG;
  return guarded([&] { s->impl.set_helmholtz(g->impl, component, nu, coef_time, flow); });
}

int lsx_set_fields(lsx_solver* s, int N, int M, int K, const double* dx, const double* dy, const double* dz,
                   const int* type, const double* thx, const double* thy, const double* thz, const double* vof, int pin,
                   int mem) {
  if (!s) return LSX_ERR_ARG;
  return guarded([&] { s->impl.set_fields(N, M, K, dx, dy, dz, type, thx, thy, thz, vof, pin, mem); });
}

int lsx_set_geometry(lsx_solver* s, const lsx_geometry* g, int pin) {
  if (!s || !g) return LSX_ERR_ARG;
  return guarded([&] { s->impl.set_geometry(g->impl, pin); });
}

}  // extern "C"
