// Geometry pipeline driver: grid setup (cell centres and upper-corner coordinates follow the
// reference's buildDeltaNodes and level-set sweep, ls_stag_uniform_mesh.h:174-185 and
// flow_sphere_task_description.h:57-74) and the ordered launch of the per-cell kernels.

#include "lsx_geometry_host.hpp"

#include <cmath>

namespace lsx {

namespace {
// cell centres: a uniform axis uses a + (k + 1/2) h; a stretched axis accumulates widths
std::vector<double> centres(const std::vector<double>& w, double a) {
  const int n = (int)w.size();
  std::vector<double> c(n);
  bool uniform = true;
  for (int i = 1; i < n; ++i) uniform = uniform && (w[i] == w[0]);
  if (uniform) {
    for (int i = 0; i < n; ++i) c[i] = a + (i + 0.5) * w[0];
  } else {
    double e = a;
    for (int i = 0; i < n; ++i) {
      c[i] = e + 0.5 * w[i];
      e += w[i];
    }
  }
  return c;
}
}  // namespace

Geometry::Geometry(int N_, int M_, int K_, const double* dx_, const double* dy_, const double* dz_, double x0_,
                   double y0_, double z0_, hipStream_t s)
    : N(N_), M(M_), K(K_), x0(x0_), y0(y0_), z0(z0_) {
  if (N < 3 || M < 3 || K < 3) throw Error(LSX_ERR_ARG, "grid dimensions must be >= 3");
  if ((long)N * M * K >= (1L << 31)) throw Error(LSX_ERR_ARG, "grid too large for 32-bit cell indices");
  if (!dx_ || !dy_ || !dz_) throw Error(LSX_ERR_ARG, "null cell width array");
  hdx.assign(dx_, dx_ + N);
  hdy.assign(dy_, dy_ + M);
  hdz.assign(dz_, dz_ + K);
  for (double w : hdx) if (!(w > 0)) throw Error(LSX_ERR_ARG, "cell widths must be positive");
  for (double w : hdy) if (!(w > 0)) throw Error(LSX_ERR_ARG, "cell widths must be positive");
  for (double w : hdz) if (!(w > 0)) throw Error(LSX_ERR_ARG, "cell widths must be positive");
  if (s) {
    stream = s;
  } else {
    LSX_HIP(hipStreamCreate(&stream));
    own_stream = true;
  }
  dx.alloc(N);
  dy.alloc(M);
  dz.alloc(K);
  LSX_HIP(hipMemcpyAsync(dx.p, hdx.data(), N * sizeof(double), hipMemcpyHostToDevice, stream));
  LSX_HIP(hipMemcpyAsync(dy.p, hdy.data(), M * sizeof(double), hipMemcpyHostToDevice, stream));
  LSX_HIP(hipMemcpyAsync(dz.p, hdz.data(), K * sizeof(double), hipMemcpyHostToDevice, stream));
  LSX_HIP(hipStreamSynchronize(stream));
}

Geometry::~Geometry() {
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

void Geometry::build(const geo::Body& body_in, geo::Box box, const double* host_ctrl) {
  if (box.hs < 1 || box.vs < 1 || box.hf > N - 1 || box.vf > M - 1 || box.hs >= box.hf || box.vs >= box.vf)
    throw Error(LSX_ERR_ARG, "geometry box must lie inside the domain's interior");
  if (body_in.kind != geo::BODY_SPHERE && body_in.kind != geo::BODY_ELLIPSOID && body_in.kind != geo::BODY_BEZIER6)
    throw Error(LSX_ERR_ARG, "unknown body kind");
  for (int q = 0; q < (body_in.kind == geo::BODY_SPHERE ? 1 : 3); ++q)
    if (!(body_in.a[q] > 0)) throw Error(LSX_ERR_ARG, "body size must be positive");
  geo::Body body = body_in;
  if (body.kind == geo::BODY_BEZIER6) {
    if (!host_ctrl) throw Error(LSX_ERR_ARG, "a Bezier body needs its 288 control values");
    for (int q = 0; q < 288; ++q)
      if (!std::isfinite(host_ctrl[q])) throw Error(LSX_ERR_ARG, "Bezier control values must be finite");
    bez.alloc(288);
    LSX_HIP(hipMemcpyAsync(bez.p, host_ctrl, 288 * sizeof(double), hipMemcpyHostToDevice, stream));
    body.ctrl = bez.p;
  }
  // upper-corner coordinates: x starts at the first box cell and steps by the next cell's width
  const std::vector<double> xc = centres(hdx, x0), yc = centres(hdy, y0), zc = centres(hdz, z0);
  std::vector<double> hcx(N, 0.0), hcy(M), hcz(K);
  double px = xc[box.hs] + 0.5 * hdx[box.hs];
  for (int x = box.hs; x < box.hf; ++x) {
    hcx[x] = px;
    if (x + 1 < N) px += hdx[x + 1];
  }
  for (int y = 0; y < M; ++y) hcy[y] = yc[y] + 0.5 * hdy[y];
  for (int z = 0; z < K; ++z) hcz[z] = zc[z] + 0.5 * hdz[z];
  const long nn = n();
  cx.alloc(N);
  cy.alloc(M);
  cz.alloc(K);
  phi.alloc(nn);
  phi_raw.alloc(nn);
  osc_tmp.alloc(nn);
  dchanged.alloc(1);
  flag.alloc(nn);
  ax.alloc(nn);
  ay.alloc(nn);
  az.alloc(nn);
  tx.alloc(nn);
  ty.alloc(nn);
  tz.alloc(nn);
  vof.alloc(nn);
  type.alloc(nn);
  sx.alloc(nn);
  sy.alloc(nn);
  sz.alloc(nn);
  LSX_HIP(hipMemcpyAsync(cx.p, hcx.data(), N * sizeof(double), hipMemcpyHostToDevice, stream));
  LSX_HIP(hipMemcpyAsync(cy.p, hcy.data(), M * sizeof(double), hipMemcpyHostToDevice, stream));
  LSX_HIP(hipMemcpyAsync(cz.p, hcz.data(), K * sizeof(double), hipMemcpyHostToDevice, stream));
  geo::level_set(cx.p, cy.p, cz.p, N, M, K, box, body, phi_raw.p, stream);
  LSX_HIP(hipGetLastError());
  osc_passes = geo::oscillation_fix(phi_raw.p, phi.p, osc_tmp.p, flag.p, dchanged.p, N, M, K, box, stream);
  if (osc_passes < 0) throw Error(LSX_ERR_HIP, "level-set oscillation fix failed");
  geo::alpha(phi.p, N, M, K, box, ax.p, ay.p, az.p, stream);
  geo::types(phi.p, N, M, K, box, type.p, sx.p, sy.p, sz.p, stream);
  geo::theta(ax.p, ay.p, az.p, sx.p, sy.p, sz.p, N, M, K, box, tx.p, ty.p, tz.p, stream);
  geo::vof(type.p, ax.p, ay.p, az.p, tx.p, ty.p, tz.p, dx.p, dy.p, dz.p, N, M, K, vof.p, stream);
  LSX_HIP(hipGetLastError());
  LSX_HIP(hipStreamSynchronize(stream));
  body_ = body;
  box_ = box;
  built = true;
  ++build_count;
}

void Geometry::stage_flow(const double* u, const double* v, const double* w, const double* u_west,
                          const double* u_out, const lsx_flow_bc& bc, int mem, FlowStage& st) const {
  if (!u || !v || !w || !u_west) throw Error(LSX_ERR_ARG, "null flow field");
  const long nu = (long)(N - 1) * M * K, nv = (long)N * (M - 1) * K, nw = (long)N * M * (K - 1);
  const long nrow = (long)M * K;
  st.args.bc = geo::FlowBc{bc.U_E, bc.V_S, bc.V_N, bc.W_B, bc.W_T, bc.pin != 0};
  if (mem == LSX_MEM_DEVICE) {
    st.args.u = u;
    st.args.v = v;
    st.args.w = w;
    st.args.u_west = u_west;
    st.args.u_out = bc.outflow ? (u_out ? u_out : u) : nullptr;
    return;
  }
  st.u.alloc(nu);
  st.v.alloc(nv);
  st.w.alloc(nw);
  st.west.alloc(nrow);
  LSX_HIP(hipMemcpyAsync(st.u.p, u, nu * sizeof(double), hipMemcpyHostToDevice, stream));
  LSX_HIP(hipMemcpyAsync(st.v.p, v, nv * sizeof(double), hipMemcpyHostToDevice, stream));
  LSX_HIP(hipMemcpyAsync(st.w.p, w, nw * sizeof(double), hipMemcpyHostToDevice, stream));
  LSX_HIP(hipMemcpyAsync(st.west.p, u_west, nrow * sizeof(double), hipMemcpyHostToDevice, stream));
  st.args.u = st.u.p;
  st.args.v = st.v.p;
  st.args.w = st.w.p;
  st.args.u_west = st.west.p;
  st.args.u_out = nullptr;
  if (bc.outflow) {
    if (u_out) {
      st.out.alloc(nu);
      LSX_HIP(hipMemcpyAsync(st.out.p, u_out, nu * sizeof(double), hipMemcpyHostToDevice, stream));
      st.args.u_out = st.out.p;
    } else {
      st.args.u_out = st.u.p;
    }
  }
}

void Geometry::Work::alloc(long n) {
  raw.alloc(n);
  phi.alloc(n);
  tmp.alloc(n);
  ax.alloc(n);
  ay.alloc(n);
  az.alloc(n);
  tx.alloc(n);
  ty.alloc(n);
  tz.alloc(n);
  vof.alloc(n);
  type.alloc(n);
  sx.alloc(n);
  sy.alloc(n);
  sz.alloc(n);
  flag.alloc(n);
}

void Geometry::rebuild_from(Work& w) {
  if (geo::oscillation_fix(w.raw.p, w.phi.p, w.tmp.p, w.flag.p, dchanged.p, N, M, K, box_, stream) < 0)
    throw Error(LSX_ERR_HIP, "level-set oscillation fix failed");
  geo::alpha(w.phi.p, N, M, K, box_, w.ax.p, w.ay.p, w.az.p, stream);
  geo::types(w.phi.p, N, M, K, box_, w.type.p, w.sx.p, w.sy.p, w.sz.p, stream);
  geo::theta(w.ax.p, w.ay.p, w.az.p, w.sx.p, w.sy.p, w.sz.p, N, M, K, box_, w.tx.p, w.ty.p, w.tz.p, stream);
  geo::vof(w.type.p, w.ax.p, w.ay.p, w.az.p, w.tx.p, w.ty.p, w.tz.p, dx.p, dy.p, dz.p, N, M, K, w.vof.p, stream);
}

double Geometry::drag(const double* p, double* dJdp, int mem) {
  if (!built) throw Error(LSX_ERR_STATE, "geometry not built");
  if (!p) throw Error(LSX_ERR_ARG, "null pressure");
  const long nn = n();
  sg_partial.alloc(256);
  const double* dp = p;
  double* dg = dJdp;
  DevBuf<double> hp, hg;
  if (mem == LSX_MEM_HOST) {
    hp.alloc(nn);
    LSX_HIP(hipMemcpyAsync(hp.p, p, nn * sizeof(double), hipMemcpyHostToDevice, stream));
    dp = hp.p;
    if (dJdp) {
      hg.alloc(nn);
      dg = hg.p;
    }
  }
  const double J = geo::drag(type.p, tx.p, dy.p, dz.p, N, M, K, box_, dp, dg, sg_partial.p, stream);
  LSX_HIP(hipGetLastError());
  if (mem == LSX_MEM_HOST && dJdp) {
    LSX_HIP(hipMemcpyAsync(dJdp, hg.p, nn * sizeof(double), hipMemcpyDeviceToHost, stream));
    LSX_HIP(hipStreamSynchronize(stream));
  }
  return J;
}

void Geometry::write_tecplot(const char* path, const double* u, const double* v, const double* w, const double* p,
                            const double* u_west, const lsx_walls& walls, bool outflow, int mem) const {
  if (!built) throw Error(LSX_ERR_STATE, "geometry not built");
  if (!path || !u || !v || !w || !p) throw Error(LSX_ERR_ARG, "null argument");
  const long nn = n(), nu = (long)(N - 1) * M * K, nv = (long)N * (M - 1) * K, nw = (long)N * M * (K - 1);
  auto get = [&](const double* a, long len) {
    std::vector<double> h(len);
    if (mem == LSX_MEM_DEVICE)
      LSX_HIP(hipMemcpy(h.data(), a, len * sizeof(double), hipMemcpyDeviceToHost));
    else
      std::copy(a, a + len, h.begin());
    return h;
  };
  std::vector<int> t(nn);
  std::vector<double> thx(nn), thy(nn), thz(nn);
  fields(t.data(), thx.data(), thy.data(), thz.data(), nullptr, nullptr);
  const std::vector<double> hu = get(u, nu), hv = get(v, nv), hw = get(w, nw), hp = get(p, nn);
  const std::vector<double> uw = u_west ? get(u_west, (long)M * K) : std::vector<double>();
  lsx::write_tecplot(path, N, M, K, centres(hdx, x0), centres(hdy, y0), centres(hdz, z0), hdx, hdy, hdz, t, thx, thy,
                     thz, hu, hv, hw, hp, uw, walls, outflow);
}

static bool device_pointer(const void* ptr) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, ptr) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

void Geometry::drag_noalloc(const double* p, double* J, double* dJdp) {
  if (!built) throw Error(LSX_ERR_STATE, "geometry not built");
  if (!p || !J) throw Error(LSX_ERR_ARG, "null pressure or J");
  if (!device_pointer(p) || !device_pointer(J) || (dJdp && !device_pointer(dJdp)))
    throw Error(LSX_ERR_ARG, "p, J and dJdp must be device pointers (the call stages nothing)");
  if (n() > DRAG_NOALLOC_MAX_CELLS)
    throw Error(LSX_ERR_ARG, "grid exceeds the 2^31 - 1 cells of the allocation-free drag");
  double* part = nullptr;
  if (geo::drag_scratch(&part) != 0) throw Error(LSX_ERR_HIP, "hipGetSymbolAddress of the drag scratch failed");
  geo::drag_noalloc(type.p, tx.p, dy.p, dz.p, N, M, K, box_, p, dJdp, part, J, stream);
  LSX_HIP(hipGetLastError());
}

void Geometry::shape_gradient(const double* p, const double* lam, int objective, const geo::RhsArgs* rhs, int pin,
                              double h_rel, double* dJdphi, double* dJdctrl, int mem) {
  if (!built) throw Error(LSX_ERR_STATE, "geometry not built");
  if (!p || !lam) throw Error(LSX_ERR_ARG, "null pressure or adjoint");
  if (objective != geo::OBJ_NONE && objective != geo::OBJ_DRAG_X) throw Error(LSX_ERR_ARG, "unknown objective");
  if (dJdctrl && body_.kind != geo::BODY_BEZIER6)
    throw Error(LSX_ERR_STATE, "control-point gradients need a Bezier body");
  if (!(h_rel > 0.0) || !(h_rel < 1e-2)) throw Error(LSX_ERR_ARG, "h_rel must lie in (0, 1e-2)");
  const long nn = n();
  double wmin = hdx[0];
  for (double v : hdx) wmin = v < wmin ? v : wmin;
  for (double v : hdy) wmin = v < wmin ? v : wmin;
  for (double v : hdz) wmin = v < wmin ? v : wmin;
  const double h = h_rel * wmin;
  const double* dp = p;
  const double* dl = lam;
  DevBuf<double> hp, hl;
  if (mem == LSX_MEM_HOST) {
    hp.alloc(nn);
    hl.alloc(nn);
    LSX_HIP(hipMemcpyAsync(hp.p, p, nn * sizeof(double), hipMemcpyHostToDevice, stream));
    LSX_HIP(hipMemcpyAsync(hl.p, lam, nn * sizeof(double), hipMemcpyHostToDevice, stream));
    dp = hp.p;
    dl = hl.p;
  }
  wp.alloc(nn);
  wm.alloc(nn);
  sg_dF.alloc(nn);
  sg_g.alloc(nn);
  LSX_HIP(hipMemsetAsync(sg_g.p, 0, nn * sizeof(double), stream));
  const int classes = 216;
  for (int cls = 0; cls < classes; ++cls) {
    geo::perturb_class(phi_raw.p, N, M, K, box_, cls, h, wp.raw.p, wm.raw.p, stream);
    rebuild_from(wp);
    rebuild_from(wm);
    geo::lagrangian_diff(wp.view(), wm.view(), dx.p, dy.p, dz.p, N, M, K, box_, pin, dp, dl, objective, rhs, h,
                         sg_dF.p, stream);
    geo::corner_gather(sg_dF.p, N, M, K, box_, cls, sg_g.p, stream);
  }
  LSX_HIP(hipGetLastError());
  if (dJdphi)
    LSX_HIP(hipMemcpyAsync(dJdphi, sg_g.p, nn * sizeof(double),
                           mem == LSX_MEM_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, stream));
  if (dJdctrl) {
    sg_f.alloc(nn);
    sg_uv.alloc(2 * nn);
    sg_w.alloc(3 * nn);
    sg_out.alloc(288);
    geo::bezier_chain(sg_g.p, cx.p, cy.p, cz.p, N, M, K, box_, body_, sg_f.p, sg_uv.p, sg_w.p, sg_out.p, stream);
    LSX_HIP(hipGetLastError());
    LSX_HIP(hipMemcpyAsync(dJdctrl, sg_out.p, 288 * sizeof(double), hipMemcpyDeviceToHost, stream));
  }
  LSX_HIP(hipStreamSynchronize(stream));
}

void Geometry::fields(int* htype, double* hthx, double* hthy, double* hthz, double* hvof, double* hphi) const {
  if (!built) throw Error(LSX_ERR_STATE, "geometry not built");
  const long nn = n();
  if (htype) LSX_HIP(hipMemcpyAsync(htype, type.p, nn * sizeof(int), hipMemcpyDeviceToHost, stream));
  if (hthx) LSX_HIP(hipMemcpyAsync(hthx, tx.p, nn * sizeof(double), hipMemcpyDeviceToHost, stream));
  if (hthy) LSX_HIP(hipMemcpyAsync(hthy, ty.p, nn * sizeof(double), hipMemcpyDeviceToHost, stream));
  if (hthz) LSX_HIP(hipMemcpyAsync(hthz, tz.p, nn * sizeof(double), hipMemcpyDeviceToHost, stream));
  if (hvof) LSX_HIP(hipMemcpyAsync(hvof, vof.p, nn * sizeof(double), hipMemcpyDeviceToHost, stream));
  if (hphi) LSX_HIP(hipMemcpyAsync(hphi, phi.p, nn * sizeof(double), hipMemcpyDeviceToHost, stream));
  LSX_HIP(hipStreamSynchronize(stream));
}

void Geometry::helmholtz7(int dir, double nu, double coef_time, int flow, double* coef_dev, hipStream_t s) const {
  if (!built) throw Error(LSX_ERR_STATE, "geometry not built");
  if (dir < 0 || dir > 2) throw Error(LSX_ERR_ARG, "velocity component must be 0 (u), 1 (v) or 2 (w)");
  if (!(nu > 0.0) || !(coef_time >= 0.0)) throw Error(LSX_ERR_ARG, "nu must be positive and coef_time >= 0");
  std::vector<int> t(n());
  LSX_HIP(hipMemcpyAsync(t.data(), type.p, n() * sizeof(int), hipMemcpyDeviceToHost, stream));
  LSX_HIP(hipStreamSynchronize(stream));
  for (int v : t)
    if (v > 0 && v != geo::FluidCell)
      throw Error(LSX_ERR_ARG, "the momentum operators are assembled for boxes without cut cells");
  geo::helmholtz7(type.p, tx.p, ty.p, tz.p, vof.p, dx.p, dy.p, dz.p, N, M, K, dir, nu, coef_time, flow, coef_dev, s);
  LSX_HIP(hipGetLastError());
}

void Geometry::laplace7(int pin, double* coef_dev, hipStream_t s) const {
  if (!built) throw Error(LSX_ERR_STATE, "geometry not built");
  // the geometry stream must have finished before another stream reads the fields
  LSX_HIP(hipStreamSynchronize(stream));
  geo::laplace7(type.p, tx.p, ty.p, tz.p, vof.p, dx.p, dy.p, dz.p, N, M, K, pin, coef_dev, s);
  LSX_HIP(hipGetLastError());
}

void Geometry::pressure_rhs(const double* u, const double* v, const double* w, const double* u_west,
                            const double* u_out, const lsx_flow_bc& bc, double* b, int mem) {
  if (!built) throw Error(LSX_ERR_STATE, "geometry not built");
  const long nn = n(), nu = (long)(N - 1) * M * K, nv = (long)N * (M - 1) * K, nw = (long)N * M * (K - 1);
  const long nrow = (long)M * K;
  geo::FlowBc fb{bc.U_E, bc.V_S, bc.V_N, bc.W_B, bc.W_T, bc.pin != 0};
  if (mem == LSX_MEM_DEVICE) {
    geo::pressure_rhs(type.p, tx.p, ty.p, tz.p, dx.p, dy.p, dz.p, N, M, K, u, v, w, u_west,
                      bc.outflow ? (u_out ? u_out : u) : nullptr, fb, b, stream);
    LSX_HIP(hipGetLastError());
    LSX_HIP(hipStreamSynchronize(stream));
    return;
  }
  DevBuf<double> du, dv, dw, dwest, dout, db;
  du.alloc(nu);
  dv.alloc(nv);
  dw.alloc(nw);
  dwest.alloc(nrow);
  db.alloc(nn);
  LSX_HIP(hipMemcpyAsync(du.p, u, nu * sizeof(double), hipMemcpyHostToDevice, stream));
  LSX_HIP(hipMemcpyAsync(dv.p, v, nv * sizeof(double), hipMemcpyHostToDevice, stream));
  LSX_HIP(hipMemcpyAsync(dw.p, w, nw * sizeof(double), hipMemcpyHostToDevice, stream));
  LSX_HIP(hipMemcpyAsync(dwest.p, u_west, nrow * sizeof(double), hipMemcpyHostToDevice, stream));
  const double* pout = nullptr;
  if (bc.outflow) {
    if (u_out) {
      dout.alloc(nu);
      LSX_HIP(hipMemcpyAsync(dout.p, u_out, nu * sizeof(double), hipMemcpyHostToDevice, stream));
      pout = dout.p;
    } else {
      pout = du.p;
    }
  }
  geo::pressure_rhs(type.p, tx.p, ty.p, tz.p, dx.p, dy.p, dz.p, N, M, K, du.p, dv.p, dw.p, dwest.p, pout, fb, db.p,
                    stream);
  LSX_HIP(hipGetLastError());
  LSX_HIP(hipMemcpyAsync(b, db.p, nn * sizeof(double), hipMemcpyDeviceToHost, stream));
  LSX_HIP(hipStreamSynchronize(stream));
}

void Geometry::correct(const double* dp, const double* us, const double* vs, const double* ws, double* u, double* v,
                       double* w, double* p, double coef, int mem) {
  if (!built) throw Error(LSX_ERR_STATE, "geometry not built");
  const long nn = n(), nu = (long)(N - 1) * M * K, nv = (long)N * (M - 1) * K, nw = (long)N * M * (K - 1);
  if (mem == LSX_MEM_DEVICE) {
    geo::correct(type.p, tx.p, ty.p, tz.p, vof.p, dx.p, dy.p, dz.p, N, M, K, dp, us, vs, ws, u, v, w, stream);
    if (p) geo::pressure_update(p, dp, coef, nn, stream);
    LSX_HIP(hipGetLastError());
    LSX_HIP(hipStreamSynchronize(stream));
    return;
  }
  DevBuf<double> ddp, dus, dvs, dws, du, dv, dw, dpp;
  ddp.alloc(nn);
  dus.alloc(nu);
  dvs.alloc(nv);
  dws.alloc(nw);
  du.alloc(nu);
  dv.alloc(nv);
  dw.alloc(nw);
  LSX_HIP(hipMemcpyAsync(ddp.p, dp, nn * sizeof(double), hipMemcpyHostToDevice, stream));
  LSX_HIP(hipMemcpyAsync(dus.p, us, nu * sizeof(double), hipMemcpyHostToDevice, stream));
  LSX_HIP(hipMemcpyAsync(dvs.p, vs, nv * sizeof(double), hipMemcpyHostToDevice, stream));
  LSX_HIP(hipMemcpyAsync(dws.p, ws, nw * sizeof(double), hipMemcpyHostToDevice, stream));
  geo::correct(type.p, tx.p, ty.p, tz.p, vof.p, dx.p, dy.p, dz.p, N, M, K, ddp.p, dus.p, dvs.p, dws.p, du.p, dv.p, dw.p,
               stream);
  if (p) {
    dpp.alloc(nn);
    LSX_HIP(hipMemcpyAsync(dpp.p, p, nn * sizeof(double), hipMemcpyHostToDevice, stream));
    geo::pressure_update(dpp.p, ddp.p, coef, nn, stream);
    LSX_HIP(hipMemcpyAsync(p, dpp.p, nn * sizeof(double), hipMemcpyDeviceToHost, stream));
  }
  LSX_HIP(hipGetLastError());
  LSX_HIP(hipMemcpyAsync(u, du.p, nu * sizeof(double), hipMemcpyDeviceToHost, stream));
  LSX_HIP(hipMemcpyAsync(v, dv.p, nv * sizeof(double), hipMemcpyDeviceToHost, stream));
  LSX_HIP(hipMemcpyAsync(w, dw.p, nw * sizeof(double), hipMemcpyDeviceToHost, stream));
  LSX_HIP(hipStreamSynchronize(stream));
}

}  // namespace lsx
