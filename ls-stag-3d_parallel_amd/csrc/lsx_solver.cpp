// Host orchestration of the pressure solve: hierarchy setup, multigrid cascade, FGMRES and
// BiCGStab loops.  Device work is queued on the handle's stream; the host only reads the
// scalars the Krylov recurrences need (one small copy per Arnoldi step for FGMRES).

#include "lsx_solver.hpp"
#include "lsx_geometry_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

namespace lsx {

namespace {
// scalar slots in `scalars`
constexpr int SC_NORM = 0;    // residual / ||w|| results
constexpr int SC_TINY = 1;    // max |diag| of a level during setup
constexpr int SC_TMP = 2;     // scratch dot results
constexpr int SC_SIGMA = 4;   // the two pinned cells' face sums (V-cycle setup)
constexpr int SC_H = 8;       // Hessenberg column entries H[hNum]
}  // namespace

Solver::Solver(const lsx_options& o, hipStream_t s) : opt(o) {
  if (opt.restart < 1 || opt.restart > 256) throw Error(LSX_ERR_ARG, "restart must be in [1, 256]");
  if (opt.max_it < 1) throw Error(LSX_ERR_ARG, "max_it must be >= 1");
  if (opt.n_post < 0) throw Error(LSX_ERR_ARG, "n_post must be >= 0");
  if (opt.krylov != LSX_KRYLOV_FGMRES && opt.krylov != LSX_KRYLOV_BICGSTAB) throw Error(LSX_ERR_ARG, "unknown krylov");
  if (opt.mg != LSX_MG_CASCADE && opt.mg != LSX_MG_VCYCLE) throw Error(LSX_ERR_ARG, "unknown multigrid");
  if (mixed() && opt.mg != LSX_MG_VCYCLE)
    throw Error(LSX_ERR_ARG, "mixed precision uses the V-cycle multigrid (LSX_MG_VCYCLE)");
  if (opt.precision != LSX_PREC_FP64 && !mixed()) throw Error(LSX_ERR_ARG, "unknown precision");
  if (opt.ortho < LSX_ORTHO_AUTO || opt.ortho > LSX_ORTHO_CGS2) throw Error(LSX_ERR_ARG, "unknown ortho");
  if (opt.mg == LSX_MG_VCYCLE && (opt.n_pre < 1 || opt.n_post < 1))
    throw Error(LSX_ERR_ARG, "the V-cycle needs n_pre >= 1 and n_post >= 1");
  if (opt.smoother != LSX_SMOOTH_RBGS && opt.smoother != LSX_SMOOTH_LINE && opt.smoother != LSX_SMOOTH_JACOBI)
    throw Error(LSX_ERR_ARG, "unknown smoother");
  if (opt.smoother == LSX_SMOOTH_JACOBI && !(opt.omega > 0.0 && opt.omega <= 1.0))
    throw Error(LSX_ERR_ARG, "Jacobi damping omega must be in (0, 1]");
  if (opt.fused_sweep < 0 || opt.fused_sweep > 2) throw Error(LSX_ERR_ARG, "fused_sweep must be 0, 1 or 2");
  if (opt.levels < 0) throw Error(LSX_ERR_ARG, "levels must be >= 0");
  if (opt.mg == LSX_MG_VCYCLE && opt.levels == 1) throw Error(LSX_ERR_ARG, "V-cycle levels must be 0 (automatic) or >= 2");
  if (opt.mg != LSX_MG_VCYCLE && opt.levels > 1)
    throw Error(LSX_ERR_ARG, "the reference cascade uses the automatic level count (levels = 0) or one level (1)");
  if (s) {
    stream = s;
  } else {
    LSX_HIP(hipStreamCreate(&stream));  // default priority
    own_stream = true;
  }
  err.alloc(1);
  scalars.alloc(SC_H + 256 * 257 / 2 + 8);
  partial.alloc(2048);
  mpartial.alloc((size_t)CGS_MAX * CGS_BLOCKS_MAX);
  cgs_h.alloc(2 * CGS_MAX);
}

Solver::~Solver() {
  clear_graphs();
  for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

// -----------------------------------------------------------------------------------------------
// kernel timing
// -----------------------------------------------------------------------------------------------
hipEvent_t Solver::ev_take() {
  if (ev_used == ev_pool.size()) {
    hipEvent_t e;
    LSX_HIP(hipEventCreate(&e));
    ev_pool.push_back(e);
  }
  return ev_pool[ev_used++];
}

void Solver::clear_graphs() {
  for (CycleGraph& g : cgraphs) (void)hipGraphExecDestroy(g.exec);
  cgraphs.clear();
}

void Solver::tbeg(int k, hipEvent_t* e) {
  (void)k;
  *e = nullptr;
  if (!timing || capturing) return;
  *e = ev_take();
  LSX_HIP(hipEventRecord(*e, stream));
}

void Solver::tend(int k, hipEvent_t b, double bytes) {
  if (!timing || !b) return;
  hipEvent_t e = ev_take();
  LSX_HIP(hipEventRecord(e, stream));
  ev_rec[k].push_back({b, e});
  ev_bytes[k] += bytes < 0 ? kernel_bytes(k) : bytes;
  if (ev_used > 4096) drain_events();
}

void Solver::drain_events() {
  LSX_HIP(hipStreamSynchronize(stream));
  for (int k = 0; k < LSX_K_COUNT; ++k) {
    for (auto& pr : ev_rec[k]) {
      float ms = 0.f;
      LSX_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
      ev_ms[k] += ms;
      ev_count[k] += 1;
    }
    ev_rec[k].clear();
  }
  ev_used = 0;
}

void Solver::set_timing(bool on) {
  if (timing) drain_events();
  timing = on;
  for (int k = 0; k < LSX_K_COUNT; ++k) {
    ev_ms[k] = 0.0;
    ev_bytes[k] = 0.0;
    ev_count[k] = 0;
  }
}

double Solver::kernel_bytes(int k) const {
  // compulsory HBM bytes per fine cell: each array read or written once (operator stored as
  // the four planes a_p, a_e, a_n, a_t of the symmetric 7-point stencil)
  const double n = (double)d0.n();
  switch (k) {
    case LSX_K_RESIDUAL: return 56.0 * n;  // c 8, coefficients 32, r read + write 16
    case LSX_K_LINE_X:
    case LSX_K_LINE_Y:
    case LSX_K_LINE_Z: return 56.0 * n;    // r 8, two factor planes 16, sub coupling 8, ch 16, c 8
    case LSX_K_OPERATOR:                   // coefficients 32, w 8, z 8 (FP64) or 4 (FP32)
      return (mixed() ? 44.0 : 48.0) * n;
    case LSX_K_ORTHO: return 32.0 * n;     // w read + write 16, v_i 8, v_i+1 8
    case LSX_K_SMOOTH:  // one red-black sweep of the finest level: x, b, 1/a_p, three faces, x written
      return (opt.precision == LSX_PREC_MIXED16 ? 22.0 : (mixed() ? 28.0 : 56.0)) * n;
    default: return 0.0;
  }
}

void Solver::kernel_stats(int k, long* launches, double* total_ms, double* bytes) {
  if (k < 0 || k >= LSX_K_COUNT) throw Error(LSX_ERR_ARG, "kernel id out of range");
  if (timing) drain_events();
  if (launches) *launches = ev_count[k];
  if (total_ms) *total_ms = ev_ms[k];
  if (bytes) *bytes = ev_count[k] ? ev_bytes[k] / ev_count[k] : kernel_bytes(k);
}

void Solver::synchronize() { LSX_HIP(hipStreamSynchronize(stream)); }

// -----------------------------------------------------------------------------------------------
// setup
// -----------------------------------------------------------------------------------------------
// cells of the next coarser level along an axis: n / 2, or (n + 1) / 2 for odd n
// (coarse_grid.h:60-67)
static int coarse_count(int n) { return (n & 1) ? (n + 1) / 2 : n / 2; }

static int level_count(int N, int M, int K) {
  // coarse_grids_hierarchy.h:73-81: halve the smallest dimension while it exceeds 2
  double aux = (double)std::min(std::min(N, M), K);
  int num = 0;
  while (aux > 2.0) {
    aux *= 0.5;
    ++num;
  }
  return num;
}

// coarsening steps of the hierarchy: the reference's automatic count, or (V-cycle) fewer when the
// options cap the number of levels (opt.levels counts the finest level too)
int Solver::coarsenings(int N, int M, int K) const {
  const int nmg = level_count(N, M, K);
  return opt.levels > 0 ? std::min(nmg, opt.levels - 1) : nmg;
}

void Solver::check_dims(int N, int M, int K) const {
  if (N < 3 || M < 3 || K < 3) throw Error(LSX_ERR_ARG, "grid dimensions must be >= 3");
  if ((long)N * M * K >= (1L << 31)) throw Error(LSX_ERR_ARG, "grid too large for 32-bit row indices");
  const int nmg = coarsenings(N, M, K);
  int n = N, m = M, k = K;
  for (int l = 0; l < nmg; ++l) {
    const int odd = (n & 1) + (m & 1) + (k & 1);
    // the reference's mixed-parity transfer classes (R = P^T with the cell-centred prolongation on
    // their even axes) widen the coarse operator beyond the 27-point stencils stored here; the
    // V-cycle coarsens any level (an odd axis ends on a coarse cell with one child)
    if (!vcyc() && (odd == 1 || odd == 2))
      throw Error(LSX_ERR_ARG,
                  "a multigrid level mixes even and odd dimensions: only even-even-even and odd-odd-odd transfers "
                  "are supported");
    n = coarse_count(n);
    m = coarse_count(m);
    k = coarse_count(k);
  }
  if (single_level()) return;  // no coarse levels: any dimensions, no coarsest solve
  // the cascade's coarsest LU holds 64 unknowns; the V-cycle solves a larger coarsest level (a
  // capped hierarchy, or the automatic one of a grid with one short axis) by its dense inverse
  if (!vcyc() && (long)n * m * k > 64) throw Error(LSX_ERR_ARG, "coarsest level exceeds 64 unknowns");
  if ((long)n * m * k > vc::DENSE_MAX)
    throw Error(LSX_ERR_ARG, "the coarsest level has more than 4096 unknowns (too few levels for the grid)");
}

void Solver::setup_levels(int N, int M, int K) {
  const int nmg = coarsenings(N, M, K);
  d0 = Dims{N, M, K};
  levels.clear();
  levels.resize(nmg + 1);
  {
    int nn = N, mm = M, kk = K;
    for (int l = 0; l <= nmg; ++l) {
      levels[l].d = Dims{nn, mm, kk};
      levels[l].npts = l == 0 ? 7 : 27;
      nn = coarse_count(nn);
      mm = coarse_count(mm);
      kk = coarse_count(kk);
    }
  }
  for (size_t l = 0; l < levels.size(); ++l) {
    MGLevel& L = levels[l];
    const long nl = L.d.n();
    if (vcyc()) {  // the V-cycle keeps its own compact levels; only the FP64 fine operator here
      if (l == 0) L.coef.alloc((size_t)L.npts * nl);
      continue;
    }
    L.coef.alloc((size_t)L.npts * nl);
    L.r.alloc(nl);
    if (l > 0) L.ch.alloc(nl);
    if (l + 1 < levels.size() || single_level())
      for (int dir = 0; dir < 3; ++dir) {
        L.dinv[dir].alloc(nl);
        L.eu[dir].alloc(nl);
      }
  }
  if (!vcyc()) cwork.alloc(d0.n());
  const long nc = levels.back().d.n();
  dense_c.release();
  dense_r.release();
  dense_part.release();
  if (nc <= 64 && !single_level()) {
    lu_scale.alloc(nc);
    lu_perm.alloc(nc);
    lu_L.alloc(nc * nc);
    lu_U.alloc(nc * nc);
    lu_col.alloc(nc);
  }
}

void Solver::set_geometry(const Geometry& g, int pin) {
  check_dims(g.N, g.M, g.K);
  ready = false;
  values_dirty = false;
  setup_levels(g.N, g.M, g.K);
  src_geo = nullptr;
  g.laplace7(pin, levels[0].coef.p, stream);
  from_csr = false;
  build_hierarchy();
  alloc_krylov();
  ready = true;
  src_geo = &g;
  src_version = g.version();
  src_pin = pin;
}

void Solver::set_helmholtz(const Geometry& g, int dir, double nu, double coef_time, int flow) {
  if (dir < 0 || dir > 2) throw Error(LSX_ERR_ARG, "velocity component must be 0 (u), 1 (v) or 2 (w)");
  const int N = g.N - (dir == 0), M = g.M - (dir == 1), K = g.K - (dir == 2);
  check_dims(N, M, K);
  src_geo = nullptr;
  ready = false;
  values_dirty = false;
  setup_levels(N, M, K);
  g.helmholtz7(dir, nu, coef_time, flow, levels[0].coef.p, stream);
  LSX_HIP(hipStreamSynchronize(stream));
  from_csr = false;
  build_hierarchy();
  alloc_krylov();
  ready = true;
}

void Solver::set_fields(int N, int M, int K, const double* dx, const double* dy, const double* dz, const int* type,
                        const double* thx, const double* thy, const double* thz, const double* vof, int pin, int mem) {
  check_dims(N, M, K);
  if (!dx || !dy || !dz || !type || !thx || !thy || !thz || !vof) throw Error(LSX_ERR_ARG, "null mesh field");
  if (mem != LSX_MEM_HOST && mem != LSX_MEM_DEVICE) throw Error(LSX_ERR_ARG, "mem must be LSX_MEM_HOST or LSX_MEM_DEVICE");
  src_geo = nullptr;
  ready = false;
  values_dirty = false;
  const long n = (long)N * M * K;
  // device copies of the fields (one staging allocation; the operator keeps nothing of them)
  DevBuf<double> w3, f4;
  DevBuf<int> t;
  w3.alloc(N + M + K);
  f4.alloc(4 * n);
  t.alloc(n);
  const hipMemcpyKind kind = mem == LSX_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  LSX_HIP(hipMemcpyAsync(w3.p, dx, N * sizeof(double), kind, stream));
  LSX_HIP(hipMemcpyAsync(w3.p + N, dy, M * sizeof(double), kind, stream));
  LSX_HIP(hipMemcpyAsync(w3.p + N + M, dz, K * sizeof(double), kind, stream));
  LSX_HIP(hipMemcpyAsync(f4.p, thx, n * sizeof(double), kind, stream));
  LSX_HIP(hipMemcpyAsync(f4.p + n, thy, n * sizeof(double), kind, stream));
  LSX_HIP(hipMemcpyAsync(f4.p + 2 * n, thz, n * sizeof(double), kind, stream));
  LSX_HIP(hipMemcpyAsync(f4.p + 3 * n, vof, n * sizeof(double), kind, stream));
  LSX_HIP(hipMemcpyAsync(t.p, type, n * sizeof(int), kind, stream));
  setup_levels(N, M, K);
  geo::laplace7(t.p, f4.p, f4.p + n, f4.p + 2 * n, f4.p + 3 * n, w3.p, w3.p + N, w3.p + N + M, N, M, K, pin,
                levels[0].coef.p, stream);
  LSX_HIP(hipGetLastError());
  LSX_HIP(hipStreamSynchronize(stream));
  from_csr = false;
  build_hierarchy();
  alloc_krylov();
  ready = true;
}

void Solver::set_csr(int N, int M, int K, const int* rp, const int* cl, const double* vl, long nnz, int mem) {
  check_dims(N, M, K);
  src_geo = nullptr;
  if (!rp || !cl || !vl || nnz <= 0) throw Error(LSX_ERR_ARG, "null CSR array or empty matrix");
  if (mem != LSX_MEM_HOST && mem != LSX_MEM_DEVICE) throw Error(LSX_ERR_ARG, "mem must be LSX_MEM_HOST or LSX_MEM_DEVICE");
  ready = false;
  values_dirty = false;
  const long n = (long)N * M * K;
  if (mem == LSX_MEM_HOST) {
    // rows must be well formed before anything touches the device
    if (rp[0] != 0 || rp[n] != nnz) throw Error(LSX_ERR_ARG, "row pointers inconsistent with nnz");
    for (long i = 0; i < n; ++i)
      if (rp[i + 1] < rp[i] || rp[i + 1] - rp[i] > 7) throw Error(LSX_ERR_ARG, "row pointers not a 7-point pattern");
    for (long j = 0; j < nnz; ++j)
      if (cl[j] < 0 || cl[j] >= n) throw Error(LSX_ERR_ARG, "column index out of range");
  }
  nnz0 = nnz;
  rowptr.alloc(n + 1);
  col.alloc(nnz);
  val.alloc(nnz);
  const hipMemcpyKind kind = mem == LSX_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  LSX_HIP(hipMemcpyAsync(rowptr.p, rp, (n + 1) * sizeof(int), kind, stream));
  LSX_HIP(hipMemcpyAsync(col.p, cl, nnz * sizeof(int), kind, stream));
  LSX_HIP(hipMemcpyAsync(val.p, vl, nnz * sizeof(double), kind, stream));
  if (mem == LSX_MEM_DEVICE) {
    // validate the device CSR on the host before any kernel indexes with it
    std::vector<int> hrp(n + 1), hcl(nnz);
    LSX_HIP(hipMemcpyAsync(hrp.data(), rp, (n + 1) * sizeof(int), hipMemcpyDeviceToHost, stream));
    LSX_HIP(hipMemcpyAsync(hcl.data(), cl, nnz * sizeof(int), hipMemcpyDeviceToHost, stream));
    LSX_HIP(hipStreamSynchronize(stream));
    if (hrp[0] != 0 || hrp[n] != nnz) throw Error(LSX_ERR_ARG, "row pointers inconsistent with nnz");
    for (long i = 0; i < n; ++i)
      if (hrp[i + 1] < hrp[i] || hrp[i + 1] - hrp[i] > 7) throw Error(LSX_ERR_ARG, "row pointers not a 7-point pattern");
    for (long j = 0; j < nnz; ++j)
      if (hcl[j] < 0 || hcl[j] >= n) throw Error(LSX_ERR_ARG, "column index out of range");
  }
  setup_levels(N, M, K);
  from_csr = true;
  build_hierarchy();
  alloc_krylov();
  ready = true;
}

void Solver::update_values(const double* vl, int mem) {
  if (!ready || !from_csr) throw Error(LSX_ERR_STATE, "update_values needs a matrix set with set_csr");
  if (!vl) throw Error(LSX_ERR_ARG, "null values");
  if (mem != LSX_MEM_HOST && mem != LSX_MEM_DEVICE) throw Error(LSX_ERR_ARG, "mem must be LSX_MEM_HOST or LSX_MEM_DEVICE");
  const hipMemcpyKind kind = mem == LSX_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  LSX_HIP(hipMemcpyAsync(val.p, vl, nnz0 * sizeof(double), kind, stream));
  LSX_HIP(hipStreamSynchronize(stream));  // the caller may reuse its host array on return
  values_dirty = true;  // the operator and hierarchy still hold the old values until rebuild()
}

void Solver::rebuild() {
  if (!ready) throw Error(LSX_ERR_STATE, "rebuild before set_csr");
  build_hierarchy();
  values_dirty = false;
}

void Solver::require_ready(const char* what) const {
  if (!ready) throw Error(LSX_ERR_STATE, std::string(what) + " before the matrix is set");
  if (values_dirty)
    throw Error(LSX_ERR_STATE, std::string(what) + " after update_values without rebuild: the operator is stale");
}

void Solver::build_hierarchy() {
  clear_graphs();
  MGLevel& F = levels[0];
  if (from_csr) {
    LSX_HIP(hipMemsetAsync(err.p, 0, sizeof(int), stream));
    csr_to_stencil7(rowptr.p, col.p, val.p, F.d, F.coef.p, err.p, stream);
    int herr = 0;
    LSX_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    LSX_HIP(hipStreamSynchronize(stream));
    if (herr & 1) throw Error(LSX_ERR_ARG, "matrix has a column outside the 7-point neighbourhood of its row");
    if (herr & 2) throw Error(LSX_ERR_ARG, "matrix has a row that names the same column twice");
  }
  sym_setup();
  if (vcyc()) {
    build_vcycle();
    return;
  }
  for (size_t l = 0; l + 1 < levels.size(); ++l) {
    MGLevel& f = levels[l];
    MGLevel& c = levels[l + 1];
    galerkin(f.coef.p, f.npts, f.d, c.coef.p, c.d, stream);
    diag_absmax(f.coef.p, f.npts, f.d, partial.p, scalars.p + SC_TINY, stream);
    for (int dir = 0; dir < 3; ++dir)
      line_factor(f.coef.p, f.npts, f.d, dir, scalars.p + SC_TINY, f.dinv[dir].p, f.eu[dir].p, stream);
  }
  if (single_level()) {  // one level: its line factors, no coarsest LU
    MGLevel& f = levels[0];
    diag_absmax(f.coef.p, f.npts, f.d, partial.p, scalars.p + SC_TINY, stream);
    for (int dir = 0; dir < 3; ++dir)
      line_factor(f.coef.p, f.npts, f.d, dir, scalars.p + SC_TINY, f.dinv[dir].p, f.eu[dir].p, stream);
  } else {
    MGLevel& C = levels.back();
    lu_factor(C.coef.p, C.d, lu_scale.p, lu_perm.p, lu_L.p, lu_U.p, lu_col.p, stream);
  }
  LSX_HIP(hipGetLastError());
  LSX_HIP(hipStreamSynchronize(stream));
}

// -----------------------------------------------------------------------------------------------
// V-cycle hierarchy: faces of the fine operator, aggregated coarse faces (half the sum of the
// four fine faces of each coarse face), Neumann diagonals, the pinned cells as Dirichlet sinks
// on coarse levels, FP64 LU on the coarsest.  Mixed precision stores the finest faces in FP32
// and coarse faces in FP16 divided by a per-level scale (stored values within [0, 16384]).
// -----------------------------------------------------------------------------------------------
void Solver::build_vcycle() {
  const int L = (int)levels.size() - 1;
  if (opt.smoother == LSX_SMOOTH_LINE && (d0.N > 512 || d0.M > 256 || d0.K > 256))
    throw Error(LSX_ERR_ARG, "line relaxation handles x lines of at most 512 cells and y and z lines of at most 256");
  if (opt.smoother == LSX_SMOOTH_LINE && d0.n() * (mixed() ? 4L : 8L) >= (1L << 32))
    throw Error(LSX_ERR_ARG, "line relaxation addresses vectors with 32-bit byte offsets: grid too large");
  vlev.clear();
  vlev.resize(L + 1);
  std::vector<DevBuf<double>> fe(2), fn(2), ft(2), dg(2);
  int cur = 0;
  {
    const long n = d0.n();
    fe[0].alloc(n);
    fn[0].alloc(n);
    ft[0].alloc(n);
    dg[0].alloc(n);
    // levels 1, 3, ... alternate into the second set (level 1 is the largest of them; a level
    // of odd size holds more than an eighth of the one above it)
    const long n1 = L >= 1 ? levels[1].d.n() : 1;
    fe[1].alloc(n1);
    fn[1].alloc(n1);
    ft[1].alloc(n1);
    dg[1].alloc(n1);
  }
  const bool m32 = mixed();
  double* sigma = scalars.p + SC_SIGMA;
  if (pins.size() > 2) throw Error(LSX_ERR_ARG, "the V-cycle takes at most two pinned (identity) rows");
  // without pinned rows the operator must carry a diagonal beyond its couplings (the mass term of
  // a momentum operator); coarse levels then take the sum of their children's excess
  const bool massive = pins.empty();
  std::vector<DevBuf<double>> ex(massive ? 2 : 0);
  if (massive) {
    ex[0].alloc(d0.n());
    ex[1].alloc(L >= 1 ? levels[1].d.n() : 1);
  }
  auto coarse_of = [&](int p, int l) {  // the level-l cell that contains fine cell p
    if (p < 0) return -1;
    const int x = p % d0.N, y = (p / d0.N) % d0.M, z = p / (d0.N * d0.M);
    const Dims& d = levels[l].d;
    return (x >> l) + d.N * ((y >> l) + d.M * (z >> l));
  };
  const int fp0 = pins.size() > 0 ? pins[0] : -1, fp1 = pins.size() > 1 ? pins[1] : -1;
  DevBuf<double> dev;  // per-row deviation from the diagonal-from-faces rule
  if (opt.diag_faces) dev.alloc(d0.n());
  double sig_h[2] = {0.0, 0.0};
  for (int l = 0; l <= L; ++l) {
    const Dims d = levels[l].d;
    const long nl = d.n();
    // the finest level keeps the pinned rows as identity rows; coarse levels take them as sinks
    const int pin0 = l == 0 ? fp0 : -1, pin1 = l == 0 ? fp1 : -1;
    VStore& S = vlev[l];
    S.L.d = d;
    S.L.pin0 = pin0;
    S.L.pin1 = pin1;
    if (l == 0) {
      vc::faces_from7(levels[0].coef.p, d, pin0, pin1, fe[cur].p, fn[cur].p, ft[cur].p, dg[cur].p, stream);
      vc::pin_sigma(fe[cur].p, fn[cur].p, ft[cur].p, d, pin0, pin1, sigma, stream);
      sig_h[0] = reduce_to_host(sigma);
      sig_h[1] = reduce_to_host(sigma + 1);
      if (massive) {
        vc::excess(fe[cur].p, fn[cur].p, ft[cur].p, dg[cur].p, d, ex[cur].p, stream);
        vc::face_absmax(ex[cur].p, ex[cur].p, ex[cur].p, nl, partial.p, scalars.p + SC_TMP, stream);
        const double emax = reduce_to_host(scalars.p + SC_TMP);
        vc::face_absmax(fe[cur].p, fn[cur].p, ft[cur].p, nl, partial.p, scalars.p + SC_TMP, stream);
        const double fmax = reduce_to_host(scalars.p + SC_TMP);
        if (!(emax > 1e-8 * fmax))
          throw Error(LSX_ERR_ARG,
                      "the V-cycle needs one or two pinned (identity) rows, or a diagonal beyond the couplings "
                      "(a mass term): the operator is singular");
      }
    } else {
      const Dims df = levels[l - 1].d;
      const int nxt = cur ^ 1;
      vc::aggregate(fe[cur].p, fn[cur].p, ft[cur].p, df, fe[nxt].p, fn[nxt].p, ft[nxt].p, d, stream);
      if (massive) vc::aggregate_cells(ex[cur].p, df, ex[nxt].p, d, stream);
      vc::diag_from_faces(fe[nxt].p, fn[nxt].p, ft[nxt].p, d, massive ? ex[nxt].p : nullptr, dg[nxt].p, stream);
      vc::pin_sink(dg[nxt].p, coarse_of(fp0, l), coarse_of(fp1, l), sigma, stream);
      cur = nxt;
    }
    // FP16 faces (mixed: coarse levels; mixed16: every level) are stored divided by a per-level
    // scale that puts the largest at 16384
    const bool f16 = m32 && !(l == 0 && opt.precision != LSX_PREC_MIXED16);
    double fscale = 1.0;
    if (l < L && f16) {
      vc::face_absmax(fe[cur].p, fn[cur].p, ft[cur].p, nl, partial.p, scalars.p + SC_TMP, stream);
      const double amax = reduce_to_host(scalars.p + SC_TMP);
      fscale = amax > 0 ? amax / 16384.0 : 1.0;
    }
    if (l < L && opt.diag_faces && nl > vc::COARSE_CELLS) {
      // a_p = -(sum of the faces) - sigma on every row but the pinned ones, rows without faces
      // holding 0 or one common value: true of every pressure operator (a Neumann Laplacian whose
      // solid cells carry a tiny diagonal) and of its aggregated levels; an operator with a mass
      // term (a momentum operator) fails the check and its levels read 1/a_p.  The small levels
      // keep 1/a_p, so that their per-level launches stay bit-identical to the one-workgroup
      // coarse cycle.
      const int c0 = l == 0 ? -1 : coarse_of(fp0, l), c1 = l == 0 ? -1 : coarse_of(fp1, l);
      double* out = scalars.p + SC_TMP;
      auto rule = [&](double dz, int mode) {
        vc::diag_rule(fe[cur].p, fn[cur].p, ft[cur].p, dg[cur].p, d, pin0, pin1, c0, c1, sig_h[0], sig_h[1], dz,
                      f16 ? 1.0 / fscale : 0.0, mode, dev.p, partial.p, out, stream);
      };
      rule(0.0, 0);
      const double dz = reduce_to_host(out);
      rule(dz, 1);
      if (reduce_to_host(out) <= 1e-12) {
        rule(dz, 2);  // -0.0 own east faces of the rows without faces whose a_p is dz
        S.L.dff = 1;
        S.L.sink0 = c0;
        S.L.sink1 = c1;
        S.L.sig0 = sig_h[0];
        S.L.sig1 = sig_h[1];
        S.L.dz = dz;
      }
    }
    if (l == L) {  // coarsest: FP64 LU of the 7-point operator, or its dense inverse (capped levels)
      if (nl <= 64) {
        DevBuf<double> c27;
        c27.alloc(27 * nl);
        vc::to27(fe[cur].p, fn[cur].p, ft[cur].p, dg[cur].p, d, pin0, pin1, c27.p, stream);
        lu_factor(c27.p, d, lu_scale.p, lu_perm.p, lu_L.p, lu_U.p, lu_col.p, stream);
        LSX_HIP(hipStreamSynchronize(stream));
      } else {
        DevBuf<double> work;
        work.alloc(2 * nl * nl + 2 * nl + 1);
        dense_c.alloc(nl * nl);
        dense_r.alloc(nl * nl);
        dense_part.alloc(nl * ((nl + 127) / 128));
        vc::dense_inverse(fe[cur].p, fn[cur].p, ft[cur].p, dg[cur].p, d, dense_c.p, dense_r.p, work.p, stream);
        LSX_HIP(hipStreamSynchronize(stream));
      }
      S.db.alloc(nl);
      S.dx.alloc(nl);
      if (m32) {
        S.fb.alloc(nl);
        S.fx.alloc(nl);
      }
      LSX_HIP(hipStreamSynchronize(stream));
      break;
    }
    if (!rb_level(l)) {  // ping-pong buffer of the line and Jacobi passes
      if (m32)
        S.fxt.alloc(nl);
      else
        S.dxt.alloc(nl);
    } else if (fused_level(l)) {  // second iterate buffer of the one-pass sweeps
      if (m32)
        S.fxt.alloc(nl);
      else
        S.dxt.alloc(nl);
    }
    if (m32) {
      S.fs.alloc(nl);
      S.fb.alloc(nl);
      if (l > 0) S.fx.alloc(nl);  // level 0 writes into the caller's output
      vc::inv_diag<float>(dg[cur].p, nl, S.fs.p, stream);
      S.L.s = S.fs.p;
    } else {
      S.ds.alloc(nl);
      if (l > 0) {  // level 0 reads the caller's input and writes the caller's output
        S.db.alloc(nl);
        S.dx.alloc(nl);
      }
      vc::inv_diag<double>(dg[cur].p, nl, S.ds.p, stream);
      S.L.s = S.ds.p;
    }
    if (!m32) {
      S.de.alloc(nl);
      S.dn.alloc(nl);
      S.dt.alloc(nl);
      vc::store_faces<double>(fe[cur].p, fn[cur].p, ft[cur].p, nl, 1.0, S.de.p, S.dn.p, S.dt.p, stream);
      S.L.ftype = vc::FACE_F64;
      S.L.f_e = S.de.p;
      S.L.f_n = S.dn.p;
      S.L.f_t = S.dt.p;
      S.L.fscale = 1.0;
    } else if (l == 0 && opt.precision != LSX_PREC_MIXED16) {
      S.fe.alloc(nl);
      S.fn.alloc(nl);
      S.ft.alloc(nl);
      vc::store_faces<float>(fe[cur].p, fn[cur].p, ft[cur].p, nl, 1.0, S.fe.p, S.fn.p, S.ft.p, stream);
      S.L.ftype = vc::FACE_F32;
      S.L.f_e = S.fe.p;
      S.L.f_n = S.fn.p;
      S.L.f_t = S.ft.p;
      S.L.fscale = 1.0;
    } else {
      const double scale = fscale;
      S.he.alloc(nl);
      S.hn.alloc(nl);
      S.ht.alloc(nl);
      vc::store_faces<_Float16>(fe[cur].p, fn[cur].p, ft[cur].p, nl, 1.0 / scale, S.he.p, S.hn.p, S.ht.p, stream);
      S.L.ftype = vc::FACE_F16;
      S.L.f_e = S.he.p;
      S.L.f_n = S.hn.p;
      S.L.f_t = S.ht.p;
      S.L.fscale = (float)scale;
    }
    LSX_HIP(hipStreamSynchronize(stream));  // the ping-pong temporaries are reused next level
  }
  LSX_HIP(hipGetLastError());
  if (m32) ztmp32.alloc(d0.n());
}

template <>
float* Solver::vx<float>(VStore& S) { return S.fx.p; }
template <>
double* Solver::vx<double>(VStore& S) { return S.dx.p; }
template <>
float* Solver::vb<float>(VStore& S) { return S.fb.p; }
template <>
double* Solver::vb<double>(VStore& S) { return S.db.p; }

template <>
float* Solver::vxt<float>(VStore& S) { return S.fxt.p; }
template <>
double* Solver::vxt<double>(VStore& S) { return S.dxt.p; }
template <>
float* Solver::vr<float>(VStore& S) { return S.fr.p; }
template <>
double* Solver::vr<double>(VStore& S) { return S.dr.p; }

template <class T>
void Solver::vcycle_apply(const double* v, T* z, const float* b32) {
  if (trans) {
    vcycle_apply_tr<T>(v, z);
    return;
  }
  // per-kernel timing records events around the launches: such cycles run eagerly
  if (!opt.graphs || capturing || timing) {
    vcycle_run<T>(v, z, b32);
    return;
  }
  // one captured graph per (v, z) pair: the Krylov loops apply the cycle to a fixed set of
  // basis slots, so a graph replays every launch of a cycle with one submission
  for (const CycleGraph& g : cgraphs)
    if (g.v == v && g.z == (const void*)z && g.b32 == b32) {
      LSX_HIP(hipGraphLaunch(g.exec, stream));
      return;
    }
  hipGraph_t graph = nullptr;
  LSX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  capturing = true;
  try {
    vcycle_run<T>(v, z, b32);
  } catch (...) {
    capturing = false;
    (void)hipStreamEndCapture(stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  capturing = false;
  LSX_HIP(hipStreamEndCapture(stream, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  LSX_HIP(ie);
  cgraphs.push_back({v, (const void*)z, b32, exec});
  LSX_HIP(hipGraphLaunch(exec, stream));
}

template <class T>
void Solver::vcycle_run(const double* v, T* z, const float* b32) {
  const int L = (int)vlev.size() - 1;
  // level 0 right-hand side: the input itself in FP64, an FP32 copy in mixed precision (b32 when
  // the caller already holds one)
  const T* b0;
  if constexpr (std::is_same<T, float>::value) {
    if (b32) {
      b0 = b32;
    } else {
      vc::convert<double, float>(v, vlev[0].fb.p, d0.n(), stream);
      b0 = vlev[0].fb.p;
    }
  } else {
    b0 = v;
  }
  // levels from `lc` down are small enough for one workgroup: a single launch does their cycle
  int lc = L + 1;
  const bool fuse_ok = opt.coarse_fuse && (!dense_c.p || vlev[L].L.d.n() <= vc::COARSE_DENSE);
  for (int l = 1; l <= L && fuse_ok; ++l)
    if (rb_level(l) && vlev[l].L.d.n() <= vc::COARSE_CELLS && L - l + 1 <= vc::COARSE_MAX) {
      lc = l;
      break;
    }
  if (lc == L) lc = L + 1;  // the LU level alone gains nothing
  const int top = lc <= L ? lc : L;
  for (int l = 0; l < top; ++l) {
    VStore& S = vlev[l];
    T* x = l == 0 ? z : vx<T>(S);
    const T* b = l == 0 ? b0 : vb<T>(S);
    smooth<T>(l, x, b, opt.n_pre, true);
    vc::resid_restrict<T>(S.L, x, b, vb<T>(vlev[l + 1]), stream);
  }
  if (lc <= L) {
    vc::CoarseLevels<T> C;
    C.count = L - lc + 1;
    for (int q = 0; q < C.count; ++q) {
      C.L[q] = vlev[lc + q].L;
      C.x[q] = vx<T>(vlev[lc + q]);
      C.b[q] = vb<T>(vlev[lc + q]);
    }
    vc::coarse_cycle<T>(C, opt.n_pre, opt.n_post, lu_scale.p, lu_perm.p, lu_L.p, lu_U.p, dense_c.p, stream);
  } else {
    coarse_solve<T>(false);
  }
  for (int l = top - 1; l >= 0; --l) {
    VStore& S = vlev[l];
    T* x = l == 0 ? z : vx<T>(S);
    const T* b = l == 0 ? b0 : vb<T>(S);
    vc::prolong_add<T>(vx<T>(vlev[l + 1]), vlev[l + 1].L.d, x, S.L.d, stream);
    smooth<T>(l, x, b, opt.n_post, false, true);  // lines z, y, x: the cycle is symmetric in its line order
  }
}

// `sweeps` smoothing sweeps of level l on x (zero_guess: x is taken as 0 and not read).
// Red-black Gauss-Seidel: red then black cells.  Line relaxation: x, y and z lines in turn,
// each pass from one buffer into the other; the passes alternate between x and the level's
// second buffer so that the last one writes x.
template <class T>
void Solver::smooth(int l, T* x, const T* b, int sweeps, bool zero_guess, bool rev_lines) {
  VStore& S = vlev[l];
  // finest-level launches are timed one by one: compulsory bytes per cell of a full red-black
  // sweep or of one line pass (x in and out, b, 1/a_p, three faces; a line pass from a zero
  // start reads no x, one that rebuilds a_p from the faces no 1/a_p)
  const double vbytes = sizeof(T), fbytes = S.L.ftype == vc::FACE_F16 ? 2.0 : (S.L.ftype == vc::FACE_F32 ? 4.0 : 8.0);
  const double pass_bytes = (4.0 * vbytes + 3.0 * fbytes) * (double)S.L.d.n();
  auto line_bytes = [&](int dir, bool zero) {
    const bool dff = S.L.dff && !trans && (dir != 0 || sizeof(T) == 4);  // as vc::line_pass
    const double v = 4.0 - (dff ? 1.0 : 0.0) - (zero ? 1.0 : 0.0);
    return (v * vbytes + 3.0 * fbytes) * (double)S.L.d.n();
  };
  hipEvent_t e0 = nullptr;
  auto t0 = [&](int k) {
    if (l == 0) tbeg(k, &e0);
  };
  auto t1 = [&](int k, double bytes) {
    if (l == 0) tend(k, e0, bytes);
  };
  if (rb_level(l) && (trans || !fused_level(l))) {
    for (int s = 0; s < sweeps; ++s) {
      const bool first = zero_guess && s == 0;
      for (int h = 0; h < 2; ++h) {
        t0(LSX_K_SMOOTH);
        if (trans)  // transposed sweeps: black then red, on A^T
          vc::rbgs_tr<T>(S.L, x, b, 1 - h, first ? 1 + h : 0, stream);
        else
          vc::rbgs<T>(S.L, x, b, h, first ? 1 + h : 0, stream);
        t1(LSX_K_SMOOTH, 0.5 * pass_bytes);
      }
    }
    return;
  }
  T* xt = vxt<T>(S);
  const long n = S.L.d.n();
  if (rb_level(l)) {  // one pass per fused red-black sweep
    T* src = x;
    if (!zero_guess && (sweeps - 1) % 2 == 0) {  // the first pass must not write the buffer it reads
      vc::copy_vec<T>(x, xt, n, stream);
      src = xt;
    }
    for (int p = 1; p <= sweeps; ++p) {
      T* dst = (sweeps - p) % 2 == 0 ? x : xt;
      t0(LSX_K_SMOOTH);
      vc::rb_sweep<T>(S.L, src, dst, b, zero_guess && p == 1, stream);
      t1(LSX_K_SMOOTH, pass_bytes);
      src = dst;
    }
    return;
  }
  if (!line_level(l)) {  // damped Jacobi, one pass per sweep (the transposed sweep: the same on A^T)
    T* src = x;
    if (!zero_guess && (sweeps - 1) % 2 == 0) {  // the first pass must not write the buffer it reads
      vc::copy_vec<T>(x, xt, n, stream);
      src = xt;
    }
    for (int p = 1; p <= sweeps; ++p) {
      T* dst = (sweeps - p) % 2 == 0 ? x : xt;
      t0(LSX_K_SMOOTH);
      vc::jacobi<T>(S.L, src, dst, b, (T)opt.omega, zero_guess && p == 1, trans ? 1 : 0, stream);
      t1(LSX_K_SMOOTH, pass_bytes);
      src = dst;
    }
    return;
  }
  const int P = 3 * sweeps;
  T* src = x;
  if (!zero_guess && (P - 1) % 2 == 0) {  // the first pass must not write the buffer it reads
    vc::copy_vec<T>(x, xt, n, stream);
    src = xt;
  }
  for (int p = 1; p <= P; ++p) {
    T* dst = (P - p) % 2 == 0 ? x : xt;
    // x, y, z lines in turn, or z, y, x (rev_lines)
    const int dir = rev_lines ? 2 - (p - 1) % 3 : (p - 1) % 3;
    t0(LSX_K_LINE_X + dir);
    vc::line_pass<T>(S.L, dir, src, dst, b, zero_guess && p == 1, trans ? 1 : 0, stream);
    t1(LSX_K_LINE_X + dir, line_bytes(dir, zero_guess && p == 1));
    src = dst;
  }
}

// The transpose of vcycle_apply: its steps in reverse order, each transposed (half-sweeps in the
// opposite colour order on A^T, restriction by P^T, prolongation by the transposed sum
// restriction, the symmetric coarsest LU unchanged).  Levels are launched one by one.
template <class T>
void Solver::vcycle_apply_tr(const double* v, T* z) {
  const int L = (int)vlev.size() - 1;
  const T* b0;
  if constexpr (std::is_same<T, float>::value) {
    vc::convert<double, float>(v, vlev[0].fb.p, d0.n(), stream);
    b0 = vlev[0].fb.p;
  } else {
    b0 = v;
  }
  for (int l = 0; l < L; ++l) {
    VStore& S = vlev[l];
    if (!vr<T>(S)) {
      if constexpr (std::is_same<T, float>::value)
        S.fr.alloc(S.L.d.n());
      else
        S.dr.alloc(S.L.d.n());
    }
  }
  for (int l = 0; l < L; ++l) {
    VStore& S = vlev[l];
    T* x = l == 0 ? z : vx<T>(S);
    const T* b = l == 0 ? b0 : vb<T>(S);
    smooth<T>(l, x, b, opt.n_post, true, false);  // transposed post-smoothing (lines x, y, z)
    vc::resid_tr<T>(S.L, x, b, vr<T>(S), stream);
    vc::restrict_pt<T>(vr<T>(S), S.L.d, vb<T>(vlev[l + 1]), stream);
  }
  coarse_solve<T>(true);
  for (int l = L - 1; l >= 0; --l) {
    VStore& S = vlev[l];
    T* x = l == 0 ? z : vx<T>(S);
    const T* b = l == 0 ? b0 : vb<T>(S);
    vc::prolong_inject_add<T>(vx<T>(vlev[l + 1]), vlev[l + 1].L.d, x, S.L.d, stream);
    smooth<T>(l, x, b, opt.n_pre, false, true);  // transposed pre-smoothing (lines z, y, x)
  }
}

// x = A_c^{-1} b on the coarsest V-cycle level (the symmetric coarse operator: its LU, or the
// dense inverse of a capped hierarchy, transposed for the transposed cycle so that it is the
// exact transpose of the forward one)
template <class T>
void Solver::coarse_solve(bool transposed) {
  VStore& C = vlev.back();
  const long nc = C.L.d.n();
  if (dense_c.p) {
    vc::dense_apply<T>(transposed ? dense_r.p : dense_c.p, (int)nc, vb<T>(C), vx<T>(C), dense_part.p, stream);
    return;
  }
  if constexpr (std::is_same<T, float>::value) {
    vc::convert<float, double>(C.fb.p, C.db.p, nc, stream);
    lu_solve(lu_scale.p, lu_perm.p, lu_L.p, lu_U.p, (int)nc, C.db.p, C.dx.p, stream);
    vc::convert<double, float>(C.dx.p, C.fx.p, nc, stream);
  } else {
    lu_solve(lu_scale.p, lu_perm.p, lu_L.p, lu_U.p, (int)nc, C.db.p, C.dx.p, stream);
  }
}

void Solver::precond(const double* v, double* z64) {
  if (!vcyc()) {
    mg_apply(v, z64);
  } else if (mixed()) {
    vcycle_apply<float>(v, ztmp32.p);
    vc::convert<float, double>(ztmp32.p, z64, d0.n(), stream);
  } else {
    vcycle_apply<double>(v, z64);
  }
}

void Solver::alloc_krylov() {
  clear_graphs();
  const long n = d0.n();
  if (opt.krylov == LSX_KRYLOV_FGMRES) {
    V.alloc((size_t)opt.restart * n);
    if (mixed())
      Z32.alloc((size_t)opt.restart * n);
    else
      Z.alloc((size_t)opt.restart * n);
    // FP32 basis copies for the projections: mixed precision with classical Gram-Schmidt (the
    // orthogonalisation fgmres() picks)
    const bool cgs = (opt.ortho == LSX_ORTHO_AUTO ? vcyc() : opt.ortho != LSX_ORTHO_MGS) && opt.restart <= CGS_MAX;
    if (mixed() && opt.dot32 && cgs)
      V32.alloc((size_t)opt.restart * n);
    else
      V32.release();
    w.alloc(n);
    H.assign(opt.restart * (opt.restart + 1) / 2, 0.0);
    cs.assign(opt.restart, 0.0);
    sn.assign(opt.restart, 0.0);
    y.assign(opt.restart, 0.0);
  } else {
    bnu.alloc(n);
    bt.alloc(n);
    by.alloc(n);
    bz.alloc(n);
    br.alloc(n);
    bp.alloc(n);
    bs.alloc(n);
    brr.alloc(n);
  }
}

// -----------------------------------------------------------------------------------------------
// fine operator: symmetric four-plane storage [a_p, a_e, a_n, a_t] when the operator is symmetric
// off its identity-like (pinned) rows, which take their faces from the coupled side; the same
// values in the same order as the 7-plane rows, so results are bit-identical.
// -----------------------------------------------------------------------------------------------
void Solver::sym_setup() {
  const long n = d0.n();
  sym_ok = false;
  sym_ex = SymExc{};
  pins.clear();
  DevBuf<unsigned char> exc;
  exc.alloc(n);
  LSX_HIP(hipMemsetAsync(err.p, 0, sizeof(int), stream));
  sym_check(levels[0].coef.p, d0, exc.p, err.p, stream);
  std::vector<unsigned char> h(n);
  int asym = 0;
  LSX_HIP(hipMemcpyAsync(h.data(), exc.p, n, hipMemcpyDeviceToHost, stream));
  LSX_HIP(hipMemcpyAsync(&asym, err.p, sizeof(int), hipMemcpyDeviceToHost, stream));
  LSX_HIP(hipStreamSynchronize(stream));
  for (long i = 0; i < n; ++i)
    if (h[i]) pins.push_back((int)i);
  if (asym || (int)pins.size() > SymExc::CAP) return;
  for (size_t q = 0; q < pins.size(); ++q) sym_ex.row[q] = pins[q];
  sym_ex.count = (int)pins.size();
  s4.alloc(4 * n);
  sym_build(levels[0].coef.p, exc.p, d0, s4.p, stream);
  LSX_HIP(hipStreamSynchronize(stream));
  sym_ok = true;
}

void Solver::fine_apply(const double* x, const double* b, double* y, int mode) {
  if (trans)
    stencil_apply_sym_t(s4.p, sym_ex, d0, x, b, y, mode, partial.p, stream);
  else if (sym_ok)
    stencil_apply_sym(s4.p, sym_ex, d0, x, b, y, mode, partial.p, stream);
  else
    stencil_apply(levels[0].coef.p, 7, d0, x, b, y, mode, partial.p, stream);
}

// -----------------------------------------------------------------------------------------------
// multigrid cascade — multigrid_solver.h:48-75
// -----------------------------------------------------------------------------------------------
void Solver::mg_apply(const double* v, double* z) {
  const int L = (int)levels.size() - 1;
  copy(v, levels[0].r.p, d0.n(), stream);
  for (int k = 0; k < L; ++k) restrict_full(levels[k].r.p, levels[k].d, levels[k + 1].r.p, levels[k + 1].d, stream);
  if (L > 0)
    lu_solve(lu_scale.p, lu_perm.p, lu_L.p, lu_U.p, (int)levels[L].d.n(), levels[L].r.p, levels[L].ch.p, stream);
  for (int k = L > 0 ? L - 1 : 0; k >= 0; --k) {
    MGLevel& lv = levels[k];
    double* ch = k == 0 ? z : lv.ch.p;
    if (L > 0) {
      prolong_copy(levels[k + 1].ch.p, levels[k + 1].d, ch, cwork.p, lv.d, stream);
    } else {  // one level: the line sweeps start from a zero correction
      fill_zero(ch, lv.d.n(), stream);
      fill_zero(cwork.p, lv.d.n(), stream);
    }
    for (int s = 0; s < opt.n_post; ++s)
      for (int dir = 2; dir >= 0; --dir) {  // z, y, x
        hipEvent_t e0 = nullptr;
        if (k == 0) tbeg(LSX_K_RESIDUAL, &e0);
        if (k == 0)
          fine_apply(cwork.p, nullptr, lv.r.p, ST_SUBR);
        else
          stencil_apply(lv.coef.p, lv.npts, lv.d, cwork.p, nullptr, lv.r.p, ST_SUBR, nullptr, stream);
        if (k == 0) {
          tend(LSX_K_RESIDUAL, e0);
          tbeg(LSX_K_LINE_X + dir, &e0);
        }
        line_solve(lv.coef.p, lv.npts, lv.d, dir, lv.dinv[dir].p, lv.eu[dir].p, lv.r.p, cwork.p, ch, opt.omega,
                   opt.z_head_fix, stream);
        if (k == 0) tend(LSX_K_LINE_X + dir, e0);
      }
  }
}

double Solver::reduce_to_host(const double* dev_scalar) {
  double h = 0.0;
  LSX_HIP(hipMemcpyAsync(&h, dev_scalar, sizeof(double), hipMemcpyDeviceToHost, stream));
  LSX_HIP(hipStreamSynchronize(stream));
  return h;
}

double Solver::dot(const double* a, const double* b, long n, double* out_dev) {
  dot_partial(a, b, n, partial.p, stream);
  finalize(partial.p, reduction_blocks(n), out_dev, 0, stream);
  return reduce_to_host(out_dev);
}

// -----------------------------------------------------------------------------------------------
// FGMRES(m) — fgmres_interface.h:78-142, solUpdate :153-168
// -----------------------------------------------------------------------------------------------
int Solver::fgmres(const double* b, double* x) {
  const long n = d0.n();
  const int m = opt.restart;
  const int m_size = m * (m + 1) / 2;
  const int nb = reduction_blocks(n);
  double* dnorm = scalars.p + SC_NORM;
  double* dH = scalars.p + SC_H;
  const double eps = opt.eps;
  const MGLevel& F = levels[0];

  const bool m32 = vcyc() && mixed();
  // orthogonalisation: the reference's MGS for the cascade; classical Gram-Schmidt (one pass, or
  // two for CGS2) for the V-cycle loops, whose windows must fit CGS_MAX
  int ortho = opt.ortho;
  if (ortho == LSX_ORTHO_AUTO) ortho = vcyc() ? LSX_ORTHO_CGS : LSX_ORTHO_MGS;
  if (ortho != LSX_ORTHO_MGS && m > CGS_MAX) ortho = LSX_ORTHO_MGS;
  const int cgs_passes = ortho == LSX_ORTHO_CGS ? 1 : (ortho == LSX_ORTHO_CGS2 ? 2 : 0);
  // The V-cycle loops with classical Gram-Schmidt keep basis vector k unnormalised, u_k = s_k v_k
  // (s_0 = 1): the update pass writes u_{k+1} = w - V h, whose norm is the Hessenberg entry
  // s_{k+1} = h_{k+1,k}, and its FP32 copy straight into the next basis slot, with no scaling
  // pass.  The cycle and the operator are linear, so z_k and A z_k carry the factor s_k, which the
  // update pass and the host remove (from w - V h, the Hessenberg column and the solution
  // coefficients).
  const bool unnorm = vcyc() && cgs_passes > 0;
  // FP32 copy of basis vector k: its own slot of V32 when the projections run over the copies,
  // else the cycle's right-hand side buffer (only the newest copy is needed)
  const bool d32 = m32 && cgs_passes > 0 && opt.dot32 && V32.p;
  auto slot32 = [&](int k) -> float* { return d32 ? V32.p + (size_t)k * n : vlev[0].fb.p; };
  std::vector<double> sc(m + 1, 1.0);
  CgsScales isq{};
  auto sol_upd = [&](int last, int lastH) {
    std::vector<double> ycoef(last + 1);
    int hNum = lastH;
    for (int k = last; k > -1; --k, --hNum) {
      const double curY = y[k] / H[hNum];
      for (int i = k - 1; i > -1; --i) {
        --hNum;
        y[i] -= curY * H[hNum];
      }
      ycoef[k] = unnorm ? curY / sc[k] : curY;
    }
    if (m32)
      sol_update_f32(x, Z32.p, n, ycoef.data(), last + 1, n, stream);
    else
      sol_update(x, Z.p, n, ycoef.data(), last + 1, n, stream);
  };

  fine_apply(x, b, w.p, ST_RESID);
  finalize(partial.p, nb, dnorm, 1, stream);
  norm_r = reduce_to_host(dnorm);
  if (!(norm_r > eps)) return 0;
  int iters = 1;
  double normW = norm_r;
  double curH = norm_r;
  std::vector<double> hcol(m + 2), hcgs(2 * CGS_MAX);
  bool resume = false;  // a V-cycle loop whose estimate converged but whose true residual did not
  for (;;) {
    // normW lives on the device for the scaling kernel
    LSX_HIP(hipMemcpyAsync(dnorm, &normW, sizeof(double), hipMemcpyHostToDevice, stream));
    for (int j = 0, hColStart = 0, hNum = 0; j < m; ++j, hColStart = hNum, ++iters) {
      double* vj = V.p + (size_t)j * n;
      double* zj = m32 ? nullptr : Z.p + (size_t)j * n;
      float* zj32 = m32 ? Z32.p + (size_t)j * n : nullptr;
      const bool scaled = !unnorm || j == 0;  // else u_j (and its FP32 copy) came from the update
      if (scaled) sc[j] = 1.0;
      if (m32) {
        // v_j = w / ||w|| and its FP32 copy (the cycle's finest right-hand side) in one pass
        if (scaled) scale_copy_f32(w.p, dnorm, vj, slot32(j), n, stream);
        vcycle_apply<float>(vj, zj32, slot32(j));
      } else {
        if (scaled) scale_copy(w.p, dnorm, vj, n, stream);
        if (vcyc())
          vcycle_apply<double>(vj, zj);
        else
          mg_apply(vj, zj);
      }
      hipEvent_t e0 = nullptr;
      tbeg(LSX_K_OPERATOR, &e0);
      if (m32) {
        if (trans)
          stencil_prod_sym_t_f32x(s4.p, sym_ex, d0, zj32, w.p, stream);
        else if (sym_ok)
          stencil_prod_sym_f32x(s4.p, sym_ex, d0, zj32, w.p, stream);
        else
          stencil7_prod_f32x(F.coef.p, d0, zj32, w.p, stream);
      } else {
        fine_apply(zj, nullptr, w.p, ST_PROD);
      }
      tend(LSX_K_OPERATOR, e0);
      if (cgs_passes) {
        // classical Gram-Schmidt: h = V^T w ; w -= V h (twice for CGS2), the last update pass
        // yielding ||w||.  Two or four streaming passes over V instead of j + 1 MGS passes.
        const int cnt = j + 1;
        const double vb = 8.0 * cnt * (double)n;
        for (int pass = 0; pass < cgs_passes; ++pass) {
          const bool last = pass + 1 == cgs_passes;
          double* hp = cgs_h.p + pass * CGS_MAX;
          tbeg(LSX_K_CGS_DOT, &e0);
          if (d32)
            cgs_dot_f32(V32.p, n, cnt, w.p, n, mpartial.p, hp, stream);
          else
            cgs_dot(V.p, n, cnt, w.p, n, mpartial.p, hp, stream);
          tend(LSX_K_CGS_DOT, e0, (d32 ? 0.5 * vb : vb) + 8.0 * n);
          tbeg(LSX_K_CGS_UPDATE, &e0);
          if (unnorm) {
            // the last pass writes u_{j+1} into the next slot (and the cycle's FP32 right-hand
            // side); the window's last step only needs the norm
            for (int k = 0; k < cnt; ++k) isq.isq[k] = 1.0 / (sc[k] * sc[k]);
            const bool to_next = last && j + 1 < m;
            double* out = to_next ? V.p + (size_t)(j + 1) * n : w.p;
            float* out32 = to_next && m32 ? slot32(j + 1) : nullptr;
            cgs_update_scaled(V.p, n, cnt, hp, isq, last ? 1.0 / sc[j] : 1.0, w.p, out, out32, n, mpartial.p,
                              last ? dnorm : nullptr, stream);
            tend(LSX_K_CGS_UPDATE, e0, vb + (out32 ? 20.0 : 16.0) * n);
          } else {
            cgs_update(V.p, n, cnt, hp, w.p, n, mpartial.p, last ? dnorm : nullptr, stream);
            tend(LSX_K_CGS_UPDATE, e0, vb + 16.0 * n);
          }
        }
        hNum = hColStart + j + 1;
        LSX_HIP(hipMemcpyAsync(hcgs.data(), cgs_h.p, 2 * CGS_MAX * sizeof(double), hipMemcpyDeviceToHost, stream));
        LSX_HIP(hipMemcpyAsync(&normW, dnorm, sizeof(double), hipMemcpyDeviceToHost, stream));
        LSX_HIP(hipStreamSynchronize(stream));
        for (int i = 0; i <= j; ++i) hcol[i] = cgs_passes == 2 ? hcgs[i] + hcgs[CGS_MAX + i] : hcgs[i];
        if (unnorm) {  // back to the orthonormal basis: h_i / (s_i s_j); u_{j+1} = ||w|| v_{j+1}
          for (int i = 0; i <= j; ++i) hcol[i] /= sc[i] * sc[j];
          sc[j + 1] = normW;
        }
      } else {
        // modified Gram-Schmidt: h_i = w . v_i ; w -= h_i v_i  (the dot of the next basis vector
        // is fused into the update pass; the last pass yields ||w||^2)
        dot_partial(w.p, V.p, n, partial.p, stream);
        finalize(partial.p, nb, dH + hColStart, 0, stream);
        for (int i = 0; i <= j; ++i) {
          const double* vi = V.p + (size_t)i * n;
          const double* vnext = i < j ? V.p + (size_t)(i + 1) * n : nullptr;
          if (vnext) tbeg(LSX_K_ORTHO, &e0);
          axpy_dot(w.p, vi, dH + hColStart + i, vnext, n, partial.p, stream);
          if (vnext) tend(LSX_K_ORTHO, e0);
          if (vnext)
            finalize(partial.p, nb, dH + hColStart + i + 1, 0, stream);
          else
            finalize(partial.p, nb, dnorm, 1, stream);
        }
        hNum = hColStart + j + 1;
        LSX_HIP(hipMemcpyAsync(hcol.data(), dH + hColStart, (j + 1) * sizeof(double), hipMemcpyDeviceToHost, stream));
        LSX_HIP(hipMemcpyAsync(&normW, dnorm, sizeof(double), hipMemcpyDeviceToHost, stream));
        LSX_HIP(hipStreamSynchronize(stream));
      }
      for (int i = 0; i <= j; ++i) H[hColStart + i] = hcol[i];
      curH = H[hColStart];
      for (int i = 0, cur = hColStart, nxt = cur + 1; i < j; ++i, cur = nxt, ++nxt) {
        const double cc = cs[i], ss = sn[i], nextH = H[nxt];
        H[cur] = cc * curH - ss * nextH;
        curH = ss * curH + cc * nextH;
      }
      double curC = curH * curH + normW * normW;
      const double curA = std::sqrt(curC);
      curC = 1.0 / curA;
      const double curS = -normW * curC;
      curC *= curH;
      H[hNum - 1] = curA;
      y[j] = curC * norm_r;
      norm_r *= curS;
      curH = std::fabs(norm_r);
      if ((curH > eps) && (iters < opt.max_it)) {
        cs[j] = curC;
        sn[j] = curS;
      } else {
        norm_r = curH;
        sol_upd(j, hNum - 1);
        if (!vcyc()) return iters;
        // the V-cycle loops confirm convergence on the true residual and carry on from it when
        // rounding in the basis has let the estimate drift below it
        fine_apply(x, b, w.p, ST_RESID);
        finalize(partial.p, nb, dnorm, 1, stream);
        norm_r = reduce_to_host(dnorm);
        if (!(norm_r > eps) || iters >= opt.max_it) return iters;
        resume = true;
        break;
      }
    }
    if (resume) {
      resume = false;
      normW = norm_r;
      ++iters;
      continue;
    }
    sol_upd(m - 1, m_size - 1);
    // w = b - A x starts the next cycle.  The reference scales it by the carried estimate
    // |gamma| and keeps the sign of gamma; the V-cycle loops restart from the true residual.
    fine_apply(x, b, w.p, ST_RESID);
    if (vcyc()) {
      finalize(partial.p, nb, dnorm, 1, stream);
      normW = reduce_to_host(dnorm);
      norm_r = normW;
    } else {
      normW = curH;
    }
  }
}

// -----------------------------------------------------------------------------------------------
// BiCGStab — bicgstab_interface.h:76-111
// -----------------------------------------------------------------------------------------------
int Solver::bicgstab(const double* b, double* x) {
  const long n = d0.n();
  const int nb = reduction_blocks(n);
  double* tmp = scalars.p + SC_TMP;
  double* dnorm = scalars.p + SC_NORM;
  int iters = 1;
  fine_apply(x, nullptr, bz.p, ST_PROD);
  sub_vec(b, bz.p, br.p, n, stream);
  norm_r = std::sqrt(dot(br.p, br.p, n, tmp));
  if (norm_r > opt.eps) {
    copy(br.p, brr.p, n, stream);
    copy(br.p, bp.p, n, stream);
    do {
      precond(bp.p, by.p);
      fine_apply(by.p, nullptr, bnu.p, ST_PROD);
      double vs = dot(brr.p, br.p, n, tmp);
      double coef = dot(brr.p, bnu.p, n, tmp);
      const double alpha = vs / coef;
      bicg_s_update(bnu.p, br.p, bs.p, alpha, n, stream);
      precond(bs.p, bz.p);
      fine_apply(bz.p, nullptr, bt.p, ST_PROD);
      double wv = dot(bt.p, bs.p, n, tmp);
      coef = dot(bt.p, bt.p, n, tmp);
      wv /= coef;
      bicg_sol_update(bs.p, by.p, bz.p, x, bt.p, br.p, alpha, wv, n, stream);
      coef = wv / alpha;
      vs *= coef;
      double beta = dot(brr.p, br.p, n, tmp);
      beta /= vs;
      bicg_p_update(bp.p, bnu.p, br.p, coef, beta, n, partial.p, stream);
      finalize(partial.p, nb, dnorm, 1, stream);
      norm_r = reduce_to_host(dnorm);
      ++iters;
    } while ((norm_r > opt.eps) && (iters < opt.max_it));
  }
  return iters;
}

// -----------------------------------------------------------------------------------------------
// public operations
// -----------------------------------------------------------------------------------------------
void Solver::solve(const double* b, double* x, int mem, int* it_out, double* nr_out) {
  require_ready("solve");
  if (!b || !x) throw Error(LSX_ERR_ARG, "null vector");
  const long n = d0.n();
  const double* db = b;
  double* dx = x;
  if (mem == LSX_MEM_HOST) {
    hb.alloc(n);
    hx.alloc(n);
    LSX_HIP(hipMemcpyAsync(hb.p, b, n * sizeof(double), hipMemcpyHostToDevice, stream));
    LSX_HIP(hipMemcpyAsync(hx.p, x, n * sizeof(double), hipMemcpyHostToDevice, stream));
    db = hb.p;
    dx = hx.p;
  }
  it = opt.krylov == LSX_KRYLOV_FGMRES ? fgmres(db, dx) : bicgstab(db, dx);
  // empty rows (closed cut cells) carry no equation; the V-cycle leaves them whatever its
  // corrections put there, the solution reports 0
  if (vcyc()) {
    if (mixed())
      vc::zero_empty<float>(vlev[0].L, dx, stream);
    else
      vc::zero_empty<double>(vlev[0].L, dx, stream);
  }
  if (mem == LSX_MEM_HOST) LSX_HIP(hipMemcpyAsync(x, dx, n * sizeof(double), hipMemcpyDeviceToHost, stream));
  LSX_HIP(hipGetLastError());
  LSX_HIP(hipStreamSynchronize(stream));
  if (it_out) *it_out = it;
  if (nr_out) *nr_out = norm_r;
}

void Solver::precondition(const double* v, double* z, int mem) {
  require_ready("precondition");
  const long n = d0.n();
  if (mem == LSX_MEM_HOST) {
    hb.alloc(n);
    hx.alloc(n);
    LSX_HIP(hipMemcpyAsync(hb.p, v, n * sizeof(double), hipMemcpyHostToDevice, stream));
    precond(hb.p, hx.p);
    LSX_HIP(hipMemcpyAsync(z, hx.p, n * sizeof(double), hipMemcpyDeviceToHost, stream));
  } else {
    precond(v, z);
  }
  LSX_HIP(hipGetLastError());
  LSX_HIP(hipStreamSynchronize(stream));
}

void Solver::apply_operator(int level, const double* x, double* yv, int mem) {
  require_ready("apply_operator");
  if (level < 0 || level >= (int)levels.size()) throw Error(LSX_ERR_ARG, "level out of range");
  if (vcyc() && level > 0) throw Error(LSX_ERR_STATE, "the V-cycle keeps coarse levels as face coefficients only");
  const MGLevel& L = levels[level];
  const long n = L.d.n();
  if (mem == LSX_MEM_HOST) {
    hb.alloc(d0.n());
    hx.alloc(d0.n());
    LSX_HIP(hipMemcpyAsync(hb.p, x, n * sizeof(double), hipMemcpyHostToDevice, stream));
    if (level == 0)
      fine_apply(hb.p, nullptr, hx.p, ST_PROD);
    else
      stencil_apply(L.coef.p, L.npts, L.d, hb.p, nullptr, hx.p, ST_PROD, nullptr, stream);
    LSX_HIP(hipMemcpyAsync(yv, hx.p, n * sizeof(double), hipMemcpyDeviceToHost, stream));
  } else if (level == 0) {
    fine_apply(x, nullptr, yv, ST_PROD);
  } else {
    stencil_apply(L.coef.p, L.npts, L.d, x, nullptr, yv, ST_PROD, nullptr, stream);
  }
  LSX_HIP(hipGetLastError());
  LSX_HIP(hipStreamSynchronize(stream));
}

void Solver::require_transpose() const {
  require_ready("adjoint operation");
  if (!vcyc()) throw Error(LSX_ERR_STATE, "adjoint operations need the V-cycle preconditioner (LSX_MG_VCYCLE)");
  if (!sym_ok)
    throw Error(LSX_ERR_STATE, "adjoint operations need an operator symmetric off its pinned rows");
}

void Solver::solve_adjoint(const double* g, double* lambda, int mem, int* it_out, double* nr_out) {
  require_transpose();
  TransScope scope(*this);
  solve(g, lambda, mem, it_out, nr_out);
}

void Solver::precondition_transpose(const double* v, double* z, int mem) {
  require_transpose();
  TransScope scope(*this);
  precondition(v, z, mem);
}

void Solver::smooth_fine(const double* b, double* x, int sweeps, int zero_guess, int mem) {
  require_ready("smooth");
  if (!vcyc()) throw Error(LSX_ERR_STATE, "the smoother belongs to the V-cycle (LSX_MG_VCYCLE)");
  if (sweeps < 1) throw Error(LSX_ERR_ARG, "sweeps must be >= 1");
  if (!b || !x) throw Error(LSX_ERR_ARG, "null vector");
  const long n = d0.n();
  const double* db = b;
  double* dx = x;
  if (mem == LSX_MEM_HOST) {
    hb.alloc(n);
    hx.alloc(n);
    LSX_HIP(hipMemcpyAsync(hb.p, b, n * sizeof(double), hipMemcpyHostToDevice, stream));
    LSX_HIP(hipMemcpyAsync(hx.p, x, n * sizeof(double), hipMemcpyHostToDevice, stream));
    db = hb.p;
    dx = hx.p;
  }
  if (mixed()) {
    vc::convert<double, float>(db, vlev[0].fb.p, n, stream);
    vc::convert<double, float>(dx, ztmp32.p, n, stream);
    smooth<float>(0, ztmp32.p, vlev[0].fb.p, sweeps, zero_guess != 0);
    vc::convert<float, double>(ztmp32.p, dx, n, stream);
  } else {
    smooth<double>(0, dx, db, sweeps, zero_guess != 0);
  }
  if (mem == LSX_MEM_HOST) LSX_HIP(hipMemcpyAsync(x, dx, n * sizeof(double), hipMemcpyDeviceToHost, stream));
  LSX_HIP(hipGetLastError());
  LSX_HIP(hipStreamSynchronize(stream));
}

void Solver::apply_operator_transpose(const double* x, double* yv, int mem) {
  require_transpose();
  TransScope scope(*this);
  apply_operator(0, x, yv, mem);
}

long Solver::level_dims(int level, int* N, int* M, int* K) const {
  if (level < 0 || level >= (int)levels.size()) throw Error(LSX_ERR_ARG, "level out of range");
  const Dims& d = levels[level].d;
  if (N) *N = d.N;
  if (M) *M = d.M;
  if (K) *K = d.K;
  return d.n();
}

void Solver::operator_diagonal(double* host) const {
  if (!ready) throw Error(LSX_ERR_STATE, "operator_diagonal before the matrix is set");
  const long n = d0.n();
  LSX_HIP(hipMemcpyAsync(host, levels[0].coef.p + 3 * n, n * sizeof(double), hipMemcpyDeviceToHost, stream));
  LSX_HIP(hipStreamSynchronize(stream));
}

int Solver::level_stencil(int level, double* host, long capacity) const {
  if (level < 0 || level >= (int)levels.size()) throw Error(LSX_ERR_ARG, "level out of range");
  if (vcyc() && level > 0) throw Error(LSX_ERR_STATE, "the V-cycle keeps coarse levels as face coefficients only");
  const MGLevel& L = levels[level];
  const long need = (long)L.npts * L.d.n();
  if (capacity < need) throw Error(LSX_ERR_ARG, "capacity too small for the level stencil");
  LSX_HIP(hipMemcpyAsync(host, L.coef.p, need * sizeof(double), hipMemcpyDeviceToHost, stream));
  LSX_HIP(hipStreamSynchronize(stream));
  return L.npts;
}

}  // namespace lsx
