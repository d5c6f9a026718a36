"""Momentum (Helmholtz) solves on the GPU through the drop-in boundary, against the oracle.

The reference solves the three momentum equations with the same Krylov and multigrid classes as
the pressure (HelmholtzDiscreteEquation with LS_STAG_FULL_MG, helmholtz_discrete_equation.h:79-345,
selecting MGFGMRES or MGBiCGStab at :285-333).  Their meshes are the staggered ones, (N-1) x M x K
for u, N x (M-1) x K for v, N x M x (K-1) for w, so one dimension is odd on every even base grid:
the library solves them with the one-level cascade (lsx_options.levels = 1, ADLJ sweeps of the
fine level from a zero correction) — the reference's odd-parity transfer operators are not rebuilt —
or with the V-cycle preconditioners, which take any dimensions (test_helmholtz_vcycle).

Each test assembles the operator with the oracle's restatement of the reference's
HelmholtzDiscreteOperatorX/Y/Z (oracle/oracle_helmholtz.cpp, pinned in test_oracle_helmholtz.py),
takes the time-derivative part of the reference's right-hand side, (-2 u^n + 0.5 u^(n-1)) M / dt
(HelmholtzDiscreteEquation::initRM, helmholtz_discrete_equation.h:70-73) with seeded velocities,
hands the CSR to lsx_set_csr exactly as the equation's solver->set(A, rhs, sol) would, and
compares the GPU solve with the oracle's solve of the same system: equal iteration counts and a
relative L2 difference of at most 1e-10."""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu

RE, DT = 300.0, 0.002
NU, COEF = 1.0 / RE, 1.5 / DT


def axis(m, ratio):
    w = ratio ** np.abs(np.arange(m) + 0.5 - m / 2)
    return w / w.sum()


def masses(g, d):
    """Staggered control-volume masses M = (V_i + V_nb) / 2 of component d (row order of its mesh)."""
    N, M, K = g.dims
    v = g.fields()["vof"].reshape(K, M, N)
    if d == 0:
        m = 0.5 * (v[:, :, :-1] + v[:, :, 1:])
    elif d == 1:
        m = 0.5 * (v[:, :-1, :] + v[:, 1:, :])
    else:
        m = 0.5 * (v[:-1, :, :] + v[1:, :, :])
    return m.ravel()


def system(dims, ratios, d, seed, flow=True, re=RE, dt=DT):
    g = O.Geometry(*[axis(m, r) for m, r in zip(dims, ratios)])
    A = O.helmholtz(g, d, 1.0 / re, 1.5 / dt, flow=flow)
    rng = np.random.default_rng(seed)
    un = 1.0 + 0.1 * rng.standard_normal(A.n)
    uprev = un + 0.01 * rng.standard_normal(A.n)
    b = (-2.0 * un + 0.5 * uprev) * masses(g, d) / dt
    return A, b


# (grid, stretching, Re, dt): the time-stepped run's Re = 300, dt = 0.002 (a mass-dominated system,
# 2 iterations) and viscous ones that exercise the Krylov loop (Re = 10 at 48x32x40: 8 to 34
# iterations; Re = 30 at 64^3: 6 to 24)
CASES = [((32, 24, 20), (1.04, 1.06, 1.02), RE, DT), ((64, 64, 64), (1.0, 1.0, 1.0), RE, DT),
         ((48, 32, 40), (1.02, 1.02, 1.02), 10.0, 0.02), ((64, 64, 64), (1.0, 1.0, 1.0), 30.0, 0.02)]


@pytest.mark.parametrize("kind", ["fgmres", "bicgstab"])
@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("dims,ratios,re,dt", CASES, ids=["32x24x20s", "64cube", "48x32x40s-re10", "64cube-re30"])
def test_helmholtz_solve_matches_oracle(lsx, dims, ratios, re, dt, d, kind):
    A, b = system(dims, ratios, d, seed=7 + d, re=re, dt=dt)
    eps = 1e-8 * np.linalg.norm(b)
    x_ref, it_ref, nr_ref = O.Solver(A, kind=kind, eps=eps, restart=30, max_it=200, precond="mg1").solve(b)
    S = (lsx.MGFGMRES if kind == "fgmres" else lsx.MGBiCGStab)(levels=1)
    x = np.zeros(A.n)
    S.set(A, b, x, lsx.Task(eps=eps, MAX_ITER=200, FGMRES_RESTART_NUM=30))
    it = S.solve()
    rel = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    res = np.linalg.norm(b - A.matvec(x))
    print(f"mesh {A.dims} component {'uvw'[d]} {kind}: it={it} it_ref={it_ref} rel={rel:.3e} "
          f"|b-Ax|={res:.3e} eps={eps:.3e}", flush=True)
    assert it == it_ref and rel <= 1e-10
    assert res <= 1.05 * eps


def test_helmholtz_one_level_preconditioner_matches_oracle(lsx):
    """One application of the one-level cascade (z, y, x line sweeps from zero) equals the oracle's."""
    A, b = system((24, 20, 16), (1.05, 1.03, 1.07), 0, seed=3)
    zr = O.Solver(A, precond="mg1").mg_apply(b)
    S = lsx.MGFGMRES(levels=1)
    S.set(A, b, np.zeros(A.n), lsx.Task(eps=1.0, FGMRES_RESTART_NUM=30))
    z = S.precondition(b)
    rel = np.linalg.norm(z - zr) / np.linalg.norm(zr)
    print(f"one-level cascade vs oracle: rel={rel:.3e}", flush=True)
    assert rel <= 1e-12


def test_one_level_rejected_for_vcycle_and_odd_grid_needs_it(lsx):
    A, b = system((16, 12, 10), (1.0, 1.0, 1.0), 0, seed=1)
    with pytest.raises(lsx.LsxError, match="even"):
        lsx.MGFGMRES().set(A, b, np.zeros(A.n), lsx.Task(eps=1e-8))
    with pytest.raises(lsx.LsxError, match="levels"):
        lsx.solver("fgmres", "fp64-vcycle", levels=1).set(A, b, np.zeros(A.n), lsx.mixed_task(1e-8))


def dense7(A):
    """The CSR operator as 7 stencil planes [b, s, w, p, e, n, t] of its mesh (level_stencil layout)."""
    N, M, K = A.dims
    n = A.n
    out = np.zeros((7, n))
    offs = {-N * M: 0, -N: 1, -1: 2, 0: 3, 1: 4, N: 5, N * M: 6}
    rows = np.repeat(np.arange(n), np.diff(A.rowptr))
    q = np.array([offs[c - r] for c, r in zip(A.col, rows)])
    out[q, rows] = A.val
    return out


@pytest.mark.parametrize("flow", [True, False])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_device_helmholtz_operator_bitwise(lsx, d, flow):
    """lsx_set_helmholtz: the momentum operator assembled on the device from a GPU geometry equals
    the oracle's restatement bit for bit, and its solve equals the oracle's."""
    dims, ratios = (24, 20, 16), (1.05, 1.03, 1.07)
    w = [axis(m, r) for m, r in zip(dims, ratios)]
    g_ref = O.Geometry(*w)  # no body
    g = lsx.Geometry(*w).build("sphere", (-5.0, -5.0, -5.0), (0.01,))  # a body outside the box
    fr, fg = g_ref.fields(), g.fields()
    for key in ("type", "thx", "thy", "thz", "vof"):
        assert np.array_equal(fr[key], fg[key]), key
    A = O.helmholtz(g_ref, d, NU, COEF, flow=flow)
    rng = np.random.default_rng(11 + d)
    b = rng.standard_normal(A.n) * 1e-3
    eps = 1e-8 * np.linalg.norm(b)
    x = np.zeros(A.n)
    S = lsx.MGFGMRES(levels=1)
    S.set_helmholtz(g, d, NU, COEF, b, x, lsx.Task(eps=eps, MAX_ITER=200, FGMRES_RESTART_NUM=30), flow=flow)
    st = S.level_stencil(0)
    ref = dense7(A)
    print(f"component {'uvw'[d]} flow={flow}: mesh {A.dims}, max |stencil - oracle| = {np.abs(st - ref).max():.3e}",
          flush=True)
    assert np.array_equal(st, ref)
    it = S.solve()
    x_ref, it_ref, _ = O.Solver(A, eps=eps, restart=30, max_it=200, precond="mg1").solve(b)
    assert it == it_ref and np.linalg.norm(x - x_ref) <= 1e-10 * np.linalg.norm(x_ref)


def test_device_helmholtz_rejects_cut_cells(lsx):
    w = [np.full(16, 1.0 / 16)] * 3
    g = lsx.Geometry(*w).build("sphere", (0.5, 0.5, 0.5), (0.2,))
    S = lsx.MGFGMRES(levels=1)
    with pytest.raises(lsx.LsxError, match="without cut cells"):
        S.set_helmholtz(g, 0, NU, COEF, np.zeros(15 * 16 * 16), np.zeros(15 * 16 * 16), lsx.Task(eps=1e-8))


@pytest.mark.parametrize("precision", ["fp64-vcycle", "mixed"])
@pytest.mark.parametrize("kind", ["fgmres", "bicgstab"])
@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("dims,ratios,re,dt", CASES, ids=["32x24x20s", "64cube", "48x32x40s-re10", "64cube-re30"])
def test_helmholtz_vcycle(lsx, dims, ratios, re, dt, d, kind, precision):
    """The V-cycle preconditioners on the momentum systems, which have no pinned rows: each coarse
    level's diagonal is the rediscretised couplings' plus the sum of its children's excess
    a_p + (sum of the six faces), the mass and wall terms (the diagonal block of P^T A P for the
    piecewise-constant prolongation).  The solve reaches the true residual eps = 1e-8 |b| in no
    more iterations than the one-level cascade and agrees with the oracle's solve to the accuracy
    that tolerance gives (relative 1e-6).  Measured: the mass-dominated systems take 1-2
    iterations, u at Re = 10 / 30 takes 7 / 5 FGMRES iterations (one-level cascade: 11 / 9), v and
    w 21-22 / 15 (30-34 / 24).  The outflow rows of v and w carry a positive excess (the outflow
    term exceeds the mass term), which puts the smallest eigenvalue of -A at 3e-5 against 9e-4
    for u (condition 1200 against 43 at 48 x 32 x 40, Re = 10); taking the wall part of a boundary
    cell's excess at half weight, as a face, measured worse (27 / 18)."""
    A, b = system(dims, ratios, d, seed=7 + d, re=re, dt=dt)
    eps = 1e-8 * np.linalg.norm(b)
    x_ref, it_ref, _ = O.Solver(A, kind=kind, eps=eps, restart=30, max_it=200, precond="mg1").solve(b)
    S = lsx.solver(kind, precision)
    x = np.zeros(A.n)
    S.set(A, b, x, lsx.mixed_task(eps, restart=30, max_it=200))
    it = S.solve()
    res = np.linalg.norm(b - A.matvec(x))
    rel = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    print(f"mesh {A.dims} component {'uvw'[d]} {kind} {precision}: it={it} (one-level cascade {it_ref}) "
          f"levels={S.num_levels()} |b-Ax|/eps={res / eps:.3f} rel={rel:.3e}", flush=True)
    assert res <= 1.05 * eps
    assert it <= it_ref
    assert rel <= 1e-6


def test_vcycle_rejects_singular_operator_without_pins(lsx):
    """A Neumann pressure operator without its pinned rows has no mass term either: singular."""
    import problems as P
    g, A = P.body_free(16, pin=False)
    S = lsx.solver("fgmres", "mixed")
    with pytest.raises(lsx.LsxError, match="singular"):
        S.set(A, np.zeros(A.n), np.zeros(A.n), lsx.mixed_task(1e-8))
        S.solve()


@pytest.mark.parametrize("precision,tol", [("fp64-vcycle", 1e-12), ("mixed", 2e-5)])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_helmholtz_vcycle_transposed(lsx, d, precision, tol):
    """The transposed V-cycle on an operator without pinned rows (aggregated mass term) is the
    transpose of the cycle: <B x, y> = <x, B^T y> (exact in FP64, FP32 rounding in mixed).  The
    momentum operators are symmetric to rounding only (their mirrored couplings differ in the
    last bit), and adjoint operations take bitwise-symmetric operators, so the test uses
    (A + A^T) / 2, which is (each pair is summed in either order to the same value)."""
    A0, _ = system((48, 32, 40), (1.02, 1.02, 1.02), d, seed=5, re=10.0, dt=0.02)
    M = A0.to_scipy()
    M = ((M + M.T) * 0.5).tocsr()
    M.sort_indices()
    A = O.Csr(A0.dims, M.indptr, M.indices, M.data)
    S = lsx.solver("fgmres", precision)
    S.set(A, np.zeros(A.n), np.zeros(A.n), lsx.mixed_task(1.0))
    rng = np.random.default_rng(9 + d)
    x, y = rng.standard_normal(A.n), rng.standard_normal(A.n)
    Bx, BTy = S.precondition(x), S.precondition_transpose(y)
    scale = np.linalg.norm(Bx) * np.linalg.norm(y)
    rel = abs(Bx @ y - x @ BTy) / scale
    print(f"mesh {A.dims} component {'uvw'[d]} {precision}: levels={S.num_levels()} transpose rel={rel:.3e}",
          flush=True)
    assert np.linalg.norm(Bx) > 0 and rel <= tol
