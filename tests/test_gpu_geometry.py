"""GPU cut-cell geometry and operator assembly against the CPU oracle (bit-exact: integer cell
and face types, and every floating-point field, are evaluated with the same formulas in the same
order without FMA contraction)."""
import numpy as np
import pytest

import oracle_py as O
import problems as P

pytestmark = pytest.mark.gpu


def axes_for(lsx, dims, ratio):
    return [lsx.stretched_axis(d, 1.0, 0.5, ratio) for d in dims]


CASES = [
    ("sphere16", (16, 16, 16), 1.0, "sphere", (0.35, 0.5, 0.5), (0.2,)),
    ("sphere32", (32, 32, 32), 1.0, "sphere", (0.35, 0.5, 0.5), (0.2,)),
    ("sphere48x32x32", (48, 32, 32), 1.0, "sphere", (0.4, 0.45, 0.55), (0.22,)),
    ("ellipsoid32_stretched", (32, 32, 32), 1.02, "ellipsoid", (0.5, 0.5, 0.5), (0.25, 0.15, 0.15)),
    ("ellipsoid64_stretched", (64, 64, 64), 1.02, "ellipsoid", (0.5, 0.5, 0.5), (0.25, 0.15, 0.15)),
]


def build_both(lsx, dims, ratio, kind, center, axes):
    w = axes_for(lsx, dims, ratio)
    g_ref = O.Geometry(*w).build_body(kind, center, axes)
    g_gpu = lsx.Geometry(*w).build(kind, center, axes)
    return g_ref, g_gpu


@pytest.mark.parametrize("name,dims,ratio,kind,center,axes", CASES, ids=[c[0] for c in CASES])
def test_geometry_fields_bitwise(lsx, name, dims, ratio, kind, center, axes):
    g_ref, g_gpu = build_both(lsx, dims, ratio, kind, center, axes)
    fr, fg = g_ref.fields(), g_gpu.fields()
    ls = g_ref.level_set()
    print(f"case={name} cut={int(((fr['type'] > 0) & (fr['type'] != 11111111)).sum())}", flush=True)
    assert np.array_equal(fg["phi"], ls["phi"])
    assert np.array_equal(fg["type"], fr["type"])
    for k in ("thx", "thy", "thz", "vof"):
        assert np.array_equal(fg[k], fr[k]), k


@pytest.mark.parametrize("name,dims,ratio,kind,center,axes", CASES, ids=[c[0] for c in CASES])
def test_operator_from_geometry_bitwise(lsx, name, dims, ratio, kind, center, axes):
    g_ref, g_gpu = build_both(lsx, dims, ratio, kind, center, axes)
    A = g_ref.laplace(pin=True)
    S = lsx.MGFGMRES()
    n = A.n
    S.set_geometry(g_gpu, np.zeros(n), np.zeros(n), lsx.Task(eps=1e-8, FGMRES_RESTART_NUM=30, MAX_ITER=500))
    st = S.level_stencil(0)
    N, M, K = dims
    dense = np.zeros((7, n))
    offs = {-N * M: 0, -N: 1, -1: 2, 0: 3, 1: 4, N: 5, N * M: 6}
    rows = np.repeat(np.arange(n), np.diff(A.rowptr))
    q = np.array([offs[c - r] for c, r in zip(A.col, rows)])
    dense[q, rows] = A.val
    assert np.array_equal(st, dense)


def test_solve_on_gpu_geometry_matches_oracle(lsx):
    dims, ratio = (32, 32, 32), 1.02
    g_ref, g_gpu = build_both(lsx, dims, ratio, "ellipsoid", (0.5, 0.5, 0.5), (0.25, 0.15, 0.15))
    A = g_ref.laplace()
    b = P.random_rhs(g_ref, seed=3, A=A)
    eps = 1e-8 * np.linalg.norm(b)
    x_ref, it_ref, _ = O.Solver(A, eps=eps, restart=30, max_it=500).solve(b)
    S = lsx.MGFGMRES()
    x = np.zeros(A.n)
    S.set_geometry(g_gpu, b, x, lsx.Task(eps=eps, FGMRES_RESTART_NUM=30, MAX_ITER=500))
    it = S.solve()
    rel = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    print(f"it={it} it_ref={it_ref} rel={rel:.3e}", flush=True)
    assert it == it_ref and rel <= 1e-10


def test_geometry_rejects_bad_box(lsx):
    w = axes_for(lsx, (16, 16, 16), 1.0)
    g = lsx.Geometry(*w)
    with pytest.raises(lsx.LsxError):
        g.build("sphere", (0.5, 0.5, 0.5), (0.2,), box=(0, 16, 1, 15))


RHS_CASES = [
    ("sphere32", (32, 32, 32), 1.0, "sphere", (0.35, 0.5, 0.5), (0.2,)),
    ("ellipsoid48x32x32_stretched", (48, 32, 32), 1.03, "ellipsoid", (0.5, 0.5, 0.5), (0.25, 0.15, 0.15)),
]


@pytest.mark.parametrize("name,dims,ratio,kind,center,axes", RHS_CASES, ids=[c[0] for c in RHS_CASES])
@pytest.mark.parametrize("outflow", [True, False])
def test_pressure_rhs_bitwise(lsx, name, dims, ratio, kind, center, axes, outflow):
    """b = D u* + flux (PoissonDiscreteEquation::recalcRightMember) on the GPU equals the oracle
    bit for bit: u* = 1 + U(-0.05, 0.05), v*, w* = U(-0.05, 0.05), a non-uniform inflow profile,
    nonzero wall velocities so every flux term is exercised."""
    g_ref, g_gpu = build_both(lsx, dims, ratio, kind, center, axes)
    N, M, K = dims
    nu, nv, nw = O.face_shapes(dims)
    rng = np.random.default_rng(11)
    u = 1.0 + rng.uniform(-0.05, 0.05, nu)
    v = rng.uniform(-0.05, 0.05, nv)
    w = rng.uniform(-0.05, 0.05, nw)
    uw = 1.0 + rng.uniform(-0.1, 0.1, M * K)
    kw = dict(U_W=uw, U_E=1.02, V_S=0.01, V_N=-0.02, W_B=0.03, W_T=-0.01, outflow=outflow)
    b_ref = O.pressure_rhs(g_ref, u, v, w, **kw)
    b_gpu = g_gpu.pressure_rhs(u, v, w, **kw)
    print(f"case={name} outflow={outflow} max|b|={np.abs(b_ref).max():.3e}", flush=True)
    assert np.array_equal(b_gpu, b_ref), np.max(np.abs(b_gpu - b_ref))


BEZIER_CASES = [
    ("bezier32", (32, 32, 32), 1.0),
    ("bezier48x32x32_stretched", (48, 32, 32), 1.03),
]


@pytest.mark.parametrize("name,dims,ratio", BEZIER_CASES, ids=[c[0] for c in BEZIER_CASES])
def test_bezier_geometry_bitwise(lsx, name, dims, ratio):
    """Six-patch Bezier body (ellipsoid fit, control points perturbed by U(-0.01, 0.01)): the
    GPU level set, types, face fractions, volumes and operator equal the oracle's bit for bit."""
    center, axes = (0.5, 0.5, 0.5), (0.25, 0.15, 0.15)
    ctrl = lsx.bezier_perturb(lsx.bezier_ellipsoid(center, axes), 0.01, seed=1)
    w = axes_for(lsx, dims, ratio)
    g_ref = O.Geometry(*w).build_bezier(ctrl, center, axes)
    g_gpu = lsx.Geometry(*w).build("bezier", center, axes, ctrl=ctrl)
    fr, fg = g_ref.fields(), g_gpu.fields()
    ls = g_ref.level_set()
    print(f"case={name} cut={int(((fr['type'] > 0) & (fr['type'] != 11111111)).sum())} "
          f"max|dphi|={np.abs(fg['phi'] - ls['phi']).max():.3e}", flush=True)
    assert np.array_equal(fg["phi"], ls["phi"])
    assert np.array_equal(fg["type"], fr["type"])
    for k in ("thx", "thy", "thz", "vof"):
        assert np.array_equal(fg[k], fr[k]), k
    A = g_ref.laplace()
    S = lsx.MGFGMRES()
    S.set_geometry(g_gpu, np.zeros(A.n), np.zeros(A.n), lsx.Task())
    x = np.random.default_rng(3).standard_normal(A.n)
    assert np.array_equal(S.apply_operator(x, 0), A.matvec(x))


@pytest.mark.parametrize("name,dims,ratio,kind,center,axes", CASES[::2], ids=[c[0] for c in CASES[::2]])
def test_operator_from_caller_fields_bitwise(lsx, name, dims, ratio, kind, center, axes):
    """lsx_set_fields with the oracle's mesh fields (the arrays a reference caller holds): the
    device-assembled operator equals the oracle's CSR operator and lsx_set_geometry's, and the
    solve equals the oracle's at an equal iteration count."""
    w = axes_for(lsx, dims, ratio)
    g_ref = O.Geometry(*w).build_body(kind, center, axes)
    fr = g_ref.fields()
    A = g_ref.laplace(pin=True)
    n = A.n
    N, M, K = dims
    b = P.random_rhs(g_ref, seed=11, A=A)
    eps = 1e-8 * np.linalg.norm(b)
    task = lsx.Task(eps=eps, FGMRES_RESTART_NUM=30, MAX_ITER=500)
    x = np.zeros(n)
    S = lsx.MGFGMRES()
    S.set_fields(dims, w, fr, b, x, task)
    st = S.level_stencil(0)
    dense = np.zeros((7, n))
    offs = {-N * M: 0, -N: 1, -1: 2, 0: 3, 1: 4, N: 5, N * M: 6}
    rows = np.repeat(np.arange(n), np.diff(A.rowptr))
    q = np.array([offs[c - r] for c, r in zip(A.col, rows)])
    dense[q, rows] = A.val
    assert np.array_equal(st, dense)
    g_gpu = lsx.Geometry(*w).build(kind, center, axes)
    S2 = lsx.MGFGMRES()
    S2.set_geometry(g_gpu, np.zeros(n), np.zeros(n), task)
    assert np.array_equal(S2.level_stencil(0), st)
    it = S.solve()
    x_ref, it_ref, _ = O.Solver(A, eps=eps, restart=30, max_it=500).solve(b)
    rel = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    print(f"case={name} it={it} it_ref={it_ref} rel={rel:.3e}", flush=True)
    assert it == it_ref and rel <= 1e-10
    with pytest.raises(lsx.LsxError, match="do not match"):
        S.set_fields((N, M, K - 1), w, fr, b, x, task)
