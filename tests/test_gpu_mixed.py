"""V-cycle preconditioner (red-black Gauss-Seidel, rediscretised face levels, pinned cells as
Dirichlet sinks on coarse levels) in FP64 and in mixed precision (FP32 values, FP16 coarse
faces) inside the FP64 flexible Krylov loop.  Both must reach the relative residual 1e-8 that
the FP64 reference path reaches, checked with the oracle's FP64 CSR operator, and the pressure
is reported against the FP64 oracle solution of the same system."""
import numpy as np
import pytest

import oracle_py as O
import problems as P

pytestmark = pytest.mark.gpu


def run_vcycle(lsx, A, b, kind, precision, restart=30, max_it=500):
    eps = 1e-8 * np.linalg.norm(b)
    S = lsx.solver(kind, precision)
    x = np.zeros(A.n)
    S.set(A, b, x, lsx.mixed_task(eps, restart=restart, max_it=max_it))
    it = S.solve()
    return S, x, it, eps


CASES = [
    ("sphere64", lambda: P.sphere(64)),
    ("ellipsoid64_stretched", lambda: P.ellipsoid_stretched(64)),
    ("ellipsoid64_stretched_1.08", lambda: P.ellipsoid_stretched(64, ratio=1.08)),
    ("cut32x16x24", lambda: P.synthetic_cut(32, 16, 24, seed=3)),
]


@pytest.mark.parametrize("precision", ["mixed", "mixed16", "fp64-vcycle"])
@pytest.mark.parametrize("kind", ["fgmres", "bicgstab"])
@pytest.mark.parametrize("name,mk", CASES, ids=[c[0] for c in CASES])
def test_vcycle_reaches_fp64_tolerance(lsx, name, mk, kind, precision):
    g, A = mk()
    b = P.random_rhs(g, seed=2, A=A)
    S, x, it, eps = run_vcycle(lsx, A, b, kind, precision)
    true_res = np.linalg.norm(b - A.matvec(x))
    x64, it64, _ = O.Solver(A, kind=kind, eps=eps, restart=30, max_it=500).solve(b)
    active = A.to_scipy().diagonal() < 0
    err = np.linalg.norm((x - x64)[active]) / np.linalg.norm(x64[active])
    print(f"case={name} kind={kind} precision={precision} it={it} fp64_cascade_it={it64} "
          f"true_rel_res={true_res / np.linalg.norm(b):.3e} p_err_vs_fp64={err:.3e}", flush=True)
    assert S.norm_r <= eps
    # BiCGStab's recursively updated residual drifts from the true one by a few percent of eps
    assert true_res <= 1.2 * eps
    assert it < 60
    # both solutions satisfy ||b - A p|| <= 1e-8 ||b||; their difference is bounded by the
    # conditioning of the pinned Neumann operator
    assert err <= 1e-3


def test_mixed_and_fp64_vcycle_agree_on_one_application(lsx):
    """One preconditioner application: the mixed V-cycle is the FP64 V-cycle up to FP32/FP16
    rounding (relative difference well below the 1e-2 a wrong level or transfer would give)."""
    g, A = P.ellipsoid_stretched(32)
    v = P.random_rhs(g, seed=4, A=A)
    z = {}
    for precision in ("mixed", "fp64-vcycle"):
        S = lsx.solver("fgmres", precision)
        S.set(A, np.zeros(A.n), np.zeros(A.n), lsx.mixed_task(1.0))
        z[precision] = S.precondition(v)
    rel = np.linalg.norm(z["mixed"] - z["fp64-vcycle"]) / np.linalg.norm(z["fp64-vcycle"])
    print(f"mixed vs fp64 V-cycle rel={rel:.3e}", flush=True)
    assert rel <= 1e-3


def test_vcycle_geometry_path_64(lsx):
    """The benchmark's construction (device geometry -> operator -> mixed solve) at 64^3."""
    w = [lsx.stretched_axis(64, 1.0, 0.5, 1.02) for _ in range(3)]
    geom = lsx.Geometry(*w).build("ellipsoid", (0.5, 0.5, 0.5), (0.25, 0.15, 0.15))
    S = lsx.solver("fgmres", "mixed")
    S.set_geometry(geom, None, None, lsx.mixed_task(1.0))
    assert S.fine_storage() == 4
    b = lsx.random_divergence_rhs(S.operator_diagonal(), geom.dims)
    S.eps = 1e-8 * np.linalg.norm(b)
    x = np.zeros(geom.n)
    S._rhs, S._sol = b, x
    it = S.solve()
    r = b - S.apply_operator(x, 0)
    print(f"it={it} rel_res={np.linalg.norm(r) / np.linalg.norm(b):.3e}", flush=True)
    assert np.linalg.norm(r) <= 1.05 * S.eps and it < 40


def test_vcycle_needs_a_pinned_row(lsx):
    g, A = P.body_free(16, pin=False)
    S = lsx.solver("fgmres", "mixed")
    with pytest.raises(lsx.LsxError):
        S.set(A, np.zeros(A.n), np.zeros(A.n), lsx.mixed_task(1.0))


@pytest.mark.parametrize("ortho", ["mgs", "cgs", "cgs2"])
def test_vcycle_orthogonalisations(lsx, ortho):
    """The V-cycle FGMRES reaches the tolerance on the true residual with each orthogonalisation;
    the reported norm is the true residual."""
    g, A = P.ellipsoid_stretched(64, ratio=1.08)
    b = P.random_rhs(g, seed=9, A=A)
    eps = 1e-8 * np.linalg.norm(b)
    code = {"mgs": lsx.ORTHO_MGS, "cgs": lsx.ORTHO_CGS, "cgs2": lsx.ORTHO_CGS2}[ortho]
    S = lsx.solver("fgmres", "mixed", ortho=code)
    x = np.zeros(A.n)
    S.set(A, b, x, lsx.mixed_task(eps, restart=30, max_it=300))
    it = S.solve()
    true_res = np.linalg.norm(b - A.matvec(x))
    print(f"ortho={ortho} it={it} true_rel={true_res / np.linalg.norm(b):.3e} norm_r={S.norm_r:.3e}", flush=True)
    assert true_res <= 1.05 * eps and it < 100
    assert abs(S.norm_r - true_res) <= 1e-3 * true_res


@pytest.mark.parametrize("restart", [30, 6])
@pytest.mark.parametrize("ortho", ["cgs", "cgs2"])
def test_projections_over_fp32_basis_copies(lsx, ortho, restart):
    """lsx_options.dot32: the Gram-Schmidt projections of the mixed FGMRES summed over the FP32
    copies of the basis.  The update and the basis stay FP64, so the Arnoldi relation holds to
    FP64 rounding and only the orthogonality of the basis degrades (to about the FP32 unit
    roundoff): the solve needs no more iterations, the reported norm is the true residual, and the
    solution differs from the FP64-projection one by no more than two converged solutions may
    (restart 6 crosses several windows, each starting from FP32 copies of a fresh basis)."""
    g, A = P.ellipsoid_stretched(64, ratio=1.08)
    b = P.random_rhs(g, seed=11, A=A)
    eps = 1e-8 * np.linalg.norm(b)
    code = {"cgs": lsx.ORTHO_CGS, "cgs2": lsx.ORTHO_CGS2}[ortho]
    out = {}
    for d32 in (True, False):
        S = lsx.solver("fgmres", "mixed", ortho=code, smoother="line", line_levels=2, dot32=d32)
        x = np.zeros(A.n)
        S.set(A, b, x, lsx.mixed_task(eps, restart=restart, max_it=400))
        it = S.solve()
        true_res = np.linalg.norm(b - A.matvec(x))
        print(f"ortho={ortho} restart={restart} dot32={d32}: it={it} true_rel={true_res / np.linalg.norm(b):.3e} "
              f"norm_r={S.norm_r:.3e}", flush=True)
        assert true_res <= 1.05 * eps
        assert abs(S.norm_r - true_res) <= 1e-3 * true_res
        out[d32] = (x, it)
    (x1, it1), (x0, it0) = out[True], out[False]
    active = A.to_scipy().diagonal() < 0
    diff = np.linalg.norm((x1 - x0)[active]) / np.linalg.norm(x0[active])
    print(f"|x_dot32 - x| / |x| = {diff:.3e}", flush=True)
    assert it1 <= it0 + 1
    assert not np.array_equal(x1, x0) and diff <= 1e-4


@pytest.mark.parametrize("precision", ["mixed", "fp64-vcycle"])
def test_coarse_levels_in_one_launch_match_per_level_launches(lsx, precision):
    """The one-workgroup cycle over the small levels performs the same per-cell operations in the
    same order as the per-level launches: one preconditioner application is bit-identical."""
    g, A = P.ellipsoid_stretched(64)
    v = P.random_rhs(g, seed=12, A=A)
    z = []
    for fuse in (True, False):
        S = lsx.solver("fgmres", precision, coarse_fuse=fuse)
        S.set(A, np.zeros(A.n), np.zeros(A.n), lsx.mixed_task(1.0))
        z.append(S.precondition(v))
    print(f"precision={precision} max|diff|={np.abs(z[0] - z[1]).max():.3e}", flush=True)
    assert np.array_equal(z[0], z[1])
