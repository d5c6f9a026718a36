"""Mixed-precision mode (FP32 V-cycle, FP16 coarse faces, FP64 flexible Krylov loop): it must
reach the same relative residual 1e-8 as the FP64 path, checked with the oracle's FP64 CSR
operator, and its pressure is reported against the FP64 oracle solution of the same system."""
import numpy as np
import pytest

import oracle_py as O
import problems as P

pytestmark = pytest.mark.gpu


def run_mixed(lsx, A, b, kind, restart=30):
    eps = 1e-8 * np.linalg.norm(b)
    S = lsx.solver(kind, "mixed")
    x = np.zeros(A.n)
    S.set(A, b, x, lsx.mixed_task(eps, restart=restart, max_it=500))
    it = S.solve()
    return S, x, it, eps


CASES = [
    ("sphere64", lambda: P.sphere(64)),
    ("ellipsoid64_stretched", lambda: P.ellipsoid_stretched(64)),
    ("cut32x16x24", lambda: P.synthetic_cut(32, 16, 24, seed=3)),
]


@pytest.mark.parametrize("kind", ["fgmres", "bicgstab"])
@pytest.mark.parametrize("name,mk", CASES, ids=[c[0] for c in CASES])
def test_mixed_reaches_fp64_tolerance(lsx, name, mk, kind):
    g, A = mk()
    b = P.random_rhs(g, seed=2, A=A)
    S, x, it, eps = run_mixed(lsx, A, b, kind)
    true_res = np.linalg.norm(b - A.matvec(x))
    x64, it64, _ = O.Solver(A, kind=kind, eps=eps, restart=30, max_it=500).solve(b)
    active = A.to_scipy().diagonal() < 0
    err = np.linalg.norm((x - x64)[active]) / np.linalg.norm(x64[active])
    print(f"case={name} kind={kind} mixed_it={it} fp64_it={it64} true_rel_res={true_res / np.linalg.norm(b):.3e} "
          f"p_err_vs_fp64={err:.3e}", flush=True)
    assert S.norm_r <= eps
    assert true_res <= 1.05 * eps
    assert it < 100
    # both solutions satisfy ||b - A p|| <= 1e-8 ||b||; their difference is bounded by the
    # conditioning of the pinned Neumann operator
    assert err <= 1e-4


def test_mixed_geometry_path_256_like_smaller(lsx):
    """The benchmark's construction (device geometry -> operator -> mixed solve) at 64^3."""
    w = [lsx.stretched_axis(64, 1.0, 0.5, 1.02) for _ in range(3)]
    geom = lsx.Geometry(*w).build("ellipsoid", (0.5, 0.5, 0.5), (0.25, 0.15, 0.15))
    S = lsx.solver("fgmres", "mixed")
    S.set_geometry(geom, None, None, lsx.mixed_task(1.0))
    b = lsx.random_divergence_rhs(S.operator_diagonal(), geom.dims)
    S.eps = 1e-8 * np.linalg.norm(b)
    x = np.zeros(geom.n)
    S._rhs, S._sol = b, x
    it = S.solve()
    r = b - S.apply_operator(x, 0)
    print(f"it={it} rel_res={np.linalg.norm(r) / np.linalg.norm(b):.3e}", flush=True)
    assert np.linalg.norm(r) <= 1.05 * S.eps and it < 60
