"""One projection step of the sphere flow (the reference's pressure-correction stage) on the GPU,
against the oracle: right-hand side b = D u* + flux, the FP64 pressure solve, and the velocity /
pressure correction.  b and the correction are bit-exact (same formulas, same order, no FMA
contraction); the solve matches at equal iteration counts to 1e-10; after the correction the
divergence of the new velocity equals the pressure residual (D inv_M G = -A)."""
import numpy as np
import pytest

import oracle_py as O
import problems as P

pytestmark = pytest.mark.gpu


def sphere_flow(lsx, n, seed=5):
    w = [np.full(n, 1.0 / n)] * 3
    g_ref = O.Geometry(*w).build_body("sphere", (0.35, 0.5, 0.5), (0.2,))
    g_gpu = lsx.Geometry(*w).build("sphere", (0.35, 0.5, 0.5), (0.2,))
    nu, nv, nw = O.face_shapes(g_ref.dims)
    rng = np.random.default_rng(seed)
    us = 1.0 + rng.uniform(-0.05, 0.05, nu)
    vs = rng.uniform(-0.05, 0.05, nv)
    ws = rng.uniform(-0.05, 0.05, nw)
    return g_ref, g_gpu, us, vs, ws


@pytest.mark.parametrize("n", [32, 64])
def test_projection_step_matches_oracle(lsx, n):
    g_ref, g_gpu, us, vs, ws = sphere_flow(lsx, n)
    b = g_gpu.pressure_rhs(us, vs, ws, U_W=1.0)
    b_ref = O.pressure_rhs(g_ref, us, vs, ws, U_W=1.0)
    assert np.array_equal(b, b_ref)
    A = g_ref.laplace()
    eps = 1e-8 * np.linalg.norm(b)
    S = lsx.MGFGMRES()
    dp = np.zeros(g_gpu.n)
    S.set_geometry(g_gpu, b, dp, lsx.Task(eps=eps, MAX_ITER=500, FGMRES_RESTART_NUM=30))
    it = S.solve()
    dp_ref, it_ref, _ = O.Solver(A, eps=eps, restart=30, max_it=500).solve(b)
    rel = np.linalg.norm(dp - dp_ref) / np.linalg.norm(dp_ref)
    p = np.zeros(g_gpu.n)
    u, v, w = g_gpu.correct(dp, us, vs, ws, p=p, coef=1.0)
    ur, vr, wr = O.velocity_correction(g_ref, dp, us, vs, ws)
    div = O.pressure_rhs(g_ref, u, v, w, U_W=1.0, outflow=True)
    res = b - A.matvec(dp)
    N, M, K = g_ref.dims
    keep = np.ones(g_ref.n, bool)
    keep[[N - 1, N * M - 1]] = False
    # the outflow flux follows the corrected u's last face, which the solve did not see
    east = np.zeros((K, M, N), bool)
    east[:, :, N - 1] = True
    keep &= ~east.ravel()
    print(f"n={n} it={it} it_ref={it_ref} rel={rel:.3e} |b|={np.linalg.norm(b):.3e} "
          f"div={np.linalg.norm(div[keep]):.3e} res={np.linalg.norm(res[keep]):.3e}", flush=True)
    assert it == it_ref and rel <= 1e-10
    assert np.array_equal(u, ur) and np.array_equal(v, vr) and np.array_equal(w, wr)
    assert np.array_equal(p, dp)
    assert np.abs(div - res)[keep].max() <= 1e-9 * np.abs(b).max()
    assert np.linalg.norm(div[keep]) <= 1.01 * eps


def test_correction_bitwise_stretched(lsx):
    w = [lsx.stretched_axis(d, 1.0, 0.5, 1.03) for d in (40, 32, 24)]
    g_ref = O.Geometry(*w).build_body("ellipsoid", (0.5, 0.5, 0.5), (0.25, 0.15, 0.15))
    g_gpu = lsx.Geometry(*w).build("ellipsoid", (0.5, 0.5, 0.5), (0.25, 0.15, 0.15))
    nu, nv, nw = O.face_shapes(g_ref.dims)
    rng = np.random.default_rng(8)
    dp = rng.standard_normal(g_ref.n)
    us, vs, ws = rng.standard_normal(nu), rng.standard_normal(nv), rng.standard_normal(nw)
    p0 = rng.standard_normal(g_ref.n)
    p = p0.copy()
    u, v, w_ = g_gpu.correct(dp, us, vs, ws, p=p, coef=0.7)
    ur, vr, wr = O.velocity_correction(g_ref, dp, us, vs, ws)
    assert np.array_equal(u, ur) and np.array_equal(v, vr) and np.array_equal(w_, wr)
    assert np.array_equal(p, p0 + dp * 0.7)
