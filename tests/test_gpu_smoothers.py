"""The V-cycle smoothers on the finest level against plain restatements on the oracle's CSR
operator A (scipy):
  red-black Gauss-Seidel: red cells (x+y+z even) then black cells, x_i = (b_i - sum_j a_ij x_j) / a_ii;
  line relaxation: for d = x, y, z in turn, x <- T_d^{-1} (b - (A - T_d) x), T_d the part of A
  coupling cells along d (the diagonal included), solved exactly per line;
  damped Jacobi: x <- x + omega D^{-1} (b - A x), D the diagonal of A.
Rows of A that are identities (pinned cells) stay identities, empty rows give 0.  FP64 results
agree to rounding; the mixed-precision smoother (FP32 values and faces) to FP32 rounding.  Then
the line-relaxation V-cycle against the red-black one on a strongly stretched grid."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import problems as P

pytestmark = pytest.mark.gpu


def direction_parts(A, dims):
    N, M, K = dims
    S = A.to_scipy().tocoo()
    n = S.shape[0]
    idx = np.arange(n)
    x, y, z = idx % N, (idx // N) % M, idx // (N * M)
    diag = S.row == S.col
    parts = []
    for off, same in ((1, lambda r, c: (y[r] == y[c]) & (z[r] == z[c])),
                      (N, lambda r, c: (x[r] == x[c]) & (z[r] == z[c])),
                      (N * M, lambda r, c: (x[r] == x[c]) & (y[r] == y[c]))):
        along = diag | ((np.abs(S.row - S.col) == off) & same(S.row, S.col))
        T = sp.csr_matrix((S.data[along], (S.row[along], S.col[along])), shape=(n, n))
        R = sp.csr_matrix((S.data[~along], (S.row[~along], S.col[~along])), shape=(n, n))
        parts.append((T, R))
    return parts


def ref_line_sweep(A, dims, b, x):
    S = A.to_scipy().tocsr()
    empty = S.diagonal() == 0
    for T, R in direction_parts(A, dims):
        T = (T + sp.diags(empty.astype(float))).tocsc()
        rhs = b - R @ x
        rhs[empty] = 0.0
        x = spl.spsolve(T, rhs)
    return x


JACOBI_OMEGA = 0.805  # the damping of lsx.mixed_task (the Task default omega)


def ref_jacobi_sweep(A, dims, b, x):
    S = A.to_scipy().tocsr()
    d = S.diagonal()
    upd = x + JACOBI_OMEGA * (b - S @ x) / np.where(d != 0, d, 1.0)
    return np.where(d != 0, upd, x)


def ref_rb_sweep(A, dims, b, x):
    N, M, K = dims
    S = A.to_scipy().tocsr()
    d = S.diagonal()
    idx = np.arange(S.shape[0])
    red = ((idx % N + (idx // N) % M + idx // (N * M)) % 2) == 0
    x = x.copy()
    for c in (red, ~red):
        r = b - S @ x + d * x
        upd = np.where(d != 0, r / np.where(d != 0, d, 1.0), 0.0)
        x[c] = upd[c]
    return x


CASES = [("sphere32", lambda: P.sphere(32)), ("ellipsoid32_stretched", lambda: P.ellipsoid_stretched(32)),
         ("cut32x16x24", lambda: P.synthetic_cut(32, 16, 24, seed=3))]


@pytest.mark.parametrize("smoother", ["line", "rbgs", "jacobi"])
@pytest.mark.parametrize("precision,tol", [("fp64-vcycle", 1e-12), ("mixed", 2e-5)])
@pytest.mark.parametrize("name,mk", CASES, ids=[c[0] for c in CASES])
def test_smoother_sweep(lsx, name, mk, precision, tol, smoother):
    g, A = mk()
    dims = g.dims
    S = lsx.solver("fgmres", precision, smoother=smoother)
    S.set(A, np.zeros(A.n), np.zeros(A.n), lsx.mixed_task(1.0))
    act = A.to_scipy().diagonal() < 0
    rng = np.random.default_rng(3)
    b = rng.standard_normal(A.n) * act
    x0 = rng.standard_normal(A.n) * act
    ref = {"line": ref_line_sweep, "rbgs": ref_rb_sweep, "jacobi": ref_jacobi_sweep}[smoother]
    x1 = S.smooth(b, x0, sweeps=1)
    x1r = ref(A, dims, b, x0)
    x0z = S.smooth(b, x0, sweeps=1, zero_guess=True)
    x0zr = ref(A, dims, b, np.zeros(A.n))
    e1 = np.linalg.norm(x1 - x1r) / np.linalg.norm(x1r)
    e0 = np.linalg.norm(x0z - x0zr) / np.linalg.norm(x0zr)
    print(f"case={name} precision={precision} smoother={smoother} rel err one sweep={e1:.3e} "
          f"from zero={e0:.3e}", flush=True)
    assert e1 <= tol and e0 <= tol


@pytest.mark.parametrize("precision", ["mixed", "fp64-vcycle"])
def test_line_vcycle_on_stretched_grid(lsx, precision):
    """Aspect ratios near 11 (ratio 1.08 over 32 cells): the line-relaxation cycle reaches the
    tolerance in fewer iterations than the red-black one, also with lines on the two finest
    levels only."""
    g, A = P.ellipsoid_stretched(64, ratio=1.08)
    b = P.random_rhs(g, seed=2, A=A)
    eps = 1e-8 * np.linalg.norm(b)
    its = {}
    for smoother in ("rbgs", "line", "line2"):
        S = lsx.solver("fgmres", precision, smoother="line" if smoother.startswith("line") else smoother,
                       line_levels=2 if smoother == "line2" else 0)
        x = np.zeros(A.n)
        S.set(A, b, x, lsx.mixed_task(eps, restart=30, max_it=500))
        its[smoother] = S.solve()
        true_res = np.linalg.norm(b - A.matvec(x))
        print(f"precision={precision} smoother={smoother} it={its[smoother]} "
              f"true_rel_res={true_res / np.linalg.norm(b):.3e}", flush=True)
        assert true_res <= 1.05 * eps
    assert its["line"] < its["rbgs"] and its["line2"] < its["rbgs"]


FUSED_CASES = [("sphere64", lambda: P.sphere(64)), ("cut32x16x24", lambda: P.synthetic_cut(32, 16, 24, seed=3)),
               ("sphere96x64x32", lambda: P.sphere(96, 64, 32))]


@pytest.mark.parametrize("precision", ["mixed", "mixed16", "fp64-vcycle"])
@pytest.mark.parametrize("name,mk", FUSED_CASES, ids=[c[0] for c in FUSED_CASES])
def test_fused_sweep_is_bitwise(lsx, name, mk, precision):
    """The one-pass red-black sweep (tiles of 64 x 16 columns, 32 planes, red values of a
    one-cell ring recomputed per tile) gives the same bits as red and black half-sweeps in
    place: one and two sweeps, from zero and from a given x, and whole solves."""
    g, A = mk()
    act = A.to_scipy().diagonal() < 0
    rng = np.random.default_rng(5)
    b = P.random_rhs(g, seed=3, A=A)
    x0 = rng.standard_normal(A.n) * act
    out = {}
    for fused in (0, 2):
        S = lsx.solver("fgmres", precision, fused_sweep=fused)
        x = np.zeros(A.n)
        S.set(A, b, x, lsx.mixed_task(1e-8 * np.linalg.norm(b), restart=30, max_it=500))
        sm = [S.smooth(b, x0, sweeps=k, zero_guess=z) for k in (1, 2) for z in (False, True)]
        it = S.solve()
        out[fused] = (sm, x.copy(), it)
    for a, c in zip(out[0][0], out[2][0]):
        assert np.array_equal(a, c), np.abs(a - c).max()
    print(f"case={name} precision={precision} it={out[0][2]}/{out[2][2]} "
          f"max|dx|={np.abs(out[0][1] - out[2][1]).max():.3e}", flush=True)
    assert out[0][2] == out[2][2]
    assert np.array_equal(out[0][1], out[2][1])


def test_line_relaxation_limits(lsx):
    """Line relaxation takes x lines of at most 512 cells and y and z lines of at most 256: a
    longer line is refused when the hierarchy is built, with either limit."""
    for dims in ((520, 4, 4), (8, 264, 4)):
        g, A = P.body_free(*dims)
        S = lsx.solver("fgmres", "mixed", smoother="line")
        with pytest.raises(lsx.LsxError):
            S.set(A, np.zeros(A.n), np.zeros(A.n), lsx.mixed_task(1.0))


@pytest.mark.parametrize("precision", ["mixed", "fp64-vcycle"])
def test_jacobi_vcycle_converges(lsx, precision):
    """The damped-Jacobi V-cycle preconditions FGMRES to the tolerance on the true residual (more
    iterations than red-black, which converges in fewer)."""
    g, A = P.sphere(64)
    b = P.random_rhs(g, seed=5, A=A)
    eps = 1e-8 * np.linalg.norm(b)
    its = {}
    for smoother in ("rbgs", "jacobi"):
        S = lsx.solver("fgmres", precision, smoother=smoother)
        x = np.zeros(A.n)
        S.set(A, b, x, lsx.mixed_task(eps, restart=30, max_it=500))
        its[smoother] = S.solve()
        true_res = np.linalg.norm(b - A.matvec(x))
        print(f"precision={precision} smoother={smoother} it={its[smoother]} "
              f"true_rel_res={true_res / np.linalg.norm(b):.3e}", flush=True)
        assert true_res <= 1.05 * eps
    assert its["rbgs"] <= its["jacobi"] < 500
