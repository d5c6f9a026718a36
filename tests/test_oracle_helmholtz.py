"""Pins of the oracle's momentum (Helmholtz) operators (oracle/oracle_helmholtz.cpp) and of the
one-level multigrid preconditioner the staggered meshes use.

* Closed form on a uniform grid of cell size h: every viscous coupling is nu h (a neighbour one
  cell away across a face of area h^2), a no-slip wall half a cell away gives 2 nu h, the diagonal
  is -(1.5/dt) h^3 minus the couplings and walls, and a flow task adds the outflow term to the last
  cell of every x row.  The expectation is stated per row from these rules plus the reference's
  listed deviations from them (no west coupling in the x-mesh's west-boundary diagonal, the east
  coupling in the y and z meshes' east-boundary diagonals, the y operator's bottom/top wall term
  scaled by dz on edges and corners, the outflow terms of helmholtz_discrete_operator_{x,y,z}.h),
  so a dropped, doubled or misplaced term of the restatement fails it.
* Symmetry of the couplings between interior (fluid) rows on a stretched grid.
* FGMRES with the one-level ADLJ preconditioner reaches the tolerance and agrees with a sparse
  direct solve (scipy).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle_py as O

RE, DT = 300.0, 0.002
NU, COEF = 1.0 / RE, 1.5 / DT

B_T = {-3: "B", -12: "B", -13: "B", -14: "B", -15: "B", -24: "B", -25: "B", -26: "B", -27: "B",
       -2: "T", -8: "T", -9: "T", -10: "T", -11: "T", -20: "T", -21: "T", -22: "T", -23: "T"}
N_S = {-4: "N", -8: "N", -12: "N", -16: "N", -17: "N", -20: "N", -21: "N", -24: "N", -25: "N",
       -5: "S", -9: "S", -13: "S", -18: "S", -19: "S", -22: "S", -23: "S", -26: "S", -27: "S"}
E_W = {-6: "E", -10: "E", -14: "E", -16: "E", -18: "E", -20: "E", -22: "E", -24: "E", -26: "E",
       -7: "W", -11: "W", -15: "W", -17: "W", -19: "W", -21: "W", -23: "W", -25: "W", -27: "W"}


def staggered_types(base, d):
    """Type of every face control volume: the smaller of its two base cells' types."""
    if d == 0:
        a, b = base[:, :, :-1], base[:, :, 1:]
    elif d == 1:
        a, b = base[:, :-1, :], base[:, 1:, :]
    else:
        a, b = base[:-1, :, :], base[1:, :, :]
    return np.minimum(a, b)


def expected_uniform(n, d, flow):
    """Rows of the uniform-grid operator (h = 1/n) from the closed-form rules."""
    h = 1.0 / n
    nh = NU * h
    g = O.Geometry(*([np.full(n, h)] * 3))
    base = g.fields()["type"].reshape(n, n, n)
    t3 = staggered_types(base, d)
    K, M, N = t3.shape
    rows = {}
    for k in range(K):
        for j in range(M):
            for i in range(N):
                t = int(t3[k, j, i])
                fluid = t == 11111111
                bt, ns, ew = B_T.get(t), N_S.get(t), E_W.get(t)
                has = {"b": fluid or bt != "B", "t": fluid or bt != "T", "s": fluid or ns != "S",
                       "n": fluid or ns != "N", "w": fluid or ew != "W", "e": fluid or ew != "E"}
                off = {q: nh for q in has if has[q]}
                diag = -COEF * h ** 3
                if fluid:
                    diag -= 6 * nh
                elif d == 0:
                    diag -= sum(off[q] for q in ("b", "s", "n", "t") if q in off)
                    diag -= nh if ew == "W" else 2 * nh          # (k_w + k_e), west: k_e only
                    if ns:
                        diag -= 2 * nh                           # no-slip wall half a cell away
                    if bt and not flow:
                        diag -= 2 * nh
                elif d == 1:
                    diag -= sum(off[q] for q in ("b", "t", "w") if q in off)
                    diag -= nh                                   # k_e, east boundary included
                    diag -= 2 * nh                               # k_s + k_n (or one and its wall)
                    if bt and not flow:
                        diag -= 2 * nh * (1.0 if t in (-2, -3) else h)  # edges/corners: x dz
                else:
                    diag -= sum(off[q] for q in ("s", "n", "w") if q in off)
                    diag -= nh                                   # k_e, east boundary included
                    diag -= 2 * nh if not (bt and flow) else nh  # k_b + k_t, or one and its wall
                    if ns:
                        diag -= 2 * nh
                if flow and i == N - 1:
                    diag += nh if d == 0 else 2 * nh             # outflow term
                rows[i + N * (j + M * k)] = (off, diag, (N, M, K))
    return rows


@pytest.mark.parametrize("flow", [True, False])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_helmholtz_uniform_closed_form(d, flow):
    n = 8
    g = O.Geometry(*([np.full(n, 1.0 / n)] * 3))
    A = O.helmholtz(g, d, NU, COEF, flow=flow)
    S = A.to_scipy().tocsr()
    exp = expected_uniform(n, d, flow)
    N, M, K = A.dims
    step = {"b": -N * M, "s": -N, "w": -1, "e": 1, "n": N, "t": N * M}
    for r, (off, diag, _) in exp.items():
        row = S.getrow(r)
        got = dict(zip(row.indices.tolist(), row.data.tolist()))
        want = {r + step[q]: v for q, v in off.items()}
        want[r] = diag
        assert set(got) == set(want), (r, sorted(got), sorted(want))
        for c, v in want.items():
            assert abs(got[c] - v) <= 1e-12 * abs(diag), (d, flow, r, c, got[c], v)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_helmholtz_interior_symmetry_stretched(d):
    w = [O_axis(m, r) for m, r in ((12, 1.05), (10, 1.08), (9, 1.03))]
    g = O.Geometry(*w)
    A = O.helmholtz(g, d, NU, COEF, flow=True)
    S = A.to_scipy().tocsr()
    N, M, K = A.dims
    base = g.fields()["type"].reshape(len(w[2]), len(w[1]), len(w[0]))
    fluid = (staggered_types(base, d) == 11111111).ravel()
    F = S[fluid][:, fluid].toarray()
    offd = F - np.diag(np.diag(F))
    assert np.count_nonzero(offd) > 0
    assert np.allclose(offd, offd.T, rtol=1e-14, atol=0)
    # strictly diagonally dominant rows (the mass term)
    D = np.abs(S.diagonal())
    R = np.asarray(abs(S).sum(axis=1)).ravel() - D
    assert np.all(D > R)


def O_axis(m, ratio):
    w = ratio ** np.abs(np.arange(m) + 0.5 - m / 2)
    return w / w.sum()


@pytest.mark.parametrize("kind", ["fgmres", "bicgstab"])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_helmholtz_one_level_solve_matches_direct(d, kind):
    w = [O_axis(m, r) for m, r in ((16, 1.04), (12, 1.06), (10, 1.02))]
    g = O.Geometry(*w)
    A = O.helmholtz(g, d, NU, COEF, flow=True)
    rng = np.random.default_rng(d)
    b = rng.standard_normal(A.n)
    eps = 1e-10 * np.linalg.norm(b)
    x, it, nr = O.Solver(A, kind=kind, eps=eps, restart=30, max_it=200, precond="mg1").solve(b)
    xd = spla.spsolve(A.to_scipy().tocsc(), b)
    rel = np.linalg.norm(x - xd) / np.linalg.norm(xd)
    print(f"d={d} {kind} it={it} norm_r={nr:.3e} rel vs direct={rel:.3e}")
    assert nr <= eps and it < 30
    assert rel <= 1e-9
