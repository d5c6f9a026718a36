"""Pin the CPU oracle against what the mathematics fixes (the reference ships no tests or golden
vectors, and its sources do not build unmodified).

* operator: closed-form 7-point finite-volume Laplacian on empty uniform and stretched grids;
  symmetry off the pinned rows and zero row sums (Neumann conservation) on cut geometries;
* transfers and Galerkin levels: the oracle's coarse operators equal R A P formed with scipy
  from Kronecker products of the 1-D cell-centred interpolation / averaging matrices;
* coarsest solve: with no smoothing, one multigrid application equals P A_c^{-1} R b with
  numpy.linalg.solve;
* ADLJ smoothing: one multigrid application with smoothing equals the same cascade written
  with scipy banded solves of each line's tridiagonal system, on an empty box and on cut
  geometries where the reference's zero-pivot rule (tridiagonal_matrix.h:45-67) decouples rows;
* FGMRES / BiCGStab: the returned residual estimate equals the true residual (Saad's
  identity for FGMRES), and the solution matches a sparse direct solve.
"""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle_py as O
import problems as P


def fv_laplacian(dx, dy, dz, pin=True):
    """Independent textbook assembly: a_nb = area / distance-between-centres."""
    N, M, K = len(dx), len(dy), len(dz)
    n = N * M * K
    rows, cols, vals = [], [], []
    idx = lambda x, y, z: x + N * (y + M * z)
    for z in range(K):
        for y in range(M):
            for x in range(N):
                i = idx(x, y, z)
                diag = 0.0
                for (xx, yy, zz, area, dist) in (
                        (x - 1, y, z, dy[y] * dz[z], 0.5 * (dx[x] + dx[x - 1]) if x > 0 else 0),
                        (x + 1, y, z, dy[y] * dz[z], 0.5 * (dx[x] + dx[x + 1]) if x < N - 1 else 0),
                        (x, y - 1, z, dx[x] * dz[z], 0.5 * (dy[y] + dy[y - 1]) if y > 0 else 0),
                        (x, y + 1, z, dx[x] * dz[z], 0.5 * (dy[y] + dy[y + 1]) if y < M - 1 else 0),
                        (x, y, z - 1, dx[x] * dy[y], 0.5 * (dz[z] + dz[z - 1]) if z > 0 else 0),
                        (x, y, z + 1, dx[x] * dy[y], 0.5 * (dz[z] + dz[z + 1]) if z < K - 1 else 0)):
                    if 0 <= xx < N and 0 <= yy < M and 0 <= zz < K:
                        # staggered mass = mean of the two cell volumes; coefficient area^2 / mass
                        vol = 0.5 * (dx[x] * dy[y] * dz[z] + dx[xx] * dy[yy] * dz[zz])
                        a = area * area / vol
                        rows.append(i); cols.append(idx(xx, yy, zz)); vals.append(a)
                        diag -= a
                rows.append(i); cols.append(i); vals.append(diag)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n)).tolil()
    if pin:
        for p in (N * M - 1, N - 1):
            A[p, :] = 0
            A[p, p] = 1.0
    return A.tocsr()


@pytest.mark.parametrize("dims", [(8, 8, 8), (12, 10, 6)])
def test_body_free_operator_closed_form(dims):
    N, M, K = dims
    g, A = P.body_free(N, M, K)
    ref = fv_laplacian(g.dx, g.dy, g.dz)
    diff = abs(A.to_scipy() - ref).max()
    assert diff <= 1e-12 * abs(ref).max()


def test_stretched_operator_closed_form():
    rng = np.random.default_rng(0)
    dx, dy, dz = (rng.uniform(0.5, 1.5, n) / n for n in (10, 8, 6))
    g = O.Geometry(dx, dy, dz)
    A = g.laplace()
    ref = fv_laplacian(dx, dy, dz)
    assert abs(A.to_scipy() - ref).max() <= 1e-12 * abs(ref).max()


def test_cut_operator_symmetry_and_conservation():
    g, A = P.synthetic_cut(16, seed=3)
    S = A.to_scipy()
    N, M, K = g.dims
    pins = {N * M - 1, N - 1}
    keep = np.array([i not in pins for i in range(A.n)])
    D = (S - S.T).tocsr()
    # asymmetry only in the pinned rows / columns
    Dk = D[keep][:, keep]
    assert abs(Dk).max() == 0.0
    t = g.fields()["type"]
    rs = np.asarray(S.sum(axis=1)).ravel()
    ok = keep & (t != 0)
    assert np.max(np.abs(rs[ok])) <= 1e-15 * abs(S).max()
    solid = t == 0
    assert np.all(S.diagonal()[solid] == 1e-16)


def kron_transfers(dims):
    def p1(nf):
        nc = nf // 2
        Pm = np.zeros((nf, nc))
        for f in range(nf):
            p = f // 2
            if f in (0, nf - 1):
                Pm[f, p] = 1.0
            elif f % 2:
                Pm[f, p], Pm[f, p + 1] = 0.75, 0.25
            else:
                Pm[f, p - 1], Pm[f, p] = 0.25, 0.75
        return Pm

    def r1(nf):
        nc = nf // 2
        Rm = np.zeros((nc, nf))
        for c in range(nc):
            Rm[c, 2 * c] = Rm[c, 2 * c + 1] = 0.5
        return Rm

    N, M, K = dims
    Pk = sp.kron(p1(K), sp.kron(p1(M), p1(N)))
    Rk = sp.kron(r1(K), sp.kron(r1(M), r1(N)))
    return Pk.tocsr(), Rk.tocsr()


@pytest.mark.parametrize("mk", [lambda: P.body_free(16), lambda: P.synthetic_cut(16, seed=5),
                                lambda: P.synthetic_cut(16, 8, 12, seed=6)])
def test_galerkin_levels_equal_RAP(mk):
    g, A = mk()
    lv = O.Solver(A, eps=1.0).levels()
    for l in range(len(lv) - 1):
        Pk, Rk = kron_transfers(lv[l].dims)
        assert np.allclose(Pk @ np.ones(Pk.shape[1]), 1.0) and np.allclose(Rk @ np.ones(Rk.shape[1]), 1.0)
        ref = (Rk @ lv[l].to_scipy() @ Pk).toarray()
        got = lv[l + 1].to_scipy().toarray()
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref)), f"level {l + 1}"


def kron_transfers_odd(dims):
    """The odd-odd-odd prolongation as a Kronecker product of 1-D node-centred interpolations
    (2n - 1 fine cells onto n coarse cells: weight 1 on the coincident fine cell, 1/2 on each
    coarse neighbour of an odd fine cell) and R = P^T."""
    def p1(nf):
        nc = (nf + 1) // 2
        Pm = np.zeros((nf, nc))
        for f in range(nf):
            if f % 2 == 0:
                Pm[f, f // 2] = 1.0
            else:
                Pm[f, (f - 1) // 2] = Pm[f, (f + 1) // 2] = 0.5
        return Pm

    N, M, K = dims
    Pk = sp.kron(p1(K), sp.kron(p1(M), p1(N))).tocsr()
    return Pk, Pk.T.tocsr()


@pytest.mark.parametrize("mk", [lambda: P.body_free(17), lambda: P.sphere(33),
                                lambda: P.synthetic_cut(17, 33, 9, seed=6)])
def test_galerkin_levels_equal_PtAP_on_odd_grids(mk):
    g, A = mk()
    lv = O.Solver(A, eps=1.0).levels()
    assert len(lv) >= 3
    for l in range(len(lv) - 1):
        dims = tuple(int(v) for v in lv[l].dims)
        assert all(v % 2 for v in dims)
        Pk, Rk = kron_transfers_odd(dims)
        assert tuple(lv[l + 1].dims) == tuple((v + 1) // 2 for v in dims)
        ref = (Rk @ lv[l].to_scipy() @ Pk).tocsr()
        got = lv[l + 1].to_scipy().tocsr()
        assert abs(got - ref).max() <= 1e-13 * abs(ref).max(), f"level {l + 1}"
        # the coarse operator keeps a 27-point stencil
        assert np.all(np.diff(lv[l + 1].rowptr) <= 27)


@pytest.mark.parametrize("kind", ["fgmres", "bicgstab"])
@pytest.mark.parametrize("mk", [lambda: P.sphere(33), lambda: P.synthetic_cut(17, 33, 9, seed=6)])
def test_odd_grid_solves_converge(mk, kind):
    g, A = mk()
    b = P.random_rhs(g, seed=3, A=A)
    S = O.Solver(A, kind=kind, eps=1e-8 * np.linalg.norm(b), max_it=200)
    x, it, nr = S.solve(b)
    assert it < 40
    assert np.linalg.norm(b - A.matvec(x)) <= 1.05e-8 * np.linalg.norm(b)


def test_mixed_parity_levels_are_rejected():
    g, A = P.body_free(17, 16, 17)
    with pytest.raises(RuntimeError, match="mixed-parity"):
        O.Solver(A, eps=1.0)


def test_coarsest_lu_with_no_smoothing():
    g, A = P.synthetic_cut(8, seed=2)
    S = O.Solver(A, eps=1.0, n_post=0)
    lv = S.levels()
    b = np.random.default_rng(1).standard_normal(A.n)
    # cascade with no smoothing: z = P0 P1 ... A_L^{-1} R ... R0 b
    r = [b]
    mats = []
    for l in range(len(lv) - 1):
        Pk, Rk = kron_transfers(lv[l].dims)
        mats.append((Pk, Rk))
        r.append(Rk @ r[-1])
    x = np.linalg.solve(lv[-1].to_scipy().toarray(), r[-1])
    for l in range(len(lv) - 2, -1, -1):
        x = mats[l][0] @ x
    z = S.mg_apply(b)
    assert np.linalg.norm(z - x) <= 1e-12 * np.linalg.norm(x)


def banded_line_solve(Ad, dims, direction, rhs, z_head_quirk=True):
    """Solve the tridiagonal line systems of one direction with scipy.linalg.solve_banded."""
    N, M, K = dims
    idx = lambda x, y, z: x + N * (y + M * z)
    out = np.zeros_like(rhs)
    if direction == 0:
        lines = [[idx(x, y, z) for x in range(N)] for z in range(K) for y in range(M)]
        heads = [l[0] for l in lines]
    elif direction == 1:
        lines = [[idx(x, y, z) for y in range(M)] for z in range(K) for x in range(N)]
        heads = [l[0] for l in lines]
    else:
        lines, heads = [], []
        for y in range(M):
            for x in range(N):
                lines.append([idx(x, y, z) for z in range(K)])
                # the reference z sweep reads its first right-hand side at y*N*K + x
                heads.append(y * N * K + x if z_head_quirk else idx(x, y, 0))
    for line, head in zip(lines, heads):
        L = len(line)
        ab = np.zeros((3, L))
        for k, i in enumerate(line):
            ab[1, k] = Ad[i, i]
            if k + 1 < L:
                ab[0, k + 1] = Ad[i, line[k + 1]]
                ab[2, k] = Ad[line[k + 1], i]
        rr = rhs[line].copy()
        rr[0] = rhs[head]
        out[line] = sl.solve_banded((1, 1), ab, rr)
    return out


def test_adlj_cascade_against_banded_solves():
    g, A = P.body_free(8)
    omega, n_post = 0.805, 2
    S = O.Solver(A, eps=1.0, n_post=n_post, omega=omega)
    lv = S.levels()
    b = np.random.default_rng(4).standard_normal(A.n)
    mats = [kron_transfers(l.dims) for l in lv[:-1]]
    r = [b.copy()]
    for l in range(len(lv) - 1):
        r.append(mats[l][1] @ r[-1])
    ch = np.linalg.solve(lv[-1].to_scipy().toarray(), r[-1])
    for l in range(len(lv) - 2, -1, -1):
        Ad = lv[l].to_scipy().toarray()
        ch = mats[l][0] @ ch
        c = ch.copy()
        for _ in range(n_post):
            for direction in (2, 1, 0):
                r[l] = r[l] - Ad @ c
                c = omega * banded_line_solve(Ad, lv[l].dims, direction, r[l])
                ch = ch + c
    z = S.mg_apply(b)
    assert np.linalg.norm(z - ch) <= 1e-12 * np.linalg.norm(ch)


@pytest.mark.parametrize("kind", ["fgmres", "bicgstab"])
@pytest.mark.parametrize("mk", [lambda: P.body_free(16), lambda: P.synthetic_cut(16, seed=7)])
def test_krylov_residual_identity_and_direct_solve(mk, kind):
    g, A = mk()
    b = P.random_rhs(g, seed=3)
    eps = 1e-10 * np.linalg.norm(b)
    x, it, nr = O.Solver(A, kind=kind, eps=eps, restart=60, max_it=500).solve(b)
    S = A.to_scipy()
    true = np.linalg.norm(b - S @ x)
    assert nr <= eps
    assert abs(true - nr) <= 1e-4 * nr + 1e-14 * np.linalg.norm(b)
    xd = spl.spsolve(S.tocsc(), b)
    t = g.fields()["type"]
    fl = t != 0  # solid-cell values are unconstrained (1e-16 diagonal, no couplings)
    assert np.linalg.norm((x - xd)[fl]) <= 1e-6 * np.linalg.norm(xd[fl])


def test_fgmres_max_it_and_zero_rhs():
    g, A = P.body_free(8)
    b = P.random_rhs(g, seed=1)
    x, it, nr = O.Solver(A, eps=1e-30, max_it=4).solve(b)
    assert it == 4 and nr > 0
    x, it, nr = O.Solver(A, eps=1e-8).solve(np.zeros(A.n))
    assert it == 0 and nr == 0.0


def line_systems(Ad, dims, direction, z_head_quirk=True):
    """Cells of every line of one direction, in line order, and the cell each line's first
    right-hand side is read from (the reference z sweep's head quirk)."""
    N, M, K = dims
    idx = lambda x, y, z: x + N * (y + M * z)
    if direction == 0:
        lines = [[idx(x, y, z) for x in range(N)] for z in range(K) for y in range(M)]
        return lines, [l[0] for l in lines]
    if direction == 1:
        lines = [[idx(x, y, z) for y in range(M)] for z in range(K) for x in range(N)]
        return lines, [l[0] for l in lines]
    lines, heads = [], []
    for y in range(M):
        for x in range(N):
            lines.append([idx(x, y, z) for z in range(K)])
            heads.append(y * N * K + x if z_head_quirk else idx(x, y, 0))
    return lines, heads


def decoupled_rows(T, tiny):
    """Rows of the tridiagonal T that the reference factorisation decouples: eliminating without
    pivoting, the pivot of row i is det(T[s:i+1, s:i+1]) / det(T[s:i, s:i]), s the first row after
    the last decoupled one; a pivot of magnitude below `tiny` is set to 0, which makes x_i = 0 and
    cuts the line there (tridiagonal_matrix.h:55-65)."""
    out, s = [], 0
    for i in range(T.shape[0]):
        den = np.linalg.det(T[s:i, s:i]) if i > s else 1.0
        piv = np.linalg.det(T[s:i + 1, s:i + 1]) / den
        if abs(piv) < tiny:
            out.append(i)
            s = i + 1
    return out


def banded_line_solve_decoupled(Ad, dims, direction, rhs, tiny, stats=None):
    """The line solves of one direction with the decoupled rows written out: such a row is an
    identity row with right-hand side 0 and its couplings dropped, the rest by solve_banded."""
    out = np.zeros_like(rhs)
    lines, heads = line_systems(Ad, dims, direction)
    for line, head in zip(lines, heads):
        L = len(line)
        T = Ad[np.ix_(line, line)]
        T = np.diag(np.diag(T)) + np.diag(np.diag(T, 1), 1) + np.diag(np.diag(T, -1), -1)
        rr = rhs[line].copy()
        rr[0] = rhs[head]
        dec = decoupled_rows(T, tiny)
        if stats is not None:
            stats.extend(line[k] for k in dec)
        for k in dec:
            T[k, :] = 0.0
            T[:, k] = 0.0
            T[k, k] = 1.0
            rr[k] = 0.0
        ab = np.zeros((3, L))
        ab[1] = np.diag(T)
        ab[0, 1:] = np.diag(T, 1)
        ab[2, :-1] = np.diag(T, -1)
        out[line] = sl.solve_banded((1, 1), ab, rr)
    return out


def cut_with_tiny_cells(n, seed):
    """synthetic_cut with three cut cells whose six faces are closed to 1e-7: rows with a tiny
    (not zero) diagonal, which the zero-pivot rule decouples like the solid rows."""
    g, _ = P.synthetic_cut(n, seed=seed)
    f = g.fields()
    N, M, K = g.dims
    t = f["type"]
    cut = [i for i in np.nonzero(t == P.CUT_CODE)[0]]
    picked = cut[::max(1, len(cut) // 3)][:3]
    thx, thy, thz = f["thx"].copy(), f["thy"].copy(), f["thz"].copy()
    for i in picked:
        thx[i] = thx[i - 1] = 1e-7
        thy[i] = thy[i - N] = 1e-7
        thz[i] = thz[i - N * M] = 1e-7
    g.set_fields(thx=thx, thy=thy, thz=thz)
    return g, g.laplace(pin=True), picked


@pytest.mark.parametrize("case", ["sphere16", "cut16_tiny"])
def test_adlj_cascade_zero_pivots_on_cut_geometry(case):
    if case == "sphere16":
        g, A = P.sphere(16)
        planted = []
    else:
        g, A, planted = cut_with_tiny_cells(16, seed=7)
    omega, n_post = 0.805, 2
    S = O.Solver(A, eps=1.0, n_post=n_post, omega=omega)
    lv = S.levels()
    b = P.random_rhs(g, seed=4, A=A) + 0.1 * np.random.default_rng(5).standard_normal(A.n)
    mats = [kron_transfers(l.dims) for l in lv[:-1]]
    r = [b.copy()]
    for l in range(len(lv) - 1):
        r.append(mats[l][1] @ r[-1])
    ch = np.linalg.solve(lv[-1].to_scipy().toarray(), r[-1])
    fine_decoupled = []
    for l in range(len(lv) - 2, -1, -1):
        Ad = lv[l].to_scipy().toarray()
        # one threshold per level: EQ_TINY2 * max|diagonal|, at least EQ_TINY (ls_stag_real_constants.h:30-31)
        tiny = max(1e-5 * np.abs(np.diag(Ad)).max(), 1e-12)
        ch = mats[l][0] @ ch
        c = ch.copy()
        for _ in range(n_post):
            for direction in (2, 1, 0):
                r[l] = r[l] - Ad @ c
                stats = fine_decoupled if l == 0 else None
                c = omega * banded_line_solve_decoupled(Ad, lv[l].dims, direction, r[l], tiny, stats)
                ch = ch + c
    z = S.mg_apply(b)
    types = g.fields()["type"]
    empty = set(int(i) for i in np.nonzero(A.to_scipy().diagonal() == 0.0)[0])  # cut cells with every face closed
    dec = set(fine_decoupled)
    n_solid = int(np.sum(types == 0))
    print(f"{case}: decoupled fine rows {len(dec)} (solid cells {n_solid}, empty rows {len(empty)}, planted tiny "
          f"cells {planted}); rel diff {np.linalg.norm(z - ch) / np.linalg.norm(ch):.3e}")
    assert n_solid > 0 and all(types[i] == 0 or i in planted or i in empty for i in dec)
    assert {int(i) for i in np.nonzero(types == 0)[0]} <= dec
    assert set(planted) <= dec
    assert np.linalg.norm(z - ch) <= 1e-12 * np.linalg.norm(ch)
