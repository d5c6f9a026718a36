"""Capped V-cycle hierarchies (lsx_options.levels): the cycle against a plain numpy restatement
of the multigrid it describes, and the benchmark workloads' level counts at their sizes (3 levels
at 64^3, 5 at 128^3; the 6-level hierarchy at 256^3 is in test_gpu_full_size.py).

Restatement (FP64, on the oracle's CSR operator A): red-black Gauss-Seidel sweeps (red cells
x+y+z even first; identity rows give x = b; empty rows keep their value), residual restricted by
summing the 8 children, coarse faces half the sum of the four fine faces they cover, coarse
diagonal minus the sum of the six faces, each pinned cell a Dirichlet sink (its fine face sum
taken off the diagonal of the coarse cell that contains it), the coarsest level solved exactly
(numpy.linalg.solve), cell-centred trilinear prolongation (weights 3/4, 1/4 per axis, 1 on the
boundary cells)."""
import numpy as np
import pytest

import oracle_py as O
import problems as P

pytestmark = pytest.mark.gpu


def fine_faces(A, dims):
    N, M, K = dims
    S = A.to_scipy().tocsr()
    n = S.shape[0]
    d = S.diagonal().copy()
    # identity rows (the reference keeps their pattern with explicit zeros: count the values)
    pins = [i for i in range(n) if S[i, i] == 1.0 and np.count_nonzero(S.getrow(i).data) == 1]
    fe, fn, ft = np.zeros(n), np.zeros(n), np.zeros(n)
    idx = np.arange(n)
    x, y, z = idx % N, (idx // N) % M, idx // (N * M)
    for f, off, has in ((fe, 1, x < N - 1), (fn, N, y < M - 1), (ft, N * M, z < K - 1)):
        for i in idx[has]:
            j = i + off
            f[i] = S[j, i] if i in pins else S[i, j]
    return fe, fn, ft, d, pins


def op_from_faces(fe, fn, ft, dg, dims):
    import scipy.sparse as sp
    N, M, K = dims
    n = N * M * K
    rows, cols, vals = [np.arange(n)], [np.arange(n)], [dg]
    for f, off in ((fe, 1), (fn, N), (ft, N * M)):
        i = np.nonzero(f)[0]
        rows += [i, i + off]
        cols += [i + off, i]
        vals += [f[i], f[i]]
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def coarse_faces(fe, fn, ft, dims):
    N, M, K = dims
    r = lambda a: a.reshape(K, M, N)
    e, nn, t = r(fe), r(fn), r(ft)
    Nc, Mc, Kc = N // 2, M // 2, K // 2
    ce = np.zeros((Kc, Mc, Nc))
    cn = np.zeros((Kc, Mc, Nc))
    ct = np.zeros((Kc, Mc, Nc))
    ce[:, :, :-1] = 0.5 * (e[0::2, 0::2, 1:-1:2] + e[0::2, 1::2, 1:-1:2] + e[1::2, 0::2, 1:-1:2] + e[1::2, 1::2, 1:-1:2])
    cn[:, :-1, :] = 0.5 * (nn[0::2, 1:-1:2, 0::2] + nn[0::2, 1:-1:2, 1::2] + nn[1::2, 1:-1:2, 0::2] + nn[1::2, 1:-1:2, 1::2])
    ct[:-1, :, :] = 0.5 * (t[1:-1:2, 0::2, 0::2] + t[1:-1:2, 0::2, 1::2] + t[1:-1:2, 1::2, 0::2] + t[1:-1:2, 1::2, 1::2])
    return ce.ravel(), cn.ravel(), ct.ravel(), (Nc, Mc, Kc)


def neumann_diag(fe, fn, ft, dims):
    N, M, K = dims
    d = -(fe + fn + ft)
    d[1:] -= np.where(np.arange(1, len(fe)) % N != 0, fe[:-1], 0.0)
    d[N:] -= np.where((np.arange(N, len(fn)) // N) % M != 0, fn[:-N], 0.0)
    d[N * M:] -= ft[:-N * M]
    return d


def rb_sweep(S, dims, b, x):
    N, M, K = dims
    d = S.diagonal()
    idx = np.arange(S.shape[0])
    red = ((idx % N + (idx // N) % M + idx // (N * M)) % 2) == 0
    x = x.copy()
    for c in (red, ~red):
        r = b - S @ x + d * x
        upd = np.where(d != 0, r / np.where(d != 0, d, 1.0), x)
        x[c] = upd[c]
    return x


def axis_weights(nf):
    W = np.zeros((nf, nf // 2))
    for f in range(nf):
        p = f >> 1
        if f == 0 or f == nf - 1:
            W[f, p] = 1.0
        elif f & 1:
            W[f, p], W[f, p + 1] = 0.75, 0.25
        else:
            W[f, p - 1], W[f, p] = 0.25, 0.75
    return W


def prolong(xc, dims):
    N, M, K = dims
    Wx, Wy, Wz = axis_weights(N), axis_weights(M), axis_weights(K)
    c = xc.reshape(K // 2, M // 2, N // 2)
    return np.einsum("kc,jb,ia,cba->kji", Wz, Wy, Wx, c).ravel()


def restrict(r, dims):
    N, M, K = dims
    a = r.reshape(K // 2, 2, M // 2, 2, N // 2, 2)
    return a.sum(axis=(1, 3, 5)).ravel()


def vcycle_ref(A, dims, levels, b, n_pre=2, n_post=2):
    """The V-cycle with `levels` levels (finest included), as the module docstring states."""
    fe, fn, ft, _, pins = fine_faces(A, dims)
    ops = [A.to_scipy().tocsr()]
    dl = [dims]
    sigma = []
    N, M, K = dims
    for p in pins:
        x, y, z = p % N, (p // N) % M, p // (N * M)
        s = fe[p] + fn[p] + ft[p]
        s += fe[p - 1] if x > 0 else 0.0
        s += fn[p - N] if y > 0 else 0.0
        s += ft[p - N * M] if z > 0 else 0.0
        sigma.append((p, s))
    f = (fe, fn, ft)
    for l in range(1, levels):
        ce, cn, ct, dc = coarse_faces(*f, dl[-1])
        dg = neumann_diag(ce, cn, ct, dc)
        for p, s in sigma:
            x, y, z = p % N, (p // N) % M, p // (N * M)
            dg[(x >> l) + dc[0] * ((y >> l) + dc[1] * (z >> l))] -= s
        ops.append(op_from_faces(ce, cn, ct, dg, dc))
        dl.append(dc)
        f = (ce, cn, ct)

    def cyc(l, bl):
        S = ops[l]
        if l == levels - 1:
            D = S.toarray()
            empty = np.diag(D) == 0
            D[empty, empty] = 1.0
            return np.linalg.solve(D, bl)
        x = np.zeros_like(bl)
        for _ in range(n_pre):
            x = rb_sweep(S, dl[l], bl, x)
        xc = cyc(l + 1, restrict(bl - S @ x, dl[l]))
        x = x + prolong(xc, dl[l])
        for _ in range(n_post):
            x = rb_sweep(S, dl[l], bl, x)
        return x

    return cyc(0, b)


@pytest.mark.parametrize("n,levels", [(16, 2), (32, 3), (16, 3)])
def test_capped_cycle_equals_restatement(lsx, n, levels):
    """levels=2 at 16^3 and 3 at 16^3: a 512- / 64-unknown coarsest level solved apart from the
    cycle; 3 at 32^3: the two coarse levels and the dense 512-unknown inverse in the one-workgroup
    coarse kernel."""
    g, A = P.synthetic_cut(n, seed=5)
    v = P.random_rhs(g, seed=6, A=A)
    S = lsx.solver("fgmres", "fp64-vcycle", levels=levels)
    S.set(A, v, np.zeros(A.n), lsx.mixed_task(1.0))
    assert S.num_levels() == levels
    z = S.precondition(v)
    zr = vcycle_ref(A, g.dims, levels, v)
    act = A.to_scipy().diagonal() != 0
    err = np.linalg.norm((z - zr)[act]) / np.linalg.norm(zr[act])
    print(f"n={n} levels={levels} rel diff vs restatement={err:.3e}", flush=True)
    assert err <= 1e-11
    # the transposed cycle is the exact transpose of the capped cycle
    rng = np.random.default_rng(7)
    x, y = rng.standard_normal(A.n) * act, rng.standard_normal(A.n) * act
    lhs = float(y @ S.precondition(x))
    rhs = float(x @ S.precondition_transpose(y))
    print(f"<Bx,y>={lhs:.15e} <x,B^T y>={rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + abs(rhs))


def test_level_cap_errors(lsx):
    g, A = P.sphere(16)
    for levels, precision in ((1, "fp64-vcycle"), (3, "fp64")):
        S = lsx.solver("fgmres", precision, levels=levels)
        with pytest.raises(lsx.LsxError):
            S.set(A, np.zeros(A.n), np.zeros(A.n), lsx.mixed_task(1.0) if precision != "fp64" else lsx.Task())
    # 2 levels at 64^3 leave 32^3 = 32768 unknowns on the coarsest level: over the dense limit
    g, A = P.body_free(64)
    S = lsx.solver("fgmres", "mixed", levels=2)
    with pytest.raises(lsx.LsxError):
        S.set(A, np.zeros(A.n), np.zeros(A.n), lsx.mixed_task(1.0))


def test_c1_three_level_rbgs(lsx):
    """C1's multigrid variant: 64^3 sphere, FGMRES(30) with a 3-level V(2,2) red-black cycle in
    FP64 (coarsest 16^3 = 4096 unknowns by the dense inverse) to 1e-8, against the oracle's FP64
    solve of the same system."""
    g, A = P.sphere(64)
    b = P.random_rhs(g, seed=2, A=A)
    eps = 1e-8 * np.linalg.norm(b)
    S = lsx.solver("fgmres", "fp64-vcycle", levels=3)
    x = np.zeros(A.n)
    S.set(A, b, x, lsx.mixed_task(eps, restart=30, max_it=300))
    it = S.solve()
    true_res = np.linalg.norm(b - A.matvec(x))
    x64, it64, _ = O.Solver(A, eps=eps, restart=30, max_it=500).solve(b)
    act = A.to_scipy().diagonal() < 0
    err = np.linalg.norm((x - x64)[act]) / np.linalg.norm(x64[act])
    print(f"C1 3-level: levels={S.num_levels()} it={it} oracle_it={it64} "
          f"true_rel_res={true_res / np.linalg.norm(b):.3e} p_err_vs_oracle={err:.3e}", flush=True)
    assert S.num_levels() == 3
    assert true_res <= 1.05 * eps
    assert err <= 1e-3


@pytest.mark.parametrize("kind", ["fgmres", "bicgstab"])
def test_c2_mixed_five_levels_128(lsx, kind):
    """C2 at its size: 128^3 sphere r = 0.2 at (0.35, 0.5, 0.5), seeded Gaussian right-hand side
    (mean removed), FP32 5-level V-cycle (128 -> 8, coarsest 512 unknowns) with FP16 coarse faces
    in FP64 FGMRES(30) to 1e-8.  Pressure and divergence error against the oracle's FP64 solve:
    the divergence error ||D u + flux|| of the corrected velocity equals ||b - A p|| with the
    oracle's FP64 operator."""
    g, A = P.sphere(128)
    b = P.random_rhs(g, seed=0, A=A)
    nb = np.linalg.norm(b)
    eps = 1e-8 * nb
    S = lsx.solver(kind, "mixed", levels=5)
    x = np.zeros(A.n)
    S.set(A, b, x, lsx.mixed_task(eps, restart=30, max_it=300))
    it = S.solve()
    x64, it64, _ = O.Solver(A, kind=kind, eps=eps, restart=30, max_it=500).solve(b)
    act = A.to_scipy().diagonal() < 0
    div = np.linalg.norm(b - A.matvec(x)) / nb
    div64 = np.linalg.norm(b - A.matvec(x64)) / nb
    err = np.linalg.norm((x - x64)[act]) / np.linalg.norm(x64[act])
    print(f"C2 {kind}: levels={S.num_levels()} it={it} oracle_it={it64} divergence error mixed={div:.3e} "
          f"oracle={div64:.3e} pressure error vs oracle={err:.3e}", flush=True)
    assert S.num_levels() == 5
    assert div <= (1.05 if kind == "fgmres" else 1.2) * 1e-8
    assert err <= 1e-3
