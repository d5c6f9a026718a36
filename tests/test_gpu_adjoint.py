"""Adjoint solves: the transposed operator A^T (pinned identity rows of A become full rows of A^T),
the transpose of the V-cycle preconditioner, and the Krylov solve on A^T.

The transposes are pinned by their defining identities <A x, y> = <x, A^T y> and
<B x, y> = <x, B^T y> on random vectors (a wrong colour order, transfer weight or pin treatment
breaks them at O(1)), A^T against the transpose of the oracle's CSR matrix, and the adjoint
solution by its residual in the oracle's CSR A^T and by the linearity of J(b) = g . A^{-1} b."""
import numpy as np
import pytest

import problems as P

pytestmark = pytest.mark.gpu


def make_solver(lsx, A, kind="fgmres", precision="fp64-vcycle", eps=1.0, max_it=500, smoother="rbgs"):
    # "line1": line relaxation on the finest level, red-black below
    S = lsx.solver(kind, precision, smoother="line" if smoother.startswith("line") else smoother,
                   line_levels=1 if smoother == "line1" else 0)
    S.set(A, np.zeros(A.n), np.zeros(A.n), lsx.mixed_task(eps, restart=30, max_it=max_it))
    return S


CASES = [
    ("sphere64", lambda: P.sphere(64)),
    ("ellipsoid32_stretched", lambda: P.ellipsoid_stretched(32)),
    ("cut32x16x24", lambda: P.synthetic_cut(32, 16, 24, seed=3)),
]


@pytest.mark.parametrize("name,mk", CASES, ids=[c[0] for c in CASES])
def test_transposed_operator(lsx, name, mk):
    g, A = mk()
    S = make_solver(lsx, A)
    assert S.fine_storage() == 4
    rng = np.random.default_rng(5)
    act = A.to_scipy().diagonal() < 0
    x, y = rng.standard_normal(A.n) * act, rng.standard_normal(A.n) * act
    ATy = S.apply_operator_transpose(y)
    ref = A.to_scipy().T @ y
    err = np.linalg.norm(ATy - ref) / np.linalg.norm(ref)
    lhs, rhs = S.apply_operator(x) @ y, x @ ATy
    print(f"case={name} A^T vs CSR^T rel={err:.3e} <Ax,y>={lhs:.15e} <x,A^T y>={rhs:.15e}", flush=True)
    assert err <= 1e-14
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(S.apply_operator(x)) * np.linalg.norm(y)
    # the pinned rows really differ between A and A^T (the operator is not symmetric there)
    assert np.linalg.norm(S.apply_operator(y) - ATy) > 1e-6 * np.linalg.norm(ATy)


@pytest.mark.parametrize("smoother", ["rbgs", "line", "line1", "jacobi"])
@pytest.mark.parametrize("precision,tol", [("fp64-vcycle", 1e-12), ("mixed", 2e-5), ("mixed16", 2e-3)])
@pytest.mark.parametrize("name,mk", CASES, ids=[c[0] for c in CASES])
def test_transposed_vcycle(lsx, name, mk, precision, tol, smoother):
    """The transposed cycle is the transpose of the cycle as a linear map (exact in FP64; in
    mixed precision up to its FP32/FP16 rounding), with either smoother."""
    g, A = mk()
    S = make_solver(lsx, A, precision=precision, smoother=smoother)
    rng = np.random.default_rng(7)
    # vectors on the active rows (solid and closed cells carry no equation)
    act = A.to_scipy().diagonal() < 0
    x, y = rng.standard_normal(A.n) * act, rng.standard_normal(A.n) * act
    Bx, BTy = S.precondition(x), S.precondition_transpose(y)
    lhs, rhs = Bx @ y, x @ BTy
    scale = np.linalg.norm(Bx) * np.linalg.norm(y)
    print(f"case={name} precision={precision} smoother={smoother} <Bx,y>={lhs:.15e} <x,B^T y>={rhs:.15e} "
          f"rel={abs(lhs - rhs) / scale:.3e}", flush=True)
    assert abs(lhs - rhs) <= tol * scale
    # B is not symmetric (trilinear prolongation against sum restriction): B^T y != B y
    assert np.linalg.norm(S.precondition(y) - BTy) > 1e-4 * np.linalg.norm(BTy)


@pytest.mark.parametrize("precision", ["mixed", "fp64-vcycle"])
@pytest.mark.parametrize("kind", ["fgmres", "bicgstab"])
@pytest.mark.parametrize("name,mk", CASES[:2], ids=[c[0] for c in CASES[:2]])
def test_adjoint_solve(lsx, name, mk, kind, precision):
    g, A = mk()
    rhs = P.random_rhs(g, seed=11, A=A)
    eps = 1e-8 * np.linalg.norm(rhs)
    S = make_solver(lsx, A, kind, precision, eps=eps)
    lam, it, nr = S.solve_adjoint(rhs)
    true_res = np.linalg.norm(rhs - A.to_scipy().T @ lam)
    # the forward solve of the same right-hand side, for the iteration count
    x = np.zeros(A.n)
    S._rhs, S._sol = rhs, x
    it_fwd = S.solve()
    print(f"case={name} kind={kind} precision={precision} adjoint it={it} forward it={it_fwd} "
          f"true_rel_res={true_res / np.linalg.norm(rhs):.3e}", flush=True)
    assert nr <= eps
    assert true_res <= 1.2 * eps
    assert it <= it_fwd + 10


def test_dJdb_is_the_adjoint_solution(lsx):
    """J(b) = g . A^{-1} b is linear in b, so J(b + d) - J(b) = lambda . d exactly for lambda =
    A^{-T} g; checked with tightly converged FP64 solves."""
    grid, A = P.ellipsoid_stretched(32)
    b = P.random_rhs(grid, seed=1, A=A)
    rng = np.random.default_rng(3)
    gvec = rng.standard_normal(A.n) * (A.to_scipy().diagonal() < 0)
    d = P.random_rhs(grid, seed=9, A=A)
    S = make_solver(lsx, A, "fgmres", "fp64-vcycle", eps=1e-13 * np.linalg.norm(b), max_it=1000)

    def J(rhs):
        x = np.zeros(A.n)
        S._rhs, S._sol = rhs, x
        S.eps = 1e-13 * np.linalg.norm(rhs)
        S.solve()
        return gvec @ x

    S.eps = 1e-13 * np.linalg.norm(gvec)
    lam, it, _ = S.solve_adjoint(gvec)
    dJ = J(b + d) - J(b)
    pred = lam @ d
    print(f"J(b+d)-J(b)={dJ:.15e} lambda.d={pred:.15e} adjoint it={it}", flush=True)
    assert abs(dJ - pred) <= 1e-9 * abs(pred)
