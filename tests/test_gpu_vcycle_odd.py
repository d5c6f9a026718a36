"""V-cycle on levels of odd size.  An axis of n cells coarsens to n / 2 cells, or to (n + 1) / 2
when n is odd, the last coarse cell then having a single child: its faces sum the fine faces
present, the restriction sums the children present, and the trilinear prolongation gives the
last fine cell its boundary weight.  Checks, on all-odd and mixed-parity grids (the reference
cascade rejects the latter):
* FGMRES with the FP64 and the mixed V-cycle, red-black or line smoothing, reaches the 1e-8 true
  relative residual in about as many iterations as on the neighbouring even grid;
* the transposed cycle is the transpose of the cycle (exact in FP64);
* the one-workgroup coarse cycle equals the per-level launches bit for bit."""
import numpy as np
import pytest

import problems as P

pytestmark = pytest.mark.gpu

# (name, odd grid, neighbouring even grid)
CASES = [
    ("sphere33", lambda: P.sphere(33), lambda: P.sphere(32)),
    ("sphere50x31x24", lambda: P.sphere(50, 31, 24), lambda: P.sphere(50, 32, 24)),
    ("cut45x38x27", lambda: P.synthetic_cut(45, 38, 27, seed=7), lambda: P.synthetic_cut(44, 38, 28, seed=7)),
]


def solve(lsx, A, b, precision, smoother):
    eps = 1e-8 * np.linalg.norm(b)
    S = lsx.solver("fgmres", precision, smoother=smoother, line_levels=2 if smoother == "line" else 0)
    x = np.zeros(A.n)
    S.set(A, b, x, lsx.mixed_task(eps, restart=30, max_it=300))
    it = S.solve()
    return it, np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b), S


@pytest.mark.parametrize("smoother", ["rbgs", "line"])
@pytest.mark.parametrize("precision", ["fp64-vcycle", "mixed"])
@pytest.mark.parametrize("name,mk,mk_even", CASES, ids=[c[0] for c in CASES])
def test_odd_levels_converge(lsx, name, mk, mk_even, precision, smoother):
    g, A = mk()
    dims = tuple(int(v) for v in A.dims)
    b = P.random_rhs(g, seed=3, A=A)
    it, rel, S = solve(lsx, A, b, precision, smoother)
    ge, Ae = mk_even()
    it_e, rel_e, _ = solve(lsx, Ae, P.random_rhs(ge, seed=3, A=Ae), precision, smoother)
    print(f"case={name} dims={dims} precision={precision} smoother={smoother}: it={it} rel={rel:.3e} "
          f"levels={S.num_levels()} (even neighbour: it={it_e} rel={rel_e:.3e})", flush=True)
    assert any(v % 2 for v in dims)
    assert rel <= 1.05e-8
    assert it <= it_e + 4


@pytest.mark.parametrize("smoother", ["rbgs", "line"])
@pytest.mark.parametrize("precision,tol", [("fp64-vcycle", 1e-12), ("mixed", 2e-5)])
@pytest.mark.parametrize("name,mk,mk_even", CASES, ids=[c[0] for c in CASES])
def test_odd_levels_transposed_cycle(lsx, name, mk, mk_even, precision, tol, smoother):
    g, A = mk()
    S = lsx.solver("fgmres", precision, smoother=smoother, line_levels=2 if smoother == "line" else 0)
    S.set(A, np.zeros(A.n), np.zeros(A.n), lsx.mixed_task(1.0))
    rng = np.random.default_rng(5)
    act = A.to_scipy().diagonal() < 0
    x, y = rng.standard_normal(A.n) * act, rng.standard_normal(A.n) * act
    Bx, BTy = S.precondition(x), S.precondition_transpose(y)
    lhs, rhs = Bx @ y, x @ BTy
    scale = np.linalg.norm(Bx) * np.linalg.norm(y)
    print(f"case={name} precision={precision} smoother={smoother}: <Bx,y>={lhs:.15e} <x,B^T y>={rhs:.15e} "
          f"rel={abs(lhs - rhs) / scale:.3e}", flush=True)
    assert abs(lhs - rhs) <= tol * scale


@pytest.mark.parametrize("precision", ["fp64-vcycle", "mixed"])
@pytest.mark.parametrize("name,mk,mk_even", CASES, ids=[c[0] for c in CASES])
def test_odd_levels_coarse_cycle_in_one_launch(lsx, name, mk, mk_even, precision):
    g, A = mk()
    v = P.random_rhs(g, seed=12, A=A)
    z = []
    for fuse in (True, False):
        S = lsx.solver("fgmres", precision, coarse_fuse=fuse)
        S.set(A, np.zeros(A.n), np.zeros(A.n), lsx.mixed_task(1.0))
        z.append(S.precondition(v))
    print(f"case={name} precision={precision}: max|diff|={np.abs(z[0] - z[1]).max():.3e}", flush=True)
    assert np.linalg.norm(z[0]) > 0 and np.array_equal(z[0], z[1])


@pytest.mark.parametrize("dims", [(5, 7, 9), (9, 5, 17), (3, 3, 3), (17, 16, 15), (9, 3, 17), (40, 3, 24)])
@pytest.mark.parametrize("precision", ["fp64-vcycle", "mixed"])
def test_small_odd_grids(lsx, dims, precision):
    """Small boxes whose hierarchies end on levels of 1-3 cells per axis converge and transpose;
    with one axis of 3 cells the automatic hierarchy stops after one coarsening, on 90 and 480
    unknowns, which the dense coarsest inverse solves.  The boxes are unit cubes, so unequal
    dimensions give anisotropic cells (40 x 3 x 24: 13:1), which the red-black smoother handles
    slowly (35 iterations there)."""
    g, A = P.body_free(*dims)
    b = P.random_rhs(g, seed=1, A=A)
    it, rel, S = solve(lsx, A, b, precision, "rbgs")
    rng = np.random.default_rng(2)
    act = A.to_scipy().diagonal() < 0
    x, y = rng.standard_normal(A.n) * act, rng.standard_normal(A.n) * act
    Bx, BTy = S.precondition(x), S.precondition_transpose(y)
    trel = abs(Bx @ y - x @ BTy) / (np.linalg.norm(Bx) * np.linalg.norm(y))
    print(f"dims={dims} precision={precision}: it={it} rel={rel:.3e} levels={S.num_levels()} transpose={trel:.3e}",
          flush=True)
    assert rel <= 1.05e-8 and it <= 40
    assert trel <= (1e-12 if precision == "fp64-vcycle" else 2e-5)


def test_cascade_still_rejects_mixed_parity(lsx):
    g, A = P.sphere(50, 31, 24)
    with pytest.raises(lsx.LsxError, match="mixes even and odd"):
        lsx.MGFGMRES().set(A, np.zeros(A.n), np.zeros(A.n), lsx.Task(eps=1.0))
