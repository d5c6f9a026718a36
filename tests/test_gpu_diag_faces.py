"""Line passes and residual restriction that rebuild a_p from the faces (lsx_options.diag_faces).

Every row of a pressure operator other than the pinned ones has a_p = -(sum of its couplings);
the aggregated coarse levels keep that rule with the pinned cells' couplings added on their sink
cells.  With the option on (the default) the forward line passes and the residual restriction of
such levels skip the 1/a_p loads.  Checks:
* one application of the FP64 V-cycle with the rule equals the one without it to rounding (1e-12
  relative), on cut geometries whose coarse levels hold the sinks, with lines on every level and
  on the finest two;
* in the mixed cycle a finest-level sweep (FP32 faces) agrees to FP32 rounding (1e-5); on the
  FP16-face levels the rebuilt a_p is the row sum of the rounded faces, which differs from the
  rounded a_p by up to the faces' rounding (2^-11 relative), so whole cycles agree to 1e-2 and the
  mixed and mixed16 solves need no more iterations with the rule;
* an operator whose rows do not sum to zero (a mass term on the diagonal) fails the set-up check
  and keeps reading 1/a_p: its line sweep still equals the scipy restatement of
  test_gpu_smoothers.py, which a diagonal rebuilt from the faces would miss by the mass term."""
import numpy as np
import pytest

import oracle_py as O
import problems as P
from test_gpu_smoothers import ref_line_sweep

pytestmark = pytest.mark.gpu

CASES = [("sphere64", lambda: P.sphere(64)), ("cut64x32x48", lambda: P.synthetic_cut(64, 32, 48, seed=5)),
         ("ellipsoid64_stretched", lambda: P.ellipsoid_stretched(64))]


@pytest.mark.parametrize("line_levels", [0, 2])
@pytest.mark.parametrize("precision,tol", [("fp64-vcycle", 1e-12), ("mixed", 1e-2)])
@pytest.mark.parametrize("name,mk", CASES, ids=[c[0] for c in CASES])
def test_cycle_equals_inverse_diagonal_cycle(lsx, name, mk, precision, tol, line_levels):
    g, A = mk()
    b = P.random_rhs(g, seed=4, A=A)
    z = {}
    for dff in (True, False):
        S = lsx.solver("fgmres", precision, smoother="line", line_levels=line_levels, diag_faces=dff)
        S.set(A, b, np.zeros(A.n), lsx.mixed_task(1.0))
        z[dff] = S.precondition(b)
    rel = np.linalg.norm(z[True] - z[False]) / np.linalg.norm(z[False])
    print(f"case={name} precision={precision} line_levels={line_levels}: |z_dff - z| / |z| = {rel:.3e}",
          flush=True)
    assert np.linalg.norm(z[False]) > 0 and rel <= tol


@pytest.mark.parametrize("name,mk", CASES, ids=[c[0] for c in CASES])
def test_mixed_finest_sweep(lsx, name, mk):
    g, A = mk()
    rng = np.random.default_rng(9)
    act = A.to_scipy().diagonal() < 0
    b = rng.standard_normal(A.n) * act
    x0 = rng.standard_normal(A.n) * act
    x1 = {}
    for dff in (True, False):
        S = lsx.solver("fgmres", "mixed", smoother="line", diag_faces=dff)
        S.set(A, b, np.zeros(A.n), lsx.mixed_task(1.0))
        x1[dff] = S.smooth(b, x0, sweeps=1)
    rel = np.linalg.norm(x1[True] - x1[False]) / np.linalg.norm(x1[False])
    print(f"case={name} mixed finest sweep: |x_dff - x| / |x| = {rel:.3e}", flush=True)
    assert rel <= 1e-5


@pytest.mark.parametrize("precision", ["mixed", "mixed16"])
@pytest.mark.parametrize("name,mk", CASES[:2], ids=[c[0] for c in CASES[:2]])
def test_mixed_iterations(lsx, name, mk, precision):
    g, A = mk()
    b = P.random_rhs(g, seed=6, A=A)
    eps = 1e-8 * np.linalg.norm(b)
    its = {}
    for dff in (True, False):
        S = lsx.solver("fgmres", precision, smoother="line", line_levels=2, diag_faces=dff)
        x = np.zeros(A.n)
        S.set(A, b, x, lsx.mixed_task(eps, restart=30, max_it=300))
        its[dff] = S.solve()
        res = np.linalg.norm(b - A.matvec(x))
        print(f"case={name} {precision} diag_faces={dff}: it={its[dff]} rel res={res / np.linalg.norm(b):.3e}",
              flush=True)
        assert res <= 1.05 * eps
    assert its[True] <= its[False] + 1


def shifted(A, frac):
    """A with frac |a_ii| added to the magnitude of every negative diagonal (a mass term; the
    pinned identity rows and the solid rows keep theirs)."""
    val = np.array(A.val, dtype=np.float64, copy=True)
    rows = np.repeat(np.arange(A.n), np.diff(A.rowptr))
    sel = (A.col == rows) & (val < 0)
    val[sel] *= 1.0 + frac
    return O.Csr(A.dims, A.rowptr, A.col, val), int(sel.sum())


@pytest.mark.parametrize("precision,tol", [("fp64-vcycle", 1e-12), ("mixed", 2e-5)])
def test_operator_with_mass_term_reads_inverse_diagonal(lsx, precision, tol):
    g, A = P.sphere(32)
    As, nsh = shifted(A, 0.05)
    assert nsh > 0.9 * A.n
    act = A.to_scipy().diagonal() < 0
    rng = np.random.default_rng(8)
    b = rng.standard_normal(A.n) * act
    x0 = rng.standard_normal(A.n) * act
    S = lsx.solver("fgmres", precision, smoother="line", diag_faces=True)
    S.set(As, np.zeros(A.n), np.zeros(A.n), lsx.mixed_task(1.0))
    x1 = S.smooth(b, x0, sweeps=1)
    x1r = ref_line_sweep(As, g.dims, b, x0)
    x1wrong = ref_line_sweep(A, g.dims, b, x0)  # what a diagonal rebuilt from the faces would give
    e = np.linalg.norm(x1 - x1r) / np.linalg.norm(x1r)
    ew = np.linalg.norm(x1wrong - x1r) / np.linalg.norm(x1r)
    print(f"precision={precision}: sweep vs restatement {e:.3e} (rule applied wrongly would be {ew:.3e})",
          flush=True)
    assert ew > 100 * tol and e <= tol
