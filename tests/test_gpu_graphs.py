"""V-cycle applications replayed from captured HIP graphs (one per Krylov basis slot) give the
same bits as eagerly launched cycles, on every replay: three solves in a row with eager work on
the same stream between them, each compared with the eager solver's result."""
import numpy as np
import pytest

import problems as P

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["mixed", "fp64-vcycle"])
@pytest.mark.parametrize("kind", ["fgmres", "bicgstab"])
def test_graph_replay_is_bitwise(lsx, kind, precision):
    g, A = P.ellipsoid_stretched(32)
    eps = 1e-8
    sol = {}
    for graphs in (False, True):
        S = lsx.solver(kind, precision, graphs=graphs)
        S.set(A, np.zeros(A.n), np.zeros(A.n), lsx.mixed_task(1.0, restart=30, max_it=500))
        runs = []
        for seed in (1, 2, 1):
            b = P.random_rhs(g, seed=seed, A=A)
            S.eps = eps * np.linalg.norm(b)
            x = np.zeros(A.n)
            S._rhs, S._sol = b, x
            it = S.solve()
            runs.append((x.copy(), it))
            S.apply_operator(x)  # eager work on the solver stream between replays
        sol[graphs] = runs
    for (xe, ie), (xg, ig) in zip(sol[False], sol[True]):
        print(f"kind={kind} precision={precision} it eager={ie} graph={ig} "
              f"max|diff|={np.abs(xe - xg).max():.3e}", flush=True)
        assert ie == ig
        assert np.array_equal(xe, xg)
