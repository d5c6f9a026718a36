"""The torch.autograd composition PressureRhs -> PressureSolve -> PressureDrag gives the same
gradients as the one-call forward/backward pair (the Lagrangian splits linearly into the three
functions' terms)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def test_autograd_chain_matches_tape(lsx):
    import lsx_pkg.autograd as AG
    n = 64
    center, axes = (0.5, 0.5, 0.5), (0.25, 0.15, 0.15)
    ctrl_np = lsx.bezier_perturb(lsx.bezier_ellipsoid(center, axes), 0.01, seed=0)
    w = [np.full(n, 1.0 / n)] * 3
    rng = np.random.default_rng(4)
    nu, nv, nw = (n - 1) * n * n, n * (n - 1) * n, n * n * (n - 1)
    flow = dict(u=1.0 + rng.uniform(-0.05, 0.05, nu), v=rng.uniform(-0.05, 0.05, nv), w=rng.uniform(-0.05, 0.05, nw))
    geom = lsx.Geometry(*w).build("bezier", center, axes, ctrl=ctrl_np)
    S = lsx.solver("fgmres", "fp64-vcycle")
    S.set_geometry(geom, np.zeros(geom.n), np.zeros(geom.n), lsx.mixed_task(1.0, max_it=3000))
    b0 = geom.pressure_rhs(**flow)
    S.eps = 1e-13 * np.linalg.norm(b0)

    ctrl = torch.tensor(np.asarray(ctrl_np).ravel(), dtype=torch.float64, requires_grad=True)
    b = AG.PressureRhs.apply(ctrl, geom, flow)
    p = AG.PressureSolve.apply(b, ctrl, S, geom)
    J = AG.PressureDrag.apply(p, ctrl, geom)
    J.backward()
    g_auto = ctrl.grad.numpy().copy()

    p2, it, tape = S.adjoint_forward(geom, b0)
    J2, dJdp = geom.drag(p2)
    S.eps = 1e-13 * np.linalg.norm(dJdp)
    _, g_tape, _ = tape.backward(dJdp, lsx.OBJ_DRAG_X, flow=flow)
    rel = np.linalg.norm(g_auto - g_tape) / np.linalg.norm(g_tape)
    print(f"J={float(J):.15e} J_tape={J2:.15e} |g|={np.linalg.norm(g_tape):.6e} rel diff={rel:.3e}", flush=True)
    assert abs(float(J) - J2) <= 1e-12 * abs(J2)
    assert rel <= 1e-7
