"""torch.autograd wrappers around the C-ABI forward/backward pair (an optional layer).

Three functions compose the pressure-drag objective of a Bezier body with control points c:
  b = PressureRhs(c)          b = D u* + flux on the geometry of c (u* fixed);
  p = PressureSolve(b, c)     A(c) p = b, solved on the GPU;
  J = PressureDrag(p, c)      J = sum over cut cells of (theta_x[i-1] - theta_x[i]) dy dz p.
Their backward passes: PressureSolve returns dJ/db = A^{-T} dJ/dp (the adjoint solve) and
-lambda^T d(A p)/dc; PressureRhs returns lambda^T db/dc; PressureDrag returns dJ/dp and the
explicit dJ/dc at fixed p.  The geometric derivatives come from lsx_geometry_shape_gradient
(central differences of the cut-cell geometry per level-set corner, chained to the control
points through the closest points).  The geometry handle must already be built from c: the
functions differentiate at that geometry and do not rebuild it."""
import numpy as np
import torch

from . import lsx


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64)


class PressureRhs(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ctrl, geom, flow):
        ctx.geom, ctx.flow = geom, flow
        return torch.from_numpy(geom.pressure_rhs(**flow))

    @staticmethod
    def backward(ctx, gb):
        zero = np.zeros(ctx.geom.n)
        _, gc = ctx.geom.shape_gradient(zero, _np(gb), lsx.OBJ_NONE, flow=ctx.flow)
        return torch.from_numpy(gc), None, None


class PressureSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, b, ctrl, solver, geom):
        bn = _np(b)
        p, it, tape = solver.adjoint_forward(geom, bn)
        ctx.tape, ctx.geom, ctx.solver = tape, geom, solver
        ctx.p = p
        # the adjoint solve uses the forward's tolerance relative to its own right-hand side
        ctx.rel = solver.eps / max(float(np.linalg.norm(bn)), 1e-300)
        return torch.from_numpy(p)

    @staticmethod
    def backward(ctx, gp):
        g = _np(gp)
        eps = ctx.solver.eps
        ctx.solver.eps = ctx.rel * float(np.linalg.norm(g))
        try:
            dJdb, _, it = ctx.tape.backward(g, lsx.OBJ_NONE, ctrl=False)
        finally:
            ctx.solver.eps = eps
        _, gc = ctx.geom.shape_gradient(ctx.p, dJdb, lsx.OBJ_NONE, flow=None)
        return torch.from_numpy(dJdb), torch.from_numpy(gc), None, None


class PressureDrag(torch.autograd.Function):
    @staticmethod
    def forward(ctx, p, ctrl, geom):
        pn = _np(p)
        J, dJdp = geom.drag(pn)
        ctx.geom, ctx.p, ctx.dJdp = geom, pn, dJdp
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        s = float(gJ)
        _, gc = ctx.geom.shape_gradient(ctx.p, np.zeros(ctx.geom.n), lsx.OBJ_DRAG_X, flow=None)
        return torch.from_numpy(s * ctx.dJdp), torch.from_numpy(s * gc), None
