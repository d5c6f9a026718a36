"""CPU reference of the shape gradient (test infrastructure): the oracle's cut-cell geometry,
operator and right-hand side, exact-arithmetic-style solves on the free block of the operator,
and the Lagrangian F(c) = J(theta(c), p) + lambda . (b(c) - A(c) p) differentiated by central
differences of whole oracle geometry rebuilds at lattice control points.

The pressure operator A is an identity row at the pinned cells and symmetric elsewhere
(A = [[I, 0], [C, A_ff]] in pinned/free order), so A p = b and A^T lambda = g reduce to solves
with the symmetric block A_ff: p_f = A_ff^-1 (b_f - C b_p), lambda_f = A_ff^-1 g_f,
lambda_p = g_p - C^T lambda_f.  These are done by Jacobi-preconditioned conjugate gradients on
-A_ff (rows with a zero diagonal, i.e. closed cut cells, and solid cells, which couple to
nothing, are left out and get 0)."""
import numpy as np
import scipy.sparse as sp

import oracle_py as O

FLUID = 11111111


def lattice_copies(faces):
    """Lattice point (3-tuple on the 4x4x4 control lattice) -> flat control indices f*16+i*4+j."""
    copies = {}
    for f, (a, s, bu, bv) in enumerate(faces):
        for i in range(4):
            for j in range(4):
                l = [0, 0, 0]
                l[a], l[bu], l[bv] = (3 if s > 0 else 0), i, j
                copies.setdefault(tuple(l), []).append(f * 16 + i * 4 + j)
    return copies


def lattice_perturb(ctrl, copies, key, comp, h):
    c = np.array(ctrl, np.float64).reshape(96, 3).copy()
    for q in copies[key]:
        c[q, comp] += h
    return c.reshape(6, 4, 4, 3)


def lattice_grad(g288, copies, key, comp):
    g = np.asarray(g288).reshape(96, 3)
    return float(sum(g[q, comp] for q in copies[key]))


def drag(g, fields, p):
    """J = sum over cut cells of (theta_x[i-1] - theta_x[i]) dy dz p[i] and its p-gradient."""
    N, M, K = g.dims
    t = fields["type"].reshape(K, M, N)
    tx = fields["thx"].reshape(K, M, N)
    dy, dz = g.dy, g.dz
    coef = np.zeros((K, M, N))
    cut = (t > 0) & (t != FLUID)
    cut[:, :, 0] = False
    coef[:, :, 1:] = ((tx[:, :, :-1] - tx[:, :, 1:]) * dy[None, :, None]) * dz[:, None, None]
    coef = np.where(cut, coef, 0.0).ravel()
    return float(coef @ p), coef


def _cg(Aop, diag, b, rtol, max_it=20000):
    x = np.zeros_like(b)
    r = b.copy()
    z = r / diag
    d = z.copy()
    rz = r @ z
    nb = np.linalg.norm(b)
    for _ in range(max_it):
        q = Aop @ d
        a = rz / (d @ q)
        x += a * d
        r -= a * q
        if np.linalg.norm(r) <= rtol * nb:
            break
        z = r / diag
        rz_new = r @ z
        d = z + (rz_new / rz) * d
        rz = rz_new
    return x


class Blocks:
    """Free/pinned split of the oracle CSR operator A."""

    def __init__(self, A):
        S = A.to_scipy().tocsr()
        n = S.shape[0]
        diag = S.diagonal()
        offd = S - sp.diags(diag)
        rownnz = np.diff(offd.tocsr().indptr)
        colnnz = np.diff(offd.tocsc().indptr)
        self.pins = np.flatnonzero((rownnz == 0) & (colnnz > 0) & (diag == 1.0))
        free = np.ones(n, bool)
        free[self.pins] = False
        free &= (diag != 0) & ((rownnz > 0) | np.isin(np.arange(n), self.pins))
        free[self.pins] = False
        self.free = np.flatnonzero(free)
        self.S = S
        self.Aff = -S[self.free][:, self.free]   # symmetric positive definite
        self.C = S[self.free][:, self.pins]       # couplings of free rows to the pins
        self.dff = self.Aff.diagonal()

    def solve(self, b, rtol=1e-13):
        x = np.zeros(self.S.shape[0])
        x[self.pins] = b[self.pins]
        rhs = b[self.free] - self.C @ b[self.pins]
        x[self.free] = _cg(self.Aff, self.dff, -rhs, rtol)
        return x

    def solve_transpose(self, g, rtol=1e-13):
        lam = np.zeros(self.S.shape[0])
        lam[self.free] = _cg(self.Aff, self.dff, -g[self.free], rtol)
        lam[self.pins] = g[self.pins] - self.C.T @ lam[self.free]
        return lam


def lagrangian(w, ctrl, center, axes, p, lam, flow, pin=True):
    """F = J(theta, p) + lambda . (b - A p) on the oracle geometry of `ctrl`, with its pieces."""
    g = O.Geometry(*w).build_bezier(ctrl, center, axes)
    A = g.laplace(pin=pin)
    f = g.fields()
    J, _ = drag(g, f, p)
    b = O.pressure_rhs(g, **flow, pin=pin)
    return J, b, A.to_scipy().tocsr(), f["type"]


def reference_gradient(w, ctrl, center, axes, flow, keys, faces, h=1e-6, rtol=1e-13):
    """(J, p, lambda, {(key, comp): dJ/dc}) with the CPU reference."""
    g = O.Geometry(*w).build_bezier(ctrl, center, axes)
    A = g.laplace(pin=True)
    B = Blocks(A)
    b = O.pressure_rhs(g, **flow, pin=True)
    p = B.solve(b, rtol)
    J, dJdp = drag(g, g.fields(), p)
    lam = B.solve_transpose(dJdp, rtol)
    copies = lattice_copies(faces)
    out = {}
    for key, comp in keys:
        Jp, bp, Ap, tp = lagrangian(w, lattice_perturb(ctrl, copies, key, comp, h), center, axes, p, lam, flow)
        Jm, bm, Am, tm = lagrangian(w, lattice_perturb(ctrl, copies, key, comp, -h), center, axes, p, lam, flow)
        dF = (Jp - Jm) + lam @ ((bp - bm) - (Ap - Am) @ p)
        out[(key, comp)] = (dF / (2 * h), bool(np.array_equal(tp, tm)))
    return J, p, lam, out
