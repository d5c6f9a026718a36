"""ctypes bindings to the CPU oracle (oracle/_build/liblsx_oracle.so).

The oracle is the checker: a CPU restatement of the reference solver path.  Only tests,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg load it.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liblsx_oracle.so")

_lib = None

PD = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
PI = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    L.lsxo_last_error.restype = C.c_char_p
    L.lsxo_geom_create.restype = C.c_void_p
    L.lsxo_geom_create.argtypes = [C.c_int, C.c_int, C.c_int, PD, PD, PD, C.c_double, C.c_double, C.c_double]
    L.lsxo_geom_destroy.argtypes = [C.c_void_p]
    L.lsxo_geom_set_fields.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.lsxo_geom_get_fields.argtypes = [C.c_void_p, PI, PD, PD, PD, PD]
    L.lsxo_geom_build_laplace.restype = C.c_long
    L.lsxo_geom_build_laplace.argtypes = [C.c_void_p, C.c_int]
    L.lsxo_geom_get_csr.argtypes = [C.c_void_p, PI, PI, PD]
    L.lsxo_solver_create.restype = C.c_void_p
    L.lsxo_solver_create.argtypes = [C.c_int, C.c_int, C.c_int, PI, PI, PD, C.c_int, C.c_int, C.c_int,
                                     C.c_double, C.c_int, C.c_int, C.c_double]
    L.lsxo_solver_destroy.argtypes = [C.c_void_p]
    L.lsxo_solver_set_eps.argtypes = [C.c_void_p, C.c_double]
    L.lsxo_solver_solve.restype = C.c_int
    L.lsxo_solver_solve.argtypes = [C.c_void_p, PD, PD, C.POINTER(C.c_double)]
    L.lsxo_mg_apply.argtypes = [C.c_void_p, PD, PD]
    L.lsxo_mg_levels.restype = C.c_int
    L.lsxo_mg_levels.argtypes = [C.c_void_p]
    L.lsxo_mg_level_info.restype = C.c_long
    L.lsxo_mg_level_info.argtypes = [C.c_void_p, C.c_int, PI]
    L.lsxo_mg_level_csr.argtypes = [C.c_void_p, C.c_int, PI, PI, PD]
    L.lsxo_mg_level_adlj.argtypes = [C.c_void_p, C.c_int, C.c_int, PD, PD, PD]
    L.lsxo_spmv.argtypes = [C.c_long, PI, PI, PD, PD, PD]
    L.lsxo_geom_build_body.argtypes = [C.c_void_p, C.c_int] + [C.c_double] * 6 + [C.c_int] * 4
    L.lsxo_geom_get_level_set.argtypes = [C.c_void_p, PD, PD, PD, PD, PI, PI, PI]
    L.lsxo_vof_table_size.restype = C.c_int
    L.lsxo_bezier_ellipsoid.argtypes = [PD, PD, PD]
    L.lsxo_bezier_phi.argtypes = [PD, PD, PD, PD, C.c_long, PD]
    L.lsxo_geom_build_bezier.argtypes = [C.c_void_p, PD, PD, PD] + [C.c_int] * 4
    L.lsxo_geom_correct.argtypes = [C.c_void_p, PD, PD, PD, PD, PD, PD, PD]
    L.lsxo_geom_rhs.argtypes = [C.c_void_p, PD, PD, PD, PD] + [C.c_double] * 5 + [C.c_void_p, C.c_int, PD,
                                                                                  C.c_void_p]
    L.lsxo_geom_build_helmholtz.restype = C.c_long
    L.lsxo_geom_build_helmholtz.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, PI]
    _lib = L
    return L


class Csr:
    def __init__(self, dims, rowptr, col, val):
        self.dims = tuple(int(d) for d in dims)
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)

    @property
    def n(self):
        return len(self.rowptr) - 1

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.n, self.n))

    def matvec(self, x):
        y = np.zeros(self.n)
        lib().lsxo_spmv(self.n, self.rowptr, self.col, self.val, np.ascontiguousarray(x, dtype=np.float64), y)
        return y


class Geometry:
    """Cut-cell geometry of the base grid (reference LSSTAGUniformMesh subset)."""

    def __init__(self, dx, dy, dz, origin=(0.0, 0.0, 0.0)):
        self.dx = np.ascontiguousarray(dx, dtype=np.float64)
        self.dy = np.ascontiguousarray(dy, dtype=np.float64)
        self.dz = np.ascontiguousarray(dz, dtype=np.float64)
        self.dims = (len(self.dx), len(self.dy), len(self.dz))
        self.h = lib().lsxo_geom_create(*self.dims, self.dx, self.dy, self.dz, *origin)
        if not self.h:
            raise RuntimeError(lib().lsxo_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().lsxo_geom_destroy(self.h)
            self.h = None

    @property
    def n(self):
        N, M, K = self.dims
        return N * M * K

    def fields(self):
        n = self.n
        t = np.zeros(n, np.int32)
        tx, ty, tz, v = (np.zeros(n) for _ in range(4))
        lib().lsxo_geom_get_fields(self.h, t, tx, ty, tz, v)
        return dict(type=t, thx=tx, thy=ty, thz=tz, vof=v)

    def set_fields(self, type=None, thx=None, thy=None, thz=None, vof=None):
        keep = []

        def ptr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a.ctypes.data

        lib().lsxo_geom_set_fields(self.h, ptr(type, np.int32), ptr(thx, np.float64), ptr(thy, np.float64),
                                   ptr(thz, np.float64), ptr(vof, np.float64))

    def build_body(self, kind, center, axes, box=None):
        """Cut-cell geometry of a sphere (kind 'sphere', axes = (r,)) or an axis-aligned
        ellipsoid (kind 'ellipsoid', axes = semi-axes) inside box (hs, hf, vs, vf)."""
        N, M, K = self.dims
        hs, hf, vs, vf = box or (1, N - 1, 1, M - 1)
        a = tuple(axes) + (axes[0],) * (3 - len(axes))
        st = lib().lsxo_geom_build_body(self.h, 0 if kind == "sphere" else 1, *center, *a, hs, hf, vs, vf)
        if st != 0:
            raise RuntimeError(lib().lsxo_last_error().decode())
        return self

    def build_bezier(self, ctrl, center, axes, box=None):
        """Cut-cell geometry of the six-patch Bezier body with control points ctrl (6, 4, 4, 3)."""
        N, M, K = self.dims
        hs, hf, vs, vf = box or (1, N - 1, 1, M - 1)
        st = lib().lsxo_geom_build_bezier(self.h, np.ascontiguousarray(ctrl, np.float64).ravel(),
                                          np.asarray(center, np.float64), np.asarray(axes, np.float64), hs, hf, vs, vf)
        if st != 0:
            raise RuntimeError(lib().lsxo_last_error().decode())
        return self

    def level_set(self):
        n = self.n
        phi, ax, ay, az = (np.zeros(n) for _ in range(4))
        sx, sy, sz = (np.zeros(n, np.int32) for _ in range(3))
        if lib().lsxo_geom_get_level_set(self.h, phi, ax, ay, az, sx, sy, sz) != 0:
            raise RuntimeError("no level set built")
        return dict(phi=phi, alpha=(ax, ay, az), side=(sx, sy, sz))

    def laplace(self, pin=True):
        nnz = lib().lsxo_geom_build_laplace(self.h, 1 if pin else 0)
        if nnz < 0:
            raise RuntimeError(lib().lsxo_last_error().decode())
        rp = np.zeros(self.n + 1, np.int32)
        col = np.zeros(nnz, np.int32)
        val = np.zeros(nnz)
        lib().lsxo_geom_get_csr(self.h, rp, col, val)
        return Csr(self.dims, rp, col, val)


def helmholtz(g, direction, nu, coef, flow=True):
    """Momentum (Helmholtz) operator of velocity component `direction` (0 u, 1 v, 2 w) on its
    staggered mesh of the body-free geometry g: nu = 1/Re, coef = COEF_INTEGR_TIME (1.5/dt);
    flow: a flow task (outflow east boundary, no bottom/top wall terms)."""
    dims = np.zeros(3, np.int32)
    nnz = lib().lsxo_geom_build_helmholtz(g.h, int(direction), float(nu), float(coef), 1 if flow else 0, dims)
    if nnz < 0:
        raise RuntimeError(lib().lsxo_last_error().decode())
    n = int(np.prod(dims))
    rp = np.zeros(n + 1, np.int32)
    col = np.zeros(nnz, np.int32)
    val = np.zeros(nnz)
    lib().lsxo_geom_get_csr(g.h, rp, col, val)
    return Csr(tuple(int(d) for d in dims), rp, col, val)


def bezier_ellipsoid(center, axes):
    """Control points (6, 4, 4, 3) of the six-patch bicubic Bezier fit of an ellipsoid."""
    ctrl = np.zeros(288)
    lib().lsxo_bezier_ellipsoid(np.asarray(center, np.float64), np.asarray(axes, np.float64), ctrl)
    return ctrl.reshape(6, 4, 4, 3)


def bezier_phi(ctrl, center, axes, pts):
    """Signed distance (positive inside) of points (n, 3) to the Bezier body."""
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    out = np.zeros(len(pts))
    lib().lsxo_bezier_phi(np.ascontiguousarray(ctrl, np.float64).ravel(), np.asarray(center, np.float64),
                          np.asarray(axes, np.float64), pts, len(pts), out)
    return out


def face_shapes(dims):
    """Sizes of the staggered face arrays (u on x faces, v on y faces, w on z faces)."""
    N, M, K = dims
    return (N - 1) * M * K, N * (M - 1) * K, N * M * (K - 1)


def pressure_rhs(g, u, v, w, U_W=1.0, U_E=1.0, V_S=0.0, V_N=0.0, W_B=0.0, W_T=0.0, outflow=True, pin=True,
                 return_flux=False):
    """b = Dx u + Dy v + Dz w + flux with b = 0 at the pinned cells (PoissonDiscreteEquation::
    recalcRightMember).  U_W: scalar or one inflow velocity per x-row (M*K); outflow: the east
    boundary flux is the last interior u face (setBaseFlux), else U_E."""
    N, M, K = g.dims
    nu, nv, nw = face_shapes(g.dims)
    u = np.ascontiguousarray(u, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    assert u.size == nu and v.size == nv and w.size == nw
    uw = np.ascontiguousarray(np.broadcast_to(np.asarray(U_W, np.float64), (M * K,)))
    b = np.zeros(g.n)
    flux = np.zeros(g.n)
    st = lib().lsxo_geom_rhs(g.h, u, v, w, uw, U_E, V_S, V_N, W_B, W_T, u.ctypes.data if outflow else None,
                             1 if pin else 0, b, flux.ctypes.data)
    if st != 0:
        raise RuntimeError(lib().lsxo_last_error().decode())
    return (b, flux) if return_flux else b


def velocity_correction(g, dp, us, vs, ws):
    """u = (G_x dp) inv_Mx + u* and likewise v, w (SetSolutions::correction)."""
    nu, nv, nw = face_shapes(g.dims)
    dp = np.ascontiguousarray(dp, np.float64)
    us, vs, ws = (np.ascontiguousarray(a, np.float64) for a in (us, vs, ws))
    assert dp.size == g.n and us.size == nu and vs.size == nv and ws.size == nw
    u, v, w = np.zeros(nu), np.zeros(nv), np.zeros(nw)
    lib().lsxo_geom_correct(g.h, dp, us, vs, ws, u, v, w)
    return u, v, w


class Solver:
    """Oracle FGMRES / BiCGStab with the reference multigrid cascade ("mg"), its fine level alone
    ("mg1": ADLJ sweeps from a zero correction, any grid dimensions) or no preconditioner."""

    def __init__(self, A, kind="fgmres", restart=30, max_it=1000, eps=1e-8, precond="mg", n_post=3, omega=0.805):
        self.A = A
        N, M, K = A.dims
        self.h = lib().lsxo_solver_create(N, M, K, A.rowptr, A.col, A.val, 0 if kind == "fgmres" else 1, restart,
                                          max_it, eps, {"mg": 1, "mg1": 2}.get(precond, 0), n_post, omega)
        if not self.h:
            raise RuntimeError(lib().lsxo_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().lsxo_solver_destroy(self.h)
            self.h = None

    def set_eps(self, eps):
        lib().lsxo_solver_set_eps(self.h, eps)

    def solve(self, b, x0=None):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
        nr = C.c_double(0.0)
        it = lib().lsxo_solver_solve(self.h, b, x, C.byref(nr))
        return x, it, nr.value

    def mg_apply(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        lib().lsxo_mg_apply(self.h, b, x)
        return x

    def levels(self):
        out = []
        for l in range(lib().lsxo_mg_levels(self.h)):
            dims = np.zeros(3, np.int32)
            nnz = lib().lsxo_mg_level_info(self.h, l, dims)
            n = int(np.prod(dims))
            rp = np.zeros(n + 1, np.int32)
            col = np.zeros(nnz, np.int32)
            val = np.zeros(nnz)
            lib().lsxo_mg_level_csr(self.h, l, rp, col, val)
            out.append(Csr(dims, rp, col, val))
        return out

    def adlj(self, level, direction):
        dims = np.zeros(3, np.int32)
        lib().lsxo_mg_level_info(self.h, level, dims)
        n = int(np.prod(dims))
        c, d, e = np.zeros(n - 1), np.zeros(n), np.zeros(n - 1)
        lib().lsxo_mg_level_adlj(self.h, level, direction, c, d, e)
        return c, d, e


def uniform_geometry(N, M, K, L=(1.0, 1.0, 1.0)):
    return Geometry(np.full(N, L[0] / N), np.full(M, L[1] / M), np.full(K, L[2] / K))
