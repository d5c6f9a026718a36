"""Python mirror of the reference solver interface over the liblsx C ABI.

``MGFGMRES`` / ``MGBiCGStab`` follow the reference's IterativeSolverInterface
(src/Base/Solvers/iterative_solver_interface.h:31-73): ``set(A, rhs, sol, task)`` stores the
matrix and builds the multigrid preconditioner, ``solve()`` updates ``sol`` in place and
returns the iteration count, ``rebuildPreconditioner()`` rebuilds the hierarchy after a
geometry change, and the public attributes ``eps``, ``norm_r`` and ``it`` mirror the
reference members.  The numerical work runs in liblsx.so (HIP kernels for gfx950); there is no
CPU fallback: loading fails loudly if the library is missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblsx.so")

K_RESIDUAL, K_LINE_X, K_LINE_Y, K_LINE_Z, K_OPERATOR, K_ORTHO, K_SMOOTH, K_CGS_DOT, K_CGS_UPDATE = range(9)
KERNEL_NAMES = {K_RESIDUAL: "residual_update", K_LINE_X: "line_solve_x", K_LINE_Y: "line_solve_y",
                K_LINE_Z: "line_solve_z", K_OPERATOR: "operator_apply", K_ORTHO: "mgs_axpy_dot",
                K_SMOOTH: "smoother_sweep", K_CGS_DOT: "cgs_dot", K_CGS_UPDATE: "cgs_update"}


def stretched_axis(n, length=1.0, center=0.5, ratio=1.0):
    """Cell widths of one axis, growing geometrically by `ratio` per cell away from `center`
    (ratio 1 gives a uniform axis of width length/n)."""
    if ratio == 1.0:
        return np.full(n, length / n)
    k = np.arange(n) + 0.5
    kc = center / length * n
    w = ratio ** np.abs(k - kc)
    return w * (length / w.sum())


def random_divergence_rhs(diag, dims, seed=0, pin=True):
    """Seeded synthetic pressure right-hand side: unit-variance Gaussian on the active cells,
    mean removed over them, zero elsewhere.  Active cells are those whose operator diagonal is
    negative (fluid and cut cells with at least one open face); solid cells (diagonal +1e-16),
    cut cells with every face closed (diagonal 0) and the pinned cells (+1) get 0, as a
    discrete divergence does."""
    N, M, K = dims
    active = diag < 0.0
    if pin:
        active[N * M - 1] = False
        active[N - 1] = False
    rng = np.random.default_rng(seed)
    b = np.zeros(diag.size)
    b[active] = rng.standard_normal(int(active.sum()))
    b[active] -= b[active].mean()
    return b


KRYLOV_FGMRES, KRYLOV_BICGSTAB = 0, 1
MG_CASCADE, MG_VCYCLE = 0, 1
PREC_FP64, PREC_MIXED, PREC_MIXED16 = 0, 1, 2
ORTHO_AUTO, ORTHO_MGS, ORTHO_CGS, ORTHO_CGS2 = 0, 1, 2, 3
MEM_HOST, MEM_DEVICE = 0, 1
OK, ERR_ARG, ERR_HIP, ERR_STATE = 0, -1, -2, -3


class LsxError(RuntimeError):
    pass


class Options(C.Structure):
    _fields_ = [
        ("krylov", C.c_int), ("restart", C.c_int), ("max_it", C.c_int), ("eps", C.c_double),
        ("mg", C.c_int), ("n_post", C.c_int), ("n_pre", C.c_int), ("omega", C.c_double),
        ("precision", C.c_int), ("levels", C.c_int), ("z_head_fix", C.c_int), ("ortho", C.c_int),
        ("coarse_fuse", C.c_int), ("graphs", C.c_int), ("smoother", C.c_int), ("fused_sweep", C.c_int),
        ("line_levels", C.c_int), ("diag_faces", C.c_int),
        ("dot32", C.c_int),
    ]


# V-cycle smoothers (lsx_options.smoother): red-black Gauss-Seidel, x/y/z line relaxation,
# damped point Jacobi (damping = the task's omega)
SMOOTHERS = {"rbgs": 0, "line": 1, "jacobi": 2}

_lib = None


def lib():
    """Load liblsx.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LsxError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.lsx_last_error.restype = C.c_char_p
    L.lsx_options_default.argtypes = [C.POINTER(Options)]
    L.lsx_options_default.restype = None
    L.lsx_create.argtypes = [C.POINTER(vp), C.POINTER(Options), vp]
    L.lsx_destroy.argtypes = [vp]
    L.lsx_set_csr.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_long, C.c_int]
    L.lsx_update_values.argtypes = [vp, vp, C.c_int]
    L.lsx_rebuild.argtypes = [vp]
    L.lsx_solve.argtypes = [vp, vp, vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.lsx_set_eps.argtypes = [vp, C.c_double]
    L.lsx_get_stats.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.lsx_precondition.argtypes = [vp, vp, vp, C.c_int]
    L.lsx_apply_operator.argtypes = [vp, C.c_int, vp, vp, C.c_int]
    L.lsx_solve_adjoint.argtypes = [vp, vp, vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.lsx_geometry_drag.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_double), vp]
    L.lsx_geometry_drag_noalloc.argtypes = [vp, vp, vp, vp]
    L.lsx_state_save.argtypes = [C.c_char_p, C.c_int, C.c_double, C.c_int, vp, vp, vp, C.c_int, C.c_int, C.c_int,
                                 vp, vp, vp, vp, vp, vp, vp, vp, C.c_int]
    L.lsx_state_load.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int, vp, vp, vp,
                                 C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, C.c_int]
    L.lsx_state_bytes.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    L.lsx_state_bytes.restype = C.c_long
    L.lsx_geometry_write_tecplot.argtypes = [vp, C.c_char_p, vp, vp, vp, vp, vp, C.POINTER(Walls), C.c_int, C.c_int]
    L.lsx_write_tecplot.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int] + [vp] * 15 + [C.POINTER(Walls), C.c_int]
    L.lsx_geometry_shape_gradient.argtypes = [vp, vp, vp, C.c_int, vp, vp, vp, vp, vp, C.POINTER(FlowBc), C.c_int,
                                              C.c_double, C.c_int, vp, vp]
    L.lsx_adjoint_forward.argtypes = [vp, vp, C.c_int, vp, vp, C.c_int, C.POINTER(vp), C.POINTER(C.c_int),
                                      C.POINTER(C.c_double)]
    L.lsx_adjoint_backward.argtypes = [vp, vp, C.c_int, vp, vp, vp, vp, vp, C.POINTER(FlowBc), C.c_double, C.c_int,
                                       vp, vp, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.lsx_tape_destroy.argtypes = [vp]
    L.lsx_precondition_transpose.argtypes = [vp, vp, vp, C.c_int]
    L.lsx_smooth.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int]
    L.lsx_apply_operator_transpose.argtypes = [vp, vp, vp, C.c_int]
    L.lsx_level_dims.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.lsx_level_dims.restype = C.c_long
    L.lsx_level_stencil.argtypes = [vp, C.c_int, vp, C.c_long]
    L.lsx_synchronize.argtypes = [vp]
    L.lsx_geometry_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_double, C.c_double,
                                      C.c_double, vp]
    L.lsx_geometry_destroy.argtypes = [vp]
    L.lsx_geometry_build.argtypes = [vp, C.POINTER(Body), C.c_int, C.c_int, C.c_int, C.c_int]
    L.lsx_geometry_fields.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.lsx_set_geometry.argtypes = [vp, vp, C.c_int]
    L.lsx_set_fields.argtypes = [vp, C.c_int, C.c_int, C.c_int] + [vp] * 8 + [C.c_int, C.c_int]
    L.lsx_set_helmholtz.argtypes = [vp, vp, C.c_int, C.c_double, C.c_double, C.c_int]
    L.lsx_geometry_rhs.argtypes = [vp, vp, vp, vp, vp, vp, C.POINTER(FlowBc), vp, C.c_int]
    L.lsx_geometry_correct.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, C.c_double, C.c_int]
    L.lsx_operator_diagonal.argtypes = [vp, vp]
    L.lsx_fine_storage.argtypes = [vp]
    L.lsx_bezier_ellipsoid.argtypes = [vp, vp, vp]
    L.lsx_kernel_timing.argtypes = [vp, C.c_int]
    L.lsx_kernel_stats.argtypes = [vp, C.c_int, C.POINTER(C.c_long), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    _lib = L
    return L


class Body(C.Structure):
    _fields_ = [("kind", C.c_int), ("center", C.c_double * 3), ("axes", C.c_double * 3),
                ("ctrl", C.c_void_p)]


BODY_SPHERE, BODY_ELLIPSOID, BODY_BEZIER6 = 0, 1, 2

# Bezier patches on the cube faces +x, -x, +y, -y, +z, -z: (normal axis, sign, u axis, v axis)
BEZIER_FACES = [(0, 1, 1, 2), (0, -1, 2, 1), (1, 1, 2, 0), (1, -1, 0, 2), (2, 1, 0, 1), (2, -1, 1, 0)]


def bezier_ellipsoid(center, axes):
    """Control points (6, 4, 4, 3) of the six-patch bicubic Bezier fit of an ellipsoid."""
    ctrl = np.zeros(288)
    c = np.ascontiguousarray(center, np.float64)
    a = np.ascontiguousarray(axes, np.float64)
    _check(lib().lsx_bezier_ellipsoid(_ptr(c), _ptr(a), _ptr(ctrl)))
    return ctrl.reshape(6, 4, 4, 3)


def bezier_perturb(ctrl, amplitude, seed):
    """Seeded U(-amplitude, amplitude) perturbation of every coordinate of the 56 distinct
    control points (lattice order), applied to all patch copies of a point so the surface stays
    closed."""
    ctrl = np.array(ctrl, np.float64).reshape(6, 4, 4, 3)
    copies = {}
    for f, (a, s, bu, bv) in enumerate(BEZIER_FACES):
        for i in range(4):
            for j in range(4):
                l = [0, 0, 0]
                l[a], l[bu], l[bv] = (3 if s > 0 else 0), i, j
                copies.setdefault(tuple(l), []).append((f, i, j))
    rng = np.random.default_rng(seed)
    for key in sorted(copies):
        d = rng.uniform(-amplitude, amplitude, 3)
        for (f, i, j) in copies[key]:
            ctrl[f, i, j] += d
    return ctrl


OBJ_NONE = 0
OBJ_DRAG_X = 1


class FlowBc(C.Structure):
    _fields_ = [("U_E", C.c_double), ("V_S", C.c_double), ("V_N", C.c_double), ("W_B", C.c_double),
                ("W_T", C.c_double), ("outflow", C.c_int), ("pin", C.c_int)]


WALL_FIELDS = ("V_W", "W_W", "U_E", "V_E", "W_E", "U_S", "V_S", "W_S", "U_N", "V_N", "W_N", "U_B", "V_B", "W_B",
               "U_T", "V_T", "W_T")


class Walls(C.Structure):
    """Wall velocities of the task (reference BaseTaskDescription U_*, V_*, W_*)."""
    _fields_ = [(n, C.c_double) for n in WALL_FIELDS]


def walls(**kw):
    """Walls with the given velocities, 0 elsewhere (e.g. walls(U_E=1.0))."""
    bad = set(kw) - set(WALL_FIELDS)
    if bad:
        raise LsxError(f"unknown wall velocities {sorted(bad)}")
    return Walls(*[float(kw.get(n, 0.0)) for n in WALL_FIELDS])


def face_sizes(N, M, K):
    """Sizes of u (x faces), v (y faces), w (z faces) and p on an N x M x K pressure grid."""
    return (N - 1) * M * K, N * (M - 1) * K, N * M * (K - 1), N * M * K


def state_bytes(N, M, K, n_bodies=0):
    r = lib().lsx_state_bytes(N, M, K, n_bodies)
    if r < 0:
        _check(r)
    return r


def state_save(path, it, time, dims, cur, prev, bodies=None):
    """Checkpoint in the reference's binary layout (SetSolutions::saveState): cur and prev are
    (u, v, w, p) host arrays; bodies = (pos (nb, 3), vel (nb, 3), loads (nb, 4)) or None."""
    N, M, K = dims
    sizes = face_sizes(N, M, K)
    arr = [np.ascontiguousarray(a, np.float64) for a in list(cur) + list(prev)]
    for a, n in zip(arr, sizes * 2):
        if a.size != n:
            raise LsxError("solution array of the wrong size")
    nb = 0
    bp = bv = bl = None
    if bodies is not None:
        bp, bv, bl = (np.ascontiguousarray(a, np.float64) for a in bodies)
        nb = bp.size // 3
        if bp.size != 3 * nb or bv.size != 3 * nb or bl.size != 4 * nb:
            raise LsxError("body arrays of the wrong size")
    _check(lib().lsx_state_save(os.fsencode(path), int(it), float(time), nb,
                                None if bp is None else _ptr(bp), None if bv is None else _ptr(bv),
                                None if bl is None else _ptr(bl), N, M, K, *[_ptr(a) for a in arr], MEM_HOST))


def state_load(path, dims, n_bodies=0):
    """Read a checkpoint written by state_save (or the reference): returns a dict with it, time,
    cur and prev (u, v, w, p) and the body arrays."""
    N, M, K = dims
    sizes = face_sizes(N, M, K)
    arr = [np.zeros(n) for n in sizes * 2]
    bp, bv, bl = np.zeros(3 * n_bodies), np.zeros(3 * n_bodies), np.zeros(4 * n_bodies)
    it = C.c_int(0)
    t = C.c_double(0.0)
    _check(lib().lsx_state_load(os.fsencode(path), C.byref(it), C.byref(t), int(n_bodies),
                                _ptr(bp) if n_bodies else None, _ptr(bv) if n_bodies else None,
                                _ptr(bl) if n_bodies else None, N, M, K, *[_ptr(a) for a in arr], MEM_HOST))
    return dict(it=it.value, time=t.value, cur=tuple(arr[:4]), prev=tuple(arr[4:]), body_pos=bp.reshape(-1, 3),
                body_vel=bv.reshape(-1, 3), loads=bl.reshape(-1, 4))


def write_tecplot(path, dims, centres, widths, fields, u, v, w, p, u_west=None, walls_=None, outflow=True):
    """Tecplot point file (reference SolTecplotPrinter) from host arrays: centres / widths are
    per-axis tuples, fields a dict with type, thx, thy, thz (as Geometry.fields returns)."""
    N, M, K = dims
    c = [np.ascontiguousarray(a, np.float64) for a in centres]
    d = [np.ascontiguousarray(a, np.float64) for a in widths]
    t = np.ascontiguousarray(fields["type"], np.int32)
    th = [np.ascontiguousarray(fields[k], np.float64) for k in ("thx", "thy", "thz")]
    f = [np.ascontiguousarray(a, np.float64) for a in (u, v, w, p)]
    for a, n in zip(f, face_sizes(N, M, K)):
        if a.size != n:
            raise LsxError("field of the wrong size")
    uw = None if u_west is None else np.ascontiguousarray(np.broadcast_to(np.asarray(u_west, np.float64), (M * K,)))
    wl = walls_ if walls_ is not None else walls()
    _check(lib().lsx_write_tecplot(os.fsencode(path), N, M, K, *[_ptr(a) for a in c + d], _ptr(t),
                                   *[_ptr(a) for a in th + f], None if uw is None else _ptr(uw), C.byref(wl),
                                   1 if outflow else 0))


class Geometry:
    """Cut-cell geometry of the pressure grid, built on the GPU (reference LSSTAGUniformMesh
    mesh functions for a static body)."""

    def __init__(self, dx, dy, dz, origin=(0.0, 0.0, 0.0), stream=None):
        self.dx = np.ascontiguousarray(dx, dtype=np.float64)
        self.dy = np.ascontiguousarray(dy, dtype=np.float64)
        self.dz = np.ascontiguousarray(dz, dtype=np.float64)
        self.dims = (len(self.dx), len(self.dy), len(self.dz))
        h = C.c_void_p()
        _check(lib().lsx_geometry_create(C.byref(h), *self.dims, _ptr(self.dx), _ptr(self.dy), _ptr(self.dz),
                                         *origin, stream))
        self.h = h

    @property
    def n(self):
        N, M, K = self.dims
        return N * M * K

    def build(self, kind, center, axes, box=None, ctrl=None):
        """kind 'sphere' (axes = (r,)), 'ellipsoid' (semi-axes) or 'bezier' (ctrl: (6, 4, 4, 3)
        control points; axes: the ellipsoid the closest-point search starts from)."""
        N, M, K = self.dims
        hs, hf, vs, vf = box or (1, N - 1, 1, M - 1)
        a = tuple(axes) + (axes[0],) * (3 - len(axes))
        code = {"sphere": BODY_SPHERE, "ellipsoid": BODY_ELLIPSOID, "bezier": BODY_BEZIER6}[kind]
        cp = None
        if code == BODY_BEZIER6:
            cp = np.ascontiguousarray(ctrl, np.float64).ravel()
            if cp.size != 288:
                raise LsxError("a Bezier body needs 288 control values")
        b = Body(code, (C.c_double * 3)(*center), (C.c_double * 3)(*a), None if cp is None else cp.ctypes.data)
        _check(lib().lsx_geometry_build(self.h, C.byref(b), hs, hf, vs, vf))
        return self

    def fields(self):
        n = self.n
        t = np.zeros(n, np.int32)
        tx, ty, tz, v, phi = (np.zeros(n) for _ in range(5))
        _check(lib().lsx_geometry_fields(self.h, _ptr(t), _ptr(tx), _ptr(ty), _ptr(tz), _ptr(v), _ptr(phi)))
        return dict(type=t, thx=tx, thy=ty, thz=tz, vof=v, phi=phi)

    def face_shapes(self):
        """Sizes of the staggered face arrays u (x faces), v (y faces), w (z faces)."""
        N, M, K = self.dims
        return (N - 1) * M * K, N * (M - 1) * K, N * M * (K - 1)

    def pressure_rhs(self, u, v, w, U_W=1.0, U_E=1.0, V_S=0.0, V_N=0.0, W_B=0.0, W_T=0.0, outflow=True, pin=True,
                     u_out=None):
        """b = Dx u + Dy v + Dz w + flux on the device (reference PoissonDiscreteEquation::
        recalcRightMember for a static body); host arrays in, host array out."""
        N, M, K = self.dims
        nu, nv, nw = self.face_shapes()
        u = np.ascontiguousarray(u, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        if u.size != nu or v.size != nv or w.size != nw:
            raise LsxError("face arrays have the wrong sizes")
        uw = np.ascontiguousarray(np.broadcast_to(np.asarray(U_W, np.float64), (M * K,)))
        uo = None if u_out is None else np.ascontiguousarray(u_out, np.float64)
        bc = FlowBc(U_E, V_S, V_N, W_B, W_T, 1 if outflow else 0, 1 if pin else 0)
        b = np.zeros(self.n)
        _check(lib().lsx_geometry_rhs(self.h, _ptr(u), _ptr(v), _ptr(w), _ptr(uw), None if uo is None else _ptr(uo),
                                      C.byref(bc), _ptr(b), MEM_HOST))
        return b

    def flow_args(self, u, v, w, U_W=1.0, U_E=1.0, V_S=0.0, V_N=0.0, W_B=0.0, W_T=0.0, outflow=True, pin=True,
                  u_out=None):
        """The flow fields of a pressure right-hand side as C arguments (arrays kept alive)."""
        M, K = self.dims[1], self.dims[2]
        nu, nv, nw = self.face_shapes()
        arr = [np.ascontiguousarray(a, np.float64) for a in (u, v, w)]
        if arr[0].size != nu or arr[1].size != nv or arr[2].size != nw:
            raise LsxError("face arrays have the wrong sizes")
        arr.append(np.ascontiguousarray(np.broadcast_to(np.asarray(U_W, np.float64), (M * K,))))
        arr.append(None if u_out is None else np.ascontiguousarray(u_out, np.float64))
        bc = FlowBc(U_E, V_S, V_N, W_B, W_T, 1 if outflow else 0, 1 if pin else 0)
        return arr, bc

    def drag(self, p):
        """Pressure drag J = sum over cut cells of (theta_x[i-1] - theta_x[i]) dy dz p[i]
        (pressure term of the reference LoadsCalculator::calc); returns (J, dJ/dp)."""
        p = np.ascontiguousarray(p, np.float64)
        if p.size != self.n:
            raise LsxError("p has the wrong size")
        J = C.c_double(0.0)
        g = np.zeros(self.n)
        _check(lib().lsx_geometry_drag(self.h, _ptr(p), MEM_HOST, C.byref(J), _ptr(g)))
        return J.value, g

    def shape_gradient(self, p, lam, objective=OBJ_NONE, flow=None, pin=True, h_rel=1e-7, ctrl=True):
        """Shape gradient at the solution p with adjoint lam = A^{-T} dJ/dp: returns
        (dJ/dphi per level-set corner, dJ/dctrl[288] or None).  flow: None for a right-hand side
        that does not depend on the geometry, else a dict of pressure_rhs arguments."""
        p = np.ascontiguousarray(p, np.float64)
        lam = np.ascontiguousarray(lam, np.float64)
        gphi = np.zeros(self.n)
        gctrl = np.zeros(288) if ctrl else None
        if flow is None:
            fa, bc = [None] * 5, None
        else:
            fa, bc = self.flow_args(**flow)
        ptrs = [None if a is None else _ptr(a) for a in fa]
        _check(lib().lsx_geometry_shape_gradient(self.h, _ptr(p), _ptr(lam), int(objective), *ptrs,
                                                 None if bc is None else C.byref(bc), 1 if pin else 0,
                                                 float(h_rel), MEM_HOST, _ptr(gphi),
                                                 None if gctrl is None else _ptr(gctrl)))
        return gphi, gctrl

    def correct(self, dp, us, vs, ws, p=None, coef=1.0):
        """Projection correction on the device (reference SetSolutions::correction): returns the new
        face velocities (u, v, w); p (host array) is updated in place by p += dp * coef."""
        nu, nv, nw = self.face_shapes()
        dp = np.ascontiguousarray(dp, np.float64)
        us, vs, ws = (np.ascontiguousarray(a, np.float64) for a in (us, vs, ws))
        if dp.size != self.n or us.size != nu or vs.size != nv or ws.size != nw:
            raise LsxError("array sizes do not match the grid")
        u, v, w = np.zeros(nu), np.zeros(nv), np.zeros(nw)
        if p is not None and not (isinstance(p, np.ndarray) and p.dtype == np.float64 and p.flags.c_contiguous):
            raise LsxError("p must be a contiguous float64 array (updated in place)")
        _check(lib().lsx_geometry_correct(self.h, _ptr(dp), _ptr(us), _ptr(vs), _ptr(ws), _ptr(u), _ptr(v), _ptr(w),
                                          None if p is None else _ptr(p), coef, MEM_HOST))
        return u, v, w

    def write_tecplot(self, path, u, v, w, p, u_west=None, walls_=None, outflow=True):
        """Tecplot point file of the solution on this grid (reference SolTecplotPrinter::printField)."""
        N, M, K = self.dims
        f = [np.ascontiguousarray(a, np.float64) for a in (u, v, w, p)]
        for a, n in zip(f, face_sizes(N, M, K)):
            if a.size != n:
                raise LsxError("field of the wrong size")
        uw = None if u_west is None else np.ascontiguousarray(np.broadcast_to(np.asarray(u_west, np.float64),
                                                                              (M * K,)))
        wl = walls_ if walls_ is not None else walls()
        _check(lib().lsx_geometry_write_tecplot(self.h, os.fsencode(path), *[_ptr(a) for a in f],
                                                None if uw is None else _ptr(uw), C.byref(wl), 1 if outflow else 0,
                                                MEM_HOST))

    def close(self):
        if self.h:
            lib().lsx_geometry_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _check(status):
    if status != 0:
        raise LsxError(f"liblsx status {status}: {lib().lsx_last_error().decode()}")


def _ptr(a):
    return C.c_void_p(a.ctypes.data)


class Task:
    """The solver fields of the reference BaseTaskDescription (base_task_description.h:41-46)."""

    def __init__(self, eps=1e-5, MAX_ITER=200, FGMRES_RESTART_NUM=13, n_post=3, omega=0.805, n_pre=0):
        self.n_pre = n_pre
        self.eps = eps
        self.MAX_ITER = MAX_ITER
        self.FGMRES_RESTART_NUM = FGMRES_RESTART_NUM
        self.n_post = n_post
        self.omega = omega


class IterativeSolver:
    """Base of MGFGMRES / MGBiCGStab (reference IterativeSolverInterface)."""

    krylov = KRYLOV_FGMRES

    def __init__(self, precision=PREC_FP64, mg=MG_CASCADE, z_head_fix=False, stream=None, ortho=0, coarse_fuse=True,
                 graphs=False, smoother="rbgs", fused_sweep=True, line_levels=0, levels=0, diag_faces=True,
                 dot32=True):
        self.levels_cap = levels
        self.diag_faces = diag_faces
        self.dot32 = dot32
        self.ortho = ortho
        self.coarse_fuse = coarse_fuse
        self.graphs = graphs
        if smoother not in SMOOTHERS:
            raise ValueError(f"unknown smoother {smoother!r}")
        self.smoother = smoother
        self.fused_sweep = fused_sweep
        self.line_levels = line_levels
        self.precision = precision
        self.mg = mg
        self.z_head_fix = z_head_fix
        self.stream = stream
        self.h = None
        self.eps = 0.0
        self.norm_r = 0.0
        self.it = 1
        self._rhs = None
        self._sol = None

    # -- reference interface --------------------------------------------------------------
    def set(self, A, rhs, sol, task):
        """A: object with dims (N, M, K) and CSR arrays rowptr/col/val (reference SparseMatrix)."""
        self._make(task)
        N, M, K = A.dims
        self._A = A
        rp = np.ascontiguousarray(A.rowptr, dtype=np.int32)
        cl = np.ascontiguousarray(A.col, dtype=np.int32)
        vl = np.ascontiguousarray(A.val, dtype=np.float64)
        _check(lib().lsx_set_csr(self.h, N, M, K, _ptr(rp), _ptr(cl), _ptr(vl), len(vl), MEM_HOST))
        self._rhs = rhs
        self._sol = sol

    def set_geometry(self, geometry, rhs, sol, task, pin=True):
        """Like set(), with the cut-cell operator built on the device from `geometry`."""
        self._make(task)
        _check(lib().lsx_set_geometry(self.h, geometry.h, 1 if pin else 0))
        self._rhs = rhs
        self._sol = sol

    def set_fields(self, dims, widths, fields, rhs, sol, task, pin=True):
        """Like set(), with the cut-cell operator assembled on the device from the caller's mesh
        fields: widths (dx, dy, dz) and a dict with type, thx, thy, thz, vof (N*M*K each)."""
        N, M, K = dims
        n = N * M * K
        w = [np.ascontiguousarray(a, np.float64) for a in widths]
        t = np.ascontiguousarray(fields["type"], np.int32)
        f = [np.ascontiguousarray(fields[k], np.float64) for k in ("thx", "thy", "thz", "vof")]
        if [a.size for a in w] != [N, M, K] or t.size != n or any(a.size != n for a in f):
            raise LsxError("mesh fields do not match the grid")
        self._make(task)
        _check(lib().lsx_set_fields(self.h, N, M, K, *[_ptr(a) for a in w], _ptr(t), *[_ptr(a) for a in f],
                                    1 if pin else 0, MEM_HOST))
        self._rhs = rhs
        self._sol = sol

    def set_helmholtz(self, geometry, component, nu, coef_time, rhs, sol, task, flow=True):
        """Like set(), with the momentum operator of velocity component 0/1/2 of `geometry` (a
        box without cut cells) assembled on the device on its staggered mesh."""
        self._make(task)
        _check(lib().lsx_set_helmholtz(self.h, geometry.h, int(component), float(nu), float(coef_time),
                                       1 if flow else 0))
        self._rhs = rhs
        self._sol = sol

    def _make(self, task):
        self.close()
        opt = Options()
        lib().lsx_options_default(C.byref(opt))
        opt.krylov = self.krylov
        opt.restart = int(task.FGMRES_RESTART_NUM)
        opt.max_it = int(task.MAX_ITER)
        opt.eps = float(task.eps)
        opt.n_post = int(task.n_post)
        opt.n_pre = int(getattr(task, "n_pre", 0))
        opt.omega = float(task.omega)
        opt.precision = self.precision
        opt.mg = self.mg
        opt.z_head_fix = 1 if self.z_head_fix else 0
        opt.ortho = int(self.ortho)
        opt.coarse_fuse = 1 if self.coarse_fuse else 0
        opt.graphs = 1 if self.graphs else 0
        opt.levels = int(getattr(self, "levels_cap", 0))
        opt.smoother = SMOOTHERS[self.smoother]
        opt.fused_sweep = int(self.fused_sweep)
        opt.line_levels = int(self.line_levels)
        opt.diag_faces = 1 if getattr(self, "diag_faces", True) else 0
        opt.dot32 = 1 if getattr(self, "dot32", True) else 0
        h = C.c_void_p()
        _check(lib().lsx_create(C.byref(h), C.byref(opt), self.stream))
        self.h = h
        self.eps = float(task.eps)

    def solve(self):
        b = np.ascontiguousarray(self._rhs, dtype=np.float64)
        if not (isinstance(self._sol, np.ndarray) and self._sol.dtype == np.float64 and self._sol.flags.c_contiguous):
            raise LsxError("sol must be a contiguous float64 array (updated in place)")
        _check(lib().lsx_set_eps(self.h, self.eps))
        it = C.c_int(0)
        nr = C.c_double(0.0)
        _check(lib().lsx_solve(self.h, _ptr(b), _ptr(self._sol), MEM_HOST, C.byref(it), C.byref(nr)))
        self.it = it.value
        self.norm_r = nr.value
        return self.it

    def rebuildPreconditioner(self):
        _check(lib().lsx_rebuild(self.h))

    # -- extras used by tests and benchmarks -----------------------------------------------
    def update_values(self, val):
        v = np.ascontiguousarray(val, dtype=np.float64)
        _check(lib().lsx_update_values(self.h, _ptr(v), MEM_HOST))

    def precondition(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        z = np.zeros_like(v)
        _check(lib().lsx_precondition(self.h, _ptr(v), _ptr(z), MEM_HOST))
        return z

    def solve_adjoint(self, g, lam=None):
        """lambda = A^{-T} g (FGMRES / BiCGStab on A^T with the transposed V-cycle); returns
        (lambda, iterations, |g - A^T lambda|).  `lam` is the initial guess (0 by default)."""
        g = np.ascontiguousarray(g, dtype=np.float64)
        lam = np.zeros_like(g) if lam is None else np.array(lam, dtype=np.float64, copy=True)
        _check(lib().lsx_set_eps(self.h, self.eps))
        it = C.c_int(0)
        nr = C.c_double(0.0)
        _check(lib().lsx_solve_adjoint(self.h, _ptr(g), _ptr(lam), MEM_HOST, C.byref(it), C.byref(nr)))
        return lam, it.value, nr.value

    def solve_device(self, b, x):
        """Solve with b and x (initial guess in, solution out) as DeviceVectors; returns it."""
        _check(lib().lsx_set_eps(self.h, self.eps))
        it = C.c_int(0)
        nr = C.c_double(0.0)
        _check(lib().lsx_solve(self.h, b.p, x.p, MEM_DEVICE, C.byref(it), C.byref(nr)))
        self.it, self.norm_r = it.value, nr.value
        return it.value

    def adjoint_forward(self, geometry, b, pin=True, p0=None):
        """Solve A(geometry) p = b and record a tape for adjoint_backward; returns (p, it, tape)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        p = np.zeros_like(b) if p0 is None else np.array(p0, dtype=np.float64, copy=True)
        _check(lib().lsx_set_eps(self.h, self.eps))
        tape = C.c_void_p()
        it = C.c_int(0)
        nr = C.c_double(0.0)
        _check(lib().lsx_adjoint_forward(self.h, geometry.h, 1 if pin else 0, _ptr(b), _ptr(p), MEM_HOST,
                                         C.byref(tape), C.byref(it), C.byref(nr)))
        self.it, self.norm_r = it.value, nr.value
        return p, it.value, Tape(tape, geometry, self)

    def use_geometry(self, geometry, pin=True):
        """Replace the operator by that of `geometry` on this handle (options unchanged)."""
        _check(lib().lsx_set_geometry(self.h, geometry.h, 1 if pin else 0))

    def smooth(self, b, x, sweeps=1, zero_guess=False):
        """`sweeps` sweeps of the V-cycle smoother on the finest level; returns the new x."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.array(x, dtype=np.float64, copy=True)
        _check(lib().lsx_smooth(self.h, _ptr(b), _ptr(x), int(sweeps), 1 if zero_guess else 0, MEM_HOST))
        return x

    def precondition_transpose(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        z = np.zeros_like(v)
        _check(lib().lsx_precondition_transpose(self.h, _ptr(v), _ptr(z), MEM_HOST))
        return z

    def apply_operator_transpose(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        _check(lib().lsx_apply_operator_transpose(self.h, _ptr(x), _ptr(y), MEM_HOST))
        return y

    def apply_operator(self, x, level=0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        N, M, K = self.level_dims(level)
        y = np.zeros(N * M * K)
        _check(lib().lsx_apply_operator(self.h, level, _ptr(x), _ptr(y), MEM_HOST))
        return y

    def operator_diagonal(self):
        N, M, K = self.level_dims(0)
        d = np.zeros(N * M * K)
        _check(lib().lsx_operator_diagonal(self.h, _ptr(d)))
        return d

    def fine_storage(self):
        """4 when the solve uses the symmetric four-plane fine operator, 7 for the general one."""
        r = lib().lsx_fine_storage(self.h)
        if r < 0:
            _check(r)
        return r

    def kernel_timing(self, enable=True):
        _check(lib().lsx_kernel_timing(self.h, 1 if enable else 0))

    def kernel_stats(self, kernel):
        """(launches, total_ms, algorithmic_bytes_per_launch) of a fine-level kernel class."""
        n, ms, by = C.c_long(), C.c_double(), C.c_double()
        _check(lib().lsx_kernel_stats(self.h, kernel, C.byref(n), C.byref(ms), C.byref(by)))
        return n.value, ms.value, by.value

    def num_levels(self):
        it, nr, lv = C.c_int(), C.c_double(), C.c_int()
        _check(lib().lsx_get_stats(self.h, C.byref(it), C.byref(nr), C.byref(lv)))
        return lv.value

    def level_dims(self, level):
        N, M, K = C.c_int(), C.c_int(), C.c_int()
        r = lib().lsx_level_dims(self.h, level, C.byref(N), C.byref(M), C.byref(K))
        if r < 0:
            _check(int(r))
        return N.value, M.value, K.value

    def level_stencil(self, level):
        N, M, K = self.level_dims(level)
        n = N * M * K
        buf = np.zeros(27 * n)
        npts = lib().lsx_level_stencil(self.h, level, _ptr(buf), len(buf))
        if npts < 0:
            _check(npts)
        return buf[: npts * n].reshape(npts, n)

    def close(self):
        if self.h:
            lib().lsx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MGFGMRES(IterativeSolver):
    """Reference MGFGMRES (src/Base/Solvers/FGMRES/mg_fgmres.h:33-52)."""
    krylov = KRYLOV_FGMRES


class MGBiCGStab(IterativeSolver):
    """Reference MGBiCGStab (src/Base/Solvers/BiCGStab/mg_bicgstab.h:33-52)."""
    krylov = KRYLOV_BICGSTAB


class DeviceVector:
    """A float64 vector in device memory (hipMalloc through the HIP runtime), for solves whose
    inputs and outputs stay on the GPU."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            h = C.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
            cls._hip = h
        return cls._hip

    def __init__(self, n):
        self.n = int(n)
        self.p = C.c_void_p()
        if self.hip().hipMalloc(C.byref(self.p), self.n * 8) != 0:
            raise LsxError("hipMalloc failed")

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        if a.size != self.n or self.hip().hipMemcpy(self.p, a.ctypes.data, self.n * 8, 1) != 0:
            raise LsxError("upload failed")

    def download(self, out=None):
        out = np.empty(self.n) if out is None else out
        if self.hip().hipMemcpy(out.ctypes.data, self.p, self.n * 8, 2) != 0:
            raise LsxError("download failed")
        return out

    def zero(self):
        if self.hip().hipMemset(self.p, 0, self.n * 8) != 0:
            raise LsxError("hipMemset failed")

    def close(self):
        if self.p:
            self.hip().hipFree(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tape:
    """Record of a forward pressure solve (lsx_adjoint_forward).  It refers to the solver handle
    and the geometry of the forward call, which must stay alive (a solver re-created by set() or
    set_geometry() invalidates the tape)."""

    def __init__(self, h, geometry, solver):
        self.h = h
        self.geometry = geometry
        self.solver = solver
        self.solver_h = solver.h.value

    def backward(self, dJdp, objective=OBJ_NONE, flow=None, h_rel=1e-7, ctrl=True):
        """dJ/db = A^{-T} dJ/dp and (ctrl) dJ/dctrl; returns (dJdb, dJdctrl or None, adjoint iterations)."""
        if self.solver.h is None or self.solver.h.value != self.solver_h:
            raise LsxError("the tape's solver handle was closed or re-created")
        _check(lib().lsx_set_eps(self.solver.h, self.solver.eps))
        g = np.ascontiguousarray(dJdp, np.float64)
        dJdb = np.zeros_like(g)
        gctrl = np.zeros(288) if ctrl else None
        if flow is None:
            fa, bc = [None] * 5, None
        else:
            fa, bc = self.geometry.flow_args(**flow)
        ptrs = [None if a is None else _ptr(a) for a in fa]
        it = C.c_int(0)
        nr = C.c_double(0.0)
        _check(lib().lsx_adjoint_backward(self.h, _ptr(g), int(objective), *ptrs, None if bc is None else C.byref(bc),
                                          float(h_rel), MEM_HOST, _ptr(dJdb), None if gctrl is None else _ptr(gctrl),
                                          C.byref(it), C.byref(nr)))
        return dJdb, gctrl, it.value

    def close(self):
        if self.h:
            lib().lsx_tape_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mixed_task(eps, restart=30, max_it=1000, n_pre=2, n_post=2):
    """Task for the mixed-precision solvers: FP32 V(n_pre, n_post) red-black Gauss-Seidel cycle."""
    return Task(eps=eps, MAX_ITER=max_it, FGMRES_RESTART_NUM=restart, n_post=n_post, n_pre=n_pre)


def solver(kind="fgmres", precision="fp64", **kw):
    """MGFGMRES / MGBiCGStab with one of four preconditioners:
    "fp64"        the reference FP64 multigrid cascade (ADLJ line smoothing, Galerkin levels);
    "fp64-vcycle" the FP64 V-cycle (red-black Gauss-Seidel, rediscretised face levels);
    "mixed"       the same V-cycle in FP32 with FP16 coarse faces;
    "mixed16"     as "mixed" with FP16 faces on the finest level too.
    The V-cycles take pressure operators (one or two pinned rows) and operators with a diagonal
    beyond their couplings (the momentum operators, whose mass term the coarse levels sum)."""
    cls = MGFGMRES if kind == "fgmres" else MGBiCGStab
    if precision == "mixed":
        return cls(precision=PREC_MIXED, mg=MG_VCYCLE, **kw)
    if precision == "mixed16":
        return cls(precision=PREC_MIXED16, mg=MG_VCYCLE, **kw)
    if precision == "fp64-vcycle":
        return cls(precision=PREC_FP64, mg=MG_VCYCLE, **kw)
    if precision != "fp64":
        raise ValueError(f"unknown precision {precision!r}")
    return cls(**kw)
