"""Test problems for the pressure solve, built with the CPU oracle.

* ``body_free``: the reference operator on an empty box (all faces open), pinned outlet rows.
* ``synthetic_cut``: cells whose centre lies inside a ball are solid, the shell of cells
  touching them gets the reference's cut-cell row shape (7 entries, theta = 0 on faces shared
  with solid cells, seeded theta in [0.05, 1] elsewhere, seeded partial volumes).  It exercises
  solid rows (diagonal PLUS_TINY), zero couplings and zero-pivot line decoupling.
* ``random_rhs``: the C2 right-hand side: unit-variance Gaussian on non-solid cells, mean
  removed, zero at solid cells and at the two pinned cells.
"""
import numpy as np

import oracle_py as O

SOLID = 0
CUT_CODE = 1111111  # a reference cut-cell code (TNE_Heptahedron3); any interior cut code
                    # takes the reference's default 7-entry row shape


def body_free(N, M=None, K=None, L=(1.0, 1.0, 1.0), pin=True):
    M = M or N
    K = K or N
    g = O.uniform_geometry(N, M, K, L)
    return g, g.laplace(pin=pin)


def synthetic_cut(N, M=None, K=None, center=(0.35, 0.5, 0.5), radius=0.2, seed=0, pin=True,
                  spacing=None):
    M = M or N
    K = K or N
    rng = np.random.default_rng(seed)
    if spacing is None:
        g = O.uniform_geometry(N, M, K)
    else:
        g = O.Geometry(*spacing)
    f = g.fields()
    xs = np.concatenate([[0.0], np.cumsum(g.dx)])
    ys = np.concatenate([[0.0], np.cumsum(g.dy)])
    zs = np.concatenate([[0.0], np.cumsum(g.dz)])
    xc = 0.5 * (xs[1:] + xs[:-1])
    yc = 0.5 * (ys[1:] + ys[:-1])
    zc = 0.5 * (zs[1:] + zs[:-1])
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    solid = (X - center[0]) ** 2 + (Y - center[1]) ** 2 + (Z - center[2]) ** 2 < radius ** 2
    interior = np.zeros_like(solid)
    interior[1:-1, 1:-1, 1:-1] = True
    solid &= interior
    near = np.zeros_like(solid)
    for ax in range(3):
        for sh in (-1, 1):
            near |= np.roll(solid, sh, axis=ax)
    cut = near & ~solid & interior
    t = f["type"].reshape(K, M, N)
    t[solid] = SOLID
    t[cut] = CUT_CODE
    thx = f["thx"].reshape(K, M, N)
    thy = f["thy"].reshape(K, M, N)
    thz = f["thz"].reshape(K, M, N)
    vof = f["vof"].reshape(K, M, N)
    # faces touching a solid cell are closed; faces between two cut cells get a seeded open
    # fraction; a face of a fully fluid cell stays open (all its corners are fluid)
    cutish = cut | solid
    for th, ax in ((thx, 2), (thy, 1), (thz, 0)):
        nb_solid = np.roll(solid, -1, axis=ax)
        nb_cut = np.roll(cutish, -1, axis=ax)
        both = cutish & nb_cut
        th[both] = rng.uniform(0.05, 1.0, size=int(both.sum()))
        th[solid | nb_solid] = 0.0
    vof[cut] *= rng.uniform(0.1, 1.0, size=int(cut.sum()))
    vof[solid] = 0.0
    g.set_fields(type=t.ravel(), thx=thx.ravel(), thy=thy.ravel(), thz=thz.ravel(), vof=vof.ravel())
    return g, g.laplace(pin=pin)


def random_rhs(g, seed=0, pin=True, A=None):
    """Gaussian on active cells (negative operator diagonal), mean removed, zero elsewhere:
    solid cells (+1e-16 diagonal), cut cells whose faces are all closed (empty rows) and the
    pinned cells get 0, so the right-hand side is consistent."""
    N, M, K = g.dims
    if A is None:
        A = g.laplace(pin=pin)
    active = A.to_scipy().diagonal() < 0.0
    if pin:
        active[N * M - 1] = False
        active[N - 1] = False
    rng = np.random.default_rng(seed)
    b = np.zeros(g.n)
    b[active] = rng.standard_normal(int(active.sum()))
    b[active] -= b[active].mean()
    return b


def sphere(N, M=None, K=None, center=(0.35, 0.5, 0.5), radius=0.2, pin=True):
    """The reference's sphere level set on a uniform grid of the unit cube (config C1 geometry)."""
    M = M or N
    K = K or N
    g = O.uniform_geometry(N, M, K)
    g.build_body("sphere", center, (radius,))
    return g, g.laplace(pin=pin)


def ellipsoid_stretched(N, ratio=1.02, axes=(0.25, 0.15, 0.15), pin=True):
    """Ellipsoid at the centre of a geometrically stretched grid (config C3 geometry)."""
    w = [stretched_axis(N, 1.0, 0.5, ratio) for _ in range(3)]
    g = O.Geometry(*w)
    g.build_body("ellipsoid", (0.5, 0.5, 0.5), axes)
    return g, g.laplace(pin=pin)


def stretched_axis(n, length=1.0, center=0.5, ratio=1.0):
    if ratio == 1.0:
        return np.full(n, length / n)
    k = np.arange(n) + 0.5
    w = ratio ** np.abs(k - center / length * n)
    return w * (length / w.sum())
