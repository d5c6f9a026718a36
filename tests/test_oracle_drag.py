"""Pin the pressure-drag functional J(p) = sum over cut cells of (theta_x[i-1] - theta_x[i]) dy dz p[i]
(the pressure term of the reference LoadsCalculator::calc, loads_calculator.h:97-116, where
n_x dS = (theta^u_{i-1} - theta^u_i) dy dz) by what the mathematics fixes, on the CPU:

* p = 1: along every x row the terms telescope to theta(first face) - theta(last face) = 1 - 1,
  and cells that are not cut have equal west and east fractions, so J(1) = 0 up to rounding;
* p = x (cell centres): summation by parts along each row gives exactly
  J(x) = - sum over x faces of (1 - theta_f) (x_{i+1} - x_i) dy dz, the solid length of each row
  times its cross-section, and this converges to -V_body under refinement: the reference's
  n dS points into the body, so J is the pressure force on the body in x and a pressure growing
  along x pushes the body towards -x.

A sign flip fails the limit -V, a wrong theta index fails the telescoping, a dropped or swapped
cell width fails the identity on a grid whose three axes have different, non-uniform widths.
The GPU functional (lsx_geometry_drag) is checked against the same properties."""
import numpy as np
import pytest

import oracle_py as O
import shape_ref as SR

C = (0.45, 0.5, 0.5)
R = 0.2


def geometry(n, stretched):
    if stretched:
        w = [P_axis(n, r) for r in (1.0, 1.04, 1.07)]
    else:
        w = [np.full(n, 1.0 / n)] * 3
    g = O.Geometry(*w)
    g.build_body("sphere", C, (R,))
    return g


def P_axis(n, ratio):
    if ratio == 1.0:
        return np.full(n, 1.0 / n)
    k = np.arange(n) + 0.5
    w = ratio ** np.abs(k - 0.5 * n)
    return w / w.sum()


def centres(w):
    e = np.concatenate([[0.0], np.cumsum(w)])
    return 0.5 * (e[1:] + e[:-1])


def solid_length_identity(g, f):
    """- sum over x faces of (1 - theta_f) (x_{i+1} - x_i) dy dz."""
    N, M, K = g.dims
    tx = f["thx"].reshape(K, M, N)[:, :, :-1]  # face between cell i and i + 1
    xc = centres(g.dx)
    dxc = np.diff(xc)
    return -float(np.sum((1.0 - tx) * dxc[None, None, :] * g.dy[None, :, None] * g.dz[:, None, None]))


@pytest.mark.parametrize("stretched", [False, True])
def test_drag_of_constant_pressure_vanishes(stretched):
    g = geometry(48, stretched)
    f = g.fields()
    J, coef = SR.drag(g, f, np.ones(g.n))
    scale = float(np.abs(coef).sum())
    print(f"stretched={stretched} J(1)={J:.3e} sum|coef|={scale:.3e}")
    assert scale > 1e-3
    assert abs(J) <= 1e-13 * scale


@pytest.mark.parametrize("stretched", [False, True])
def test_drag_of_linear_pressure_is_minus_solid_length(stretched):
    g = geometry(48, stretched)
    f = g.fields()
    N, M, K = g.dims
    x = np.broadcast_to(centres(g.dx)[None, None, :], (K, M, N)).ravel()
    J, _ = SR.drag(g, f, x)
    ident = solid_length_identity(g, f)
    print(f"stretched={stretched} J(x)={J:.15e} identity={ident:.15e}")
    assert abs(J - ident) <= 1e-13 * abs(ident)


def test_drag_of_linear_pressure_converges_to_minus_volume():
    V = 4.0 / 3.0 * np.pi * R ** 3
    errs = []
    for n in (32, 64, 128):
        g = geometry(n, False)
        f = g.fields()
        N, M, K = g.dims
        x = np.broadcast_to(centres(g.dx)[None, None, :], (K, M, N)).ravel()
        J, _ = SR.drag(g, f, x)
        errs.append(abs(J + V) / V)
        print(f"n={n} J(x)={J:.6e} -V={-V:.6e} rel err {errs[-1]:.3e}")
    assert errs[2] < errs[1] < errs[0]
    assert errs[2] <= 0.01


@pytest.mark.gpu
@pytest.mark.parametrize("stretched", [False, True])
def test_gpu_drag_properties(lsx, stretched):
    """The GPU functional on the GPU-built geometry: J(1) = 0 and the solid-length identity."""
    n = 48
    w = [P_axis(n, r) for r in ((1.0, 1.04, 1.07) if stretched else (1.0, 1.0, 1.0))]
    g = lsx.Geometry(*w).build("sphere", C, (R,))
    f = g.fields()
    J1, coef = g.drag(np.ones(g.n))
    N, M, K = g.dims
    x = np.broadcast_to(centres(w[0])[None, None, :], (K, M, N)).ravel()
    Jx, _ = g.drag(x)
    ident = solid_length_identity(O.Geometry(*w), f)
    print(f"GPU stretched={stretched} J(1)={J1:.3e} J(x)={Jx:.15e} identity={ident:.15e}", flush=True)
    assert abs(J1) <= 1e-13 * float(np.abs(coef).sum())
    assert abs(Jx - ident) <= 1e-13 * abs(ident)
