"""Pins the oracle's pressure right-hand side (b = D u* + flux, PoissonDiscreteEquation::
recalcRightMember) against closed forms of the discrete divergence:
  * a uniform stream u = U_W through an empty box with an outflow east boundary has b = 0
    exactly (every interior face cancels; inflow and outflow fluxes cancel the boundary faces);
  * for any face field, the sum of b over the cells telescopes to the boundary fluxes
    (sum over x-rows of (u_last - U_W) dy dz, walls closed), up to rounding;
  * a single nonzero face puts +/- theta*area on exactly its two cells;
  * the pinned cells get b = 0.
"""
import numpy as np
import pytest

import oracle_py as O
import problems as P


def test_uniform_stream_is_divergence_free():
    g = O.uniform_geometry(12, 10, 8)
    nu, nv, nw = O.face_shapes(g.dims)
    b = O.pressure_rhs(g, np.full(nu, 1.0), np.zeros(nv), np.zeros(nw), U_W=1.0)
    assert np.all(b == 0.0)


def test_rhs_sum_telescopes_to_boundary_flux():
    N, M, K = 16, 12, 10
    w = [P.stretched_axis(N, 1.0, 0.5, 1.05), P.stretched_axis(M, 1.0, 0.5, 1.03), P.stretched_axis(K, 1.0, 0.5, 1.0)]
    g = O.Geometry(*w)
    rng = np.random.default_rng(1)
    nu, nv, nw = O.face_shapes(g.dims)
    u, v, wz = rng.standard_normal(nu), rng.standard_normal(nv), rng.standard_normal(nw)
    uw = rng.standard_normal(M * K)
    b = O.pressure_rhs(g, u, v, wz, U_W=uw, pin=False)
    dy, dz = w[1], w[2]
    rows = np.arange(M * K)
    y, z = rows % M, rows // M
    u_last = u.reshape(M * K, N - 1)[:, -1]
    expect = np.sum((u_last - uw) * dy[y] * dz[z])
    assert abs(b.sum() - expect) <= 1e-12 * np.abs(b).sum()


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_single_face_stencil(axis):
    g, A = P.sphere(16)
    f = g.fields()
    N, M, K = g.dims
    nu, nv, nw = O.face_shapes(g.dims)
    arrs = [np.zeros(nu), np.zeros(nv), np.zeros(nw)]
    # a face between two fluid cells away from the body and the walls
    x, y, zc = 13, 5, 4
    i = x + N * (y + M * zc)
    if axis == 0:
        arrs[0][(y + M * zc) * (N - 1) + x] = 1.0
        j, th, area = i + 1, f["thx"][i], g.dy[y] * g.dz[zc]
    elif axis == 1:
        arrs[1][zc * N * (M - 1) + y * N + x] = 1.0
        j, th, area = i + N, f["thy"][i], g.dx[x] * g.dz[zc]
    else:
        arrs[2][i] = 1.0
        j, th, area = i + N * M, f["thz"][i], g.dx[x] * g.dy[y]
    b = O.pressure_rhs(g, *arrs, U_W=0.0, outflow=False, U_E=0.0)
    assert th == 1.0
    nz = np.nonzero(b)[0]
    assert list(nz) == [i, j]
    assert b[i] == th * area and b[j] == -th * area


def test_pins_are_zero_and_flux_on_boundaries():
    g, A = P.sphere(16)
    N, M, K = g.dims
    nu, nv, nw = O.face_shapes(g.dims)
    rng = np.random.default_rng(2)
    u = 1.0 + rng.uniform(-0.05, 0.05, nu)
    b, flux = O.pressure_rhs(g, u, rng.uniform(-0.05, 0.05, nv), rng.uniform(-0.05, 0.05, nw), return_flux=True)
    assert b[N - 1] == 0.0 and b[N * M - 1] == 0.0
    fl = flux.reshape(K, M, N)
    assert np.all(fl[:, :, 1:N - 1] == 0.0)          # static body: no interior flux
    assert np.all(fl[:, :, 0] < 0.0)                 # inflow
    east = u.reshape(K, M, N - 1)[:, :, -1] * g.dy[None, :, None][..., 0] * g.dz[:, None]
    assert np.array_equal(fl[:, :, N - 1], east)     # outflow = last interior face


def test_correction_of_constant_pressure_is_identity():
    g, A = P.sphere(16)
    nu, nv, nw = O.face_shapes(g.dims)
    rng = np.random.default_rng(3)
    us, vs, ws = rng.standard_normal(nu), rng.standard_normal(nv), rng.standard_normal(nw)
    u, v, w = O.velocity_correction(g, np.full(g.n, 2.5), us, vs, ws)
    assert np.array_equal(u, us) and np.array_equal(v, vs) and np.array_equal(w, ws)


@pytest.mark.parametrize("case", ["free", "sphere", "stretched_ellipsoid"])
def test_corrected_divergence_equals_pressure_residual(case):
    """D inv_M G = -A off the pinned rows, so after the correction the new velocity's divergence
    (D u + flux) equals the pressure residual b - A dp: the identity behind the reference's
    divergence error.  Faces with theta below 1e-5 enter A but not the gradient (their control
    volume is typed solid), which moves the identity by O(theta^2)."""
    if case == "free":
        g, A = P.body_free(12, 10, 8)
    elif case == "sphere":
        g, A = P.sphere(16)
    else:
        g, A = P.ellipsoid_stretched(16)
    N, M, K = g.dims
    nu, nv, nw = O.face_shapes(g.dims)
    rng = np.random.default_rng(4)
    us, vs, ws = 1.0 + rng.uniform(-0.05, 0.05, nu), rng.uniform(-0.05, 0.05, nv), rng.uniform(-0.05, 0.05, nw)
    b = O.pressure_rhs(g, us, vs, ws, outflow=False, U_W=1.0, U_E=1.0)
    dp = rng.standard_normal(g.n)
    u, v, w = O.velocity_correction(g, dp, us, vs, ws)
    div = O.pressure_rhs(g, u, v, w, outflow=False, U_W=1.0, U_E=1.0)
    res = b - A.matvec(dp)
    keep = np.ones(g.n, bool)
    keep[[N - 1, N * M - 1]] = False
    # the pinned cells' neighbours see the pinned value through A's columns and through G alike
    err = np.abs(div - res)[keep].max()
    print(case, err, np.abs(res).max())
    assert err <= 1e-9 * np.abs(res).max()
