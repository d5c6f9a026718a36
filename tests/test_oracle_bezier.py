"""Pins the oracle's six-patch bicubic Bezier body (the pressure benchmark's body; the reference
has no 3D Bezier surface) against independent evaluations:
  * the control net is closed: copies of a shared cube edge or corner point are identical;
  * the fitted patches interpolate the ellipsoid's cube-sphere points at u, v in {0,1/3,2/3,1}
    (a numpy Bernstein evaluation of the control net, written here, reproduces them);
  * a sphere fit stays within its measured interpolation error of r - |x - c|;
  * |phi| equals the brute-force distance to a dense sampling of the six patches (numpy), and
    its sign is + inside, - outside.
"""
import numpy as np
import pytest

import oracle_py as O

FACES = [(0, 1.0, 1, 2), (0, -1.0, 2, 1), (1, 1.0, 2, 0), (1, -1.0, 0, 2), (2, 1.0, 0, 1), (2, -1.0, 1, 0)]


def bernstein(t):
    t = np.asarray(t, float)[..., None]
    k = np.arange(4)
    binom = np.array([1, 3, 3, 1])
    return binom * t ** k * (1 - t) ** (3 - k)


def patch_points(ctrl, f, us, vs):
    bu, bv = bernstein(us), bernstein(vs)           # (nu, 4), (nv, 4)
    return np.einsum("ai,cj,ijk->ack", bu, bv, ctrl[f])


def cube_sphere(center, axes, f, u, v):
    a, s, bu, bv = FACES[f]
    q = np.zeros(3)
    q[a], q[bu], q[bv] = s, 2 * u - 1, 2 * v - 1
    return np.asarray(center) + np.asarray(axes) * q / np.linalg.norm(q)


CENTER, AXES = (0.5, 0.5, 0.5), (0.25, 0.15, 0.15)


def test_control_net_is_closed():
    ctrl = O.bezier_ellipsoid(CENTER, AXES)
    seen = {}
    for f, (a, s, bu, bv) in enumerate(FACES):
        for i in range(4):
            for j in range(4):
                l = [0, 0, 0]
                l[a], l[bu], l[bv] = (3 if s > 0 else 0), i, j
                key = tuple(l)
                if key in seen:
                    assert np.array_equal(seen[key], ctrl[f, i, j])
                seen[key] = ctrl[f, i, j]
    assert len(seen) == 56


def test_patches_interpolate_the_ellipsoid_nodes():
    ctrl = O.bezier_ellipsoid(CENTER, AXES)
    nodes = np.array([0, 1 / 3, 2 / 3, 1])
    for f in range(6):
        pts = patch_points(ctrl, f, nodes, nodes)
        for ia, u in enumerate(nodes):
            for ic, v in enumerate(nodes):
                assert np.allclose(pts[ia, ic], cube_sphere(CENTER, AXES, f, u, v), atol=1e-14)
    # the oracle's distance to those nodes is zero
    pts = np.array([cube_sphere(CENTER, AXES, f, u, v) for f in range(6) for u in nodes for v in nodes])
    assert np.abs(O.bezier_phi(ctrl, CENTER, AXES, pts)).max() < 1e-12


def dense_surface(ctrl, m=160):
    t = np.linspace(0, 1, m)
    return np.concatenate([patch_points(ctrl, f, t, t).reshape(-1, 3) for f in range(6)])


def test_sphere_fit_error_and_level_set():
    r = 0.2
    ctrl = O.bezier_ellipsoid((0.5, 0.5, 0.5), (r, r, r))
    surf = dense_surface(ctrl)
    dev = np.abs(np.linalg.norm(surf - 0.5, axis=1) - r).max()
    assert dev < 0.05 * r                          # node interpolation of the cube-sphere map
    # near the surface the two level sets differ by at most the Hausdorff distance of the surfaces;
    # sample directions, so the sphere-side term of that distance is measured too
    rng = np.random.default_rng(1)
    dirs = rng.normal(size=(400, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    on_sphere = 0.5 + r * dirs
    phi_on = O.bezier_phi(ctrl, (0.5, 0.5, 0.5), (r, r, r), on_sphere)
    haus = max(dev, np.abs(phi_on).max())
    assert haus < 0.05 * r
    pts = 0.5 + dirs * (r + rng.uniform(-0.03, 0.03, (400, 1)))
    phi = O.bezier_phi(ctrl, (0.5, 0.5, 0.5), (r, r, r), pts)
    exact = r - np.linalg.norm(pts - 0.5, axis=1)
    assert np.abs(phi - exact).max() <= haus + 1e-9


@pytest.mark.parametrize("perturb", [0.0, 0.01])
def test_distance_matches_brute_force(perturb):
    ctrl = O.bezier_ellipsoid(CENTER, AXES)
    if perturb:
        # perturb the 56 distinct control points, keeping shared copies equal
        rng = np.random.default_rng(7)
        keys = {}
        for f, (a, s, bu, bv) in enumerate(FACES):
            for i in range(4):
                for j in range(4):
                    l = [0, 0, 0]
                    l[a], l[bu], l[bv] = (3 if s > 0 else 0), i, j
                    keys.setdefault(tuple(l), []).append((f, i, j))
        for key in sorted(keys):
            dlt = rng.uniform(-perturb, perturb, 3)
            for (f, i, j) in keys[key]:
                ctrl[f, i, j] += dlt
    m = 240
    grid = np.stack([patch_points(ctrl, f, np.linspace(0, 1, m), np.linspace(0, 1, m)) for f in range(6)])
    surf = grid.reshape(-1, 3)
    # largest distance from a surface point to the nearest sample: half the diagonal of a grid cell
    du = np.linalg.norm(np.diff(grid, axis=1), axis=-1).max()
    dv = np.linalg.norm(np.diff(grid, axis=2), axis=-1).max()
    delta = 0.5 * np.hypot(du, dv)
    rng = np.random.default_rng(2)
    # points within a few cells of the surface, where the cut-cell geometry reads phi
    base = surf[rng.integers(0, len(surf), 300)]
    pts = base + rng.normal(0, 0.01, (300, 3))
    phi = O.bezier_phi(ctrl, CENTER, AXES, pts)
    brute = np.array([np.sqrt(((surf - p) ** 2).sum(1)).min() for p in pts])
    # the sampled minimum overestimates the true distance d by at most sqrt(d^2 + delta^2) - d
    lower = np.sqrt(np.maximum(brute ** 2 - delta ** 2, 0.0))
    assert np.all(np.abs(phi) <= brute + 1e-12)    # Newton never ends farther than the samples
    assert np.all(np.abs(phi) >= lower - 1e-12)
    inside = O.bezier_phi(ctrl, CENTER, AXES, np.array([CENTER]))
    outside = O.bezier_phi(ctrl, CENTER, AXES, np.array([[0.05, 0.05, 0.05]]))
    assert inside[0] > 0 and outside[0] < 0
