"""The library's six-patch Bezier fit of an ellipsoid (host code, no GPU): a closed control net
whose surface stays closer to the ellipsoid than the node-interpolating fit of the oracle, and
a seeded perturbation that moves shared copies together."""
import numpy as np

import oracle_py as O
from test_oracle_bezier import FACES, patch_points

CENTER, AXES = (0.5, 0.5, 0.5), (0.25, 0.15, 0.15)


def max_deviation(ctrl, center, axes, m=120):
    t = np.linspace(0, 1, m)
    pts = np.concatenate([patch_points(ctrl, f, t, t).reshape(-1, 3) for f in range(6)])
    q = (pts - np.asarray(center)) / np.asarray(axes)
    # radial deviation in units of the smallest semi-axis
    return np.abs(np.linalg.norm(q, axis=1) - 1.0).max() * min(axes)


def lattice_copies():
    copies = {}
    for f, (a, s, bu, bv) in enumerate(FACES):
        for i in range(4):
            for j in range(4):
                l = [0, 0, 0]
                l[a], l[bu], l[bv] = (3 if s > 0 else 0), i, j
                copies.setdefault(tuple(l), []).append((f, i, j))
    return copies


def test_fit_is_closed_and_closer_than_node_interpolation(lsx):
    ctrl = lsx.bezier_ellipsoid(CENTER, AXES)
    for key, cps in lattice_copies().items():
        for (f, i, j) in cps[1:]:
            assert np.array_equal(ctrl[f, i, j], ctrl[cps[0]])
    dev = max_deviation(ctrl, CENTER, AXES)
    dev_nodes = max_deviation(O.bezier_ellipsoid(CENTER, AXES), CENTER, AXES)
    print(f"least-squares fit {dev:.3e}, node interpolation {dev_nodes:.3e}")
    assert dev < 0.5 * dev_nodes and dev < 0.02 * min(AXES)


def test_perturbation_keeps_the_net_closed(lsx):
    ctrl = lsx.bezier_perturb(lsx.bezier_ellipsoid(CENTER, AXES), 0.01, seed=3)
    base = lsx.bezier_ellipsoid(CENTER, AXES)
    moved = np.abs(ctrl - base)
    assert moved.max() <= 0.01 and moved.max() > 0.005
    for key, cps in lattice_copies().items():
        for (f, i, j) in cps[1:]:
            assert np.array_equal(ctrl[f, i, j], ctrl[cps[0]])
