"""Shape gradient of the pressure drag through the pressure solve on the GPU.

J = sum over cut cells of (theta_x[i-1] - theta_x[i]) dy dz p (the pressure term of the
reference LoadsCalculator::calc), p = A(c)^-1 b(c) for a right-hand side assembled from a
predicted velocity field, c the Bezier control points.  The GPU gradient (adjoint solve on A^T,
geometric derivatives per level-set corner, chained to the control points through the closest
points) is checked against
  - central differences of J itself (h = 1e-6 on lattice control points, every evaluation a
    GPU geometry rebuild and a tightly converged FP64 solve) at 128^3;
  - the CPU reference gradient (oracle geometry, direct block solves, Lagrangian differences of
    whole oracle rebuilds) at 64^3.
A step that changes any cell type straddles a discontinuity of the discrete geometry and is
retried with a ten times smaller step."""
import numpy as np
import pytest

import oracle_py as O
import shape_ref as R

pytestmark = pytest.mark.gpu

CENTER, AXES = (0.5, 0.5, 0.5), (0.25, 0.15, 0.15)


def case(lsx, n, seed=0):
    ctrl = lsx.bezier_perturb(lsx.bezier_ellipsoid(CENTER, AXES), 0.01, seed=seed)
    w = [np.full(n, 1.0 / n)] * 3
    rng = np.random.default_rng(4)
    nu, nv, nw = (n - 1) * n * n, n * (n - 1) * n, n * n * (n - 1)
    flow = dict(u=1.0 + rng.uniform(-0.05, 0.05, nu), v=rng.uniform(-0.05, 0.05, nv), w=rng.uniform(-0.05, 0.05, nw))
    copies = R.lattice_copies(lsx.BEZIER_FACES)
    keys = sorted(copies)
    sel = [(keys[k], int(c)) for k, c in zip(rng.choice(len(keys), 8, replace=False), rng.integers(0, 3, 8))]
    return ctrl, w, flow, copies, sel


def solve_drag(lsx, w, ctrl, flow, rtol=1e-13):
    geom = lsx.Geometry(*w).build("bezier", CENTER, AXES, ctrl=ctrl)
    S = lsx.solver("fgmres", "fp64-vcycle")
    n = geom.n
    S.set_geometry(geom, np.zeros(n), np.zeros(n), lsx.mixed_task(1.0, max_it=3000))
    b = geom.pressure_rhs(**flow)
    S.eps = rtol * np.linalg.norm(b)
    p = np.zeros(n)
    S._rhs, S._sol = b, p
    S.solve()
    J, dJdp = geom.drag(p)
    return geom, S, p, J, dJdp


def gpu_gradient(lsx, w, ctrl, flow):
    geom, S, p, J, dJdp = solve_drag(lsx, w, ctrl, flow)
    S.eps = 1e-13 * np.linalg.norm(dJdp)
    lam, it, _ = S.solve_adjoint(dJdp)
    gphi, g288 = geom.shape_gradient(p, lam, lsx.OBJ_DRAG_X, flow=flow)
    return J, p, lam, gphi, g288, it


def test_drag_matches_oracle(lsx):
    ctrl, w, flow, _, _ = case(lsx, 48)
    geom = lsx.Geometry(*w).build("bezier", CENTER, AXES, ctrl=ctrl)
    g_ref = O.Geometry(*w).build_bezier(ctrl, CENTER, AXES)
    p = np.random.default_rng(2).standard_normal(geom.n)
    J, dJdp = geom.drag(p)
    J_ref, dJdp_ref = R.drag(g_ref, g_ref.fields(), p)
    print(f"J={J:.15e} J_ref={J_ref:.15e}", flush=True)
    assert np.array_equal(dJdp, dJdp_ref)
    assert abs(J - J_ref) <= 1e-13 * np.abs(dJdp * p).sum()


def test_shape_gradient_vs_finite_differences_128(lsx):
    ctrl, w, flow, copies, sel = case(lsx, 128)
    J, p, lam, gphi, g288, it = gpu_gradient(lsx, w, ctrl, flow)
    print(f"J={J:.12e} adjoint it={it} corners with dF/dphi != 0: {int((gphi != 0).sum())}", flush=True)
    rows = []
    for key, comp in sel:
        for h in (1e-6, 1e-7):
            Jp = solve_drag(lsx, w, R.lattice_perturb(ctrl, copies, key, comp, h), flow)
            Jm = solve_drag(lsx, w, R.lattice_perturb(ctrl, copies, key, comp, -h), flow)
            smooth = np.array_equal(Jp[0].fields()["type"], Jm[0].fields()["type"])
            if smooth:
                break
        fd = (Jp[3] - Jm[3]) / (2 * h)
        ga = R.lattice_grad(g288, copies, key, comp)
        print(f"lattice={key} comp={comp} h={h:g} fd={fd:.10e} adjoint={ga:.10e} "
              f"rel={abs(ga - fd) / max(abs(fd), 1e-300):.2e} types_equal={smooth}", flush=True)
        if smooth:
            rows.append((fd, ga))
    assert len(rows) >= 6
    fd = np.array([r[0] for r in rows])
    ga = np.array([r[1] for r in rows])
    rel = np.linalg.norm(ga - fd) / np.linalg.norm(fd)
    print(f"relative error over {len(rows)} control values: {rel:.3e}", flush=True)
    assert rel <= 1e-5


def test_shape_gradient_vs_cpu_reference_64(lsx):
    ctrl, w, flow, copies, sel = case(lsx, 64)
    J, p, lam, gphi, g288, it = gpu_gradient(lsx, w, ctrl, flow)
    J_ref, p_ref, lam_ref, ref = R.reference_gradient(w, ctrl, CENTER, AXES, flow, sel, lsx.BEZIER_FACES)
    act = np.abs(p_ref) > 0
    print(f"J={J:.15e} J_ref={J_ref:.15e} p rel={np.linalg.norm(p - p_ref) / np.linalg.norm(p_ref):.2e} "
          f"lambda rel={np.linalg.norm(lam - lam_ref) / np.linalg.norm(lam_ref):.2e}", flush=True)
    assert abs(J - J_ref) <= 1e-10 * abs(J_ref)
    rows = []
    for key, comp in sel:
        g_ref, smooth = ref[(key, comp)]
        ga = R.lattice_grad(g288, copies, key, comp)
        print(f"lattice={key} comp={comp} cpu={g_ref:.10e} gpu={ga:.10e} "
              f"rel={abs(ga - g_ref) / max(abs(g_ref), 1e-300):.2e} types_equal={smooth}", flush=True)
        if smooth:
            rows.append((g_ref, ga))
    assert len(rows) >= 6
    a = np.array(rows)
    assert np.linalg.norm(a[:, 1] - a[:, 0]) <= 1e-5 * np.linalg.norm(a[:, 0])


def test_tape_forward_backward(lsx):
    """The forward/backward pair gives the separate calls' results, also when the solver's
    operator was replaced between the two (the backward call then rebuilds the hierarchy)."""
    ctrl, w, flow, copies, _ = case(lsx, 64)
    J, p, lam, gphi, g288, it = gpu_gradient(lsx, w, ctrl, flow)
    geom = lsx.Geometry(*w).build("bezier", CENTER, AXES, ctrl=ctrl)
    other = lsx.Geometry(*w).build("ellipsoid", CENTER, (0.2, 0.2, 0.1))
    S = lsx.solver("fgmres", "fp64-vcycle")
    n = geom.n
    S.set_geometry(other, np.zeros(n), np.zeros(n), lsx.mixed_task(1.0, max_it=3000))
    b = geom.pressure_rhs(**flow)
    S.eps = 1e-13 * np.linalg.norm(b)
    p2, it_f, tape = S.adjoint_forward(geom, b)
    J2, dJdp = geom.drag(p2)
    S.use_geometry(other)
    S.eps = 1e-13 * np.linalg.norm(dJdp)
    dJdb, g2, it_b = tape.backward(dJdp, lsx.OBJ_DRAG_X, flow=flow)
    print(f"J={J:.15e} J_tape={J2:.15e} forward it={it_f} backward it={it_b}", flush=True)
    assert abs(J2 - J) <= 1e-11 * abs(J)
    assert np.linalg.norm(dJdb - lam) <= 1e-9 * np.linalg.norm(lam)
    assert np.linalg.norm(g2 - g288) <= 1e-7 * np.linalg.norm(g288)
    tape.close()


def test_oscillation_fix_on_a_thin_body(lsx):
    """A pancake thinner than a cell: many level-set nodes inside have both z-neighbours outside,
    so the ordered oscillation fix changes them, and the GPU's repeated parallel passes must
    reproduce the oracle's sequential sweep bit for bit."""
    n = 48
    w = [np.full(n, 1.0 / n)] * 3
    g_ref = O.Geometry(*w).build_body("ellipsoid", CENTER, (0.3, 0.25, 0.012))
    g_gpu = lsx.Geometry(*w).build("ellipsoid", CENTER, (0.3, 0.25, 0.012))
    fr, fg = g_ref.fields(), g_gpu.fields()
    phi_ref = g_ref.level_set()["phi"]
    # the raw level set of the ellipsoid at the upper cell corners, for the count of changed nodes
    c = (np.arange(n) + 1.0) / n
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    s = min(0.3, 0.25, 0.012)
    raw = s * (1 - np.sqrt(((x - 0.5) / 0.3) ** 2 + ((y - 0.5) / 0.25) ** 2 + ((z - 0.5) / 0.012) ** 2))
    inside = np.zeros((n, n, n), bool)
    inside[:, 1:n - 1, 1:n - 1] = True
    changed = int((np.abs(raw.ravel() - fg["phi"]) > 1e-12)[inside.ravel()].sum())
    print(f"nodes changed by the oscillation fix: {changed}", flush=True)
    assert changed > 10
    assert np.array_equal(fg["phi"], phi_ref)
    assert np.array_equal(fg["type"], fr["type"])
    for k in ("thx", "thy", "thz", "vof"):
        assert np.array_equal(fg[k], fr[k]), k
