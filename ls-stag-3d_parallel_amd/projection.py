"""Time-stepped pressure projection on one GPU: the projection half of the reference's time
loop (TimeSteppingMethodInterface, time_stepping_method_interface.h:180-215: right-hand side
from the predicted velocity, pressure-correction solve, velocity and pressure correction),
repeated over steps with the operator and its multigrid hierarchy built once and reused.

The momentum predictor is outside this library.  The predicted velocity of step k is a seeded
synthetic one: the corrected velocity of step k-1 plus dt times a uniform random forcing of
amplitude `forcing` (seed, step) on every face, so each step's right-hand side carries a fresh
divergence for the solve to remove.

Per step: the mixed-precision solve (time, iterations), the same right-hand side solved by the
FP64 V-cycle (time, iterations, pressure difference), and the divergence left after the
correction, |D u + flux| over the active cells with the step's own outflow flux, which equals
the solve's residual (and, for reference, with the outflow flux re-read from the corrected
velocity, as the next step's right-hand side sees it).  After the
last step: the pressure drag of the final pressure and one adjoint solve for its gradient."""
import time

import numpy as np

from . import lsx


def sphere_domain(dims=(384, 256, 256), length=(1.5, 1.0, 1.0)):
    return [np.full(d, l / d) for d, l in zip(dims, length)]


def run(dims=(384, 256, 256), length=(1.5, 1.0, 1.0), center=(0.5, 0.5, 0.5), radius=0.1, re=300.0, dt=0.002,
        steps=20, forcing=0.5, seed=0, graphs=True, compare_fp64=True, adjoint=True, log=None):
    """Run `steps` projection steps; returns a dict of per-step records and the final adjoint."""
    widths = sphere_domain(dims, length)
    geom = lsx.Geometry(*widths).build("sphere", center, (radius,))
    n = geom.n
    nu_f, nv_f, nw_f = geom.face_shapes()
    task = lsx.mixed_task(1.0, restart=30, max_it=1000)
    t0 = time.perf_counter()
    S = lsx.solver("fgmres", "mixed", graphs=graphs)
    S.set_geometry(geom, np.zeros(n), np.zeros(n), task)
    setup_mixed = time.perf_counter() - t0
    S64 = None
    if compare_fp64:
        t0 = time.perf_counter()
        S64 = lsx.solver("fgmres", "fp64-vcycle", graphs=graphs)
        S64.set_geometry(geom, np.zeros(n), np.zeros(n), task)
        setup_fp64 = time.perf_counter() - t0
    act = S.operator_diagonal() < 0
    # the solves read b and write dp in device memory; host copies are outside their timing
    db, ddp = lsx.DeviceVector(n), lsx.DeviceVector(n)
    u = np.ones(nu_f)
    v = np.zeros(nv_f)
    w = np.zeros(nw_f)
    p = np.zeros(n)
    rng = np.random.default_rng(seed)
    records = []
    for k in range(steps):
        us = u + dt * rng.uniform(-forcing, forcing, nu_f)
        vs = v + dt * rng.uniform(-forcing, forcing, nv_f)
        ws = w + dt * rng.uniform(-forcing, forcing, nw_f)
        b = geom.pressure_rhs(us, vs, ws)
        nb = float(np.linalg.norm(b))
        rec = {"step": k, "rhs_norm": nb}
        db.upload(b)
        ddp.zero()
        S.eps = 1e-8 * nb
        t0 = time.perf_counter()
        rec["iterations"] = S.solve_device(db, ddp)
        rec["solve_s"] = time.perf_counter() - t0
        dp = ddp.download()
        if S64 is not None:
            ddp.zero()
            S64.eps = 1e-8 * nb
            t0 = time.perf_counter()
            rec["fp64_iterations"] = S64.solve_device(db, ddp)
            rec["fp64_solve_s"] = time.perf_counter() - t0
            dp64 = ddp.download()
            rec["dp_rel_diff_vs_fp64"] = float(np.linalg.norm((dp - dp64)[act]) / max(np.linalg.norm(dp64[act]), 1e-300))
        u, v, w = geom.correct(dp, us, vs, ws, p=p, coef=1.0)
        # D u + flux with the step's own outflow flux (the last interior u face of u*): the
        # continuity residual the solve removed, equal to b - A dp
        div = geom.pressure_rhs(u, v, w, u_out=us)
        rec["divergence_after"] = float(np.linalg.norm(div[act]))
        # with the outflow flux re-read from the corrected velocity (the next step's flux)
        div_next = geom.pressure_rhs(u, v, w)
        rec["divergence_after_outflow_updated"] = float(np.linalg.norm(div_next[act]))
        rec["divergence_before"] = float(np.linalg.norm(b[act]))
        records.append(rec)
        if log:
            log(rec)
    out = {"dims": list(dims), "cells": n, "steps": steps, "dt": dt, "re": re, "radius": radius,
           "predictor": f"synthetic: previous corrected velocity + dt * U(-{forcing}, {forcing}) forcing (seed {seed})",
           "setup_mixed_s": setup_mixed, "records": records, "graphs": graphs}
    if S64 is not None:
        out["setup_fp64_s"] = setup_fp64
    if adjoint:
        J, dJdp = geom.drag(p)
        S.eps = 1e-8 * float(np.linalg.norm(dJdp))
        t0 = time.perf_counter()
        lam, it, nr = S.solve_adjoint(dJdp)
        out["adjoint"] = {"drag": J, "iterations": it, "solve_s": time.perf_counter() - t0,
                          "residual_rel": nr / max(float(np.linalg.norm(dJdp)), 1e-300)}
    return out


def summary(out):
    r = out["records"]
    s = {"mean_solve_ms": 1e3 * float(np.mean([x["solve_s"] for x in r])),
         "iterations": [x["iterations"] for x in r],
         "max_divergence_after": max(x["divergence_after"] for x in r)}
    if "fp64_solve_s" in r[0]:
        s["mean_fp64_solve_ms"] = 1e3 * float(np.mean([x["fp64_solve_s"] for x in r]))
        s["fp64_iterations"] = [x["fp64_iterations"] for x in r]
        s["max_dp_rel_diff_vs_fp64"] = max(x["dp_rel_diff_vs_fp64"] for x in r)
    return s
