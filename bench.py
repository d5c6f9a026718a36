#!/usr/bin/env python3
"""Pressure-solve throughput on one MI355X.

A step is one pressure Poisson solve of the cut-cell operator to relative residual 1e-8
(absolute eps = 1e-8 ||b||, x0 = 0), FGMRES(30) (or BiCGStab) preconditioned by geometric
multigrid.  Workload: 256^3 cells on the unit cube, per-axis geometric stretching ratio 1.02
towards the body centre, an ellipsoidal body (semi-axes 0.25/0.15/0.15) at the centre, the
reference's pinned-outlet pressure boundary conditions, a seeded unit-variance Gaussian
right-hand side on fluid cells with its mean removed.  Geometry, operator and multigrid
hierarchy are built on the GPU before the timed region; b and x live in HBM.

Prints one JSON line (metric, value = solves/s, roofline of the dominant kernel measured with
HIP events on the solver stream during the timed region, CPU baseline of the oracle).
"""
import argparse
import ctypes as C
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "ls-stag-3d_parallel_amd")
HBM_PEAK_GBS = 8000.0
METRIC = "pressure solves/sec at 256³ to rel-res 1e-8 (mixed vs FP64 GPU); % HBM peak"


def load_lsx():
    name = "lsx_pkg"
    if name in sys.modules:
        return sys.modules[name].lsx
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod.lsx


class Hip:
    """Minimal HIP runtime access for device-resident benchmark inputs."""

    def __init__(self):
        self.l = C.CDLL("libamdhip64.so")
        self.l.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.l.hipFree.argtypes = [C.c_void_p]
        self.l.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.l.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]

    def check(self, st, what):
        if st != 0:
            raise RuntimeError(f"{what} failed with HIP status {st}")

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.l.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        return p

    def h2d(self, dst, arr):
        self.check(self.l.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, 1), "hipMemcpy H2D")

    def d2h(self, arr, src):
        self.check(self.l.hipMemcpy(arr.ctypes.data, src, arr.nbytes, 2), "hipMemcpy D2H")

    def zero(self, p, nbytes):
        self.check(self.l.hipMemset(p, 0, nbytes), "hipMemset")

    def sync(self):
        self.check(self.l.hipDeviceSynchronize(), "hipDeviceSynchronize")


def workload(n, ratio=1.02, center=(0.5, 0.5, 0.5), axes=(0.25, 0.15, 0.15)):
    lsx = load_lsx()
    widths = [lsx.stretched_axis(n, 1.0, center[q], ratio) for q in range(3)]
    return widths, center, axes


def cpu_baseline(sample_n, gpu_it, gpu_n, restart, krylov):
    """The CPU oracle timed on this host on the same workload at sample_n^3 (solve only), scaled to
    one gpu_n^3 solve by the cell ratio and the iteration-count ratio."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O
    lsx = load_lsx()
    widths, center, axes = workload(sample_n)
    g = O.Geometry(*widths)
    g.build_body("ellipsoid", center, axes)
    A = g.laplace()
    b = lsx.random_divergence_rhs(A.to_scipy().diagonal(), g.dims)
    eps = 1e-8 * float(np.linalg.norm(b))
    S = O.Solver(A, kind=krylov, eps=eps, restart=restart, max_it=1000)
    t0 = time.perf_counter()
    x, it, nr = S.solve(b)
    t = time.perf_counter() - t0
    scale = (gpu_n / sample_n) ** 3 * (gpu_it / max(it, 1))
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": 1.0 / (t * scale), "unit": "solves/s", "cores": cores, "kind": "port",
        "sample": (f"oracle {krylov.upper()} solve at {sample_n}^3 ({it} iterations, {t:.2f} s, "
                   f"{cores} OpenMP threads), scaled to {gpu_n}^3 by cells x{(gpu_n / sample_n) ** 3:.0f} "
                   f"and iterations {gpu_it}/{it}"),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64"], default="fp64")
    ap.add_argument("--krylov", choices=["fgmres", "bicgstab"], default="fgmres")
    ap.add_argument("--restart", type=int, default=30)
    ap.add_argument("--cpu-sample-grid", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the solve runs on a single "
              "MI355X; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    lsx = load_lsx()
    hip = Hip()
    n = args.grid
    widths, center, axes = workload(n)
    t_setup = time.perf_counter()
    geom = lsx.Geometry(*widths)
    geom.build("ellipsoid", center, axes)
    dims = geom.dims
    S = lsx.MGFGMRES() if args.krylov == "fgmres" else lsx.MGBiCGStab()
    S.set_geometry(geom, None, None, lsx.Task(eps=1.0, MAX_ITER=1000, FGMRES_RESTART_NUM=args.restart))
    b = lsx.random_divergence_rhs(S.operator_diagonal(), dims)
    nb = float(np.linalg.norm(b))
    eps = 1e-8 * nb
    lsx._check(lsx.lib().lsx_set_eps(S.h, eps))
    ncell = geom.n
    db = hip.malloc(ncell * 8)
    dx = hip.malloc(ncell * 8)
    hip.h2d(db, b)
    hip.sync()
    setup_s = time.perf_counter() - t_setup
    L = lsx.lib()
    it = C.c_int()
    nr = C.c_double()

    def step():
        hip.zero(dx, ncell * 8)
        st = L.lsx_solve(S.h, db, dx, lsx.MEM_DEVICE, C.byref(it), C.byref(nr))
        if st != 0:
            raise RuntimeError(L.lsx_last_error().decode())

    for _ in range(args.warmup):
        step()
    hip.sync()
    S.kernel_timing(True)
    hip.sync()
    t0 = time.perf_counter()
    its = []
    for _ in range(args.steps):
        step()
        its.append(it.value)
    hip.sync()
    elapsed = time.perf_counter() - t0
    stats = {k: S.kernel_stats(k) for k in lsx.KERNEL_NAMES}
    S.kernel_timing(False)

    # true residual of the last solve, checked on the device operator
    x = np.zeros(ncell)
    hip.d2h(x, dx)
    r = b - S.apply_operator(x, 0)
    rel_res = float(np.linalg.norm(r) / nb)

    dom = max(stats, key=lambda k: stats[k][1])
    launches, tot_ms, bytes_per = stats[dom]
    avg_s = tot_ms / launches / 1e3
    achieved = bytes_per / avg_s / 1e9
    kernels = {lsx.KERNEL_NAMES[k]: {"launches": v[0], "avg_us": (v[1] / v[0] * 1e3) if v[0] else None,
                                     "GBs": (v[2] / (v[1] / v[0] / 1e3) / 1e9) if v[0] else None,
                                     "share_of_step": v[1] / 1e3 / elapsed}
               for k, v in stats.items() if v[0]}
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get(f"{lsx.KERNEL_NAMES[dom]}@{n}")
        except (ValueError, OSError):
            traffic = None

    out = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{n}^3 stretched grid (ratio 1.02 to centre), ellipsoid body 0.25/0.15/0.15 "
                         f"(level set of the ellipsoid the Bezier body approximates), "
                         f"{args.krylov.upper()}({args.restart}) + reference multigrid cascade (ADLJ), FP64, "
                         f"rel tol 1e-8"),
            "precision": args.precision, "krylov": args.krylov, "cells": ncell,
            "iterations": its, "true_rel_residual": rel_res, "setup_s": setup_s,
            "mg_levels": S.num_levels(),
        },
        "roofline": {"bound": "hbm", "kernel": lsx.KERNEL_NAMES[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": bytes_per, "avg_launch_us": avg_s * 1e6},
        "kernels": kernels,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(args.cpu_sample_grid, its[-1], n, args.restart, args.krylov)
        except Exception as e:  # the GPU measurement stands on its own
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
