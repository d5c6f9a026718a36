#!/usr/bin/env python3
"""Pressure-solve throughput on one MI355X.

A step is one pressure Poisson solve of the cut-cell operator to relative residual 1e-8
(absolute eps = 1e-8 ||b||, x0 = 0), FGMRES(30) (or BiCGStab) preconditioned by a geometric
multigrid V-cycle: mixed precision by default (FP32 values, FP16 coarse faces; x/y/z line
relaxation on the two finest levels and red-black Gauss-Seidel below, or red-black on every
level with --smoother rbgs), compared with the same V-cycle in FP64 (the plain FP64 GPU solve) and with the
reference's FP64 multigrid cascade, which is reported at an iteration cap because it does not
reach 1e-8 on this stretched grid.  Workload: 256^3 cells on the unit cube, per-axis geometric
stretching ratio 1.02 towards the body centre, a six-patch bicubic Bezier body fitted to the
ellipsoid with semi-axes 0.25/0.15/0.15 at the centre (control points moved by a seeded
U(-0.01, 0.01)), the reference's pinned-outlet pressure boundary conditions, a seeded unit-variance Gaussian
right-hand side on fluid cells with its mean removed.  Geometry, operator and multigrid
hierarchy are built on the GPU before the timed region; b and x live in HBM.

Prints one JSON line (metric, value = solves/s, roofline of the dominant kernel measured with
HIP events on the solver stream in extra solves after the timed region, CPU baseline of the
oracle).  The FP64 comparison is the same V-cycle and smoothers in FP64, both with the same
kernels' optimisations.
"""
import argparse
import ctypes as C
import importlib.util
import json
import os
import subprocess
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

    def free(self, p):
        self.check(self.l.hipFree(p), "hipFree")

    def sync(self):
        self.check(self.l.hipDeviceSynchronize(), "hipDeviceSynchronize")


def workload(n, ratio=1.02, center=(0.5, 0.5, 0.5), axes=(0.25, 0.15, 0.15)):
    lsx = load_lsx()
    widths = [lsx.stretched_axis(n, 1.0, center[q], ratio) for q in range(3)]
    return widths, center, axes


def body_ctrl(center, axes, seed=0):
    """Control points of the benchmark body: six-patch bicubic Bezier fit of the ellipsoid, every
    coordinate of its 56 distinct control points moved by a seeded U(-0.01, 0.01)."""
    lsx = load_lsx()
    return lsx.bezier_perturb(lsx.bezier_ellipsoid(center, axes), 0.01, seed)


def build_body_geometry(geom, body, center, axes):
    if body == "bezier":
        if hasattr(geom, "build_bezier"):  # oracle geometry
            return geom.build_bezier(body_ctrl(center, axes), center, axes)
        return geom.build("bezier", center, axes, ctrl=body_ctrl(center, axes))
    if hasattr(geom, "build_body"):
        return geom.build_body("ellipsoid", center, axes)
    return geom.build("ellipsoid", center, axes)


def cpu_problem(case, restart, krylov, body):
    """The oracle's (the reference FP64 cascade algorithm's) problem of one workload: C1 is the 64^3
    sphere projection step (b = D u* + flux of u* = 1 + U(-0.05, 0.05), v*, w* = U(-0.05, 0.05),
    seed 5), C2 the 128^3 sphere with a unit Gaussian right-hand side, C3@128 the 256^3 workload's
    stretched Bezier body at 128^3 (its per-iteration cost scales with the cells)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O
    lsx = load_lsx()
    if case in ("C1", "C2"):
        n = 64 if case == "C1" else 128
        g = O.Geometry(*([np.full(n, 1.0 / n)] * 3)).build_body("sphere", (0.35, 0.5, 0.5), (0.2,))
        A = g.laplace()
        if case == "C1":
            nu, nv, nw = O.face_shapes(g.dims)
            rng = np.random.default_rng(5)
            b = O.pressure_rhs(g, 1.0 + rng.uniform(-0.05, 0.05, nu), rng.uniform(-0.05, 0.05, nv),
                               rng.uniform(-0.05, 0.05, nw), U_W=1.0)
        else:
            b = lsx.random_divergence_rhs(A.to_scipy().diagonal(), g.dims)
    else:
        n = 128
        widths, center, axes = workload(n)
        g = O.Geometry(*widths)
        build_body_geometry(g, body, center, axes)
        A = g.laplace()
        b = lsx.random_divergence_rhs(A.to_scipy().diagonal(), g.dims)
    eps = 1e-8 * float(np.linalg.norm(b))
    return O.Solver(A, kind=krylov, eps=eps, restart=restart, max_it=1000), b, n


def cpu_worker(case, restart, krylov, body):
    """One timed oracle solve (setup excluded), printed as one JSON line; run in a child process
    whose OMP_NUM_THREADS / OMP_PROC_BIND set its threads."""
    S, b, n = cpu_problem(case, restart, krylov, body)
    t0 = time.perf_counter()
    x, it, nr = S.solve(b)
    t = time.perf_counter() - t0
    print(json.dumps({"case": case, "n": n, "it": it, "seconds": t, "norm_r": nr,
                      "threads": int(os.environ.get("OMP_NUM_THREADS", "0"))}), flush=True)


def c2_against_oracle(krylov, restart):
    """The C2 workload (128^3 sphere r = 0.2 at (0.35, 0.5, 0.5), seeded Gaussian right-hand side,
    FP32 5-level V-cycle with FP16 coarse faces in FP64 Krylov, rel 1e-8) on the GPU through the
    drop-in CSR boundary, against the oracle's FP64 solve of the same system (the reference
    cascade): pressure error on the active cells and divergence error ||b - A p|| / ||b|| (the
    divergence of the corrected velocity) of both."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O
    lsx = load_lsx()
    n = 128
    g = O.Geometry(*([np.full(n, 1.0 / n)] * 3)).build_body("sphere", (0.35, 0.5, 0.5), (0.2,))
    A = g.laplace()
    b = lsx.random_divergence_rhs(A.to_scipy().diagonal(), g.dims)
    nb = float(np.linalg.norm(b))
    eps = 1e-8 * nb
    S = lsx.solver(krylov, "mixed", levels=5)
    x = np.zeros(A.n)
    S.set(A, b, x, lsx.mixed_task(eps, restart=restart, max_it=300))
    S.solve()  # warm-up
    x[:] = 0.0
    t0 = time.perf_counter()
    it = S.solve()
    t_gpu = time.perf_counter() - t0
    x64, it64, _ = O.Solver(A, kind=krylov, eps=eps, restart=restart, max_it=500).solve(b)
    act = A.to_scipy().diagonal() < 0
    return {"workload": "128^3 sphere, Gaussian right-hand side, mixed 5-level V-cycle vs the oracle's FP64 cascade",
            "iterations": it, "oracle_iterations": it64, "gpu_solve_s_host_arrays": t_gpu,
            "pressure_rel_error_vs_oracle": float(np.linalg.norm((x - x64)[act]) / np.linalg.norm(x64[act])),
            "divergence_rel_error": float(np.linalg.norm(b - A.matvec(x)) / nb),
            "oracle_divergence_rel_error": float(np.linalg.norm(b - A.matvec(x64)) / nb)}


def physical_cores():
    """Physical cores (distinct socket/core pairs from lscpu) among the CPUs this process may use,
    and the machine's total."""
    mine = os.sched_getaffinity(0)
    pairs, total = set(), set()
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, timeout=30).stdout
        for line in out.splitlines():
            if line.startswith("#") or not line.strip():
                continue
            cpu, core, sock = (int(q) for q in line.split(",")[:3])
            total.add((sock, core))
            if cpu in mine:
                pairs.add((sock, core))
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return (len(pairs) or len(mine)), (len(total) or len(mine))


def cpu_baseline(gpu_n, restart, krylov, body="bezier"):
    """The CPU oracle (the reference's FP64 cascade algorithm, C++/OpenMP) timed on this host, in
    child processes pinned with OMP_PROC_BIND=close: C1 (64^3) and C2 (128^3) measured directly on
    1 thread and on every physical core this process may use; the headline value is the 256^3
    workload extrapolated from its 128^3 solve (same body and stretching) by the cell count at the
    same iteration count.  The reference algorithm does not reach 1e-8 at 256^3 at all, so that
    extrapolated rate is an upper bound on the CPU path there."""
    phys, phys_total = physical_cores()
    cap = int(os.environ.get("OMP_NUM_THREADS", phys) or phys)
    nthreads = max(1, min(phys, cap))
    runs = {}
    for case, threads in (("C1", 1), ("C1", nthreads), ("C2", 1), ("C2", nthreads), ("C3@128", nthreads)):
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-worker", case, "--restart",
                            str(restart), "--krylov", krylov, "--body", body], env=env, capture_output=True,
                           text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"cpu worker {case} failed: {r.stderr[-400:]}")
        res = json.loads(r.stdout.strip().splitlines()[-1])
        runs[f"{case}@{threads}t"] = res
        print(f"[bench] cpu oracle {case} on {threads} thread(s): {res['it']} its in {res['seconds']:.3f} s",
              file=sys.stderr, flush=True)
    c3 = runs[f"C3@128@{nthreads}t"]
    scale = (gpu_n / c3["n"]) ** 3
    measured = {k: {"grid": f"{v['n']}^3", "threads": v["threads"], "iterations": v["it"], "seconds": v["seconds"],
                    "solves_per_s": 1.0 / v["seconds"]} for k, v in runs.items()}
    return {
        "value": 1.0 / (c3["seconds"] * scale), "unit": "solves/s", "cores": nthreads, "kind": "port",
        "physical_cores_available": phys, "physical_cores_machine": phys_total, "omp_proc_bind": "close",
        "sample": (f"oracle (reference FP64 cascade, C++/OpenMP) {krylov.upper()}({restart}) solve of the {gpu_n}^3 "
                   f"workload's body and stretching at 128^3 to 1e-8: {c3['it']} iterations in {c3['seconds']:.2f} s "
                   f"on {nthreads} threads (one per physical core, OMP_PROC_BIND=close), extrapolated to {gpu_n}^3 "
                   f"by cells x{scale:.0f} at the same iteration count (an upper bound: the reference algorithm "
                   f"does not reach 1e-8 at {gpu_n}^3); C1 and C2 measured directly on 1 and {nthreads} threads "
                   f"in 'measured'"),
        "measured": measured,
    }


def run_solver(lsx, hip, geom, precision, krylov, restart, steps, warmup, max_it=1000, smooth=(2, 2), graphs=False,
               smoother="rbgs", line_levels=0, levels=0, profile_steps=0, fused_sweep=1, diag_faces=True,
               dot32=True):
    """Set up one solver on the device geometry and time `steps` solves (x0 = 0 each); then
    `profile_steps` more solves with per-kernel event timing."""
    S = lsx.solver(krylov, precision, graphs=graphs, smoother=smoother, line_levels=line_levels, levels=levels,
                   fused_sweep=fused_sweep, diag_faces=diag_faces, dot32=dot32) \
        if precision != "fp64" else lsx.solver(krylov, precision)
    task = lsx.Task(eps=1.0, MAX_ITER=max_it, FGMRES_RESTART_NUM=restart) if precision == "fp64" else \
        lsx.mixed_task(1.0, restart=restart, max_it=max_it, n_pre=smooth[0], n_post=smooth[1])
    t_setup = time.perf_counter()
    S.set_geometry(geom, None, None, task)
    hip.sync()
    setup_s = time.perf_counter() - t_setup
    b = lsx.random_divergence_rhs(S.operator_diagonal(), geom.dims)
    nb = float(np.linalg.norm(b))
    lsx._check(lsx.lib().lsx_set_eps(S.h, 1e-8 * nb))
    ncell = geom.n
    db = hip.malloc(ncell * 8)
    dx = hip.malloc(ncell * 8)
    hip.h2d(db, b)
    hip.sync()
    L = lsx.lib()
    it = C.c_int()
    nr = C.c_double()

    def step():
        hip.zero(dx, ncell * 8)
        st = L.lsx_solve(S.h, db, dx, lsx.MEM_DEVICE, C.byref(it), C.byref(nr))
        if st != 0:
            raise RuntimeError(L.lsx_last_error().decode())

    for _ in range(warmup):
        step()
    hip.sync()
    t0 = time.perf_counter()
    its = []
    for _ in range(steps):
        step()
        its.append(it.value)
    hip.sync()
    elapsed = time.perf_counter() - t0
    x = np.zeros(ncell)
    hip.d2h(x, dx)
    # per-kernel durations from HIP events on the solver stream, in separate solves after the timed
    # region (the events and their synchronisation stay out of the headline time)
    stats = {}
    if profile_steps > 0:
        S.kernel_timing(True)
        for _ in range(profile_steps):
            step()
        hip.sync()
        stats = {k: S.kernel_stats(k) for k in lsx.KERNEL_NAMES}
        S.kernel_timing(False)
    r = b - S.apply_operator(x, 0)
    hip.free(db)
    hip.free(dx)
    return dict(S=S, elapsed=elapsed, its=its, stats=stats, setup_s=setup_s,
                rel_res=float(np.linalg.norm(r) / nb), x=x, levels=S.num_levels())


def copy_calibration(gib):
    """Five device-to-device copies of `gib` GiB (torch's elementwise copy kernel): the known
    traffic against which the FETCH_SIZE / WRITE_SIZE counters of the solver kernels are scaled."""
    import torch
    n = int(gib * (1 << 30)) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "copies": 5}), flush=True)


def projection_run(steps):
    """Time-stepped projection: 384x256x256 uniform grid on [0,1.5]x[0,1]^2, sphere r = 0.1 at
    (0.5, 0.5, 0.5), Re = 300, dt = 0.002, operator and hierarchy built once, V-cycles replayed
    from captured HIP graphs, mixed solve per step with the FP64 V-cycle on the same right-hand
    side, one adjoint solve of the final drag."""
    load_lsx()
    proj = sys.modules["lsx_pkg"].projection

    def log(r):
        print(f"[projection] step {r['step']}: it {r['iterations']} ({r['solve_s'] * 1e3:.1f} ms), fp64 it "
              f"{r.get('fp64_iterations')} ({r.get('fp64_solve_s', 0) * 1e3:.1f} ms), divergence "
              f"{r['divergence_before']:.3e} -> {r['divergence_after']:.3e}", file=sys.stderr, flush=True)

    out = proj.run(steps=steps, graphs=True, log=log)
    out["summary"] = proj.summary(out)
    print(json.dumps(out), flush=True)


def momentum_run(n, steps):
    """Momentum (Helmholtz) solves u, v, w on the n^3 stretched grid of the headline workload
    without its body (the device assembly takes boxes without cut cells), assembled on the device
    (lsx_set_helmholtz), FGMRES(30) to 1e-8 |b| from x = 0 with b a seeded Gaussian, at the
    time-stepped run's Re = 300, dt = 0.002; the one-level cascade (the reference's cascade has no
    transfers for the staggered meshes' one odd dimension) against the mixed V-cycle.  (At
    Re = 30, dt = 0.02 on this stretched grid the v and w operators are indefinite from 128^3 up:
    their outflow term outweighs the mass term on the last x plane.)  One JSON line."""
    lsx = load_lsx()
    hip = Hip()
    widths, _, _ = workload(n)
    geom = lsx.Geometry(*widths).build("sphere", (-5.0, -5.0, -5.0), (0.01,))  # the body lies outside the box
    hip.sync()
    L = lsx.lib()
    out = []
    for re_, dt in ((300.0, 0.002),):
        for d in range(3):
            dims = [n, n, n]
            dims[d] -= 1
            m = dims[0] * dims[1] * dims[2]
            b = np.random.default_rng(17 + d).standard_normal(m) * 1e-3
            nb = float(np.linalg.norm(b))
            db, dx = hip.malloc(m * 8), hip.malloc(m * 8)
            hip.h2d(db, b)
            for name, make, task in (
                    ("one-level cascade", lambda: lsx.MGFGMRES(levels=1), lsx.Task(eps=1.0, MAX_ITER=500)),
                    ("mixed V-cycle", lambda: lsx.solver("fgmres", "mixed", graphs=True, smoother="line",
                                                         line_levels=2), lsx.mixed_task(1.0, max_it=500))):
                S = make()
                S.set_helmholtz(geom, d, 1.0 / re_, 1.5 / dt, None, None, task)
                lsx._check(L.lsx_set_eps(S.h, 1e-8 * nb))
                it, nr = C.c_int(), C.c_double()

                def step():
                    hip.zero(dx, m * 8)
                    st = L.lsx_solve(S.h, db, dx, lsx.MEM_DEVICE, C.byref(it), C.byref(nr))
                    if st != 0:
                        raise RuntimeError(L.lsx_last_error().decode())

                step()
                hip.sync()
                t0 = time.perf_counter()
                for _ in range(steps):
                    step()
                hip.sync()
                sec = (time.perf_counter() - t0) / steps
                x = np.zeros(m)
                hip.d2h(x, dx)
                r = b - S.apply_operator(x, 0)
                rec = {"component": "uvw"[d], "mesh": dims, "re": re_, "dt": dt, "preconditioner": name,
                       "iterations": it.value, "ms_per_solve": sec * 1e3, "solves_per_s": 1.0 / sec,
                       "true_rel_residual": float(np.linalg.norm(r) / nb)}
                print(f"[momentum] {rec}", file=sys.stderr, flush=True)
                out.append(rec)
                S.close()
            hip.free(db)
            hip.free(dx)
    print(json.dumps({"momentum": out, "grid": n, "steps": steps, "data": "synthetic"}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="only run 5 device copies of this many GiB (counter calibration) and exit")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--precision", choices=["mixed", "mixed16", "fp64-vcycle", "fp64"], default="mixed")
    ap.add_argument("--krylov", choices=["fgmres", "bicgstab"], default="fgmres")
    ap.add_argument("--restart", type=int, default=30)
    ap.add_argument("--body", choices=["bezier", "ellipsoid"], default="bezier")
    ap.add_argument("--smooth", default="2,2", help="V-cycle pre,post red-black sweeps")
    ap.add_argument("--levels", type=int, default=0,
                    help="V-cycle levels, the finest included (0 = automatic: halve to 2 cells)")
    ap.add_argument("--no-fp64-compare", action="store_true")
    ap.add_argument("--cascade-max-it", type=int, default=200,
                    help="iteration cap of the reference-cascade comparison solve (0 = skip it)")
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", dest="graphs", action="store_false",
                    help="launch every V-cycle kernel from the host instead of replaying the cycles from captured "
                         "HIP graphs (one graph per Krylov basis slot, captured on first use)")
    ap.add_argument("--smoother", choices=["rbgs", "line", "jacobi"], default="line",
                    help="V-cycle smoother: red-black Gauss-Seidel, x/y/z line relaxation or damped Jacobi")
    ap.add_argument("--line-levels", type=int, default=2,
                    help="with --smoother line: line relaxation on the finest this many levels, red-black "
                         "below (0 = every level)")
    ap.add_argument("--fused-sweep", type=int, default=1, choices=[0, 1, 2],
                    help="red-black sweeps with both colours in one launch: 1 on a finest level of >= 2^23 "
                         "cells, 2 on every level, 0 never")
    ap.add_argument("--no-diag-faces", action="store_true",
                    help="V-cycle line passes and residual restriction read 1/a_p instead of rebuilding a_p "
                         "from the faces")
    ap.add_argument("--no-c3-legs", action="store_true",
                    help="skip the BiCGStab and 6-level solves reported next to the headline")
    ap.add_argument("--no-dot32", action="store_true",
                    help="mixed FGMRES: Gram-Schmidt projections over the FP64 basis instead of its FP32 copies")
    ap.add_argument("--projection-steps", type=int, default=0,
                    help="instead of the solve benchmark: this many time steps of the 384x256x256 sphere "
                         "projection run (one JSON line of per-step results)")
    ap.add_argument("--momentum", type=int, default=0,
                    help="instead of the solve benchmark: this many timed solves of each momentum (u, v, w) "
                         "system on the grid without the body, one-level cascade against the mixed V-cycle "
                         "(one JSON line)")
    args = ap.parse_args()
    if args.cpu_worker:
        cpu_worker(args.cpu_worker, args.restart, args.krylov, args.body)
        return
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.projection_steps > 0:
        projection_run(args.projection_steps)
        return
    if args.momentum > 0:
        momentum_run(args.grid, args.momentum)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the solve runs on a single "
              "MI355X; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    lsx = load_lsx()
    hip = Hip()
    n = args.grid
    widths, center, axes = workload(n)
    t_geo = time.perf_counter()
    geom = lsx.Geometry(*widths)
    build_body_geometry(geom, args.body, center, axes)
    hip.sync()
    geo_s = time.perf_counter() - t_geo
    print(f"[bench] geometry {n}^3 built in {geo_s:.2f} s", file=sys.stderr, flush=True)

    smooth = tuple(int(q) for q in args.smooth.split(","))
    main_run = run_solver(lsx, hip, geom, args.precision, args.krylov, args.restart, args.steps, args.warmup,
                          smooth=smooth, graphs=args.graphs, smoother=args.smoother, line_levels=args.line_levels,
                          levels=args.levels, profile_steps=max(1, min(3, args.steps)),
                          fused_sweep=args.fused_sweep, diag_faces=not args.no_diag_faces,
                          dot32=not args.no_dot32)
    print(f"[bench] {args.precision}: {args.steps} solves in {main_run['elapsed']:.3f} s, its {main_run['its']}",
          file=sys.stderr, flush=True)
    fp64 = None
    if args.precision in ("mixed", "mixed16") and not args.no_fp64_compare:
        fp64 = run_solver(lsx, hip, geom, "fp64-vcycle", args.krylov, args.restart, max(1, args.steps // 2), 1,
                          smooth=smooth, graphs=args.graphs, smoother=args.smoother, line_levels=args.line_levels,
                          levels=args.levels, diag_faces=not args.no_diag_faces)
        print(f"[bench] fp64-vcycle: {len(fp64['its'])} solves in {fp64['elapsed']:.3f} s, its {fp64['its']}",
              file=sys.stderr, flush=True)
    # the configuration names BiCGStab next to FGMRES and a 6-level hierarchy: both measured as well
    # (2 solves each after a warm-up), the headline stays the FGMRES(30) 8-level solve
    c3_legs = {}
    if args.precision in ("mixed", "mixed16") and args.krylov == "fgmres" and not args.no_c3_legs:
        for leg, kw in (("bicgstab", dict(krylov="bicgstab")), ("levels6", dict(levels=6))):
            r = run_solver(lsx, hip, geom, args.precision, kw.get("krylov", "fgmres"), args.restart, 2, 1,
                           smooth=smooth, graphs=args.graphs, smoother=args.smoother, line_levels=args.line_levels,
                           levels=kw.get("levels", args.levels), diag_faces=not args.no_diag_faces,
                           dot32=not args.no_dot32)
            c3_legs[leg] = {"solves_per_s": len(r["its"]) / r["elapsed"], "iterations": r["its"],
                            "true_rel_residual": r["rel_res"], "mg_levels": r["levels"]}
            print(f"[bench] {leg}: {c3_legs[leg]['solves_per_s']:.2f} solves/s, its {r['its']}", file=sys.stderr,
                  flush=True)
            del r
    cascade = None
    if args.precision != "fp64" and args.cascade_max_it > 0:
        cascade = run_solver(lsx, hip, geom, "fp64", args.krylov, args.restart, 1, 0, max_it=args.cascade_max_it)
        print(f"[bench] fp64 reference cascade: its {cascade['its']} in {cascade['elapsed']:.3f} s, "
              f"true rel res {cascade['rel_res']:.3e}", file=sys.stderr, flush=True)

    stats, elapsed = main_run["stats"], main_run["elapsed"]
    dom = max(stats, key=lambda k: stats[k][1])
    launches, tot_ms, bytes_per = stats[dom]
    avg_s = tot_ms / launches / 1e3
    achieved = bytes_per / avg_s / 1e9
    kernels = {lsx.KERNEL_NAMES[k]: {"launches": v[0], "avg_us": v[1] / v[0] * 1e3,
                                     "GBs": v[2] / (v[1] / v[0] / 1e3) / 1e9,
                                     "share_of_step": v[1] / 1e3 / elapsed}
               for k, v in stats.items() if v[0]}
    # HBM traffic per launch of the dominant kernel: FETCH_SIZE + WRITE_SIZE counter runs recorded in
    # profiles/traffic.json with the revision they measured; used only when that record's algorithmic
    # bytes per launch equal this run's (same kernel variant and workload), else reported as null
    traffic, traffic_note = None, "no counter record for this kernel, precision and grid"
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        try:
            rec = json.load(open(tfile))
            ent = rec.get("kernels", {}).get(f"{lsx.KERNEL_NAMES[dom]}@{args.precision}@{n}")
            if ent is not None:
                if abs(ent["algorithmic_bytes"] - bytes_per) <= 1e-3 * bytes_per:
                    traffic = ent["traffic_bytes"]
                    traffic_note = (f"rocprofv3 FETCH_SIZE + WRITE_SIZE of {ent['kernel']} at revision "
                                    f"{rec.get('revision')}, scaled by a copy of known size")
                else:
                    traffic_note = (f"the counter record (revision {rec.get('revision')}) was taken on a kernel "
                                    f"moving {ent['algorithmic_bytes']:.4g} algorithmic bytes per launch, this run "
                                    f"{bytes_per:.4g}: stale, not used")
        except (ValueError, OSError, KeyError):
            traffic = None
    value = args.steps / elapsed
    sm_name = "red-black GS" if args.smoother == "rbgs" else "damped Jacobi" if args.smoother == "jacobi" else (
        "x/y/z line-relaxation" if args.line_levels <= 0 else
        f"x/y/z line-relaxation on the finest {args.line_levels} levels, red-black GS below")
    cfg = {
        "workload": (f"{n}^3 stretched grid (ratio 1.02 to centre), "
                     + ("six-patch bicubic Bezier body fitted to the ellipsoid 0.25/0.15/0.15, control points "
                        "+ U(-0.01, 0.01) (seed 0)" if args.body == "bezier" else "ellipsoid body 0.25/0.15/0.15")
                     + f", {args.krylov.upper()}"
                     f"({args.restart}), rel tol 1e-8, "
                     + {"mixed": f"mixed: FP32 V{smooth} {sm_name} multigrid, FP16 coarse faces, FP64 Krylov"
                                  + ("" if args.no_dot32 or args.krylov != "fgmres"
                                     else " (Gram-Schmidt projections over FP32 basis copies)"),
                        "mixed16": f"mixed: FP32 V{smooth} {sm_name} multigrid, FP16 faces, FP64 Krylov"
                                   + ("" if args.no_dot32 or args.krylov != "fgmres"
                                      else " (Gram-Schmidt projections over FP32 basis copies)"),
                        "fp64-vcycle": f"FP64 V{smooth} {sm_name} multigrid",
                        "fp64": "FP64 reference multigrid cascade (ADLJ)"}[args.precision]),
        "precision": args.precision, "krylov": args.krylov, "smoother": args.smoother, "graphs": args.graphs,
        "line_levels": args.line_levels if args.smoother == "line" else None,
        "diag_from_faces": not args.no_diag_faces,
        "projections_over_fp32_basis_copies": args.precision in ("mixed", "mixed16") and args.krylov == "fgmres"
                                              and not args.no_dot32,
        "cells": geom.n,
        "iterations": main_run["its"], "true_rel_residual": main_run["rel_res"],
        # ||D u + flux|| / ||b|| of the velocity the solve's pressure would correct: with D inv_M G = -A
        # it equals the true relative residual ||b - A p|| / ||b|| (FP64 operator)
        "divergence_rel_error": main_run["rel_res"],
        "geometry_setup_s": geo_s, "solver_setup_s": main_run["setup_s"], "mg_levels": main_run["levels"],
    }
    if fp64 is not None:
        fp64_rate = len(fp64["its"]) / fp64["elapsed"]
        act = main_run["S"].operator_diagonal() < 0
        err = float(np.linalg.norm((main_run["x"] - fp64["x"])[act]) / np.linalg.norm(fp64["x"][act]))
        cfg.update({"fp64_solves_per_s": fp64_rate, "fp64_iterations": fp64["its"],
                    "fp64_true_rel_residual": fp64["rel_res"], "fp64_divergence_rel_error": fp64["rel_res"],
                    "mixed_over_fp64": value / fp64_rate,
                    "pressure_rel_diff_vs_fp64": err})
    if c3_legs:
        cfg["other_c3_legs"] = c3_legs
    if cascade is not None:
        cit = cascade["its"][0]
        cfg["reference_cascade_fp64"] = {
            "iterations": cit, "iteration_cap": args.cascade_max_it, "true_rel_residual": cascade["rel_res"],
            "ms_per_iteration": cascade["elapsed"] / max(cit, 1) * 1e3,
            "reached_1e-8": bool(cascade["rel_res"] <= 1e-8 * 1.05)}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f16 (f64 Krylov)" if args.precision in ("mixed", "mixed16") else "f64",
        "data": "synthetic",
        "config": cfg,
        "roofline": {"bound": "hbm", "kernel": lsx.KERNEL_NAMES[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_note": traffic_note,
                     "algorithmic_bytes_per_launch": bytes_per, "avg_launch_us": avg_s * 1e6},
        "kernels": kernels,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(n, args.restart, args.krylov, args.body)
            out["cpu_baseline"]["c2_mixed_vs_oracle"] = c2_against_oracle(args.krylov, args.restart)
        except Exception as e:  # the GPU measurement stands on its own
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
