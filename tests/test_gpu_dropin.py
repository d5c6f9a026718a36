"""The drop-in boundary used the way a reference-side wrapper of the C ABI uses it for the
reference's IterativeSolverInterface (iterative_solver_interface.h:46-55, mg_fgmres.h:35-39):

* lsx_set_csr with device pointers (the reference's CUDA build hands over device arrays) gives
  the same solve, bit for bit, as the host-pointer call;
* lsx_update_values + lsx_rebuild (rebuildPreconditioner) after new values with the same pattern
  equals a fresh lsx_set_csr of those values, and a solve between the two calls is refused;
* the C++ mirror class LsxIterativeSolver (include/lsx_iterative_solver.hpp) solves on the GPU
  with the same result as the Python mirror, including its rebuildPreconditioner path."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O
import problems as P

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "ls-stag-3d_parallel_amd")


class DeviceMem:
    """hipMalloc'd buffers through the HIP runtime (test plumbing only)."""

    def __init__(self):
        self.l = C.CDLL("libamdhip64.so")
        self.l.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.l.hipFree.argtypes = [C.c_void_p]
        self.l.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.bufs = []

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = C.c_void_p()
        assert self.l.hipMalloc(C.byref(p), max(8, arr.nbytes)) == 0
        assert self.l.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0
        self.bufs.append(p)
        return p

    def get(self, p, n, dtype=np.float64):
        assert self.l.hipDeviceSynchronize() == 0
        out = np.zeros(n, dtype)
        assert self.l.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.l.hipFree(p)
        self.bufs = []


def options(lsx, precision, eps):
    L = lsx.lib()
    o = lsx.Options()
    L.lsx_options_default(C.byref(o))
    o.restart, o.max_it, o.eps = 30, 500, eps
    if precision != "fp64":
        o.mg, o.n_pre, o.n_post = lsx.MG_VCYCLE, 2, 2
        o.precision = lsx.PREC_MIXED if precision == "mixed" else lsx.PREC_FP64
    return o


def solve_with(lsx, o, A, b, mem, dev=None):
    """Create, set_csr (host or device arrays), solve from x0 = 0; returns (x, it, norm_r)."""
    L = lsx.lib()
    h = C.c_void_p()
    lsx._check(L.lsx_create(C.byref(h), C.byref(o), None))
    try:
        N, M, K = A.dims
        rp = np.ascontiguousarray(A.rowptr, np.int32)
        cl = np.ascontiguousarray(A.col, np.int32)
        vl = np.ascontiguousarray(A.val, np.float64)
        it, nr = C.c_int(), C.c_double()
        if mem == lsx.MEM_HOST:
            lsx._check(L.lsx_set_csr(h, N, M, K, rp.ctypes.data, cl.ctypes.data, vl.ctypes.data, len(vl), mem))
            x = np.zeros(A.n)
            lsx._check(L.lsx_solve(h, b.ctypes.data, x.ctypes.data, mem, C.byref(it), C.byref(nr)))
        else:
            drp, dcl, dvl = dev.put(rp), dev.put(cl), dev.put(vl)
            db, dx = dev.put(b), dev.put(np.zeros(A.n))
            lsx._check(L.lsx_set_csr(h, N, M, K, drp, dcl, dvl, len(vl), mem))
            lsx._check(L.lsx_solve(h, db, dx, mem, C.byref(it), C.byref(nr)))
            x = dev.get(dx, A.n)
        return x, it.value, nr.value
    finally:
        L.lsx_destroy(h)


@pytest.mark.parametrize("precision", ["fp64", "fp64-vcycle", "mixed"])
def test_device_csr_equals_host_csr(lsx, precision):
    g, A = P.sphere(32)
    b = P.random_rhs(g, seed=11, A=A)
    o = options(lsx, precision, 1e-8 * np.linalg.norm(b))
    dev = DeviceMem()
    try:
        xh, ith, nrh = solve_with(lsx, o, A, b, lsx.MEM_HOST)
        xd, itd, nrd = solve_with(lsx, o, A, b, lsx.MEM_DEVICE, dev)
    finally:
        dev.free()
    print(f"precision={precision} host it={ith} norm_r={nrh:.6e} device it={itd} norm_r={nrd:.6e}", flush=True)
    assert ith == itd and nrh == nrd
    assert np.array_equal(xh, xd)


def same_pattern_pair():
    """Two operators with one pattern: the empty box on a uniform and on a stretched grid."""
    N = 32
    g1 = O.uniform_geometry(N, N, N)
    w = P.stretched_axis(N, 1.0, 0.5, 1.05)
    g2 = O.Geometry(w, w[::-1].copy(), w)
    A1, A2 = g1.laplace(pin=True), g2.laplace(pin=True)
    assert np.array_equal(A1.rowptr, A2.rowptr) and np.array_equal(A1.col, A2.col)
    assert not np.array_equal(A1.val, A2.val)
    return g1, A1, g2, A2


@pytest.mark.parametrize("mem", ["host", "device"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_update_values_rebuild_equals_fresh_set(lsx, precision, mem):
    g1, A1, g2, A2 = same_pattern_pair()
    b = P.random_rhs(g2, seed=12, A=A2)
    o = options(lsx, precision, 1e-8 * np.linalg.norm(b))
    L = lsx.lib()
    x_fresh, it_fresh, nr_fresh = solve_with(lsx, o, A2, b, lsx.MEM_HOST)
    h = C.c_void_p()
    lsx._check(L.lsx_create(C.byref(h), C.byref(o), None))
    dev = DeviceMem()
    try:
        N, M, K = A1.dims
        rp = np.ascontiguousarray(A1.rowptr, np.int32)
        cl = np.ascontiguousarray(A1.col, np.int32)
        v1 = np.ascontiguousarray(A1.val, np.float64)
        v2 = np.ascontiguousarray(A2.val, np.float64)
        lsx._check(L.lsx_set_csr(h, N, M, K, rp.ctypes.data, cl.ctypes.data, v1.ctypes.data, len(v1), lsx.MEM_HOST))
        if mem == "host":
            lsx._check(L.lsx_update_values(h, v2.ctypes.data, lsx.MEM_HOST))
        else:
            lsx._check(L.lsx_update_values(h, dev.put(v2), lsx.MEM_DEVICE))
        x = np.zeros(A1.n)
        it, nr = C.c_int(), C.c_double()
        st = L.lsx_solve(h, b.ctypes.data, x.ctypes.data, lsx.MEM_HOST, C.byref(it), C.byref(nr))
        assert st == lsx.ERR_STATE, "a solve between update_values and rebuild must be refused"
        assert b"stale" in L.lsx_last_error()
        lsx._check(L.lsx_rebuild(h))
        lsx._check(L.lsx_solve(h, b.ctypes.data, x.ctypes.data, lsx.MEM_HOST, C.byref(it), C.byref(nr)))
    finally:
        L.lsx_destroy(h)
        dev.free()
    print(f"precision={precision} values from {mem}: rebuilt it={it.value} fresh it={it_fresh}", flush=True)
    assert it.value == it_fresh and nr.value == nr_fresh
    assert np.array_equal(x, x_fresh)


def test_duplicate_column_rejected(lsx):
    g, A = P.body_free(8)
    rp = np.array(A.rowptr, np.int32)
    cl = np.array(A.col, np.int32)
    vl = np.array(A.val, np.float64)
    i = 100  # an interior row: name its diagonal column twice (nnz unchanged)
    lo, hi = rp[i], rp[i + 1]
    d = lo + int(np.nonzero(cl[lo:hi] == i)[0][0])
    cl[d + 1 if d + 1 < hi else d - 1] = i
    L = lsx.lib()
    o = options(lsx, "fp64", 1e-8)
    h = C.c_void_p()
    lsx._check(L.lsx_create(C.byref(h), C.byref(o), None))
    try:
        st = L.lsx_set_csr(h, 8, 8, 8, rp.ctypes.data, cl.ctypes.data, vl.ctypes.data, len(vl), lsx.MEM_HOST)
        print("duplicate column:", st, L.lsx_last_error(), flush=True)
        assert st == lsx.ERR_ARG and b"twice" in L.lsx_last_error()
    finally:
        L.lsx_destroy(h)


CPP_SOLVE = r"""
// Reads a CSR matrix and a right-hand side, solves through LsxIterativeSolver, then changes the
// values in place, calls rebuildPreconditioner() and solves again (the reference call sequence
// set / solve / rebuildPreconditioner / solve).
#include <cstdio>
#include <vector>
#include "lsx_iterative_solver.hpp"

template <class T> static std::vector<T> rd(const char* path, long n) {
  std::vector<T> v(n);
  FILE* f = std::fopen(path, "rb");
  if (!f || std::fread(v.data(), sizeof(T), n, f) != (size_t)n) { std::fprintf(stderr, "read %s\n", path); std::exit(2); }
  std::fclose(f);
  return v;
}
int main(int argc, char** argv) {
  const int N = std::atoi(argv[1]);
  const long n = (long)N * N * N, nnz = std::atol(argv[2]);
  const int precision = std::atoi(argv[3]);
  const double eps = std::atof(argv[4]);
  auto rp = rd<int>("rowptr.bin", n + 1);
  auto cl = rd<int>("col.bin", nnz);
  auto v = rd<double>("val1.bin", nnz);
  auto v2 = rd<double>("val2.bin", nnz);
  auto b = rd<double>("rhs.bin", n);
  std::vector<double> x(n, 0.0);
  lsx_options o = lsx_cpp::mg_options(LSX_KRYLOV_FGMRES, 30, 500, eps, 3, precision);
  try {
    lsx_cpp::LsxIterativeSolver s(o);
    lsx_cpp::CsrView A{N, N, N, rp.data(), cl.data(), v.data(), nnz, LSX_MEM_HOST};
    s.set(A, b.data(), x.data());
    const int it1 = s.solve();
    FILE* f = std::fopen("x1.bin", "wb");
    std::fwrite(x.data(), sizeof(double), n, f);
    std::fclose(f);
    std::printf("it1 %d %.17g\n", it1, s.norm_r);
    for (long j = 0; j < nnz; ++j) v[j] = v2[j];  // the equation changed its matrix in place
    s.rebuildPreconditioner();
    for (long i = 0; i < n; ++i) x[i] = 0.0;
    const int it2 = s.solve();
    f = std::fopen("x2.bin", "wb");
    std::fwrite(x.data(), sizeof(double), n, f);
    std::fclose(f);
    std::printf("it2 %d %.17g\n", it2, s.norm_r);
  } catch (const lsx_cpp::lsx_exception& e) {
    std::fprintf(stderr, "lsx: %s\n", e.what());
    return 3;
  }
  return 0;
}
"""


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cpp_mirror_solves_on_gpu(lsx, tmp_path, precision):
    g1, A1, g2, A2 = same_pattern_pair()
    b = P.random_rhs(g1, seed=13, A=A1)
    eps = 1e-8 * np.linalg.norm(b)
    prec = lsx.PREC_FP64 if precision == "fp64" else lsx.PREC_MIXED
    np.asarray(A1.rowptr, np.int32).tofile(tmp_path / "rowptr.bin")
    np.asarray(A1.col, np.int32).tofile(tmp_path / "col.bin")
    np.asarray(A1.val, np.float64).tofile(tmp_path / "val1.bin")
    np.asarray(A2.val, np.float64).tofile(tmp_path / "val2.bin")
    b.tofile(tmp_path / "rhs.bin")
    src = tmp_path / "solve.cpp"
    src.write_text(CPP_SOLVE)
    exe = tmp_path / "solve"
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-L", LIBDIR,
                        "-llsx", f"-Wl,-rpath,{LIBDIR}", "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe), "32", str(len(A1.val)), str(prec), f"{float(eps):.17g}"], capture_output=True, text=True,
                       timeout=120, cwd=tmp_path)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    lines = dict(l.split(" ", 1) for l in r.stdout.strip().splitlines())
    it1, nr1 = lines["it1"].split()
    it2, nr2 = lines["it2"].split()
    x1 = np.fromfile(tmp_path / "x1.bin")
    x2 = np.fromfile(tmp_path / "x2.bin")
    o = options(lsx, "fp64" if precision == "fp64" else "mixed", eps)
    o.n_post = 2 if precision == "mixed" else 3
    ref1 = solve_with(lsx, o, A1, b, lsx.MEM_HOST)
    ref2 = solve_with(lsx, o, A2, b, lsx.MEM_HOST)
    assert int(it1) == ref1[1] and float(nr1) == ref1[2] and np.array_equal(x1, ref1[0])
    assert int(it2) == ref2[1] and float(nr2) == ref2[2] and np.array_equal(x2, ref2[0])
    assert np.linalg.norm(b - A2.matvec(x2)) <= 1.05 * eps
