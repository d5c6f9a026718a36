"""The allocation-free drag entry point (lsx_geometry_drag_noalloc) against lsx_geometry_drag.

Both are called through the C ABI on the same geometry and pressure: J and dJ/dp must be
bit-identical.  Two calls in a row, the second on a smaller grid, check that the library's
shared scratch array carries nothing from one call into the next.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sync(lsx):
    h = lsx.DeviceVector.hip()
    assert h.hipDeviceSynchronize() == 0


def _compare(lsx, n, seed):
    L = lsx.lib()
    w = [np.full(n, 1.0 / n)] * 3
    g = lsx.Geometry(*w).build("sphere", (0.45, 0.5, 0.5), (0.23,))
    p = np.random.default_rng(seed).standard_normal(g.n)
    J_host, g_host = g.drag(p)  # existing entry point, host arrays
    dp = lsx.DeviceVector(g.n)
    dp.upload(p)
    # existing entry point on device arrays
    J_dev_ref = C.c_double(0.0)
    dg_ref = lsx.DeviceVector(g.n)
    lsx._check(L.lsx_geometry_drag(g.h, dp.p, lsx.MEM_DEVICE, C.byref(J_dev_ref), dg_ref.p))
    # allocation-free variant
    dJ = lsx.DeviceVector(1)
    dJ.upload(np.array([np.nan]))
    dg = lsx.DeviceVector(g.n)
    dg.upload(np.full(g.n, np.nan))
    print(f"noalloc drag: n={n}^3 seed={seed}", flush=True)
    lsx._check(L.lsx_geometry_drag_noalloc(g.h, dp.p, dJ.p, dg.p))
    _sync(lsx)
    J = float(dJ.download()[0])
    gg = dg.download()
    # dJdp may be NULL
    dJ2 = lsx.DeviceVector(1)
    lsx._check(L.lsx_geometry_drag_noalloc(g.h, dp.p, dJ2.p, None))
    _sync(lsx)
    J2 = float(dJ2.download()[0])
    print(f"  J={J:.17e} J_host={J_host:.17e} J_dev={J_dev_ref.value:.17e} J(no dJdp)={J2:.17e} "
          f"cut cells={int(np.count_nonzero(g_host))}", flush=True)
    assert np.count_nonzero(g_host) > 0
    assert J == J_host == J_dev_ref.value == J2
    assert np.array_equal(gg, g_host)
    assert np.array_equal(dg_ref.download(), g_host)
    # host memory is refused before anything runs
    Jh = np.zeros(1)
    st = L.lsx_geometry_drag_noalloc(g.h, C.c_void_p(p.ctypes.data), dJ.p, None)
    assert st == lsx.ERR_ARG and b"device pointers" in L.lsx_last_error()
    st = L.lsx_geometry_drag_noalloc(g.h, dp.p, C.c_void_p(Jh.ctypes.data), None)
    assert st == lsx.ERR_ARG
    for v in (dp, dg_ref, dJ, dg, dJ2):
        v.close()
    g.close()
    return J


def test_noalloc_drag_matches_existing_entry(lsx):
    J_big = _compare(lsx, 48, seed=3)
    J_small = _compare(lsx, 24, seed=4)  # smaller grid: nothing of the first call may leak
    assert J_big != J_small
