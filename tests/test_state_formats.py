"""State formats: the reference's binary checkpoint and its Tecplot point file.

Checkpoint (set_solutions.h:317-339 / 247-269, solution.h:76-81, base_task_description.h:151-160):
the byte layout is read back field by field with `struct` at the offsets the reference's writes
produce (int32, then doubles, no padding) and the round trip through lsx_state_load is exact.

Tecplot (sol_tecplot_printer.h:254-347): every point of the (N+2)(M+2)(K+2) zone is compared with
its value stated point by point from the printer's rules (walls take the wall velocities, cells
average their faces, cut cells weight them by their face fractions, solid cells are 0, the west
column and the edge rows use the wall values beside them), on a random grid with every kind of
cell, with and without an outflow boundary.
"""
import os
import struct

import numpy as np
import pytest

from conftest import load_lsx

lsx = load_lsx()

FLUID, SOLID = 11111111, 0


def random_state(rng, dims):
    return tuple(rng.standard_normal(n) for n in lsx.face_sizes(*dims))


def test_checkpoint_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    dims = (7, 5, 4)
    cur, prev = random_state(rng, dims), random_state(rng, dims)
    pos, vel, loads = rng.standard_normal((2, 3)), rng.standard_normal((2, 3)), rng.standard_normal((2, 4))
    path = str(tmp_path / "continueSolveTask.dat")
    lsx.state_save(path, 123, 0.25, dims, cur, prev, bodies=(pos, vel, loads))
    raw = open(path, "rb").read()
    nu, nv, nw, npc = lsx.face_sizes(*dims)
    assert len(raw) == lsx.state_bytes(*dims, 2) == 4 + 8 + 2 * 6 * 8 + 2 * 4 * 8 + 2 * 8 * (nu + nv + nw + npc)
    assert struct.unpack_from("<i", raw, 0)[0] == 123
    assert struct.unpack_from("<d", raw, 4)[0] == 0.25
    off = 12
    for b in range(2):  # per body: centre x, y, z then centre velocity x, y, z
        assert struct.unpack_from("<6d", raw, off) == tuple(pos[b]) + tuple(vel[b])
        off += 48
    for b in range(2):  # the task's loads F_aer_x, F_aer_y, F_aer_z, moment_aer
        assert struct.unpack_from("<4d", raw, off) == tuple(loads[b])
        off += 32
    for a in cur + prev:  # current u, v, w, p then previous u, v, w, p
        got = np.frombuffer(raw, "<f8", a.size, off)
        assert np.array_equal(got, a)
        off += 8 * a.size
    assert off == len(raw)
    st = lsx.state_load(path, dims, n_bodies=2)
    assert st["it"] == 123 and st["time"] == 0.25
    assert np.array_equal(st["body_pos"], pos) and np.array_equal(st["body_vel"], vel)
    assert np.array_equal(st["loads"], loads)
    for a, b in zip(st["cur"] + st["prev"], cur + prev):
        assert np.array_equal(a, b)


def test_checkpoint_without_bodies_and_size_check(tmp_path):
    rng = np.random.default_rng(1)
    dims = (6, 6, 6)
    cur, prev = random_state(rng, dims), random_state(rng, dims)
    path = str(tmp_path / "s.dat")
    lsx.state_save(path, 1, 0.002, dims, cur, prev)
    assert os.path.getsize(path) == lsx.state_bytes(*dims)
    with pytest.raises(lsx.LsxError, match="size does not match"):
        lsx.state_load(path, (6, 6, 5))
    with pytest.raises(lsx.LsxError, match="size does not match"):
        lsx.state_load(path, dims, n_bodies=1)
    st = lsx.state_load(path, dims)
    assert st["it"] == 1 and all(np.array_equal(a, b) for a, b in zip(st["cur"], cur))


# ---- Tecplot --------------------------------------------------------------------------------


def read_tecplot(path, dims):
    N, M, K = dims
    lines = open(path).read().splitlines()
    assert lines[:10] == ['TITLE     = "Velocity Vectors" ', 'VARIABLES = "x" ', '"y"', '"z"', '"P"', '"SPD_x"',
                          '"SPD_y"', '"SPD_z"', '"SPD"', 'ZONE T="Data" ']
    assert lines[10] == f"I={N + 2}, J={M + 2}, K={K + 2}, F=POINT"
    data = np.array([[float(t) for t in l.split(" ")] for l in lines[11:]])
    assert data.shape == ((N + 2) * (M + 2) * (K + 2), 8)
    return data.reshape(K + 2, M + 2, N + 2, 8)


def expected_points(dims, xc, yc, zc, dx, dy, dz, f, u, v, w, p, uw, wl, outflow):
    """Each point's value from the printer's rules (not its traversal)."""
    N, M, K = dims
    uu = u.reshape(K, M, N - 1)
    vv = v.reshape(K, M - 1, N)
    ww = w.reshape(K - 1, M, N)
    pp = p.reshape(K, M, N)
    t = f["type"].reshape(K, M, N)
    tx, ty, tz = (f[k].reshape(K, M, N) for k in ("thx", "thy", "thz"))
    X = np.concatenate([[0.0], xc, [xc[-1] + 0.5 * dx[-1]]])
    Y = np.concatenate([[0.0], yc, [yc[-1] + 0.5 * dy[-1]]])
    Z = np.concatenate([[0.0], zc, [zc[-1] + 0.5 * dz[-1]]])

    def wgt(ta, va, tb, vb):
        s = ta + tb
        return (ta * va + tb * vb) / s if s > 1e-5 else 0.0

    def w_lo(k, j, i):  # z velocity below cell (the bottom wall's for k = 0)
        return wl.W_B if k == 0 else ww[k - 1, j, i]

    def w_hi(k, j, i):
        return wl.W_T if k == K - 1 else ww[k, j, i]

    out = np.zeros((K + 2, M + 2, N + 2, 8))
    for kp in range(K + 2):
        k = min(max(kp - 1, 0), K - 1)
        wall_plane = (kp == 0 or kp == K + 1) and not outflow
        for jp in range(M + 2):
            j = min(max(jp - 1, 0), M - 1)
            for ip in range(N + 2):
                i = min(max(ip - 1, 0), N - 1)
                P = pp[k, j, i]
                if wall_plane:
                    vel = (wl.U_B, wl.V_B, wl.W_B) if kp == 0 else (wl.U_T, wl.V_T, wl.W_T)
                elif jp == 0:
                    vel = (wl.U_S, wl.V_S, wl.W_S)
                elif jp == M + 1:
                    vel = (wl.U_N, wl.V_N, wl.W_N)
                else:
                    edge = j == 0 or j == M - 1
                    vs = wl.V_S if j == 0 else vv[k, j - 1, i]   # south face or wall
                    vn = wl.V_N if j == M - 1 else vv[k, j, i]   # north face or wall
                    # an edge row carries the z velocities of its first cell
                    wi = 0 if edge else i
                    wlo, whi = w_lo(k, j, wi), w_hi(k, j, wi)
                    if ip == 0:
                        vel = (uw[j + M * k] if uw is not None else 0.0, wl.V_W, wl.W_W)
                    elif ip == N + 1:
                        vel = (uu[k, j, N - 2], 0.5 * (vn + vs), 0.5 * (whi + wlo)) if outflow else \
                            (wl.U_E, wl.V_E, wl.W_E)
                    elif i == 0:
                        ux = 0.5 * (uu[k, j, 0] + uw[j + M * k]) if uw is not None else 0.0
                        vel = (ux, 0.5 * (vn + vs), 0.5 * (whi + wlo))
                    elif i == N - 1:
                        ux = uu[k, j, N - 2] if outflow else 0.5 * (uu[k, j, N - 2] + wl.U_E)
                        vel = (ux, 0.5 * (vn + vs), 0.5 * (whi + wlo))
                    elif edge or t[k, j, i] == FLUID:
                        vel = (0.5 * (uu[k, j, i] + uu[k, j, i - 1]), 0.5 * (vn + vs), 0.5 * (whi + wlo))
                    elif t[k, j, i] == SOLID:
                        vel = (0.0, 0.0, 0.0)
                    else:  # cut or boundary-typed cell: face-fraction weights
                        ux = wgt(tx[k, j, i], uu[k, j, i], tx[k, j, i - 1], uu[k, j, i - 1])
                        uy = wgt(ty[k, j, i], vv[k, j, i], ty[k, j - 1, i], vv[k, j - 1, i])
                        if k == 0:
                            uz = wgt(tz[k, j, i], ww[k, j, i], tz[k, j, i], wl.W_B)
                        elif k == K - 1:
                            uz = wgt(tz[k, j, i], ww[k - 1, j, i], tz[k, j, i], wl.W_T)
                        else:
                            uz = wgt(tz[k, j, i], ww[k, j, i], tz[k - 1, j, i], ww[k - 1, j, i])
                        vel = (ux, uy, uz)
                spd = float(np.sqrt(sum(c * c for c in vel)))
                out[kp, jp, ip] = (X[ip], Y[jp], Z[kp], P) + tuple(vel) + (spd,)
    return out


def synthetic_fields(rng, dims):
    N, M, K = dims
    n = N * M * K
    codes = np.array([FLUID, FLUID, FLUID, SOLID, 11110000, 10111111, -3, -2])
    t = rng.choice(codes, n).astype(np.int32)
    th = {k: rng.uniform(0.0, 1.0, n) for k in ("thx", "thy", "thz")}
    th["thx"][rng.random(n) < 0.2] = 0.0
    return dict(type=t, **th)


@pytest.mark.parametrize("outflow", [True, False])
@pytest.mark.parametrize("profile", [True, False])
def test_tecplot_points(tmp_path, outflow, profile):
    rng = np.random.default_rng(7 + outflow + 2 * profile)
    dims = (7, 6, 5)
    N, M, K = dims
    dx, dy, dz = (rng.uniform(0.5, 1.5, m) / m for m in dims)
    xc, yc, zc = ((np.cumsum(d) - 0.5 * d) for d in (dx, dy, dz))
    f = synthetic_fields(rng, dims)
    u, v, w, p = (rng.standard_normal(m) for m in lsx.face_sizes(*dims))
    uw = rng.uniform(0.5, 1.5, M * K) if profile else None
    wl = lsx.walls(**{k: rng.uniform(-1, 1) for k in lsx.WALL_FIELDS})
    path = str(tmp_path / "Sol000001.dat")
    lsx.write_tecplot(path, dims, (xc, yc, zc), (dx, dy, dz), f, u, v, w, p, u_west=uw, walls_=wl, outflow=outflow)
    got = read_tecplot(path, dims)
    exp = expected_points(dims, xc, yc, zc, dx, dy, dz, f, u, v, w, p, uw, wl, outflow)
    # six significant digits in the file
    bad = ~np.isclose(got, exp, rtol=6e-6, atol=1e-300)
    if bad.any():
        idx = np.argwhere(bad)[0]
        raise AssertionError(f"point {tuple(idx[:3])} var {idx[3]}: file {got[tuple(idx)]} expected {exp[tuple(idx)]}")
    if outflow:  # flow tasks repeat the first and last cell planes at the walls
        assert np.array_equal(got[0, :, :, 3:], got[1, :, :, 3:])
        assert np.array_equal(got[K + 1, :, :, 3:], got[K, :, :, 3:])


def test_tecplot_uniform_flow_is_uniform(tmp_path):
    """A uniform stream u = 1 through an empty box prints SPD_x = 1 at every cell point."""
    dims = (8, 6, 5)
    N, M, K = dims
    h = [np.full(m, 1.0 / m) for m in dims]
    c = [(np.arange(m) + 0.5) / m for m in dims]
    n = N * M * K
    f = dict(type=np.full(n, FLUID, np.int32), thx=np.ones(n), thy=np.ones(n), thz=np.ones(n))
    nu, nv, nw, npc = lsx.face_sizes(*dims)
    path = str(tmp_path / "u.dat")
    lsx.write_tecplot(path, dims, c, h, f, np.ones(nu), np.zeros(nv), np.zeros(nw), np.zeros(npc), u_west=1.0,
                      walls_=lsx.walls(U_E=1.0), outflow=True)
    got = read_tecplot(path, dims)
    cells = got[1:K + 1, 1:M + 1, :, :]  # every row between the south and north walls
    assert np.all(cells[..., 4] == 1.0) and np.all(cells[..., 5:7] == 0.0) and np.all(cells[..., 7] == 1.0)


@pytest.mark.gpu
@pytest.mark.parametrize("outflow", [True, False])
def test_geometry_tecplot_equals_host_writer(tmp_path, outflow):
    """lsx_geometry_write_tecplot (fields of a GPU-built cut-cell geometry) writes the same bytes as
    lsx_write_tecplot given that geometry's fields, centres and widths."""
    dims = (24, 20, 16)
    w = [lsx.stretched_axis(d, 1.0, 0.5, 1.03) for d in dims]
    g = lsx.Geometry(*w).build("sphere", (0.45, 0.5, 0.5), (0.2,))
    f = g.fields()
    assert np.count_nonzero((f["type"] > 0) & (f["type"] != FLUID)) > 0
    rng = np.random.default_rng(3)
    u, v, w_, p = (rng.standard_normal(m) for m in lsx.face_sizes(*dims))
    uw = rng.uniform(0.9, 1.1, dims[1] * dims[2])
    wl = lsx.walls(U_E=1.0, V_S=0.1, W_T=-0.2)
    a, b = str(tmp_path / "gpu.dat"), str(tmp_path / "host.dat")
    g.write_tecplot(a, u, v, w_, p, u_west=uw, walls_=wl, outflow=outflow)
    centres = [np.cumsum(d) - 0.5 * d for d in w]
    lsx.write_tecplot(b, dims, centres, w, f, u, v, w_, p, u_west=uw, walls_=wl, outflow=outflow)
    assert open(a, "rb").read() == open(b, "rb").read()
    read_tecplot(a, dims)


@pytest.mark.gpu
def test_checkpoint_device_arrays(tmp_path):
    """Device-resident solutions are written and read back through the same layout."""
    dims = (16, 12, 10)
    rng = np.random.default_rng(9)
    cur, prev = random_state(rng, dims), random_state(rng, dims)
    L = lsx.lib()
    dv = [lsx.DeviceVector(a.size) for a in cur + prev]
    for d, a in zip(dv, cur + prev):
        d.upload(a)
    path = str(tmp_path / "dev.dat")
    lsx._check(L.lsx_state_save(path.encode(), 7, 1.5, 0, None, None, None, *dims, *[d.p for d in dv],
                                lsx.MEM_DEVICE))
    st = lsx.state_load(path, dims)
    assert st["it"] == 7 and st["time"] == 1.5
    assert all(np.array_equal(x, y) for x, y in zip(st["cur"] + st["prev"], cur + prev))
    for d in dv:
        d.zero()
    it, t = lsx.C.c_int(0), lsx.C.c_double(0.0)
    lsx._check(L.lsx_state_load(path.encode(), lsx.C.byref(it), lsx.C.byref(t), 0, None, None, None, *dims,
                                *[d.p for d in dv], lsx.MEM_DEVICE))
    assert it.value == 7 and t.value == 1.5
    assert all(np.array_equal(d.download(), a) for d, a in zip(dv, cur + prev))
