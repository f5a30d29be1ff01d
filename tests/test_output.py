"""Output cadences (srcs/main.f90:706-743): VTK files, j-slices, interface
positions, centre/velocity monitor files.

CPU: the Fortran host's writers fed with vertex values / crossings computed
here (numpy, the reference's sum order) produce byte-identical VTK files
and identical monitor text to the reference's own mkvtk_phi, mkvtk_p and
find_interface_positions_upper/_lower, compiled from the reference sources
into oracle/_ref.  GPU: the device reductions (felgpu_vtk_points_,
mk_slice_j_, felgpu_interface_crossings_) equal the numpy restatements bit
for bit, and a run's VTK files equal the reference writer's on the same
field.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

import felharness as H
from felharness import D, I, Grid

need_ref = pytest.mark.skipif(not H.ref_available(), reason="oracle/_ref (compiled reference) not built")


def _drv():
    from felmpf import _lib
    return _lib.driver()


def vtk_points_np(g, q, kind):
    """The writers' ghost fills and vertex averages (srcs/mkvtk_phi.f:15-80,
    srcs/mkvtk_p.f:15-84), in the reference's sum order."""
    ni, nj, nk = g.ni, g.nj, g.nk
    ix = np.arange(0, ni + 2)
    mx = np.where(ix == 0, ni, np.where(ix == ni + 1, 1, ix))
    jy = np.arange(0, nj + 2)
    my = np.where(jy == 0, 1, np.where(jy == nj + 1, nj, jy))
    kz = np.arange(0, nk + 2)
    if kind == 0:
        mz = np.where(kz == 0, 1, np.where(kz == nk + 1, nk, kz))
    else:
        mz = np.where(kz == 0, nk, np.where(kz == nk + 1, 1, kz))
    f = q[np.ix_(mz + 2, my + 2, mx + 2)]  # filled (0..n+1)^3 block, (k, j, i)
    a = f[:-1, :-1, :-1]
    out = (((((((a + f[:-1, :-1, 1:]) + f[:-1, 1:, :-1]) + f[:-1, 1:, 1:]) + f[1:, :-1, :-1])
             + f[1:, :-1, 1:]) + f[1:, 1:, :-1]) + f[1:, 1:, 1:]) * 0.125
    return np.ascontiguousarray(out)


def crossings_np(g, q, jrow, dx, xl, first_only):
    x = np.zeros((g.nk, g.ni))
    cnt = np.zeros(g.nk, dtype=np.int32)
    for k in range(1, g.nk + 1):
        n = 0
        for i in range(1, g.ni + 1):
            p1, p2 = q[k + 2, jrow + 2, i + 2], q[k + 2, jrow + 2, i + 3]
            if p1 > 0.5 and p2 < 0.5:
                x[k - 1, n] = (i - 0.5) * dx + (0.5 - p1) * dx / (p2 - p1) - xl * 0.75
                n += 1
                if first_only:
                    break
        cnt[k - 1] = n
    return x, cnt


def _field(g, seed):
    rng = np.random.default_rng(seed)
    q = H.smooth_phi(g, rng)
    return q


def _write_vtk(path, g, step, name, vals):
    p = path.encode()
    nm = name.encode()
    _drv().fel_vtk_write(ctypes.c_char_p(p), ctypes.c_int(len(p)), ctypes.c_int(g.ni),
                         ctypes.c_int(g.nj), ctypes.c_int(g.nk), ctypes.c_int(step),
                         ctypes.c_double(g.dx), ctypes.c_double(g.dy), ctypes.c_double(g.dz),
                         ctypes.c_char_p(nm), ctypes.c_int(len(nm)),
                         vals.ctypes.data_as(ctypes.c_void_p))


@need_ref
@pytest.mark.parametrize("kind,name,ref", [(0, "phi", "mkvtk_phi"), (1, "p", "mkvtk_p")])
def test_vtk_files_equal_the_reference_writers(tmp_path, monkeypatch, kind, name, ref):
    monkeypatch.chdir(tmp_path)
    g = Grid(10, 6, 8, 0.1, 0.05, 0.125)
    q = _field(g, 31)
    vals = vtk_points_np(g, q, kind)
    _write_vtk(str(tmp_path / "mine.vtk"), g, 37, name, vals)
    svall = np.array([g.ni, g.nj, g.nk], dtype=np.int32)
    H.call_ref(ref, svall, I(37), D(g.dx), D(g.dy), D(g.dz), q.copy())
    theirs = (tmp_path / f"{name}_0000037.vtk").read_bytes()
    assert (tmp_path / "mine.vtk").read_bytes() == theirs


@need_ref
def test_interface_monitor_lines_equal_the_reference(capfd):
    g = Grid(24, 6, 5, 1.0 / 12, 0.1, 0.2)
    q = np.ones(g.shape)
    xs = (np.arange(-2, g.ni + 4) - 0.5) * g.dx
    for k in range(g.nk + 6):
        # a liquid band (phi ~ 0) of half width 0.4 around x = 1 + 0.03 k
        q[k, :, :] = (1.0 / (1.0 + np.exp(-(np.abs(xs - 1.0 - 0.03 * k) - 0.4) / 0.05)))[None, :]
    q[4, 3, :] = 1.0  # k = 2 on the bottom row: no crossing, counted as 0
    xl = 2.0
    drv = _drv()
    drv.fel_drv_flush()  # earlier monitor lines still in the Fortran runtime's buffer
    H.ref().felref_flush()
    for jrow, first, ref in ((g.nj, 0, "find_interface_positions_upper"),
                             (1, 1, "find_interface_positions_lower")):
        capfd.readouterr()
        H.call_ref(ref, I(g.ni), I(g.nj), I(g.nk), q, D(g.dx), D(g.dy), D(g.dz), D(xl))
        # the compiled reference binds its own Fortran runtime, which buffers unit 6
        H.ref().felref_flush()
        theirs = capfd.readouterr().out
        x, cnt = crossings_np(g, q, jrow, g.dx, xl, first)
        xt = np.ascontiguousarray(x)  # (nk, ni) C order = x(ni, nk) column-major
        drv.fel_interface_print(ctypes.c_int(0 if first == 0 else 1), ctypes.c_int(g.ni),
                                ctypes.c_int(g.nk), ctypes.c_double(g.dz),
                                xt.ctypes.data_as(ctypes.c_void_p),
                                cnt.ctypes.data_as(ctypes.c_void_p))
        drv.fel_drv_flush()
        mine = capfd.readouterr().out
        assert mine == theirs and len(theirs) > 0, (ref, mine, theirs)


# --------------------------------------------------------------------------
# GPU
# --------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("kind", [0, 1])
def test_vtk_points_on_device(kind):
    g = Grid(20, 12, 16, 0.1, 0.1, 0.1)
    q = _field(g, 41)
    out = np.zeros((g.nk + 1) * (g.nj + 1) * (g.ni + 1))
    H.call_gpu("felgpu_vtk_points", I(g.ni), I(g.nj), I(g.nk), q, I(kind), out)
    want = vtk_points_np(g, q, kind).ravel()
    assert np.array_equal(out, want)


def slice_np(g, q, j, ip, jp, kp):
    out = np.zeros((g.nk + 1, g.ni + 1))
    for k in range(g.nk + 1):
        for i in range(g.ni + 1):
            v = lambda a, b, c: q[c + 2, b + 2, a + 2]  # noqa: E731
            out[k, i] = (v(i, j, k) + v(i + ip, j, k) + v(i, j + jp, k) + v(i + ip, j + jp, k) +
                         v(i, j, k + kp) + v(i + ip, j, k + kp) + v(i, j + jp, k + kp) +
                         v(i + ip, j + jp, k + kp)) * 0.125
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("offs", [(0, 1, 1), (1, 1, 0), (1, 1, 1)])
def test_mk_slice_j_on_device(offs):
    g = Grid(12, 8, 10, 0.1, 0.1, 0.1)
    rng = np.random.default_rng(43)
    q = rng.standard_normal(g.shape)
    wrk = np.full((g.nk + 1, g.ni + 1), 7.0)
    out = np.full((g.nk + 1, g.ni + 1), 9.0)
    nid = np.full(6, -1, dtype=np.int32)
    H.call_gpu("mk_slice_j", I(0), I(0), nid, I(1), I(g.ni), I(g.nj), I(g.nk), I(3), I(offs[0]),
               I(offs[1]), I(offs[2]), wrk, q, out)
    want = slice_np(g, q, 3, *offs)
    assert np.array_equal(out, want) and np.array_equal(wrk, want)


@pytest.mark.gpu
@pytest.mark.parametrize("jrow,first", [(6, 0), (1, 1)])
def test_interface_crossings_on_device(jrow, first):
    g = Grid(32, 6, 8, 1.0 / 16, 0.1, 0.1)
    rng = np.random.default_rng(47)
    q = np.clip(0.5 + 0.8 * np.sin(np.arange(g.ni + 6) * 0.7)[None, None, :]
                + 0.1 * rng.standard_normal(g.shape), 0.0, 1.0)
    x = np.zeros(g.ni * g.nk)
    cnt = np.zeros(g.nk, dtype=np.int32)
    H.call_gpu("felgpu_interface_crossings", I(g.ni), I(g.nj), I(g.nk), q, I(jrow), D(g.dx),
               D(2.0), I(first), x, cnt)
    wx, wc = crossings_np(g, q, jrow, g.dx, 2.0, first)
    assert np.array_equal(cnt, wc)
    for k in range(g.nk):
        assert np.array_equal(x.reshape(g.nk, g.ni)[k, :wc[k]], wx[k, :wc[k]])


@pytest.mark.gpu
@need_ref
def test_run_vtk_files_equal_the_reference_writer(tmp_path, monkeypatch):
    from felmpf import config as C
    from felmpf.driver import GpuRun
    monkeypatch.chdir(tmp_path)
    cfg = C.sessile_droplet().scaled(16, 16, 16)
    run = GpuRun(cfg)
    try:
        run.steps(2)
        run.output("vtk")
        run.output("uwphi")
        run.output("center_vel")
        phi0, p = run.field("phi0"), run.field("p")
    finally:
        run.close()
    svall = np.array([cfg.ni, cfg.nj, cfg.nk], dtype=np.int32)
    os.rename("phi_0000002.vtk", "mine_phi.vtk")
    os.rename("p_0000002.vtk", "mine_p.vtk")
    H.call_ref("mkvtk_phi", svall, I(2), D(cfg.dx), D(cfg.dy), D(cfg.dz), phi0)
    H.call_ref("mkvtk_p", svall, I(2), D(cfg.dx), D(cfg.dy), D(cfg.dz), p)
    assert (tmp_path / "mine_phi.vtk").read_bytes() == (tmp_path / "phi_0000002.vtk").read_bytes()
    assert (tmp_path / "mine_p.vtk").read_bytes() == (tmp_path / "p_0000002.vtk").read_bytes()
    lines = (tmp_path / "uwphi0000002").read_text().split("\n")
    assert len([ln for ln in lines if ln.strip()]) == (cfg.ni + 1) * (cfg.nk + 1)
    assert len((tmp_path / "center_01.dat").read_text().split("\n")) == 2
