"""Checkpoint/restart files in the reference's format (srcs/dataou.f,
srcs/datain.f: one unformatted sequential record of u, v, w, p, uo, vo, wo,
po, phi(...,0:nbub), time, nstep).

CPU: the Fortran host's writer and reader (fel_ckpt_write/fel_ckpt_read,
host arrays only) exchange files with the reference's own dataou/datain,
compiled from the reference sources into oracle/_ref, in both directions,
bit for bit.  GPU: a run restarted from a checkpoint holds the checkpointed
state and continues the uninterrupted run, in the FP64 and the FP32 mode
(see test_restart_matches_the_oracle_restart for what the reference's
restart recomputes) and continues as the oracle restarted from the same file.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import felharness as H
from felharness import D, I, Grid


def _drv():
    from felmpf import _lib
    return _lib.driver()


def _state(g, seed):
    rng = np.random.default_rng(seed)
    fields = [rng.standard_normal(g.shape) for _ in range(8)]
    phi = rng.random((2,) + g.shape)
    return fields, phi


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ckpt(fn, path, g, step, t, fields, phi):
    fn(ctypes.c_char_p(path), ctypes.c_int(len(path)), ctypes.c_int(g.ni), ctypes.c_int(g.nj),
       ctypes.c_int(g.nk), ctypes.c_int(1), ctypes.byref(step), ctypes.byref(t),
       *[_ptr(a) for a in fields], _ptr(phi))


@pytest.mark.skipif(not H.ref_available(), reason="oracle/_ref (compiled reference) not built")
def test_files_are_interchangeable_with_the_reference(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    g = Grid(6, 5, 4)
    fields, phi = _state(g, 11)
    drv = _drv()
    # this host's writer -> the reference's datain (which reads "x.inp")
    _ckpt(drv.fel_ckpt_write, b"x.inp", g, ctypes.c_int(17), ctypes.c_double(0.125), fields, phi)
    got = [g.zeros() for _ in range(8)]
    gphi = np.zeros_like(phi)
    step, t = I(0), D(0.0)
    H.call_ref("datain", I(0), I(0), I(g.ni), I(g.nj), I(g.nk), I(1), step, t, *got, gphi)
    assert step.value == 17 and t.value == 0.125
    for a, b in zip(fields, got):
        assert np.array_equal(a, b)
    assert np.array_equal(phi, gphi)
    # the reference's dataou (writes "x<nstep:7>") -> this host's reader
    fields2, phi2 = _state(g, 12)
    H.call_ref("dataou", I(0), I(0), I(g.ni), I(g.nj), I(g.nk), I(1), I(42), D(3.5), *fields2, phi2)
    back = [g.zeros() for _ in range(8)]
    bphi = np.zeros_like(phi2)
    s2, t2 = ctypes.c_int(0), ctypes.c_double(0.0)
    _ckpt(drv.fel_ckpt_read, b"x0000042", g, s2, t2, back, bphi)
    assert s2.value == 42 and t2.value == 3.5
    for a, b in zip(fields2, back):
        assert np.array_equal(a, b)
    assert np.array_equal(phi2, bphi)
    # identical bytes for identical contents
    _ckpt(drv.fel_ckpt_write, b"mine", g, ctypes.c_int(42), ctypes.c_double(3.5), fields2, phi2)
    assert (tmp_path / "mine").read_bytes() == (tmp_path / "x0000042").read_bytes()


def test_record_layout(tmp_path):
    """One record: 4-byte length markers around 10 padded images, a real*8
    and a default integer."""
    g = Grid(4, 4, 4)
    fields, phi = _state(g, 13)
    path = str(tmp_path / "ck").encode()
    _ckpt(_drv().fel_ckpt_write, path, g, ctypes.c_int(5), ctypes.c_double(2.0), fields, phi)
    raw = (tmp_path / "ck").read_bytes()
    n = 10 * g.n3 * 8 + 8 + 4
    assert len(raw) == n + 8
    head = np.frombuffer(raw[:4], dtype=np.int32)[0]
    tail = np.frombuffer(raw[-4:], dtype=np.int32)[0]
    assert head == tail == n
    body = raw[4:-4]
    u = np.frombuffer(body[:g.n3 * 8], dtype=np.float64).reshape(g.shape)
    assert np.array_equal(u, fields[0])  # column-major image = C-order (nk, nj, ni) array
    assert np.frombuffer(body[-12:-4], dtype=np.float64)[0] == 2.0
    assert np.frombuffer(body[-4:], dtype=np.int32)[0] == 5


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
def test_restart_matches_the_oracle_restart(tmp_path, precision):
    """datain + the reference's restart fills reproduce the checkpointed
    state except the wall ghost rows (j <= 0, j >= nj+1) of u, v, w, uo, vo,
    wo and phi(...,0), which the fills recompute: bndu with the restored
    phi, where the uninterrupted run holds ghosts made with the previous
    step's phi (srcs/main.f90:379-380 vs 650; with l1 = l2 they agree to
    rounding), and contact-angle ghosts of phi(...,0) where it holds Neumann
    ghosts (main.f90:399-402 vs 509-510).  The checkpoint does not hold
    phin, whose stale ghost layer the next step's truncation reads
    (srcs/solphi_mthinc3.f:43-83), so a restarted run departs from the
    uninterrupted one at the 1e-6 level -- as the reference's restart does.
    The restarted GPU run is therefore held to the oracle restarted from the
    same file: the whole-step bars, 1e-10 after one step and 1e-8 after
    three (FP32 mode: 1e-5 relative L2, the FP32 mode's bar)."""
    from felmpf import config as C
    from felmpf.driver import GpuRun
    from oracle_run import OracleRun, rel_err
    cfg = C.droplet_impact().scaled(64, 16, 32)
    names = ["u", "v", "w", "p", "uo", "vo", "wo", "po", "phi0", "phi1"]
    ck = str(tmp_path / "x0000004")
    a = GpuRun(cfg, precision=precision)
    try:
        a.steps(4)
        a.dataou(ck)
        at_ck = {f: a.field(f) for f in names}
    finally:
        a.close()
    b = GpuRun(cfg, precision=precision)
    o = OracleRun(cfg)
    try:
        b.datain(ck)
        o.datain(ck)
        assert b.state()["nstep"] == 4
        inner = (slice(None), slice(3, cfg.nj + 3), slice(None))  # rows j = 1..nj
        for f in names:
            got = b.field(f)
            assert np.array_equal(got[inner], at_ck[f][inner]), f
            d = np.max(np.abs(got - at_ck[f]))
            assert d <= 1e-14 * max(np.max(np.abs(at_ck[f])), 1e-300) or f == "phi0", (f, d)
        for n, tol in ((1, 1e-10), (2, 1e-8)):
            b.steps(n)
            o.steps(n)
            assert b.state()["time"] == o.state()["time"]
            for f in ("u", "v", "w", "p", "phi1"):
                x, y = o.field(f)[3:-3, 3:-3, 3:-3], b.field(f)[3:-3, 3:-3, 3:-3]
                if precision == 64:
                    e, bar = rel_err(x, y), tol  # the whole-step bar (max norm)
                else:
                    e, bar = np.linalg.norm(x - y) / np.linalg.norm(x), 1e-5  # FP32 bar (L2)
                assert e <= bar, (f, n, e)
    finally:
        b.close()
        o.close()
