"""Deferred chains of entry points on device-resident arrays.

calsij_ -> cal_arith_tau_ -> cal_div_tensor_ (srcs/main.f90:537-540) is
recorded by the first two calls and run by the third in one kernel that
writes s, tau and the divergence.  Every array must equal the oracle's three
separate calls bit for bit -- including the cells calsij and arith_tau leave
untouched, which start with random contents here -- and any other access
in between must run what has been recorded at that point.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import felharness as H
from felharness import D, I, Grid
from test_gpu_reuse import Registered, same_bits

pytestmark = pytest.mark.gpu


def visc_setup(g, seed):
    r = np.random.default_rng(seed)
    u, v, w = (H.smooth_vel(g, r) for _ in range(3))
    rmu = r.uniform(0.01, 0.2, g.shape)
    s = r.standard_normal((6,) + g.shape)    # untouched cells keep these
    tau = r.standard_normal((6,) + g.shape)
    vis = [r.standard_normal(g.shape) for _ in range(3)]
    return [u, v, w, rmu, s, tau] + vis


def oracle_chain(g, arrs, interrupt=None):
    u, v, w, rmu, s, tau, vu, vv, vw = [a.copy() for a in arrs]
    dims = [I(g.ni), I(g.nj), I(g.nk)]
    inv = [D(1 / g.dx), D(1 / g.dy), D(1 / g.dz)]
    H.call_oracle("calsij", *dims, *inv, u, v, w, s)
    if interrupt == "rmu":
        rmu[...] = rmu * 1.5
    if interrupt == "u":
        u[...] = u * 0.5
    H.call_oracle("cal_arith_tau", *dims, s, rmu, tau)
    if interrupt == "rmu_after_tau":
        rmu[...] = rmu * 1.5
    H.call_oracle("cal_div_tensor", *dims, *inv, tau, vu, vv, vw)
    return [u, v, w, rmu, s, tau, vu, vv, vw]


def rescale_upload(reg, a, f):
    reg.download(a)
    a[...] = a * f
    H.gpu().felgpu_upload_(ctypes.c_void_p(a.ctypes.data))


@pytest.mark.parametrize("interrupt", [None, "read_s", "rmu", "u", "read_tau", "rmu_after_tau"])
def test_viscous_chain_equals_the_three_calls(interrupt):
    g = Grid(20, 12, 16, 0.1, 0.05, 0.08)
    host = visc_setup(g, 50)
    want = oracle_chain(g, host, interrupt)
    reg = Registered()
    try:
        dev = [reg.add(a.copy()) for a in host]
        u, v, w, rmu, s, tau, vu, vv, vw = dev
        dims = [I(g.ni), I(g.nj), I(g.nk)]
        inv = [D(1 / g.dx), D(1 / g.dy), D(1 / g.dz)]
        H.call_gpu("calsij", *dims, *inv, u, v, w, s)
        if interrupt == "read_s":
            reg.download(s)          # runs calsij here
        if interrupt == "rmu":       # the viscosity changes before arith_tau
            rescale_upload(reg, rmu, 1.5)
        if interrupt == "u":         # a source of the recorded s changes: s is made first
            rescale_upload(reg, u, 0.5)
        H.call_gpu("cal_arith_tau", *dims, s, rmu, tau)
        if interrupt == "read_tau":
            reg.download(tau)        # runs calsij and arith_tau here
        if interrupt == "rmu_after_tau":  # a source of the recorded tau changes
            rescale_upload(reg, rmu, 1.5)
        H.call_gpu("cal_div_tensor", *dims, *inv, tau, vu, vv, vw)
        names = ["u", "v", "w", "rmu", "s", "tau", "vis_u", "vis_v", "vis_w"]
        for name, x, y in zip(names, dev, want):
            assert same_bits(reg.download(x), y), name
    finally:
        reg.close()


def test_viscous_chain_sources_rewritten_after_the_chain():
    """A velocity rewritten after the chain does not disturb its results."""
    g = Grid(16, 8, 12, 0.1, 0.1, 0.1)
    host = visc_setup(g, 51)
    want = oracle_chain(g, host)
    reg = Registered()
    try:
        dev = [reg.add(a.copy()) for a in host]
        u, v, w, rmu, s, tau, vu, vv, vw = dev
        dims = [I(g.ni), I(g.nj), I(g.nk)]
        inv = [D(1 / g.dx), D(1 / g.dy), D(1 / g.dz)]
        H.call_gpu("calsij", *dims, *inv, u, v, w, s)
        H.call_gpu("cal_arith_tau", *dims, s, rmu, tau)
        H.call_gpu("cal_div_tensor", *dims, *inv, tau, vu, vv, vw)
        H.call_gpu("cpy", *dims, rmu, u)
        for x, y in zip([s, tau, vu, vv, vw], want[4:]):
            assert same_bits(reg.download(x), y)
    finally:
        reg.close()


NID = np.full(6, -1, dtype=np.int32)


def test_interior_write_of_a_deferred_copy():
    """pn deferred on p (cpy_ pn -> p) and then written by corunp_explicit_
    on the interior only: its ghost cells keep pn's contents, made from the
    deferred copy without copying the interior."""
    g = Grid(16, 10, 12, 0.1, 0.05, 0.08)
    r = np.random.default_rng(52)
    host = [r.standard_normal(g.shape) for _ in range(7)]  # dp, phat, un, vn, wn, pn, p
    want = [a.copy() for a in host]
    args = [NID.copy(), I(g.ni), I(g.nj), I(g.nk), D(0.81), D(1 / g.dx), D(1 / g.dy),
            D(1 / g.dz), D(0.003)]
    H.call_oracle("corunp_explicit", *args, *want[:6])
    want[6] = host[5].copy()
    reg = Registered()
    try:
        dev = [reg.add(a.copy()) for a in host]
        H.call_gpu("cpy", I(g.ni), I(g.nj), I(g.nk), dev[5], dev[6])  # p <- pn (pn deferred)
        H.call_gpu("corunp_explicit", *args, *dev[:6])
        for name, x, y in zip(["dp", "phat", "un", "vn", "wn", "pn", "p"], dev, want):
            assert same_bits(reg.download(x), y), name
    finally:
        reg.close()


def test_coefficients_after_their_viscosity_is_overwritten():
    """cal_arith_coef_vis_ records its 45 outputs as a function of the
    viscosity; cal_arith_mean_ then overwrites the viscosity as a whole (the
    record takes over the old buffer); a later read of a coefficient array
    makes them from the old viscosity."""
    g = Grid(12, 8, 10, 0.1, 0.05, 0.08)
    r = np.random.default_rng(53)
    rmu = r.uniform(0.01, 0.2, g.shape)
    phi = r.random(g.shape)
    dims = [I(g.ni), I(g.nj), I(g.nk)]
    inv = [D(1 / g.dx), D(1 / g.dy), D(1 / g.dz)]
    want = [g.zeros() for _ in range(45)]
    H.call_oracle("cal_arith_coef_vis", *dims, *inv, rmu.copy(), *want)
    want_mu = g.zeros()
    H.call_oracle("cal_arith_mean", *dims, phi.copy(), D(1e-3), D(1e-5), want_mu)
    reg = Registered()
    try:
        dmu = reg.add(rmu.copy())
        dphi = reg.add(phi.copy())
        coef = [reg.add(r.standard_normal(g.shape)) for _ in range(45)]
        H.call_gpu("cal_arith_coef_vis", *dims, *inv, dmu, *coef)
        H.call_gpu("cal_arith_mean", *dims, dphi, D(1e-3), D(1e-5), dmu)
        assert same_bits(reg.download(dmu), want_mu)
        for m in (0, 7, 22, 44):
            assert same_bits(reg.download(coef[m]), want[m]), m
    finally:
        reg.close()
