"""caldt_ (srcs/caldt.f) on the GPU against the oracle, and a run with the
adaptive step (fel_cfg.cfl > 0, the call srcs/main.f90:479 leaves commented
out) against the oracle's timestep with the same setting.  The largest
advective rate is a maximum, exact in any order, so dt and time agree bit
for bit on the same fields (and after one step); later the fields differ
at the whole-step bars (1e-10 after one step, 1e-8 after 20) and dt
follows them."""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import felharness as H
from felharness import D, I, Grid
from felmpf import config as C
from felmpf.driver import GpuRun
from oracle_run import OracleRun, rel_err

pytestmark = pytest.mark.gpu


def _call(fn, g, nid, u, v, w, cfl, props, time):
    dt, t = D(0.0), D(time)
    args = [I(0), nid, I(0), I(1), I(g.ni), I(g.nj), I(g.nk), I(3), I(0), D(1 / g.dx), D(1 / g.dy),
            D(1 / g.dz), D(cfl), *[D(x) for x in props], u, v, w, dt, t]
    if fn == "oracle":
        H.call_oracle("caldt", *args, np.zeros(3))
    else:
        H.call_gpu("caldt", *args)
    return dt.value, t.value


@pytest.mark.parametrize("walls", [True, False])
@pytest.mark.parametrize("scale", [1.0, 1e-6, 0.0])
def test_caldt_matches_oracle(walls, scale):
    g = Grid(40, 24, 16, 0.05, 0.05, 0.1)
    rng = np.random.default_rng(90)
    u, v, w = (scale * rng.standard_normal(g.shape) for _ in range(3))
    v[:, g.nj + 2, :] = 50.0 * scale  # the row j = nj, excluded below a Y_PLUS wall
    nid = np.full(6, -1 if walls else 0, dtype=np.int32)
    props = (1.0, 0.1, 1e-3, 1e-5, 1e-2)
    want = _call("oracle", g, nid, u, v, w, 0.05, props, 0.25)
    got = _call("gpu", g, nid, u, v, w, 0.05, props, 0.25)
    assert got == want


@pytest.mark.parametrize("nsteps,tol", [(1, 1e-10), (20, 1e-8)])
def test_adaptive_step_run_matches_oracle(nsteps, tol):
    cfg = dataclasses.replace(C.droplet_impact().scaled(64, 16, 32), cfl=0.05)
    g = GpuRun(cfg)
    o = OracleRun(cfg)
    try:
        g.steps(nsteps)
        o.steps(nsteps)
        sg, so = g.state(), o.state()
        if nsteps == 1:
            assert sg["dt"] == so["dt"] and sg["time"] == so["time"]
        assert abs(sg["dt"] - so["dt"]) <= tol * so["dt"]
        assert abs(sg["time"] - so["time"]) <= tol * so["time"]
        assert sg["dt"] != cfg.dt  # the adaptive step is in use
        for f in ("u", "v", "w", "p", "phi1"):
            e = rel_err(o.field(f)[3:-3, 3:-3, 3:-3], g.field(f)[3:-3, 3:-3, 3:-3])
            assert e <= tol, (f, e)
    finally:
        g.close()
        o.close()
