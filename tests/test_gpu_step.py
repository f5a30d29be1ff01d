"""Whole-timestep parity: the Fortran host driving the GPU library against
the oracle's timestep, on the benchmark configurations at reduced grids
(same physics, cubic cells).  Bar: u, v, w, p, phi within 1e-10 relative
(max norm) after one step and within 1e-8 after 20; the SOR iteration count
agrees; phi stays in [0,1]; the liquid volume is conserved to 1e-12.

The Couette configuration saturates the dynamic contact angle at 0/180 deg
from the first step (zeta*u_cl/sigma ~ 3 > 1), where the reference's
contact-angle ghost divides by sin(theta) + 1e-20
(srcs/bnd_contact_angle.f90:34-46) and amplifies last-bit differences of
sin/cos/acos by ~1e16 in the wall ghosts of phi; it is held to the bar for
one step and to the invariants over 20."""
import numpy as np
import pytest

import felmpf
from felmpf import config as C
from oracle_run import OracleRun, rel_err

pytestmark = pytest.mark.gpu

FIELDS = ["u", "v", "w", "p", "phi0", "phi1"]

CASES = {
    "droplet_impact": (lambda: C.droplet_impact().scaled(64, 16, 32), 20),
    "sessile_droplet": (lambda: C.sessile_droplet().scaled(16, 16, 16), 20),
    "couette_stripe": (lambda: C.couette_stripe().scaled(64, 32, 16), 1),
    "patterned_wall": (lambda: C.patterned_wall().scaled(32, 16, 32), 20),
}


def interior(cfg, a):
    return a[3:cfg.nk + 3, 3:cfg.nj + 3, 3:cfg.ni + 3]


@pytest.mark.parametrize("case", list(CASES))
def test_step_parity(case):
    make, horizon = CASES[case]
    cfg = make()
    g = felmpf.GpuRun(cfg)
    o = OracleRun(cfg)
    vol0 = interior(cfg, o.field("phi1")).sum()
    try:
        done = 0
        for n, tol in ((1, 1e-10), (19, 1e-8)):
            g.steps(n)
            o.steps(n)
            done += n
            sg, so = g.state(), o.state()
            if done <= horizon:
                assert sg["sor_lp"] == so["sor_lp"], (sg, so)
                for f in FIELDS:
                    a, b = o.field(f), g.field(f)
                    e = rel_err(interior(cfg, a), interior(cfg, b))
                    assert e <= tol, (case, f, e)
            phi = interior(cfg, g.field("phi1"))
            assert phi.min() >= 0.0 and phi.max() <= 1.0
            assert abs(phi.sum() - vol0) <= 1e-12 * vol0
    finally:
        g.close()
        o.close()


# Stated sizes: the sessile droplet at its configured 64^3, the droplet
# impact at 128x32x64 (1/64 of the benchmark's cells) with result reuse on
# and off, each for 20 steps.  The sessile droplet is held to 20 of its 100
# steps: the reference's own program blows up on it at step 25
# (tests/golden/traj_sessile_64cube_restart.npz, test_gpu_trajectory.py), as
# this timestep does.  Bars: 1e-10 after one step, 1e-8 at the last; the
# sweep count agrees at each check.
LONG = {
    "sessile_droplet_64cube": (lambda: C.sessile_droplet(), (1, 20), True),
    "droplet_impact_128x32x64_reuse": (lambda: C.droplet_impact().scaled(128, 32, 64), (1, 20), True),
    "droplet_impact_128x32x64_no_reuse": (lambda: C.droplet_impact().scaled(128, 32, 64), (1, 20),
                                          False),
    # 32 droplets on x-striped walls with the CFL step (caldt every step)
    "scale_run_64x32x64": (lambda: C.scale_run().scaled(64, 32, 64), (1, 20), True),
}


@pytest.mark.parametrize("case", list(LONG))
def test_step_parity_at_stated_size(case):
    make, checks, reuse = LONG[case]
    cfg = make()
    print(f"case {case} grid {cfg.ni}x{cfg.nj}x{cfg.nk}", flush=True)
    g = felmpf.GpuRun(cfg, reuse=reuse)
    o = OracleRun(cfg)
    vol0 = interior(cfg, o.field("phi1")).sum()
    try:
        done = 0
        for n in checks:
            g.steps(n - done)
            o.steps(n - done)
            done = n
            assert g.state()["sor_lp"] == o.state()["sor_lp"]
            tol = 1e-10 if n == 1 else 1e-8
            for f in FIELDS:
                e = rel_err(interior(cfg, o.field(f)), interior(cfg, g.field(f)))
                print(f"  step {n} {f} {e:.2e}", flush=True)
                assert e <= tol, (case, n, f, e)
            phi = interior(cfg, g.field("phi1"))
            assert phi.min() >= 0.0 and phi.max() <= 1.0
            assert abs(phi.sum() - vol0) <= 1e-12 * vol0
    finally:
        g.close()
        o.close()
