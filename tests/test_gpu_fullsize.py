"""The timestep at full size: droplet impact, 512x256x128
(reference i,j,k = 512x128x256, 16.8 M cells), through properties that hold
at any size -- phi stays in [0,1], the liquid volume is conserved by the
mass correction (srcs/solphi_mthinc3.f:101-160) to 1e-12, the projected
velocity is divergence-free to the Poisson solver's precision
(srcs/corunp_explicit.f after srcs/solp_fft_tdma1-3.f), every field stays
finite, and runs with and without result reuse are bitwise equal; the FP32
mode is held to its own bars, also on the scale run (512x512x256, 67 M
cells).  The oracle is too slow at these sizes;
tests/test_gpu_step.py holds the GPU path to it on smaller grids.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "phi1")


def _run(steps, reuse=True, precision=64, fp32_modes=False, config="droplet_impact"):
    from felmpf import config as C
    from felmpf.driver import GpuRun
    from felmpf.longrun import diagnostics
    cfg = C.CONFIGS[config]()
    run = GpuRun(cfg, reuse=reuse, precision=precision, fp32_modes=fp32_modes)
    try:
        d0 = diagnostics(cfg, run)
        run.steps(steps)
        d = diagnostics(cfg, run)
        fields = {f: run.field(f) for f in FIELDS}
        return cfg, d0, d, fields
    finally:
        run.close()


def test_full_size_step_invariants():
    cfg, d0, d, fields = _run(5)
    assert (cfg.ni, cfg.nj, cfg.nk) == (512, 128, 256)
    phi = fields["phi1"][3:-3, 3:-3, 3:-3]
    assert phi.min() >= 0.0 and phi.max() <= 1.0
    assert abs(d["volume"] - d0["volume"]) <= 1e-12 * d0["volume"], (d["volume"], d0["volume"])
    # max |div u| after the projection; the velocity is of order 1 and
    # dx = 1/128, so 1e-11 leaves a wide margin over FP64 rounding (~1e-14)
    assert d["max_div"] <= 1e-11, d["max_div"]
    for f, a in fields.items():
        assert np.all(np.isfinite(a)), f


def test_full_size_reuse_is_bitwise_neutral():
    _, _, _, on = _run(2, reuse=True)
    _, _, _, off = _run(2, reuse=False)
    for f in FIELDS:
        assert np.array_equal(on[f].view(np.int64), off[f].view(np.int64)), f


def test_full_size_fp32_mode_invariants():
    """The FP32 mode (FP32+refine) at full size: phi in [0,1], volume within
    the FP32 mode's 1e-10 bar (the volume fraction and its mass correction
    stay FP64), max |div u| at FP32 rounding of order-1 velocities (1.4e-5
    measured over 1,000 steps), every field finite."""
    _, d0, d, fields = _run(5, precision=32, fp32_modes=True)
    phi = fields["phi1"][3:-3, 3:-3, 3:-3]
    assert phi.min() >= 0.0 and phi.max() <= 1.0
    assert abs(d["volume"] - d0["volume"]) <= 1e-10 * d0["volume"], (d["volume"], d0["volume"])
    assert d["max_div"] <= 1e-4, d["max_div"]
    for f, a in fields.items():
        assert np.all(np.isfinite(a)), f


def test_scale_run_fp32_refine_full_size():
    """The scale run at its stated 512x512x256 (reference 512x256x512, 67 M
    cells: 32 seeded droplets, bottom x-stripes of 60/100 deg, CFL step) in
    the FP32+refine mode for 3 steps: phi in [0,1], volume within 1e-10,
    max |div u| at FP32 rounding, every field finite."""
    print("scale_run 512x256x512 fp32+refine, 3 steps", flush=True)
    cfg, d0, d, fields = _run(3, precision=32, fp32_modes=True, config="scale_run")
    assert (cfg.ni, cfg.nj, cfg.nk) == (512, 256, 512)
    phi = fields["phi1"][3:-3, 3:-3, 3:-3]
    assert phi.min() >= 0.0 and phi.max() <= 1.0
    assert abs(d["volume"] - d0["volume"]) <= 1e-10 * d0["volume"], (d["volume"], d0["volume"])
    assert d["max_div"] <= 1e-4, d["max_div"]
    for f, a in fields.items():
        assert np.all(np.isfinite(a)), f


def test_fp32_refine_tracks_fp64_over_1000_steps():
    """FP32+refine against the FP64 path over 1,000 steps of the droplet
    impact at full size (dt 1e-4): the velocities' relative L2 difference
    stays <= 1e-5 at every checkpoint (4e-7 measured, not growing), the
    FP32 run's volume drift <= 1e-10 and its max |div u| at FP32 rounding
    (1.5e-5 measured)."""
    from felmpf.longrun import mixed_precision
    print("droplet_impact 512x256x128, 1000 steps, FP64 and FP32+refine", flush=True)
    out = mixed_precision("droplet_impact", steps=1000, every=500)
    for row in out["checkpoints"]:
        for f, e in row["velocity_rel_l2"].items():
            assert e <= 1e-5, (row["step"], f, e)
        assert row["volume_drift"]["fp32"] <= 1e-10, row
        assert row["max_div"]["fp32"] <= 1e-4, row
