"""The trajectory fixtures written by the reference's own program
(tests/golden/make_trajectory.py, cases of oracle/ref_main.py) and the run
configurations that reproduce them (felmpf.config.reference_program).
Test infrastructure."""
from __future__ import annotations

import json
import os

import numpy as np

from felmpf import config as C

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# fixture case -> keyword arguments of config.reference_program
CASES = {
    "couette_32x8x32": dict(),
    "variable_32x8x32": dict(rhol=1.0, rhog=0.1, rmul=1.95, rmug=0.195, uwall=0.5,
                             theta_0_a=75.0, theta_0_b=55.0, period=4, ratio_a=1, grv=0.2),
    "thin_32x8x32": dict(yl=3.4, zl=4.25, steps=20),
}

FIELDS = ("u", "v", "w", "p", "phi0", "phi1")

# Droplet configurations the reference's program ran from a restart record
# of their initial state (tests/golden/make_trajectory.py, RESTART): case ->
# the step at which the reference blows up (err0 of its Gauss-Seidel monitor
# >= 1e3 or NaN), or None.
RESTART_CASES = {"sessile_32cube_restart": None, "sessile_64cube_restart": 25,
                 "wall120_64x32x64_restart": 16}
RESTART_FIELDS = ("u", "v", "w", "p", "phi1")


def restart_config(fx) -> C.RunConfig:
    """The run configuration a restart fixture was written from, with the
    wall maps the reference's program builds from its wall parameters
    (config.with_reference_walls: x-stripes of the top wall)."""
    from dataclasses import replace
    spec = json.loads(str(fx["spec"]))
    cfg = C.CONFIGS[spec["config"]]().scaled(*spec["grid"])
    if "bottom_theta" in spec:
        cfg = replace(cfg, bottom=C.WallSide(spec["bottom_theta"], cfg.bottom.slip))
    return C.with_reference_walls(cfg)


def write_restart_record(path: str, cfg: C.RunConfig):
    """The restart record the reference's program started from (its format,
    srcs/datain.f:20-39, from oracle/ref_main.py)."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(GOLDEN), "..", "oracle"))
    import ref_main
    ref_main.write_restart(path, C.initial_state(cfg))


def fixture(case: str):
    return np.load(os.path.join(GOLDEN, f"traj_{case}.npz"))


def config(case: str, imon_t: int = 0) -> C.RunConfig:
    return C.reference_program(case, imon_t=imon_t, **CASES[case])


def fortran_real(text: str) -> float:
    return float(text.lower().replace("d", "e"))


def fixture_params(fx) -> dict:
    return json.loads(str(fx["params"]))


def interior(cfg, a):
    return a[3:cfg.nk + 3, 3:cfg.nj + 3, 3:cfg.ni + 3]


def max_rel(a, b):
    """max |a - b| / max |b| (b the reference's values)."""
    s = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b))) / (s if s > 0 else 1.0)


def parse_monitor_lines(text: str, nmax: int) -> dict:
    """The four monitor lines of each step (AV_PHI, Err_solu, Err_solp,
    Err_div) as arrays indexed by step - 1, the form of the fixture's
    mon_* entries."""
    keys = ("av_phi", "sor_lp", "sor_ratio", "sor_err", "sor_err0", "solp_ratio", "solp_err",
            "solp_err0", "div_ratio", "div_err", "div_err0")
    mon = {k: np.full(nmax, np.nan) for k in keys}
    for line in text.splitlines():
        t = line.split()
        if len(t) < 3:
            continue
        if t[0] == "AV_PHI":
            mon["av_phi"][int(t[1]) - 1] = float(t[2])
        elif t[0] == "Err_solu":
            n = int(t[1]) - 1
            mon["sor_lp"][n] = int(t[2])
            mon["sor_ratio"][n], mon["sor_err"][n], mon["sor_err0"][n] = map(float, t[3:6])
        elif t[0] == "Err_solp":
            n = int(t[1]) - 1
            mon["solp_ratio"][n], mon["solp_err"][n], mon["solp_err0"][n] = map(float, t[2:5])
        elif t[0] == "Err_div":
            n = int(t[1]) - 1
            mon["div_ratio"][n], mon["div_err"][n], mon["div_err0"][n] = map(float, t[2:5])
    return mon
