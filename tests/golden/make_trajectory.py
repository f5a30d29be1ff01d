"""Writes trajectory fixtures from the reference's own program.

Each case of oracle/ref_main.py (the reference's main.f90 with its parameter
assignments set to the case's values) is built and run with one MPI rank and
one thread (serial sums), writing its restart record every step
(srcs/dataou.f) and its four monitor lines every step (imon_t = 1):
AV_PHI (srcs/main.f90:519-524), Err_solu (srcs/solu_sor4.f:298-305),
Err_solp (srcs/solp_fft_tdma4.f:88-93) and Err_div (srcs/main.f90:675-684).
traj_<case>.npz holds

  params      JSON: the case's parameter assignments (Fortran text)
  steps       the steps whose fields are stored
  s<n>_<f>    field f in {u, v, w, p, phi0, phi1} after step n, interior
              cells only, shape (nk, nj, ni)
  mon_<name>  per-step monitor values (index = step - 1): av_phi,
              sor_lp, sor_ratio, sor_err, sor_err0, solp_ratio, solp_err,
              solp_err0, div_ratio, div_err, div_err0 (10 significant digits,
              as printed)

Cases of RESTART (droplet configurations of felmpf.config, which the
reference's program reaches through its restart path) additionally hold
`spec` (JSON: the configuration) and store u, v, w, p and phi(.,1).

Regenerate with

  make -C oracle ref && python tests/golden/make_trajectory.py
"""
from __future__ import annotations

import json
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.join(REPO, "fel-multi-phase-flow_amd"))

import ref_main as R  # noqa: E402

STORED = {"couette_32x8x32": (1, 20, 100), "variable_32x8x32": (1, 20, 100),
          "thin_32x8x32": (1, 20)}

# Droplet configurations of felmpf.config run by the reference's program from
# a restart record of their initial state (irestart = 1; ref_main.py's
# restart_params and write_restart): name -> (configuration spec, steps,
# stored steps).  The spec is the factory name, the grid and, optionally,
# a uniform bottom-wall angle; the last two cases blow up in the reference
# (steps 25 and 16), so only their monitor lines are stored.
RESTART = {
    "sessile_32cube_restart": ({"config": "sessile_droplet", "grid": [32, 32, 32]}, 60, (1, 60)),
    "sessile_64cube_restart": ({"config": "sessile_droplet", "grid": [64, 64, 64]}, 28, ()),
    "wall120_64x32x64_restart": ({"config": "patterned_wall", "grid": [64, 32, 64],
                                  "bottom_theta": 120.0}, 20, ()),
}


def restart_config(spec):
    """The run configuration of a restart case (also used by the tests)."""
    from dataclasses import replace
    from felmpf import config as C
    cfg = C.CONFIGS[spec["config"]]().scaled(*spec["grid"])
    if "bottom_theta" in spec:
        cfg = replace(cfg, bottom=C.WallSide(spec["bottom_theta"], cfg.bottom.slip))
    return cfg


FIELDS = ("u", "v", "w", "p", "uo", "vo", "wo", "po")


def grid_of(params):
    nsv = int(params["nsv"])
    ev = {"nsv": nsv}
    return tuple(int(eval(params[f"svall({m})"], {}, ev)) for m in (1, 2, 3))


def read_record(path, ni, nj, nk, nbub=1):
    """One restart record of srcs/dataou.f:24-38 (unformatted sequential:
    4-byte length markers around u, v, w, p, uo, vo, wo, po, phi(0:nbub),
    time, nstep)."""
    n3 = (ni + 6) * (nj + 6) * (nk + 6)
    raw = open(path, "rb").read()
    length = np.frombuffer(raw[:4], dtype=np.int32)[0]
    assert length == 8 * n3 * (8 + nbub + 1) + 12, (length, n3)
    body = raw[4:4 + length]
    a = np.frombuffer(body[:8 * n3 * (9 + nbub)], dtype=np.float64)
    shape = (nk + 6, nj + 6, ni + 6)
    out = {f: a[m * n3:(m + 1) * n3].reshape(shape) for m, f in enumerate(FIELDS)}
    for l in range(nbub + 1):
        out[f"phi{l}"] = a[(8 + l) * n3:(9 + l) * n3].reshape(shape)
    out["time"] = np.frombuffer(body[8 * n3 * (9 + nbub):8 * n3 * (9 + nbub) + 8], np.float64)[0]
    out["nstep"] = int(np.frombuffer(body[-4:], np.int32)[0])
    return out


def parse_monitors(lines, nmax, allow_nan=False):
    mon = {k: np.full(nmax, np.nan) for k in (
        "av_phi", "sor_lp", "sor_ratio", "sor_err", "sor_err0", "solp_ratio", "solp_err",
        "solp_err0", "div_ratio", "div_err", "div_err0")}
    for line in lines:
        t = line.split()
        if not t:
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
    if not allow_nan:
        for k, v in mon.items():
            assert not np.isnan(v).any(), f"monitor {k} missing for some step"
    return mon


def make(case):
    params = R.case_params(case)
    ni, nj, nk = grid_of(params)
    nmax = int(params["nmax"])
    R.build(case)
    wd = tempfile.mkdtemp(prefix=f"traj_{case}_")
    try:
        lines = [l for _, l in R.run(case, wd, threads=1)]
        mon = parse_monitors(lines, nmax)
        arrs = {f"mon_{k}": v for k, v in mon.items()}
        for s in STORED[case]:
            rec = read_record(os.path.join(wd, f"x{s:07d}"), ni, nj, nk)
            assert rec["nstep"] == s
            for f in ("u", "v", "w", "p", "phi0", "phi1"):
                arrs[f"s{s}_{f}"] = rec[f][3:nk + 3, 3:nj + 3, 3:ni + 3].copy()
    finally:
        shutil.rmtree(wd, ignore_errors=True)
    path = os.path.join(HERE, f"traj_{case}.npz")
    np.savez_compressed(path, params=np.array(json.dumps(params)),
                        steps=np.array(STORED[case]), **arrs)
    print(f"{path}: {os.path.getsize(path) / 1024:.0f} KiB, lp {sorted(set(mon['sor_lp']))}")


def make_restart(case):
    from felmpf.config import initial_state
    spec, nmax, stored = RESTART[case]
    cfg = restart_config(spec)
    # restart records every step: the fields of the stored steps
    params = dict(R.restart_params(cfg, steps=nmax), idout="1")
    R.build(case, params=params)
    wd = tempfile.mkdtemp(prefix=f"traj_{case}_")
    try:
        R.write_restart(os.path.join(wd, "x.inp"), initial_state(cfg))
        lines = [l for _, l in R.run(case, wd, threads=1)]
        mon = parse_monitors(lines, nmax, allow_nan=True)
        arrs = {f"mon_{k}": v for k, v in mon.items()}
        for s in stored:
            rec = read_record(os.path.join(wd, f"x{s:07d}"), cfg.ni, cfg.nj, cfg.nk)
            assert rec["nstep"] == s
            for f in ("u", "v", "w", "p", "phi1"):
                arrs[f"s{s}_{f}"] = rec[f][3:cfg.nk + 3, 3:cfg.nj + 3, 3:cfg.ni + 3].copy()
    finally:
        shutil.rmtree(wd, ignore_errors=True)
    path = os.path.join(HERE, f"traj_{case}.npz")
    np.savez_compressed(path, params=np.array(json.dumps(params)), spec=np.array(json.dumps(spec)),
                        steps=np.array(stored, dtype=np.int64), **arrs)
    blow = np.flatnonzero(~(mon["sor_err0"] < 1e3))
    print(f"{path}: {os.path.getsize(path) / 1024:.0f} KiB, "
          f"first step with err0 >= 1e3 or NaN: {blow[0] + 1 if blow.size else None}")


def main(argv):
    for case in (argv or list(STORED) + list(RESTART)):
        if case in RESTART:
            make_restart(case)
        else:
            make(case)


if __name__ == "__main__":
    main(sys.argv[1:])
