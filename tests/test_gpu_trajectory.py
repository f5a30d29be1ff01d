"""The GPU timestep against the reference's own program.

The fixtures tests/golden/traj_*.npz were written by the reference's main.f90
(tests/golden/make_trajectory.py): its restart record after steps 1, 20 and
100 and its monitor lines of every step.  The Fortran host
(host/fel_driver.f90) drives the GPU library from the same initial state with
the monitors on every step (imon_t = 1) and must reproduce:

  * u, v, w, p, phi(.,0) and phi(.,1) within 1e-10 relative (max norm) after
    step 1 and within 1e-8 after 20 and 100 steps;
  * every monitor line (srcs/main.f90:519-524,675-684, srcs/solu_sor4.f:298-305,
    srcs/solp_fft_tdma4.f:88-93): the Gauss-Seidel sweep count exactly, and
    AV_PHI, err/err0 of the sweeps, the RMS divergence before the projection
    (err_div0 = err0 of Err_solp) within 1e-8 relative of the printed values;
    the residuals left by the solves (err of Err_solp and Err_div, rounding
    noise of 1e-15 to 1.4e-12 of err0 in the reference's lines) below 1e-10
    of err0 in both;
  * the monitor lines in the reference's format (Fortran E20.10, 0.dddE+xx).

Each case runs with result reuse on and with it off.
"""
import ctypes
import re

import numpy as np
import pytest

import felmpf
import reference_cases as RC

pytestmark = pytest.mark.gpu

E20 = r" +-?0\.\d{10}E[+-]\d\d"
LINE = {
    "Err_solu": re.compile(rf"^Err_solu +\d+ +\d+(?:{E20}){{3}}$"),
    "Err_solp": re.compile(rf"^Err_solp +\d+(?:{E20}){{3}}$"),
}


def flush_all(run):
    run.drv.fel_drv_flush()
    ctypes.CDLL(None).fflush(None)


@pytest.mark.parametrize("reuse", [True, False], ids=["reuse", "no_reuse"])
@pytest.mark.parametrize("case", list(RC.CASES))
def test_gpu_follows_the_reference_program(case, reuse, capfd):
    fx = RC.fixture(case)
    cfg = RC.config(case, imon_t=1)
    stored = [int(s) for s in fx["steps"]]
    nmax = stored[-1]
    print(f"case {case} reuse={reuse} grid {cfg.ni}x{cfg.nj}x{cfg.nk}", flush=True)
    g = felmpf.GpuRun(cfg, reuse=reuse)
    try:
        capfd.readouterr()
        done = 0
        for n in stored:
            g.steps(n - done)
            done = n
            tol = 1e-10 if n == 1 else 1e-8
            for f in RC.FIELDS:
                e = RC.max_rel(RC.interior(cfg, g.field(f)), fx[f"s{n}_{f}"])
                assert e <= tol, (case, n, f, e)
        flush_all(g)
        text = capfd.readouterr().out
    finally:
        g.close()
    for line in text.splitlines():
        for key, pat in LINE.items():
            if line.startswith(key):
                assert pat.match(line), line
    mon = RC.parse_monitor_lines(text, nmax)
    for k, v in mon.items():
        assert not np.isnan(v).any(), f"monitor {k} not printed every step"
    assert np.array_equal(mon["sor_lp"], fx["mon_sor_lp"])
    for k in ("av_phi", "sor_ratio", "sor_err", "sor_err0", "solp_err0", "div_err0"):
        e = np.max(np.abs(mon[k] / fx[f"mon_{k}"] - 1.0))
        assert e <= 1e-8, (k, e)
    for k, k0 in (("solp_err", "solp_err0"), ("div_err", "div_err0")):
        assert np.all(mon[k] <= 1e-10 * mon[k0]), k
        assert np.all(fx[f"mon_{k}"] <= 1e-10 * fx[f"mon_{k0}"]), k


@pytest.mark.parametrize("case", list(RC.RESTART_CASES))
def test_gpu_follows_the_reference_program_from_a_restart(case, tmp_path, capfd):
    """Droplet configurations started, like the reference's program, from a
    restart record (fel_drv_datain, srcs/main.f90:375-404): the fields of the
    stored steps within 1e-10 (step 1) and 1e-8, every monitor line before
    the reference blows up within 1e-8 (sweep counts exact; the converged
    last sweep's err and err/err0, rounding-level here, within 1e-6), and
    the blow-up
    at the reference's step (err0 of the Gauss-Seidel monitor >= 1e3 or NaN
    in both)."""
    fx = RC.fixture(case)
    cfg = RC.restart_config(fx)
    cfg.imon_t = 1
    blow = RC.RESTART_CASES[case]
    nsteps = blow if blow else len(fx["mon_sor_lp"])
    stored = [int(s) for s in fx["steps"]]
    path = str(tmp_path / "x.inp")
    RC.write_restart_record(path, cfg)
    print(f"case {case} grid {cfg.ni}x{cfg.nj}x{cfg.nk} steps {nsteps}", flush=True)
    g = felmpf.GpuRun(cfg)
    try:
        g.datain(path)
        flush_all(g)
        capfd.readouterr()
        done = 0
        for n in stored + [nsteps]:
            if n <= done:
                continue
            g.steps(n - done)
            done = n
            if n in stored:
                tol = 1e-10 if n == 1 else 1e-8
                for f in RC.RESTART_FIELDS:
                    e = RC.max_rel(RC.interior(cfg, g.field(f)), fx[f"s{n}_{f}"])
                    assert e <= tol, (case, n, f, e)
        flush_all(g)
        text = capfd.readouterr().out
    finally:
        g.close()
    mon = RC.parse_monitor_lines(text, nsteps)
    good = slice(0, nsteps - 1) if blow else slice(0, nsteps)
    assert np.array_equal(mon["sor_lp"][good], fx["mon_sor_lp"][good])
    for k in ("av_phi", "sor_ratio", "sor_err", "sor_err0", "solp_err0", "div_err0"):
        e = np.max(np.abs(mon[k][good] / fx[f"mon_{k}"][good] - 1.0))
        # err of the converged last sweep is ~1e-9 of the velocities here:
        # rounding-level, so it and err/err0 agree to ~1e-7
        assert e <= (1e-6 if k in ("sor_ratio", "sor_err") else 1e-8), (k, e)
    if blow:
        assert not (mon["sor_err0"][blow - 1] < 1e3), mon["sor_err0"][blow - 1]
        assert not (fx["mon_sor_err0"][blow - 1] < 1e3)
