"""Pins the oracle's timestep (oracle/fel_oracle_step.c) to the reference's
own program.

The fixtures tests/golden/traj_*.npz were written by the reference's
main.f90 (compiled with its makefile sources, parameter assignments set per
case, one rank, one thread; tests/golden/make_trajectory.py).  The oracle
starts from the same configuration -- its initial state from the
restatements of srcs/init_phi_legendre.f90 and srcs/init_array_x_stripe.f90
in felmpf.config, themselves pinned here against the compiled reference --
and must follow the reference's trajectory: u, v, w, p and phi within 1e-12
relative (max norm) after step 1 and 1e-9 after 20 and 100 steps, the same
Gauss-Seidel sweep count every step and its err/err0 within the 10 digits
the reference prints.  The only differences allowed are those of the
reference's FFT (srcs/fourn.f's trigonometric recurrence against the
oracle's directly evaluated DFT, ~1e-15 per transform).
"""
import numpy as np
import pytest

import felharness as H
import reference_cases as RC
from felharness import D, I
from felmpf import config as C
from oracle_run import OracleRun

# parameter names of srcs/main.f90 -> (config attribute, parse)
PARAMS = {"xl": "xl", "yl": "yl", "zl": "zl", "rhol": "rhol", "rhog": "rhog", "rmul": "rmul",
          "rmug": "rmug", "uwall": "uwall", "surface_tension": "sigma"}


@pytest.mark.parametrize("case", list(RC.CASES))
def test_configuration_is_the_fixtures(case):
    """The run configuration restates the case's parameter assignments."""
    fx = RC.fixture(case)
    p = RC.fixture_params(fx)
    cfg = RC.config(case)
    nsv = int(p["nsv"])
    assert (cfg.ni, cfg.nj, cfg.nk) == tuple(
        int(eval(p[f"svall({m})"], {}, {"nsv": nsv})) for m in (1, 2, 3))
    assert cfg.params["nsv"] == nsv
    assert cfg.dt == 32.0e-2 / nsv
    defaults = C.reference_program()
    for name, attr in PARAMS.items():
        want = RC.fortran_real(p[name]) if name in p else getattr(defaults, attr)
        assert getattr(cfg, attr) == want, name
    xs = cfg.params["x_stripe"]
    if "theta_0_a" in p:
        assert xs["theta"][0] == RC.fortran_real(p["theta_0_a"])
    if "theta_0_b" in p:
        assert xs["theta"][1] == RC.fortran_real(p["theta_0_b"])
        assert xs["theta"][2] == 180.0 - RC.fortran_real(p["theta_0_b"])
    for key in ("period", "ratio_a"):
        if key in p:
            assert xs[key] == int(p[key])
    if "grv" in p:
        assert cfg.grvp == -RC.fortran_real(p["grv"])
    assert cfg.steps == int(p["nmax"])


@pytest.mark.skipif(not H.ref_available(), reason="compiled reference not built")
@pytest.mark.parametrize("period,ratio_a", [(1, 1), (4, 1), (3, 2)])
def test_initialisers_equal_the_reference(period, ratio_a):
    """felmpf.config's x-stripe maps and stripe phase equal the compiled
    srcs/init_array_x_stripe.f90 and srcs/init_phi_legendre.f90."""
    cfg = C.reference_program(period=period, ratio_a=ratio_a)
    q = np.zeros(cfg.shape)
    H.call_ref("init_array_x_stripe", I(cfg.ni), I(cfg.nj), I(cfg.nk), D(1.5), D(2.5), D(3.5), q,
               I(period), I(ratio_a))
    assert np.array_equal(q, C.x_stripe_map(cfg, 1.5, 2.5, 3.5, period, ratio_a))
    phi = np.zeros((2,) + cfg.shape)
    H.call_ref("init_phi", I(1), np.array([cfg.ni, cfg.nj, cfg.nk], dtype=np.int32), phi, I(1),
               I(1), I(cfg.params["nsv"]))
    assert np.array_equal(phi[1], C.initial_state(cfg)["phi1"])
    assert not phi[0].any()


@pytest.mark.parametrize("case", list(RC.CASES))
def test_oracle_follows_the_reference_program(case):
    fx = RC.fixture(case)
    cfg = RC.config(case)
    o = OracleRun(cfg)
    stored = [int(s) for s in fx["steps"]]
    try:
        for n in range(1, stored[-1] + 1):
            o.steps(1)
            st = o.state()
            assert st["sor_lp"] == fx["mon_sor_lp"][n - 1], (n, st)
            assert abs(st["sor_ratio"] / fx["mon_sor_ratio"][n - 1] - 1) <= 1e-9, (n, st)
            if n in stored:
                tol = 1e-12 if n == 1 else 1e-9
                for f in RC.FIELDS:
                    e = RC.max_rel(RC.interior(cfg, o.field(f)), fx[f"s{n}_{f}"])
                    assert e <= tol, (case, n, f, e)
    finally:
        o.close()


@pytest.mark.parametrize("case", list(RC.RESTART_CASES))
def test_oracle_follows_the_reference_program_from_a_restart(case, tmp_path):
    """Droplet configurations, started like the reference's program from a
    restart record (srcs/main.f90:375-404): the same sweep count and err/err0
    (to 1e-6: the last sweep's change is rounding-level here) every step
    before the reference blows up, the fields of the stored
    steps within 1e-12 (step 1) and 1e-9, and the blow-up at the
    reference's step, where the velocity jumps by six orders of magnitude
    or becomes NaN."""
    fx = RC.fixture(case)
    cfg = RC.restart_config(fx)
    blow = RC.RESTART_CASES[case]
    path = str(tmp_path / "x.inp")
    RC.write_restart_record(path, cfg)
    o = OracleRun(cfg)
    o.datain(path)
    stored = [int(s) for s in fx["steps"]]
    last = blow if blow else len(fx["mon_sor_lp"])
    try:
        umax_prev = None
        for n in range(1, last + 1):
            o.steps(1)
            st = o.state()
            u = RC.interior(cfg, o.field("u"))
            umax = float(np.max(np.abs(u)))
            if n == blow:
                assert not (umax < 1e6 * umax_prev), (n, umax, umax_prev)
                break
            assert st["sor_lp"] == fx["mon_sor_lp"][n - 1], (n, st)
            # err of the converged last sweep (~1e-9 of the velocities) is
            # rounding-level: its ratio to err0 agrees to ~1e-7
            assert abs(st["sor_ratio"] / fx["mon_sor_ratio"][n - 1] - 1) <= 1e-6, (n, st)
            umax_prev = umax
            if n in stored:
                tol = 1e-12 if n == 1 else 1e-9
                for f in RC.RESTART_FIELDS:
                    e = RC.max_rel(RC.interior(cfg, o.field(f)), fx[f"s{n}_{f}"])
                    assert e <= tol, (case, n, f, e)
    finally:
        o.close()
