"""The tangent mode (forward-mode differentiation of the discrete timestep).

Pins:
* values: a tangent run computes the values with the FP64 path's
  operations, so every field is bitwise equal to an FP64 run's, and a run
  with zero seeds has exactly zero derivatives;
* derivatives: the derivative of the fields after a window of steps along a
  wall-parameter direction (contact angle of a patch group, slip length)
  agrees with finite differences of FP64 runs to 1e-5 relative (L2 over the
  interior, against the derivative's norm).  The differences are central
  and Richardson-extrapolated, (4 D(h) - D(2h))/3 with D(h) = (f(p+h) -
  f(p-h))/2h, whose error is O(h^4) plus the rounding error ~1e-16/h of the
  fields: h = 0.005 deg for an angle and 1% of the slip length keep both
  below the bar even for the volume fraction, whose interior derivative is
  small next to its O(1) values.  (Much larger steps cross kinks of the
  discrete map -- the sign of the contact-line velocity in GNBC, clamps --
  where no derivative exists; the window is short enough that h = 0.005
  deg crosses none);
* the autograd harness: gradients of droplet observables (liquid centroid,
  wetted area and its second moment) through StepWindow equal the central
  differences of the observables, to 1e-5 relative.
"""
from __future__ import annotations

import numpy as np
import pytest

from felmpf import config as C
from felmpf import sensitivity as S
from felmpf.driver import GpuRun

pytestmark = pytest.mark.gpu

FIELDS = ["u", "v", "w", "p", "phi1"]


def _interior(cfg, a):
    return a[3:cfg.nk + 3, 3:cfg.nj + 3, 3:cfg.ni + 3]


def _sessile():
    return C.sessile_droplet().scaled(32, 32, 32)


def _patterned():
    return C.patterned_wall().scaled(32, 16, 32)


def test_tangent_values_are_the_fp64_values():
    cfg = _sessile()
    ref = GpuRun(cfg)
    ref.steps(4)
    want = {f: ref.field(f) for f in FIELDS + ["dp", "theta_array"]}
    ref.close()
    run = GpuRun(cfg, tangent=True)
    try:
        run.steps(4)
        for f, a in want.items():
            assert np.array_equal(run.field(f), a), f
            assert not np.any(run.tangent(f)), f
    finally:
        run.close()


def _perturbed(cfg, seed, h, nsteps, read):
    run = GpuRun(cfg)
    try:
        S.apply_perturbation(run, seed, h)
        run.steps(nsteps)
        return read(run)
    finally:
        run.close()


def richardson(cfg, seed, h, nsteps, read):
    """(4 D(h) - D(2h))/3 of read(run) (a dict of arrays or an array)."""
    f = {s: _perturbed(cfg, seed, s * h, nsteps, read) for s in (1, -1, 2, -2)}

    def comb(a1, m1, a2, m2):
        return (4.0 * (a1 - m1) / (2.0 * h) - (a2 - m2) / (4.0 * h)) / 3.0

    if isinstance(f[1], dict):
        return {k: comb(f[1][k], f[-1][k], f[2][k], f[-2][k]) for k in f[1]}
    return comb(f[1], f[-1], f[2], f[-2])


def step_h(cfg, seed):
    return 0.005 if "theta_0_array" in seed else 0.01 * cfg.bottom.slip


def _fd_fields(cfg, seed, h, nsteps):
    return richardson(cfg, seed, h, nsteps,
                      lambda run: {f: _interior(cfg, run.field(f)) for f in FIELDS})


@pytest.mark.parametrize("which", ["angle", "slip"])
@pytest.mark.parametrize("make", [_sessile, _patterned], ids=["sessile", "patterned"])
def test_tangent_matches_central_differences(make, which):
    cfg = make()
    seeds = S.parameter_seeds(cfg)
    seed = seeds[0] if which == "angle" else seeds[-1]
    h = step_h(cfg, seed)
    nsteps = 5
    fd = _fd_fields(cfg, seed, h, nsteps)
    run = GpuRun(cfg, tangent=True)
    try:
        for name, dirn in seed.items():
            run.seed(name, dirn)
        run.steps(nsteps)
        for f in FIELDS:
            t = _interior(cfg, run.tangent(f))
            scale = np.linalg.norm(t)
            assert scale > 0.0, (f, "derivative vanishes")
            err = np.linalg.norm(t - fd[f]) / scale
            assert err <= 1e-5, (cfg.name, which, f, err, scale)
    finally:
        run.close()


def test_observable_gradients_through_autograd():
    torch = pytest.importorskip("torch")
    cfg = _patterned()
    seeds = S.parameter_seeds(cfg)          # two angle groups and the slip length
    nsteps = 5
    fn = S.step_window()
    params = torch.zeros(len(seeds), dtype=torch.float64, requires_grad=True)
    phi = fn.apply(params, cfg, nsteps, seeds)
    obs = S.observables(cfg, phi)
    grads = []
    for m in range(obs.numel()):
        (g,) = torch.autograd.grad(obs[m], params, retain_graph=True)
        grads.append(g.numpy())
    grads = np.array(grads)                  # (observables, parameters)
    obs0 = obs.detach().numpy()
    for p, seed in enumerate(seeds):
        h = step_h(cfg, seed)
        fd = richardson(cfg, seed, h, nsteps,
                        lambda run: S.observables_np(cfg, _interior(cfg, run.field("phi1"))))
        for m in range(len(fd)):
            # rounding floor of the differences: ~1e-15 of the observable over h
            floor = 1e-15 * abs(obs0[m]) / h
            assert abs(grads[m, p] - fd[m]) <= 1e-5 * max(abs(fd[m]), abs(grads[m, p])) + floor, \
                (p, m, grads[m, p], fd[m], floor)
