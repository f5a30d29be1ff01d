"""Sensitivities of droplet observables to the wall parameters, from the
library's tangent mode (forward-mode differentiation of the discrete
timestep).

Parameters are wall maps of the bottom wall (row j = 1):
  * contact-angle patches: one parameter per set of wall cells sharing a
    static angle (a uniform wall has one; the checkerboard of
    WallSide(theta0, theta_alt, patch) has two, or one per patch with
    per_patch=True), in degrees;
  * the slip length l1 = l2 = ls, with the GNBC friction zeta following it
    (zeta = 0.21*6*(mu_l/l1 + mu_g/l2), srcs/main.f90:172).

Observables (liquid is phi = 0, the reference's phase convention):
  * the liquid centroid (x, y, z) over the interior cells;
  * the wetted area A = sum over the wall row of (1 - phi) dx dz and its
    second moment about the wetted centroid, sum (1 - phi) r^2 dx dz.

StepWindow is a torch.autograd.Function over a window of n timesteps:
forward(params) runs one tangent run per parameter (each run also gives the
values) and returns the final liquid fraction of the interior; backward
contracts the incoming gradient with the stored tangents.  The cost is one
tangent run per parameter; the runs start from the same state.
AdjointWindow has the same contract with the gradient from one reverse
sweep of the whole step (felmpf.adjoint.FullWindow), whatever the number of
parameters.
"""
from __future__ import annotations

import numpy as np

from .config import RunConfig, initial_state, zeta_of
from .driver import GpuRun


def _interior(cfg: RunConfig, a):
    return a[3:cfg.nk + 3, 3:cfg.nj + 3, 3:cfg.ni + 3]


def patch_masks(cfg: RunConfig, per_patch: bool = False) -> list[np.ndarray]:
    """Seed masks (full padded shape) of the bottom wall's contact-angle
    parameters: cells of row j = 1 grouped by static angle, or by patch."""
    th = initial_state(cfg)["theta_0_array"][:, 3, :]
    masks = []
    side = cfg.bottom
    if per_patch and side.patch > 0:
        ii = ((np.arange(-2, cfg.ni + 4) - 1) % cfg.ni) // side.patch
        kk = ((np.arange(-2, cfg.nk + 4) - 1) % cfg.nk) // side.patch
        lab = kk[:, None] * (cfg.ni // side.patch + 1) + ii[None, :]
        groups = [lab == v for v in np.unique(lab)]
    else:
        groups = [th == v for v in np.unique(th)]
    for g in groups:
        m = np.zeros(cfg.shape)
        m[:, 3, :] = g.astype(np.float64)
        masks.append(m)
    return masks


def slip_seeds(cfg: RunConfig) -> dict[str, np.ndarray]:
    """Seeds of d/d(ls) on the bottom wall: l1 = l2 = ls and zeta(ls)."""
    ls = cfg.bottom.slip
    row = np.zeros(cfg.shape)
    row[:, 3, :] = 1.0
    dz = -0.21 * 6.0 * (cfg.rmul + cfg.rmug) / ls ** 2
    assert abs(zeta_of(cfg.rmul, cfg.rmug, ls, ls) - 0.21 * 6.0 * (cfg.rmul + cfg.rmug) / ls) \
        <= 1e-12 * abs(zeta_of(cfg.rmul, cfg.rmug, ls, ls))
    return {"l1_array": row, "l2_array": row.copy(), "zeta_array": row * dz}


def parameter_seeds(cfg: RunConfig, per_patch: bool = False, slip: bool = True):
    """List of seed dicts, one per parameter: contact-angle groups, then the
    slip length."""
    seeds = [{"theta_0_array": m} for m in patch_masks(cfg, per_patch)]
    if slip:
        seeds.append(slip_seeds(cfg))
    return seeds


def apply_perturbation(run: GpuRun, seed: dict[str, np.ndarray], h: float):
    """Adds h * seed to the host maps of a (non-tangent) run and uploads
    them: the finite-difference counterpart of GpuRun.seed."""
    for name, dirn in seed.items():
        v = run.view(name)
        v += h * dirn
        run.upload(name)


def observables_np(cfg: RunConfig, phi: np.ndarray) -> np.ndarray:
    """[x_c, y_c, z_c, A, M] of the interior liquid fraction 1 - phi (numpy)."""
    import torch
    return observables(cfg, torch.from_numpy(np.ascontiguousarray(phi))).numpy()


def observables(cfg: RunConfig, phi_int):
    """Observables of the interior volume fraction phi_int (torch tensor of
    shape (nk, nj, ni), float64) as a tensor [x_c, y_c, z_c, A, M] in
    reference coordinates (x along i, y along j = wall normal, z along k)."""
    import torch
    liq = 1.0 - phi_int
    x = (torch.arange(1, cfg.ni + 1, dtype=torch.float64) - 0.5) * cfg.dx
    y = (torch.arange(1, cfg.nj + 1, dtype=torch.float64) - 0.5) * cfg.dy
    z = (torch.arange(1, cfg.nk + 1, dtype=torch.float64) - 0.5) * cfg.dz
    vol = liq.sum()
    xc = (liq * x[None, None, :]).sum() / vol
    yc = (liq * y[None, :, None]).sum() / vol
    zc = (liq * z[:, None, None]).sum() / vol
    wall = liq[:, 0, :]
    da = cfg.dx * cfg.dz
    area = wall.sum() * da
    wx = (wall * x[None, :]).sum() * da / area
    wz = (wall * z[:, None]).sum() * da / area
    r2 = (x[None, :] - wx) ** 2 + (z[:, None] - wz) ** 2
    mom = (wall * r2).sum() * da
    return torch.stack([xc, yc, zc, area, mom])


def _make_function():
    import torch

    class StepWindow(torch.autograd.Function):
        """phi_final = StepWindow.apply(params, cfg, nsteps, seeds, state).

        params: 1-D float64 tensor, one entry per seed dict; its values are
        offsets added to the configuration's parameters (zero = the
        configuration itself; the derivative is taken there).  Returns the
        interior volume fraction (nk, nj, ni) after nsteps.
        """

        @staticmethod
        def forward(ctx, params, cfg, nsteps, seeds, state=None):
            if params.dim() != 1 or params.numel() != len(seeds):
                raise ValueError("one parameter per seed")
            if torch.any(params != 0):
                raise ValueError("StepWindow differentiates at the configuration's parameters: "
                                 "pass zero offsets")
            tangents, value = [], None
            for seed in seeds:
                run = GpuRun(cfg, state=state, tangent=True)
                try:
                    for name, dirn in seed.items():
                        run.seed(name, dirn)
                    run.steps(nsteps)
                    phi = _interior(cfg, run.field("phi1"))
                    if value is None:
                        value = phi
                    elif not np.array_equal(value, phi):
                        raise RuntimeError("tangent runs from one state disagree in value")
                    tangents.append(_interior(cfg, run.tangent("phi1")).copy())
                finally:
                    run.close()
            ctx.save_for_backward(torch.from_numpy(np.stack(tangents)))
            return torch.from_numpy(value.copy())

        @staticmethod
        def backward(ctx, grad_phi):
            (tan,) = ctx.saved_tensors
            g = (tan * grad_phi.unsqueeze(0)).flatten(1).sum(1)
            return g, None, None, None, None

    return StepWindow


_MAP_KEYS = {"theta_0_array": "theta0", "zeta_array": "zeta", "l1_array": "l1",
             "l2_array": "l2"}


def driver_start(cfg: RunConfig, state=None):
    """The host driver's state after its start-up (fills of the initial
    fields; phin(.,1) as the start-up left it) and its wall maps, as
    FullWindow takes them."""
    from .adjoint import FULL_STATE
    run = GpuRun(cfg, state=state)
    try:
        st = {k: run.field(k) for k in FULL_STATE + ("phin1",)}
        maps = {v: run.field(k) for k, v in _MAP_KEYS.items()}
    finally:
        run.close()
    return st, maps


def observable_weights(cfg: RunConfig, phi: np.ndarray, which: str):
    """Value and gradient (full padded shape, nonzero on the interior) of a
    droplet observable of the volume fraction phi (liquid is phi = 0):
    "area", the wetted area sum over the bottom wall row of (1 - phi) dx dz;
    "height", the liquid centroid height sum (1 - phi) y / sum (1 - phi)."""
    g = np.zeros(cfg.shape)
    if which == "area":
        da = cfg.dx * cfg.dz
        g[3:cfg.nk + 3, 3, 3:cfg.ni + 3] = -da
        return float(np.sum(1.0 - phi[3:cfg.nk + 3, 3, 3:cfg.ni + 3]) * da), g
    if which == "height":
        liq = 1.0 - _interior(cfg, phi)
        y = (np.arange(1, cfg.nj + 1) - 0.5) * cfg.dy
        vol = float(np.sum(liq))
        yc = float(np.sum(liq * y[None, :, None])) / vol
        _interior(cfg, g)[...] = -(y[None, :, None] - yc) / vol
        return yc, g
    raise ValueError(which)


def observable_gradients(cfg: RunConfig, nsteps: int, seeds, which: str = "area",
                         checkpoint_every=None, state=None, resident: bool = True,
                         on_step=None) -> dict:
    """Gradient of a droplet observable (observable_weights) after `nsteps`
    steps from the driver's start state, with respect to each parameter
    direction in `seeds` (parameter_seeds), from one reverse sweep of the
    whole step (FullWindow, checkpoints every `checkpoint_every` steps).
    Returns the observable, the gradients and the host wall times of the
    forward window and of the reverse."""
    import time
    from .adjoint import FULL_STATE, DeviceFullWindow, FullStep, FullWindow
    st, maps = driver_start(cfg, state)
    step = FullStep(cfg, maps, resident=resident)
    try:
        # device-resident: the states and adjoints stay in device memory
        cls = DeviceFullWindow if resident else FullWindow
        win = cls(step, nsteps, checkpoint_every=checkpoint_every, on_step=on_step)
        t0 = time.perf_counter()
        out = win.forward(st)
        t1 = time.perf_counter()
        value, w = observable_weights(cfg, out["phi1"], which)
        y = {k: np.zeros(cfg.shape) for k in FULL_STATE}
        y["phi1"] = w
        bx = win.reverse(y)
        t2 = time.perf_counter()
    finally:
        step.close()
    grads = [sum(float(np.sum(d * bx[_MAP_KEYS[n]])) for n, d in seed.items()) for seed in seeds]
    return {"value": value, "gradients": grads, "forward_s": t1 - t0, "reverse_s": t2 - t1,
            "forward_steps_run": win.forward_steps, "phi1": out["phi1"], "weights": w}


def _make_adjoint_function():
    import torch

    class AdjointWindow(torch.autograd.Function):
        """phi_final = AdjointWindow.apply(params, cfg, nsteps, seeds, state).

        The same contract as StepWindow, with the gradient from one reverse
        sweep of the whole step (felmpf.adjoint.FullWindow: every step
        recomputed from its stored input, then reversed through the
        library's adjoint entry points) instead of one tangent run per
        parameter: backward costs one window's reverse whatever the number
        of parameters.
        """

        @staticmethod
        def forward(ctx, params, cfg, nsteps, seeds, state=None):
            from .adjoint import FullStep, FullWindow
            if params.dim() != 1 or params.numel() != len(seeds):
                raise ValueError("one parameter per seed")
            if torch.any(params != 0):
                raise ValueError("AdjointWindow differentiates at the configuration's "
                                 "parameters: pass zero offsets")
            st, maps = driver_start(cfg, state)
            # host arrays (nothing registered): backward may run once per
            # torch.autograd.grad call on the same graph
            win = FullWindow(FullStep(cfg, maps), nsteps)
            out = win.forward(st)
            ctx.window, ctx.seeds, ctx.cfg = win, seeds, cfg
            return torch.from_numpy(_interior(cfg, out["phi1"]).copy())

        @staticmethod
        def backward(ctx, grad_phi):
            from .adjoint import FULL_STATE
            cfg = ctx.cfg
            y = {k: np.zeros(cfg.shape) for k in FULL_STATE}
            _interior(cfg, y["phi1"])[...] = grad_phi.detach().cpu().numpy()
            bx = ctx.window.reverse(y)
            g = [sum(float(np.sum(d * bx[_MAP_KEYS[n]])) for n, d in seed.items())
                 for seed in ctx.seeds]
            return torch.tensor(g, dtype=torch.float64), None, None, None, None

    return AdjointWindow


def adjoint_window():
    """The reverse-sweep torch.autograd.Function (torch imported on first
    use)."""
    global _ADJOINT_WINDOW
    try:
        return _ADJOINT_WINDOW
    except NameError:
        _ADJOINT_WINDOW = _make_adjoint_function()
        return _ADJOINT_WINDOW


def step_window():
    """The torch.autograd.Function (torch is imported on first use)."""
    global _STEP_WINDOW
    try:
        return _STEP_WINDOW
    except NameError:
        _STEP_WINDOW = _make_function()
        return _STEP_WINDOW
