"""Long runs of the timestep and the quantities compared between them.

``directional_check`` checks tangent-mode derivatives of the droplet
observables along random directions of the wall parameters (every patch
angle of a patterned wall and the slip length) against central differences
of FP64 runs over a window of steps.

``mixed_precision`` runs one configuration twice from the same initial
state -- the FP64 path and the FP32 mode (FP32+refine: FP32 flow fields,
stencils and pressure-solve modes, FP64 refinement of the pressure
correction; or FP64 solve modes) -- and reports, at checkpoints, the
relative L2 difference of the velocities, the largest divergence of each
run's corrected velocity, the wall contact-line radius and the volume
drift of each run.

    python -m felmpf.longrun [--config droplet_impact] [--steps 1000]
                             [--every 250] [--fp64-modes]
    python -m felmpf.longrun --directions 8 [--config patterned_wall]
                             [--steps 400] [--h 1e-6]

print one JSON object.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from .config import CONFIGS, RunConfig
from .driver import GpuRun


def _interior(a):
    return a[3:-3, 3:-3, 3:-3]


def diagnostics(cfg: RunConfig, run: GpuRun) -> dict:
    """Largest |div| of the velocity (srcs/caldiv.f:24-41 on the interior),
    liquid volume (1 - phi), and the radius of the wetted wall area (the
    bottom row's liquid fraction), each from the device state."""
    u, v, w = (run.field(f) for f in ("u", "v", "w"))
    div = ((_interior(u) - u[3:-3, 3:-3, 2:-4]) / cfg.dx +
           (_interior(v) - v[3:-3, 2:-4, 3:-3]) / cfg.dy +
           (_interior(w) - w[2:-4, 3:-3, 3:-3]) / cfg.dz)
    liq = 1.0 - _interior(run.field("phi1"))
    area = float(liq[:, 0, :].sum()) * cfg.dx * cfg.dz
    return {"max_div": float(np.abs(div).max()),
            "volume": float(liq.sum()) * cfg.dx * cfg.dy * cfg.dz,
            "contact_radius": math.sqrt(area / math.pi),
            "velocity": {f: _interior(a).copy() for f, a in (("u", u), ("v", v), ("w", w))}}


def _run(cfg, checkpoints, **kw):
    run = GpuRun(cfg, **kw)
    out = {}
    try:
        d0 = diagnostics(cfg, run)
        done, t = 0, 0.0
        for n in checkpoints:
            run.sync()
            t0 = time.perf_counter()
            run.steps(n - done)
            run.sync()
            t += time.perf_counter() - t0
            done = n
            out[n] = diagnostics(cfg, run)
            print(f"longrun: {cfg.name} {kw} step {n}", file=sys.stderr, flush=True)
        return d0, out, t
    finally:
        run.close()


def mixed_precision(name: str = "droplet_impact", steps: int = 1000, every: int = 250,
                    fp32_modes: bool = True, cfl: float = 0.0) -> dict:
    """FP64 and FP32 runs of the same configuration from the same initial
    state; cfl > 0 takes each step's dt from caldt (srcs/caldt.f) instead of
    the configuration's fixed dt."""
    from dataclasses import replace
    cfg = CONFIGS[name]()
    if cfl > 0:
        cfg = replace(cfg, cfl=cfl)
    checkpoints = list(range(every, steps + 1, every))
    if not checkpoints or checkpoints[-1] != steps:
        checkpoints.append(steps)
    a0, ref, t64 = _run(cfg, checkpoints, precision=64)
    b0, mix, t32 = _run(cfg, checkpoints, precision=32, fp32_modes=fp32_modes)
    rows = []
    for n in checkpoints:
        r, m = ref[n], mix[n]
        rel = {f: float(np.linalg.norm(m["velocity"][f] - r["velocity"][f]) /
                        np.linalg.norm(r["velocity"][f])) for f in ("u", "v", "w")}
        rows.append({
            "step": n,
            "velocity_rel_l2": rel,
            "max_div": {"fp64": r["max_div"], "fp32": m["max_div"]},
            # zero until the droplet wets the wall
            "contact_radius": {"fp64": r["contact_radius"], "fp32": m["contact_radius"],
                               "abs_diff": abs(m["contact_radius"] - r["contact_radius"])},
            "volume_drift": {"fp64": abs(r["volume"] - a0["volume"]) / a0["volume"],
                             "fp32": abs(m["volume"] - b0["volume"]) / b0["volume"]},
        })
    return {"config": f"{cfg.name} {cfg.ni}x{cfg.nk}x{cfg.nj} (ref i,j,k = {cfg.ni}x{cfg.nj}x{cfg.nk})",
            "dt": f"caldt, cfl {cfg.cfl}" if cfg.cfl > 0 else cfg.dt,
            "steps": steps, "fp32_variant": "FP32+refine" if fp32_modes else "FP32, FP64 solve modes",
            "ms_per_step": {"fp64": 1e3 * t64 / steps, "fp32": 1e3 * t32 / steps},
            "checkpoints": rows}


def directional_check(name: str = "patterned_wall", steps: int = 400, ndir: int = 8,
                      h: float = 1e-6, seed: int = 2024, scale=None) -> dict:
    """Derivatives of the observables [x_c, y_c, z_c, A, M] (see
    felmpf.sensitivity.observables) after `steps` steps along `ndir` random
    unit directions of the parameters p = (patch angles in radians, slip
    length / its configured value): one tangent run per direction against
    Richardson-extrapolated central differences (4 D(h) - D(2h))/3 of FP64
    runs.  Returns the relative error of each observable's derivative.
    `scale` = (ni, nj, nk) runs the configuration on a coarser grid."""
    import torch

    from . import sensitivity as S
    cfg = CONFIGS[name]()
    if scale:
        cfg = cfg.scaled(*scale)
    masks = S.patch_masks(cfg, per_patch=True)
    slip = S.slip_seeds(cfg)
    rng = np.random.default_rng(seed)
    deg = 180.0 / math.pi

    def observables(phi_int):
        return S.observables(cfg, torch.from_numpy(np.ascontiguousarray(phi_int)))

    def observables_t(phi_t):
        return S.observables(cfg, phi_t)

    def interior(a):
        return a[3:cfg.nk + 3, 3:cfg.nj + 3, 3:cfg.ni + 3]

    def fd_value(dirn, t):
        run = GpuRun(cfg)
        try:
            S.apply_perturbation(run, dirn, t)
            run.steps(steps)
            return observables(interior(run.field("phi1"))).numpy()
        finally:
            run.close()

    rows = []
    for n in range(ndir):
        v = rng.standard_normal(len(masks) + 1)
        v /= np.linalg.norm(v)
        th = np.zeros(cfg.shape)
        for p, m in enumerate(masks):
            th += v[p] * deg * m                 # degrees per radian of the parameter
        dirn = {"theta_0_array": th}
        for k, a in slip.items():
            dirn[k] = v[-1] * cfg.bottom.slip * a  # per unit of ls / ls0
        run = GpuRun(cfg, tangent=True)
        try:
            for k, a in dirn.items():
                run.seed(k, a)
            run.steps(steps)
            phi = interior(run.field("phi1"))
            dphi = interior(run.tangent("phi1"))
        finally:
            run.close()
        _, jvp = torch.autograd.functional.jvp(
            observables_t, torch.from_numpy(np.ascontiguousarray(phi)),
            torch.from_numpy(np.ascontiguousarray(dphi)))
        tan = jvp.numpy()
        f = {s_: fd_value(dirn, s_ * h) for s_ in (1, -1, 2, -2)}
        fd = (4.0 * (f[1] - f[-1]) / (2.0 * h) - (f[2] - f[-2]) / (4.0 * h)) / 3.0
        err = np.abs(tan - fd) / np.maximum(np.abs(fd), 1e-300)
        rows.append({"direction": n, "tangent": tan.tolist(), "central_difference": fd.tolist(),
                     "rel_err": err.tolist()})
        print(f"longrun: direction {n} rel err {err.max():.2e}", file=sys.stderr, flush=True)
    return {"config": f"{cfg.name} {cfg.ni}x{cfg.nk}x{cfg.nj} (ref i,j,k = {cfg.ni}x{cfg.nj}x{cfg.nk})",
            "steps": steps, "parameters": len(masks) + 1, "h": h,
            "observables": ["x_c", "y_c", "z_c", "wetted_area", "wetted_moment"],
            "directions": rows,
            "max_rel_err": max(max(r["rel_err"]) for r in rows)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="droplet_impact")
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--every", type=int, default=250)
    ap.add_argument("--fp64-modes", action="store_true",
                    help="FP32 mode with FP64 pressure-solve modes instead of FP32+refine")
    ap.add_argument("--directions", type=int, default=0,
                    help="> 0: the directional-derivative check instead (patterned wall)")
    ap.add_argument("--h", type=float, default=1e-6)
    ap.add_argument("--cfl", type=float, default=0.0,
                    help="> 0: dt from caldt each step (CFL number) instead of the fixed dt")
    ap.add_argument("--scale", default="", help="ni,nj,nk (reference i,j,k) of a coarser grid")
    args = ap.parse_args()
    if args.directions > 0:
        name = "patterned_wall" if args.config == "droplet_impact" else args.config
        scale = tuple(int(x) for x in args.scale.split(",")) if args.scale else None
        rep = directional_check(name, args.steps, args.directions, args.h, scale=scale)
        print(json.dumps(rep), flush=True)
        return
    rep = mixed_precision(args.config, args.steps, args.every, not args.fp64_modes, args.cfl)
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
