"""Long runs of the timestep and the quantities compared between them.

``mixed_precision`` runs one configuration twice from the same initial
state -- the FP64 path and the FP32 mode (FP32+refine: FP32 flow fields,
stencils and pressure-solve modes, FP64 refinement of the pressure
correction; or FP64 solve modes) -- and reports, at checkpoints, the
relative L2 difference of the velocities, the largest divergence of each
run's corrected velocity, the wall contact-line radius and the volume
drift of each run.

    python -m felmpf.longrun [--config droplet_impact] [--steps 1000]
                             [--every 250] [--fp64-modes]

prints one JSON object.
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
                    fp32_modes: bool = True) -> dict:
    cfg = CONFIGS[name]()
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
            "steps": steps, "fp32_variant": "FP32+refine" if fp32_modes else "FP32, FP64 solve modes",
            "ms_per_step": {"fp64": 1e3 * t64 / steps, "fp32": 1e3 * t32 / steps},
            "checkpoints": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="droplet_impact")
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--every", type=int, default=250)
    ap.add_argument("--fp64-modes", action="store_true",
                    help="FP32 mode with FP64 pressure-solve modes instead of FP32+refine")
    args = ap.parse_args()
    rep = mixed_precision(args.config, args.steps, args.every, not args.fp64_modes)
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
