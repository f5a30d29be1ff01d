#!/usr/bin/env python3
"""Benchmark: Mcell-steps/s of the FP64 GPU timestep on the droplet-impact
configuration, 512x256x128 cells (reference orientation 512x128x256, wall-
normal j), one MI355X.

One step = one call sequence of srcs/main.f90:473-748 through the drop-in
entry points, driven by the Fortran host (host/fel_driver.f90), all fields
device-resident before the timed region.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME]
                  [--no-cpu-baseline]

The roofline object describes the dominant kernel (the colour-pair
Gauss-Seidel update of the viscous solve): algorithmic bytes per launch
divided by its average duration, measured with HIP events recorded on the
library's stream around each launch inside the timed region.  The CPU
baseline is the oracle (a serial C restatement of the reference) timed on a
bounded sample of the same configuration on this host.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "fel-multi-phase-flow_amd"))

METRIC = "Mcell-steps/s at 512x256x128 (FP64 vs FP32+refine); adjoint/forward time"
HBM_PEAK_GBS = 8000.0

# event-timer ids recorded by the library (csrc: prof_begin/prof_end)
TIMERS = {1: "mthinc_flux", 2: "weno5", 3: "sor_pair", 4: "sor_bc", 5: "fft_forward",
          6: "tdma", 7: "fft_inverse"}


def sor_pair_bytes(cfg) -> float:
    """Algorithmic HBM bytes of one colour-pair launch, averaged over the three
    components: every element the launch must read or write, counted once, at
    N = ni*nj*nk interior cells.  The pair updates N/4 cells of the swept
    component; their stencils touch every cell of the swept component, of the
    other two components and of the viscosity (coefficients are recomputed
    from it), N reads each; the source is read and the swept component
    written at the N/4 updated cells; rho and rho^{n+1} are read at each
    updated cell and its partner in the component direction, which covers
    N/2 cells for u and w and all N cells for v.  Hence 5.5 N (u, w) and
    6.5 N (v) elements of 8 bytes."""
    n = float(cfg.cells)
    return 8.0 * n * (5.5 + 6.5 + 5.5) / 3.0


def measured_traffic(cfg):
    """HBM bytes per colour-pair launch from the committed rocprofv3 PMC
    summary of this workload (FETCH_SIZE and WRITE_SIZE runs, calibrated),
    or None when no summary exists for the configuration run."""
    path = os.path.join(REPO, "profiles",
                        f"hbm_traffic_pmc_{cfg.name}_{cfg.ni}x{cfg.nk}x{cfg.nj}.md")
    try:
        for line in open(path):
            if "traffic_bytes_per_launch:" in line:
                return float(line.split("traffic_bytes_per_launch:")[1].split("`")[0]), \
                    os.path.relpath(path, REPO)
    except OSError:
        pass
    return None, None


def cpu_baseline(cfg_full, seconds=15.0):
    """Times the oracle's timestep (serial C restatement, 1 thread) on the same
    configuration with the grid coarsened 4x in every direction (1/64 of the
    cells, same domain and physics), for about `seconds`."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from oracle_run import OracleRun  # test infrastructure: the checker, timed as baseline
    cfg = cfg_full.scaled(cfg_full.ni // 4, cfg_full.nj // 4, cfg_full.nk // 4)
    o = OracleRun(cfg)
    o.steps(2)  # start-up transient
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        o.steps(1)
        n += 1
    dt = time.perf_counter() - t0
    o.close()
    return {"value": round(cfg.cells * n / dt / 1e6, 4), "unit": "Mcell-steps/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle timestep, {cfg.name} at {cfg.ni}x{cfg.nj}x{cfg.nk} (ref i,j,k) "
                      f"= {cfg.cells} cells, {n} steps after 2 warm-up steps, {dt:.1f} s, 1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="droplet_impact")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is described in DESIGN.md, not built)", file=sys.stderr)
        sys.exit(2)

    import felmpf
    from felmpf import _lib

    cfg = felmpf.CONFIGS[args.config]()
    run = felmpf.GpuRun(cfg)
    gpu = _lib.gpu()
    try:
        import torch  # device-wide synchronisation around the timed region
        sync_dev = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # noqa: BLE001
        sync_dev = lambda: None  # noqa: E731

    run.steps(args.warmup)
    run.sync()
    sync_dev()
    on = ctypes.c_int(1)
    gpu.felgpu_prof_enable_(ctypes.byref(on))
    gpu.felgpu_prof_reset_()
    run.sync()
    sync_dev()
    t0 = time.perf_counter()
    run.steps(args.steps)
    run.sync()
    sync_dev()
    t1 = time.perf_counter()
    off = ctypes.c_int(0)
    gpu.felgpu_prof_enable_(ctypes.byref(off))
    elapsed = t1 - t0
    st = run.state()

    timers = {}
    for tid, name in TIMERS.items():
        ms, cnt = ctypes.c_double(0), ctypes.c_int64(0)
        gpu.felgpu_prof_read_(ctypes.byref(ctypes.c_int(tid)), ctypes.byref(ms), ctypes.byref(cnt))
        timers[name] = (ms.value, cnt.value)
    ms_sor, n_sor = timers["sor_pair"]
    avg_s = ms_sor / 1e3 / max(n_sor, 1)
    b = sor_pair_bytes(cfg)
    achieved = b / avg_s / 1e9 if n_sor else 0.0
    value = cfg.cells * args.steps / elapsed / 1e6
    traffic, traffic_src = measured_traffic(cfg)
    phi = run.field("phi1")[3:-3, 3:-3, 3:-3]
    run.close()

    cpu = None if args.no_cpu_baseline else cpu_baseline(cfg, args.cpu_seconds)
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{cfg.name} {cfg.ni}x{cfg.nk}x{cfg.nj} (ref i,j,k = {cfg.ni}x{cfg.nj}x{cfg.nk}), "
                               "FP64 drop-in timestep driven by the Fortran host",
                   "cells": cfg.cells, "precision": "fp64",
                   "sor_iterations_last_step": st["sor_lp"] + 1},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "sor_pair_kernel (colour-pair Gauss-Seidel update)",
                     "bytes_per_launch": b, "avg_launch_us": round(avg_s * 1e6, 2),
                     "launches": n_sor, "traffic_source": traffic_src},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    detail = {k: {"ms_total": round(v[0], 3), "launches": v[1],
                  "ms_per_step": round(v[0] / args.steps, 3)} for k, v in timers.items()}
    print(json.dumps({"kernel_timers": detail, "phi_min": float(phi.min()),
                      "phi_max": float(phi.max())}), file=sys.stderr)


if __name__ == "__main__":
    main()
