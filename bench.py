#!/usr/bin/env python3
"""Benchmark: Mcell-steps/s of the GPU timestep on the droplet-impact
configuration, 512x256x128 cells (reference orientation 512x128x256, wall-
normal j), one MI355X.

One step = one call sequence of srcs/main.f90:473-748 through the drop-in
entry points, driven by the Fortran host (host/fel_driver.f90), all fields
device-resident before the timed region.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME]
                  [--no-cpu-baseline] [--precision 64|32]

By default four runs start from the same initial state: `value` is the
FP64 path (results equal to the reference's FP64 arithmetic); `fp32_refine`
holds the FP32 mode's throughput (FP32 pressure-solve modes with one FP64
refinement step), its speed-up and the relative L2 errors of u, v, w, p
against the FP64 run after the same number of steps; `fp32_fp64_poisson`
the same for the FP32 mode with FP64 pressure-solve modes;
`tangent` holds the time of the tangent mode (the same step computing the
FP64 values and their derivative along the bottom wall's contact angle:
forward-mode differentiation, one parameter direction per run) relative to
the FP64 step, and whether its values equal the FP64 run's bit for bit;
`adjoint` the time of the reverse-mode sweep of the velocity-pressure
update at a fixed phase field against the same update's forward.

The roofline object describes the dominant kernel (the colour-pair
Gauss-Seidel update of the viscous solve): algorithmic bytes per launch
divided by its average duration, measured with HIP events on the library's
stream inside the timed region.  The launches go through
hipExtLaunchKernelGGL, which stamps an event with its own launch's start or
end, so the timing adds no packets between launches: the three pairs of a
sweep that run back to back are one timed span (start of the first to the
end of the third, divided by three), the pair before a ghost-refresh launch
is not timed.  The CPU
baseline is the reference's own program (srcs/main.f90 and its makefile
sources, built with OpenMP by oracle/ref_main.py) timed on this host's cores
on a bounded sample of the same configuration: half the resolution in every
direction, three steps after the start-up step.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "fel-multi-phase-flow_amd"))

METRIC = "Mcell-steps/s at 512x256x128 (FP64 vs FP32+refine); adjoint/forward time"
HBM_PEAK_GBS = 8000.0

# event-timer ids recorded by the library (csrc: prof_launch_events)
TIMERS = {3: "sor_pair", 4: "sor_pair12"}


def sor_pair_bytes(cfg) -> float:
    """Algorithmic HBM bytes of one colour-pair launch, averaged over the three
    components: every element the launch must read or write, counted once, at
    N = ni*nj*nk interior cells.  The pair updates N/4 cells of the swept
    component; their stencils touch every cell of the swept component, of the
    other two components and of the viscosity (coefficients are recomputed
    from it), N reads each; the source and the face density (computed once
    per solve) are read and the swept component written at the N/4 updated
    cells.  Hence 4.75 N elements of 8 bytes."""
    return 8.0 * float(cfg.cells) * 4.75


def sor_pair12_bytes(cfg) -> float:
    """Algorithmic HBM bytes of one launch of colour pairs 1 and 2 together
    (the pair-2 cells' x neighbours are the pair-1 cells of the same row, so
    one launch of whole rows does both): the three components and the
    viscosity are read once (4 N), the sources and face densities of both
    pairs (4 x N/4) and the swept component written at both pairs' cells
    (2 x N/4).  Hence 5.5 N elements of 8 bytes."""
    return 8.0 * float(cfg.cells) * 5.5


def measured_traffic(cfg, key="traffic_bytes_per_launch"):
    """HBM bytes per launch of the colour-pair kernel (key
    traffic_bytes_per_launch) or of the pair-1/2 kernel
    (traffic12_bytes_per_launch) from the committed rocprofv3 PMC summary of
    this workload (FETCH_SIZE and WRITE_SIZE runs, calibrated), or None when
    no summary exists for the configuration run."""
    path = os.path.join(REPO, "profiles",
                        f"hbm_traffic_pmc_{cfg.name}_{cfg.ni}x{cfg.nk}x{cfg.nj}.md")
    try:
        for line in open(path):
            if key + ":" in line:
                return float(line.split(key + ":")[1].split("`")[0]), \
                    os.path.relpath(path, REPO)
    except OSError:
        pass
    return None, None


def host_threads() -> int:
    """Threads for the CPU baseline: OMP_NUM_THREADS where it is set (the
    machine's share of cores), else the cores this process may run on."""
    try:
        return max(1, int(os.environ["OMP_NUM_THREADS"]))
    except (KeyError, ValueError):
        return len(os.sched_getaffinity(0))


# Sample of the reference program's timing: the configuration at half the
# resolution in every direction (1/8 of the cells, same domain and physics),
# one start-up step and REF_TIMED_STEPS timed steps.
REF_FACTOR = 2
REF_TIMED_STEPS = 3


def reference_case(cfg_full):
    """(case name, run configuration, parameter assignments) of the reference
    program's timing sample (oracle/ref_main.py builds it)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import ref_main as R  # the checker's recipe for the reference's own program
    f = REF_FACTOR
    cfg = cfg_full.scaled(cfg_full.ni // f, cfg_full.nj // f, cfg_full.nk // f)
    case = f"bench_{cfg.name}_{cfg.ni}x{cfg.nj}x{cfg.nk}"
    return case, cfg, R.restart_params(cfg, steps=1 + REF_TIMED_STEPS)


def cpu_baseline_reference(cfg_full):
    """Times the reference's own program (srcs/main.f90 with its makefile
    sources, amdflang -O2 -fopenmp, built by oracle/ref_main.py) on this
    host's cores: the configuration at half resolution, started from a
    restart record of the same initial state; steps 2..4 are timed from the
    arrival of step 2's first monitor line to that of step 4's last (the
    program flushes its output at both points, srcs/main.f90:487,700).
    Returns None when the program is not built."""
    import tempfile
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import ref_main as R
    from felmpf.config import initial_state
    case, cfg, _ = reference_case(cfg_full)
    if not os.path.exists(R.exe_path(case, omp=True)):
        return None
    threads = host_threads()
    wd = tempfile.mkdtemp(prefix="felref_bench_")
    try:
        R.write_restart(os.path.join(wd, "x.inp"), initial_state(cfg))
        out = R.run(case, wd, threads=threads, omp=True, timeout=600)
    finally:
        import shutil
        shutil.rmtree(wd, ignore_errors=True)
    t_start = [t for t, line in out if line.startswith("nstep=") and int(line.split()[1]) == 2]
    t_end = [t for t, line in out if line.startswith("Err_div") and
             int(line.split()[1]) == 1 + REF_TIMED_STEPS]
    if not t_start or not t_end:
        raise RuntimeError("reference program ended early:\n" +
                           "\n".join(line for _, line in out[-20:]))
    dt = t_end[0] - t_start[0]
    sweeps = [int(line.split()[2]) for _, line in out if line.startswith("Err_solu")]
    return {"value": round(cfg.cells * REF_TIMED_STEPS / dt / 1e6, 4), "unit": "Mcell-steps/s",
            "cores": threads, "kind": "reference",
            "sample": f"reference program srcs/main.f90 (makefile sources, amdflang -O2 -fopenmp, "
                      f"1 MPI rank, {threads} OpenMP threads) on {cfg.name} at "
                      f"{cfg.ni}x{cfg.nj}x{cfg.nk} (ref i,j,k) = {cfg.cells} cells (1/8 of the "
                      f"metric's), from a restart record of the same initial state; steps 2-"
                      f"{1 + REF_TIMED_STEPS} timed ({dt:.1f} s), step 1 is the start-up transient; "
                      f"monitors every step; Gauss-Seidel sweeps per step {sweeps}"}


def cpu_baseline_port(cfg_full, seconds=15.0):
    """Fallback where the reference program is not built: the oracle's
    timestep (serial C restatement, 1 thread) on the configuration with the
    grid coarsened 4x in every direction, for about `seconds`."""
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


def measure(cfg, precision, steps, warmup, keep_fields=False, tangent=False, fp32_modes=True,
            refine=(3, 1.0e-7)):
    """Runs warmup + steps timesteps at the given precision (or in the
    tangent mode, seeded along the bottom wall's contact angle); returns the
    timing, the event timers and (optionally) the final fields.  In the
    FP32 mode, fp32_modes selects FP32 pressure-solve modes with FP64
    refinement (True) or FP64 modes (False); refine = (at most so many
    refinement steps, relative residual tolerance)."""
    import felmpf
    from felmpf import _lib
    gpu = _lib.gpu()
    gpu.felgpu_set_refinement_(ctypes.byref(ctypes.c_int(refine[0])),
                               ctypes.byref(ctypes.c_double(refine[1])))
    run = felmpf.GpuRun(cfg, precision=precision, tangent=tangent, fp32_modes=fp32_modes)
    if tangent:
        seed = np.zeros(cfg.shape)
        seed[:, 3, :] = 1.0  # d/d(theta_0) of the whole bottom wall row
        run.seed("theta_0_array", seed)
    try:
        import torch  # device-wide synchronisation around the timed region
        sync_dev = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # noqa: BLE001
        sync_dev = lambda: None  # noqa: E731
    vol0 = run.field("phi1")[3:-3, 3:-3, 3:-3].sum()
    run.steps(warmup)
    run.sync()
    sync_dev()
    gpu.felgpu_prof_enable_(ctypes.byref(ctypes.c_int(1)))
    gpu.felgpu_prof_reset_()
    calls, sweeps = ctypes.c_longlong(0), ctypes.c_longlong(0)
    gpu.felgpu_sor_stats_(ctypes.byref(ctypes.c_int(1)), ctypes.byref(calls), ctypes.byref(sweeps))
    psolves, pwarm, prefine = ctypes.c_longlong(0), ctypes.c_longlong(0), ctypes.c_longlong(0)
    pfirst, pfinal = ctypes.c_double(0), ctypes.c_double(0)
    pstats = (ctypes.byref(psolves), ctypes.byref(pwarm), ctypes.byref(prefine),
              ctypes.byref(pfirst), ctypes.byref(pfinal))
    gpu.felgpu_poisson_stats_(ctypes.byref(ctypes.c_int(1)), *pstats)
    run.sync()
    sync_dev()
    t0 = time.perf_counter()
    run.steps(steps)
    run.sync()
    sync_dev()
    t1 = time.perf_counter()
    gpu.felgpu_prof_enable_(ctypes.byref(ctypes.c_int(0)))
    gpu.felgpu_sor_stats_(ctypes.byref(ctypes.c_int(0)), ctypes.byref(calls), ctypes.byref(sweeps))
    gpu.felgpu_poisson_stats_(ctypes.byref(ctypes.c_int(0)), *pstats)
    st = run.state()
    timers = {}
    for tid, name in TIMERS.items():
        ms, cnt = ctypes.c_double(0), ctypes.c_int64(0)
        gpu.felgpu_prof_read_(ctypes.byref(ctypes.c_int(tid)), ctypes.byref(ms), ctypes.byref(cnt))
        timers[name] = (ms.value, cnt.value)
    phi = run.field("phi1")[3:-3, 3:-3, 3:-3]
    res = {"elapsed": t1 - t0, "state": st, "timers": timers,
           "sweeps_mean": sweeps.value / max(calls.value, 1),
           "sweeps_last": min(st["sor_lp"] + 1, 21),
           "phi_min": float(phi.min()), "phi_max": float(phi.max()),
           "volume_drift": float(abs(phi.sum() - vol0) / vol0)}
    if psolves.value:
        res["poisson_refine"] = {
            "solves": psolves.value, "warm_started": pwarm.value,
            "refine_steps_per_solve": round(prefine.value / psolves.value, 3),
            "rel_residual_after_first_solve_last_step": float(f"{pfirst.value:.3e}"),
            "rel_residual_last_step": float(f"{pfinal.value:.3e}")}
    if keep_fields:
        res["fields"] = {f: run.field(f)[3:-3, 3:-3, 3:-3] for f in ("u", "v", "w", "p")}
    run.close()
    return res


def _rep_times(fn, reps, sync):
    """Per-repetition wall times (ms) of fn() followed by a device sync."""
    out = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        sync()
        out.append((time.perf_counter() - t0) * 1e3)
    return out


def measure_adjoint(cfg, warmup, reps, step_ms):
    """The reverse of the velocity-pressure update at a fixed phase field
    (felmpf.adjoint.MomentumStep, every array device-resident) against its
    forward, both on the state and phase of a run after `warmup` steps, with
    result reuse off so that each pass computes every routine.  The reverse
    timing includes restoring the four output adjoints (device copies).
    Each figure is the median of `reps` synchronised repetitions (all
    reported)."""
    import felmpf
    from felmpf import _lib, adjoint
    gpu = _lib.gpu()
    run = felmpf.GpuRun(cfg)
    run.steps(warmup)
    names = {"rho": "rho", "rhon": "rhon", "rmu": "rmu", "rmun": "rmun", "phi": "phi1",
             "l1": "l1_array", "l2": "l2_array", "fst_u": "sum_fst_u", "fst_v": "sum_fst_v",
             "fst_w": "sum_fst_w", "fst_un": "sum_fst_un", "fst_vn": "sum_fst_vn",
             "fst_wn": "sum_fst_wn"}
    fixed = {k: run.field(v) for k, v in names.items()}
    state = {k: run.field(k) for k in adjoint.STATE}
    run.close()
    reuse_before = int(gpu.felgpu_result_reuse_())
    gpu.felgpu_set_result_reuse_(ctypes.byref(ctypes.c_int(0)))
    step = adjoint.MomentumStep(cfg, fixed, resident=True)
    try:
        rng = np.random.default_rng(7)
        ybuf = {k: step.array(rng.standard_normal(cfg.shape)) for k in adjoint.OUT}
        g3 = [adjoint.I(cfg.ni), adjoint.I(cfg.nj), adjoint.I(cfg.nk)]
        def rev_once():
            for k in adjoint.OUT:
                adjoint.gpu_call("cpy", *g3, ybuf[k], step.b[k])
            step.reverse_set()

        step.set_state(state)
        step.forward_set()
        gpu.felgpu_sync_()
        fwd_t = _rep_times(step.forward_set, reps, gpu.felgpu_sync_)
        rev_once()
        gpu.felgpu_sync_()
        rev_t = _rep_times(rev_once, reps, gpu.felgpu_sync_)
        lp = step.lp
    finally:
        step.close()
        gpu.felgpu_set_result_reuse_(ctypes.byref(ctypes.c_int(reuse_before)))
    fwd, rev = float(np.median(fwd_t)), float(np.median(rev_t))
    return {"scope": "reverse of the velocity-pressure update at a fixed phase field "
                     "(felmpf.adjoint.MomentumStep: WENO5 of u and uo, pressure predictor, "
                     "viscous stress, momentum source, Gauss-Seidel solve, projection, ghost "
                     "refreshes) through the *_b_ entry points",
            "forward_ms": round(fwd, 3), "reverse_ms": round(rev, 3),
            "reverse_vs_forward": round(rev / fwd, 3),
            "reverse_vs_fp64_step": round(rev / step_ms, 3) if step_ms == step_ms else None,
            "sor_sweeps": min(lp + 1, 21), "reps": reps,
            "forward_ms_reps": [round(t, 3) for t in fwd_t],
            "reverse_ms_reps": [round(t, 3) for t in rev_t],
            "held_fixed": "volume fraction, densities, viscosities and surface-tension "
                          "forces (the full_step leg reverses their dependence too)"}


def measure_phase_adjoint(cfg, warmup, reps, step_ms):
    """The reverse of the volume-fraction update at a fixed velocity
    (felmpf.adjoint.PhaseStep, device-resident) against its forward, on the
    state and wall data of a run after `warmup` steps, result reuse off.
    The reverse timing includes restoring the output adjoint (device
    copies of both slices)."""
    import felmpf
    from felmpf import _lib, adjoint
    gpu = _lib.gpu()
    run = felmpf.GpuRun(cfg)
    run.steps(warmup)
    params = {"u": run.field("u"), "v": run.field("v"), "w": run.field("w"),
              "theta0": run.field("theta_0_array"), "zeta": run.field("zeta_array")}
    phi = run.field("phi1")
    run.close()
    reuse_before = int(gpu.felgpu_result_reuse_())
    gpu.felgpu_set_result_reuse_(ctypes.byref(ctypes.c_int(0)))
    step = adjoint.PhaseStep(cfg, params, resident=True)
    try:
        rng = np.random.default_rng(8)
        ybuf = step.array(rng.standard_normal((2,) + cfg.shape))
        g3 = [adjoint.I(cfg.ni), adjoint.I(cfg.nj), adjoint.I(cfg.nk)]
        def rev_once():
            for l in (0, 1):
                adjoint.gpu_call("cpy", *g3, ybuf[l], step.b["phin"][l])
            step.reverse_set()

        step.set_state({"phi": phi})
        step.forward_set()
        gpu.felgpu_sync_()
        fwd_t = _rep_times(step.forward_set, reps, gpu.felgpu_sync_)
        rev_once()
        gpu.felgpu_sync_()
        rev_t = _rep_times(rev_once, reps, gpu.felgpu_sync_)
    finally:
        step.close()
        gpu.felgpu_set_result_reuse_(ctypes.byref(ctypes.c_int(reuse_before)))
    fwd, rev = float(np.median(fwd_t)), float(np.median(rev_t))
    return {"scope": "reverse of the volume-fraction update at a fixed velocity "
                     "(felmpf.adjoint.PhaseStep: vertex gradients, MTHINC fluxes with the "
                     "Newton offset by the implicit-function theorem, update, truncation and "
                     "mass correction, gnbc, contact-angle ghosts, fills, summation) through "
                     "the *_b_ entry points; gives d/d(phi, u, v, w, theta_0, zeta, uwall)",
            "forward_ms": round(fwd, 3), "reverse_ms": round(rev, 3),
            "reverse_vs_forward": round(rev / fwd, 3),
            "reverse_vs_fp64_step": round(rev / step_ms, 3) if step_ms == step_ms else None, "reps": reps,
            "forward_ms_reps": [round(t, 3) for t in fwd_t],
            "reverse_ms_reps": [round(t, 3) for t in rev_t]}


def measure_full_adjoint(cfg, warmup, reps, step_ms):
    """The reverse of the whole step (felmpf.adjoint.FullStep,
    device-resident) against its forward through the same drop-in routines,
    on the state and wall maps of a run after `warmup` steps, result reuse
    off.  The reverse gives the adjoints of the whole state and of the wall
    maps theta_0, zeta, l1, l2 and uwall; its timing includes restoring the
    output adjoints (device copies)."""
    import felmpf
    from felmpf import _lib, adjoint
    gpu = _lib.gpu()
    run = felmpf.GpuRun(cfg)
    run.steps(warmup)
    state = {k: run.field(k) for k in adjoint.FULL_STATE}
    maps = {"theta0": run.field("theta_0_array"), "zeta": run.field("zeta_array"),
            "l1": run.field("l1_array"), "l2": run.field("l2_array")}
    run.close()
    reuse_before = int(gpu.felgpu_result_reuse_())
    gpu.felgpu_set_result_reuse_(ctypes.byref(ctypes.c_int(0)))
    step = adjoint.FullStep(cfg, maps, resident=True)
    try:
        mo, ph = step.mo, step.ph
        rng = np.random.default_rng(9)
        ybuf = {k: mo.array(rng.standard_normal(cfg.shape)) for k in adjoint.OUT}
        ybuf["phin"] = mo.array(rng.standard_normal((2,) + cfg.shape))
        g3 = [adjoint.I(cfg.ni), adjoint.I(cfg.nj), adjoint.I(cfg.nk)]

        def restore():
            for k in adjoint.OUT:
                adjoint.gpu_call("cpy", *g3, ybuf[k], mo.b[k])
            for l in (0, 1):
                adjoint.gpu_call("cpy", *g3, ybuf["phin"][l], ph.b["phin"][l])

        def rev_once():
            restore()
            step.reverse_set()

        step.set_state(state)
        step.forward_set()
        gpu.felgpu_sync_()
        fwd_t = _rep_times(step.forward_set, reps, gpu.felgpu_sync_)
        rev_once()
        gpu.felgpu_sync_()
        rev_t = _rep_times(rev_once, reps, gpu.felgpu_sync_)
        lp = step.lp
    finally:
        step.close()
        gpu.felgpu_set_result_reuse_(ctypes.byref(ctypes.c_int(reuse_before)))
    fwd, rev = float(np.median(fwd_t)), float(np.median(rev_t))
    return {"scope": "reverse of the whole step (felmpf.adjoint.FullStep: phase update, "
                     "contact-angle fill of phi(.,0), material properties, surface tension, "
                     "velocity-pressure update with the density, viscosity, force, phi and "
                     "slip-length adjoints) through the *_b_/_bp_ entry points; gives "
                     "d/d(state, theta_0, zeta, l1, l2, uwall)",
            "forward_ms": round(fwd, 3), "reverse_ms": round(rev, 3),
            "reverse_vs_forward": round(rev / fwd, 3),
            "reverse_vs_fp64_step": round(rev / step_ms, 3) if step_ms == step_ms else None,
            "sor_sweeps": min(lp + 1, 21), "reps": reps,
            "forward_ms_reps": [round(t, 3) for t in fwd_t],
            "reverse_ms_reps": [round(t, 3) for t in rev_t]}


def adjoint_window_run(nsteps, every):
    """The scale run's FP64 adjoint: the scale-run setup (32 droplets,
    x-stripes of 60/100 deg, slip 0.02) on 256x256x128 (reference axes), dt
    frozen at the CFL step of srcs/caldt.f at the start, a window of nsteps
    whole steps reversed with checkpoints every `every` steps; gradients of
    the wetted area and of the liquid centroid height with respect to the
    two stripe angles and the slip length (one reverse sweep each), each
    against the tangent mode's derivative along the same direction (one
    tangent run per parameter)."""
    import dataclasses
    import felmpf
    from felmpf import sensitivity as S
    base = felmpf.CONFIGS["scale_run"]().scaled(256, 128, 256)
    run = felmpf.GpuRun(base)
    run.steps(1)
    dt = run.state()["dt"]
    run.close()
    cfg = dataclasses.replace(base, cfl=0.0, dt=dt)
    seeds = S.parameter_seeds(cfg)
    theta = S.initial_state(cfg)["theta_0_array"]
    names = [("theta %g deg" % float(np.max(s["theta_0_array"] * theta)))
             if "theta_0_array" in s else "slip length" for s in seeds]

    def progress(phase, n):
        if n % 20 == 0:
            print(f"adjoint window: {phase} step {n}", file=sys.stderr, flush=True)

    obs = {}
    phi_end = None
    for which in ("area", "height"):
        r = S.observable_gradients(cfg, nsteps, seeds, which, checkpoint_every=every,
                                   on_step=progress)
        phi_end = r.pop("phi1")
        obs[which] = r
    tangents = {which: [] for which in obs}
    t_s = []
    values_equal = True
    for seed in seeds:
        t0 = time.perf_counter()
        tr = felmpf.GpuRun(cfg, tangent=True)
        try:
            for name, dirn in seed.items():
                tr.seed(name, dirn)
            tr.steps(nsteps)
            tp = tr.tangent("phi1")
            values_equal = values_equal and bool(np.allclose(tr.field("phi1"), phi_end, rtol=0,
                                                             atol=1e-9))
        finally:
            tr.close()
        t_s.append(round(time.perf_counter() - t0, 2))
        for which, r in obs.items():
            tangents[which].append(float(np.sum(r["weights"] * tp)))
    out = {"adjoint_window": "scale_run",
           "grid": "256x256x128 (ni, nj, nk = %d, %d, %d)" % (cfg.ni, cfg.nj, cfg.nk),
           "steps": nsteps, "checkpoint_every": every, "dt": dt, "parameters": names,
           "values_equal_driver": values_equal, "tangent_run_s": t_s}
    for which, r in obs.items():
        sc = max(abs(t) for t in tangents[which])
        out[which] = {"value": r["value"], "gradients_reverse": r["gradients"],
                      "gradients_tangent": tangents[which],
                      "abs_diff_over_max": [abs(a - b) / sc for a, b in
                                            zip(r["gradients"], tangents[which])],
                      "forward_window_s": round(r["forward_s"], 2),
                      "reverse_window_s": round(r["reverse_s"], 2),
                      "forward_steps_run": r["forward_steps_run"]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="droplet_impact")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--precision", type=int, default=0, choices=(0, 32, 64),
                    help="64 or 32: time that path only; default: FP64, FP32 and tangent "
                         "(value = FP64)")
    ap.add_argument("--adjoint", choices=("momentum", "phase", "full_step"), default=None,
                    help="time one reverse-mode leg only (after --warmup steps, --steps "
                         "repetitions) and print its JSON object instead of the bench line")
    ap.add_argument("--adjoint-window", action="store_true",
                    help="the scale run's FP64 adjoint: 256x256x128, a --window-steps window "
                         "with checkpoints every --checkpoint steps, gradients of the wetted "
                         "area and the centroid height with respect to the two stripe angles "
                         "and the slip length, each checked against a tangent-mode run; "
                         "prints one JSON object")
    ap.add_argument("--window-steps", type=int, default=200)
    ap.add_argument("--checkpoint", type=int, default=20)
    args = ap.parse_args()
    if args.adjoint_window:
        print(json.dumps(adjoint_window_run(args.window_steps, args.checkpoint)))
        return
    if args.adjoint is not None:
        import felmpf
        cfg = felmpf.CONFIGS[args.config]()
        fn = {"momentum": measure_adjoint, "phase": measure_phase_adjoint,
              "full_step": measure_full_adjoint}[args.adjoint]
        res = fn(cfg, args.warmup, args.steps, float("nan"))
        print(json.dumps({"adjoint": args.adjoint, "config": args.config, **res}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design only, not built)", file=sys.stderr)
        sys.exit(2)

    import felmpf

    cfg = felmpf.CONFIGS[args.config]()
    # the CPU baseline runs first: it starts the reference's program as a
    # child process, before this process touches the GPU
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_reference(cfg) or cpu_baseline_port(cfg, args.cpu_seconds)
    head_prec = 32 if args.precision == 32 else 64
    both = args.precision == 0
    main_run = measure(cfg, head_prec, args.steps, args.warmup, keep_fields=both)
    fp32 = fp32_64 = None

    def fp32_summary(r32, dtype):
        errs = {}
        for f in ("u", "v", "w", "p"):
            a, b = main_run["fields"][f], r32["fields"][f]
            errs[f] = float(np.linalg.norm(a - b) / np.linalg.norm(a))
        v32 = cfg.cells * args.steps / r32["elapsed"] / 1e6
        v64 = cfg.cells * args.steps / main_run["elapsed"] / 1e6
        return {"value": round(v32, 3), "unit": "Mcell-steps/s",
                "ms_per_step": round(r32["elapsed"] / args.steps * 1e3, 3),
                "speedup_vs_fp64": round(v32 / v64, 3),
                "rel_l2_error_vs_fp64_after_steps": args.warmup + args.steps,
                "rel_l2_error_vs_fp64": {k: float(f"{v:.3e}") for k, v in errs.items()},
                "volume_drift": float(f"{r32['volume_drift']:.3e}"),
                "sor_sweeps_last_step": r32["sweeps_last"],
                "sor_sweeps_per_step_mean": round(r32["sweeps_mean"], 3),
                **({"poisson_refine": r32["poisson_refine"]} if "poisson_refine" in r32 else {}),
                "dtype": dtype}

    if both:
        fp32 = fp32_summary(
            measure(cfg, 32, args.steps, args.warmup, keep_fields=True, fp32_modes=True),
            "f32 flow fields, stencils and pressure-solve modes; f64 volume fraction, mass "
            "sums, contact-line velocity and pressure correction (warm-started solve, FP64 "
            "residual refined to 1e-7)")
        # the same mode with the residual refined to FP64 level, and unrefined
        fp32["variants"] = {}
        for key, lim in (("refined_to_1e-10", (3, 1.0e-10)), ("unrefined", (0, 1.0e-7))):
            v = fp32_summary(measure(cfg, 32, args.steps, args.warmup, keep_fields=True,
                                     fp32_modes=True, refine=lim), "")
            fp32["variants"][key] = {k: v[k] for k in (
                "ms_per_step", "speedup_vs_fp64", "rel_l2_error_vs_fp64", "poisson_refine")
                if k in v}
        fp32_64 = fp32_summary(
            measure(cfg, 32, args.steps, args.warmup, keep_fields=True, fp32_modes=False),
            "f32 flow fields and stencils; f64 volume fraction, mass sums, contact-line "
            "velocity, pressure-solve modes and pressure correction (no refinement needed)")
    tangent = None
    if both:
        rt = measure(cfg, 64, args.steps, args.warmup, keep_fields=True, tangent=True)
        same = all(np.array_equal(main_run["fields"][f], rt["fields"][f]) for f in ("u", "v", "w", "p"))
        tangent = {"ms_per_step": round(rt["elapsed"] / args.steps * 1e3, 3),
                   "time_vs_fp64_forward": round(rt["elapsed"] / main_run["elapsed"], 3),
                   "directions": 1,
                   "values_equal_fp64": bool(same),
                   "kind": "forward-mode derivative of the step (FP64 value + derivative per "
                           "element)"}
    adjoint_leg = None
    if both:
        step_ms = main_run["elapsed"] / args.steps * 1e3
        adjoint_leg = {"momentum": measure_adjoint(cfg, args.warmup, 5, step_ms),
                       "phase": measure_phase_adjoint(cfg, args.warmup, 5, step_ms)}
        try:
            adjoint_leg["full_step"] = measure_full_adjoint(cfg, args.warmup, 3, step_ms)
        except Exception as e:  # noqa: BLE001 - reported, the main line stands
            adjoint_leg["full_step"] = {"error": f"{type(e).__name__}: {e}"}

    timers = main_run["timers"]
    ms_sor, n_sor = timers["sor_pair"]
    avg_s = ms_sor / 1e3 / max(n_sor, 1)
    b = sor_pair_bytes(cfg) * (0.5 if head_prec == 32 else 1.0)
    achieved = b / avg_s / 1e9 if n_sor else 0.0
    elapsed = main_run["elapsed"]
    value = cfg.cells * args.steps / elapsed / 1e6
    traffic, traffic_src = measured_traffic(cfg) if head_prec == 64 else (None, None)
    ms12, n12 = timers.get("sor_pair12", (0.0, 0))
    roof12 = None
    if n12:
        avg12 = ms12 / 1e3 / n12
        b12 = sor_pair12_bytes(cfg) * (0.5 if head_prec == 32 else 1.0)
        a12 = b12 / avg12 / 1e9
        t12 = measured_traffic(cfg, "traffic12_bytes_per_launch")[0] if head_prec == 64 else None
        roof12 = {"bound": "hbm", "achieved": round(a12, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": round(a12 / HBM_PEAK_GBS, 4), "traffic": t12,
                  "kernel": "sor_pair12_kernel (colour pairs 1 and 2 in one launch)",
                  "bytes_per_launch": b12, "avg_launch_us": round(avg12 * 1e6, 2),
                  "launches": n12,
                  "pair_equivalent_frac": round(2 * b / avg12 / 1e9 / HBM_PEAK_GBS, 4)}

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
        "dtype": "f64" if head_prec == 64 else "f32+f64-refine",
        "data": "synthetic",
        "config": {"workload": f"{cfg.name} {cfg.ni}x{cfg.nk}x{cfg.nj} (ref i,j,k = {cfg.ni}x{cfg.nj}x{cfg.nk}), "
                               "drop-in timestep driven by the Fortran host",
                   "cells": cfg.cells, "precision": "fp64" if head_prec == 64 else "fp32+refine",
                   "sor_sweeps_last_step": main_run["sweeps_last"],
                   "sor_sweeps_per_step_mean": round(main_run["sweeps_mean"], 3),
                   **({"poisson_refine": main_run["poisson_refine"]}
                      if "poisson_refine" in main_run else {}),
                   "volume_drift": float(f"{main_run['volume_drift']:.3e}")},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "sor_pair_kernel (colour-pair Gauss-Seidel update: pairs 0 and 3 "
                               "of each sweep; timed on pair 3, the same kernel with in-launch "
                               "ghost images)",
                     "bytes_per_launch": b, "avg_launch_us": round(avg_s * 1e6, 2),
                     "launches": n_sor, "traffic_source": traffic_src},
        "roofline_pair12": roof12,
        "fp32_refine": fp32,
        "fp32_fp64_poisson": fp32_64,
        "tangent": tangent,
        "adjoint": adjoint_leg,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    detail = {k: {"ms_total": round(v[0], 3), "launches": v[1],
                  "ms_per_step": round(v[0] / args.steps, 3)} for k, v in timers.items()}
    print(json.dumps({"kernel_timers": detail, "phi_min": main_run["phi_min"],
                      "phi_max": main_run["phi_max"]}), file=sys.stderr)


if __name__ == "__main__":
    main()
