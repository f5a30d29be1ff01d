"""Builds and runs the reference's own program (srcs/main.f90) on a chosen
configuration.  Test infrastructure only (a checker): it is used to write the
trajectory fixtures under tests/golden/ and to time the reference on the host
for the benchmark's CPU baseline; no product code calls it.

The program is the reference's: every makefile source (makefile:7-83) is
compiled where it lies under the reference tree by oracle/Makefile (target
`ref`), and main.f90 is compiled from a copy in oracle/_ref/main_<case>/ with
these changes only:

  * parameter assignments of srcs/main.f90:91-249 -- grid (nsv, svall),
    restart flag, domain and physics, output cadences -- set to the case's
    values.  Each assignment named by a case must occur exactly once in those
    lines; nothing outside them is touched except:
  * the documented out-of-bounds slice: after a `do l=1,nbub` loop the loop
    variable is nbub+1 = 2, one past phi(...,0:1), and that slice is passed to
    bndu and solu_sor4 (srcs/main.f90:337-339,623,625,650) and, with `l`
    unset, in the restart branch (:379-380).  Those seven arguments read
    phi(-2,-2,-2,1).

The time step stays the reference's dt = 0.32/nsv (srcs/main.f90:481).

  python oracle/ref_main.py build <case> [--omp]
  python oracle/ref_main.py run <case> <workdir> [--threads N] [--omp]
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys
import threading
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.environ.get("FEL_REFERENCE", "/root/reference")
MAIN_SRC = os.path.join(REF, "srcs", "main.f90")
PARAM_LINES = (91, 249)
SLICE_LINES = (337, 338, 339, 379, 380, 623, 625, 650)

_TRAJ = {"nmax": "100", "idout": "1", "imkuvp": "1000000", "imkvtk": "1000000", "imon_t": "1"}

# name -> parameter assignments (Fortran source text of the right-hand side)
CASES = {
    # The reference's own configuration (Couette flow past a stripe of the
    # "g" phase, bottom wall 90 deg, top wall x-stripes of 64/116 deg, unit
    # density and viscosity ratios) on a 32x8x32 grid of cubic cells (a
    # quarter of the reference's x length, 0.31 of its y length, 4x its z
    # length); 100 steps, checkpoint and monitors every step.
    "couette_32x8x32": {
        "nsv": "8", "svall(1)": "nsv*4", "svall(2)": "nsv", "svall(3)": "32",
        "xl": "17.0d0", "yl": "4.25d0", "zl": "17.0d0", **_TRAJ,
    },
    # Unequal phases: the stripe (phi = 1, the "g" phase) is 10x lighter and
    # less viscous than its surroundings, gravity along -j, contact angles
    # 75 deg (bottom) and 55/125 deg x-stripes of period 4 (top), walls at
    # -/+0.5.  Same grid and cadences.
    "variable_32x8x32": {
        "nsv": "8", "svall(1)": "nsv*4", "svall(2)": "nsv", "svall(3)": "32",
        "xl": "17.0d0", "yl": "4.25d0", "zl": "17.0d0",
        "rhol": "1.0d0", "rhog": "1.0d-1", "rmul": "1.95d0", "rmug": "1.95d-1",
        "uwall": "0.5d0", "theta_0_a": "75.0d0", "theta_0_b": "55.0d0",
        "period": "4", "ratio_a": "1", "grv": "2.0d-1", **_TRAJ,
    },
    # The reference's lengths in y and z on the same grid: cells 8x thinner
    # in z than in x, where the Gauss-Seidel runs its 21 sweeps every step
    # without reaching its tolerance.  20 steps.
    "thin_32x8x32": {
        "nsv": "8", "svall(1)": "nsv*4", "svall(2)": "nsv", "svall(3)": "32",
        "xl": "17.0d0", "yl": "3.4d0", "zl": "4.25d0", **dict(_TRAJ, nmax="20"),
    },
}


def fortran_real(v: float) -> str:
    """A REAL(8) literal that reads back as exactly v."""
    return repr(float(v)).replace("e", "d") + ("" if "e" in repr(float(v)) else "d0")


def restart_params(cfg, steps: int) -> dict[str, str]:
    """Parameter assignments that run the reference's program on a run
    configuration of felmpf.config from a restart record (irestart = 1,
    srcs/main.f90:375-404): grid, lengths, phases, gravity, the bottom wall
    as the uniform "a" side and the top wall as the "b" side of
    init_array_x_stripe (its "c" side mirrors the angle and swaps the slip
    lengths, srcs/main.f90:179-182), the step dt = 0.32/nsv nearest to the
    configuration's, `steps` steps with the monitors on every step and no
    file output beyond the start-up's."""
    nsv = max(1, round(0.32 / cfg.dt))
    return {
        "nsv": str(nsv), "svall(1)": str(cfg.ni), "svall(2)": str(cfg.nj),
        "svall(3)": str(cfg.nk), "irestart": "1",
        "xl": fortran_real(cfg.xl), "yl": fortran_real(cfg.yl), "zl": fortran_real(cfg.zl),
        "rhol": fortran_real(cfg.rhol), "rhog": fortran_real(cfg.rhog),
        "rmul": fortran_real(cfg.rmul), "rmug": fortran_real(cfg.rmug),
        "surface_tension": fortran_real(cfg.sigma), "uwall": fortran_real(cfg.uwall),
        "l1_a": fortran_real(cfg.bottom.slip), "theta_0_a": fortran_real(cfg.bottom.theta0),
        "l1_b": fortran_real(cfg.top.slip), "theta_0_b": fortran_real(cfg.top.theta0),
        "grv": fortran_real(-cfg.grvp), "bet_mthinc": fortran_real(cfg.bet),
        "nmax": str(steps), "idout": "1000000", "imkuvp": "1000000", "imkvtk": "1000000",
        "imon_t": "1",
    }


def write_restart(path: str, fields: dict, time: float = 0.0, nstep: int = 0):
    """The restart record srcs/datain.f:20-39 reads: one unformatted
    sequential record (4-byte length markers) of u, v, w, p, uo, vo, wo, po
    and phi(...,0:1) in the padded column-major image, time and nstep.
    `fields` holds u, v, w, p and phi1 as (nk+6, nj+6, ni+6) arrays; uo..po
    are zero (srcs/init.f:24-41) and phi(...,0) = phi(...,1) (one phase)."""
    import numpy as np
    u = fields["u"]
    zero = np.zeros_like(u)
    order = [u, fields["v"], fields["w"], fields["p"], zero, zero, zero, zero,
             fields["phi1"], fields["phi1"]]
    n = sum(a.nbytes for a in order) + 12
    if n >= 2 ** 31:
        raise ValueError("record of 2 GiB or more (needs subrecords)")
    marker = np.array([n], dtype=np.int32).tobytes()
    with open(path, "wb") as f:
        f.write(marker)
        for a in order:
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        f.write(np.array([time], dtype=np.float64).tobytes())
        f.write(np.array([nstep], dtype=np.int32).tobytes())
        f.write(marker)


def configured_main(params: dict[str, str]) -> str:
    """The text of main.f90 with the case's assignments and the slice fix."""
    with open(MAIN_SRC) as f:
        lines = f.read().split("\n")
    lo, hi = PARAM_LINES
    for name, rhs in params.items():
        pat = re.compile(r"^(\s*)" + re.escape(name) + r"\s*=(.*)$")
        hits = [n for n in range(lo - 1, hi) if pat.match(lines[n])]
        if len(hits) != 1:
            raise ValueError(f"{name}: {len(hits)} assignments in main.f90:{lo}-{hi}")
        n = hits[0]
        lines[n] = pat.match(lines[n]).group(1) + f"{name}={rhs}"
    for ln in SLICE_LINES:
        s = lines[ln - 1]
        if s.count("phi(-2,-2,-2,l)") != 1:
            raise ValueError(f"main.f90:{ln} does not pass phi(-2,-2,-2,l): {s!r}")
        lines[ln - 1] = s.replace("phi(-2,-2,-2,l)", "phi(-2,-2,-2,1)")
    return "\n".join(lines)


def case_params(case: str) -> dict[str, str]:
    if case not in CASES:
        raise KeyError(f"unknown case {case!r} (known: {', '.join(CASES)})")
    return CASES[case]


def build(case: str, omp: bool = False, params: dict[str, str] | None = None) -> str:
    """Writes the configured main.f90 and links the program; returns its path."""
    d = os.path.join(HERE, "_ref", f"main_{case}")
    os.makedirs(d, exist_ok=True)
    text = configured_main(params if params is not None else case_params(case))
    path = os.path.join(d, "main.f90")
    old = open(path).read() if os.path.exists(path) else None
    if old != text:
        with open(path, "w") as f:
            f.write(text)
    target = os.path.join("_ref", f"main_{case}", "a.out.omp" if omp else "a.out")
    subprocess.run(["make", "--no-print-directory", "-s", target], cwd=HERE, check=True)
    return os.path.join(HERE, target)


def exe_path(case: str, omp: bool = False) -> str:
    return os.path.join(HERE, "_ref", f"main_{case}", "a.out.omp" if omp else "a.out")


def run(case: str, workdir: str, threads: int = 1, omp: bool = False, timeout: float = 3600,
        on_line=None) -> list[tuple[float, str]]:
    """Runs the built program in `workdir` (one MPI rank, `threads` OpenMP
    threads) and returns its stdout as (arrival time, line) pairs.  `on_line`
    may return True to stop the run early (the process is then killed)."""
    exe = exe_path(case, omp)
    if not os.path.exists(exe):
        raise FileNotFoundError(f"{exe} is not built (python oracle/ref_main.py build {case})")
    os.makedirs(workdir, exist_ok=True)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
    out: list[tuple[float, str]] = []
    p = subprocess.Popen([exe], cwd=workdir, env=env, stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True, bufsize=1)
    stop = threading.Event()

    def watchdog():
        if not stop.wait(timeout):
            p.kill()

    threading.Thread(target=watchdog, daemon=True).start()
    try:
        for line in p.stdout:
            out.append((time.monotonic(), line.rstrip("\n")))
            if on_line is not None and on_line(line):
                p.kill()
                break
    finally:
        stop.set()
        p.wait()
    return out


def main(argv):
    if len(argv) < 2 or argv[0] not in ("build", "run"):
        sys.exit(__doc__)
    omp = "--omp" in argv
    threads = 1
    if "--threads" in argv:
        threads = int(argv[argv.index("--threads") + 1])
    if argv[0] == "build":
        print(build(argv[1], omp))
    else:
        wd = argv[2]
        if os.path.isdir(wd):
            shutil.rmtree(wd)
        for t, line in run(argv[1], wd, threads, omp):
            print(f"{t:12.3f} {line}")


if __name__ == "__main__":
    main(sys.argv[1:])
