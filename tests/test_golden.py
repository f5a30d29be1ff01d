"""Replays the golden vectors (tests/golden/*.npz, written from the compiled
reference sources by tests/golden/make_golden.py) through the oracle and,
on a GPU, through the GPU library.

The oracle must reproduce the reference bit for bit, except the FFT stages,
whose reference (srcs/fourn.f) uses a trigonometric recurrence (1e-13 of the
largest mode).  The GPU library must reproduce them bit for bit for
arithmetic-only routines and to the stated relative tolerance where it
evaluates transcendentals (exp, cosh, acos) with the device's libm.
"""
from __future__ import annotations

import glob
import json
import os

import numpy as np
import pytest

import felharness as H
from felharness import D, I

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))
# traj_*.npz are trajectory fixtures of the reference's program (replayed by
# test_oracle_trajectory.py and test_gpu_trajectory.py)
CASES = [os.path.splitext(os.path.basename(p))[0] for p in GOLDEN
         if not os.path.basename(p).startswith("traj_")]

# relative tolerances (of each output's max) where bit equality is not expected
ORACLE_TOL = {"solp_fft_tdma1": 1e-13, "solp_fft_tdma3": 1e-13}
# the GPU library sums in a fixed tree order (mass correction, means,
# divergence norm, centroid and velocity sums) and evaluates cos, sin, exp,
# cosh and acos with the device's libm
GPU_TOL = {"solp_fft_tdma1": 1e-13, "solp_fft_tdma3": 1e-13, "solphi_mthinc1": 1e-13,
           "gnbc": 1e-14, "solphi_mthinc3": 1e-14, "bnd_avzero": 1e-14,
           "bnd_contact_angle": 1e-13, "cal_center": 1e-13, "cal_vel": 1e-13}
GPU_SCALAR_TOL = {"caldiv": 1e-13, "cal_av": 1e-13, "cal_center": 1e-13}


def load(case):
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", case + ".npz"))
    spec = json.loads(str(z["spec"]))
    args, expect = [], {}
    for k, (kind, val) in enumerate(spec["args"]):
        if kind in ("i", "d"):
            args.append(I(val) if kind == "i" else D(val))
            if f"y{k}" in z:
                expect[k] = z[f"y{k}"].item()
        else:
            a = z[f"a{k}"]
            args.append(a.astype(np.int32 if kind == "ia" else np.float64).copy())
            if f"x{k}" in z:
                expect[k] = z[f"x{k}"]
    return spec["routine"], args, expect


def edge_mask(args):
    """bnd_contact_angle: every cell but the i = -2, ni+3 and k = -2, nk+3
    columns, where the reference reads outside its row
    (srcs/bnd_contact_angle.f90:42-44,75-77) and the oracle and the GPU read
    0; the following periodic fill rewrites those columns."""
    ni, nj, nk = (a.value for a in args[1:4])
    m = np.ones((nk + 6, nj + 6, ni + 6), bool)
    m[:, :, [0, -1]] = False
    m[[0, -1], :, :] = False
    return m.ravel()


def compare(routine, args, expect, tol, singular_mode=None, scalar_tol=None):
    assert expect, "a golden case must change at least one argument"
    if routine == "bnd_contact_angle":
        singular_mode = ~edge_mask(args)
    for k, want in expect.items():
        got = args[k]
        if isinstance(got, H.Ref):
            t = (scalar_tol or {}).get(routine, 0.0)
            if t == 0.0:
                assert got.value == want, (routine, k, got.value, want)
            else:
                assert abs(got.value - want) <= t * abs(want), (routine, k, got.value, want)
            continue
        if singular_mode is not None:
            got, want = got.copy().ravel(), want.copy().ravel()
            got[singular_mode] = want[singular_mode] = 0.0
        t = tol.get(routine, 0.0)
        if t == 0.0:
            bad = np.flatnonzero(got != want)
            assert bad.size == 0, f"{routine} arg {k}: {bad.size} values differ, " \
                                  f"max {np.max(np.abs(got - want))}"
        else:
            assert np.max(np.abs(got - want)) <= t * max(np.max(np.abs(want)), 1e-300), routine


def test_golden_vectors_present():
    assert len(CASES) >= 30


@pytest.mark.parametrize("case", CASES)
def test_oracle_reproduces_reference(case):
    routine, args, expect = load(case)
    # orc_caldt takes one more argument than srcs/caldt.f: the three limits
    # (not wanted here: NULL)
    extra = [0] if routine == "caldt" else []
    H.call_oracle(routine, *args, *extra)
    compare(routine, args, expect, ORACLE_TOL)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_gpu_reproduces_reference(case):
    routine, args, expect = load(case)
    H.call_gpu(routine, *args)
    singular = None
    if routine == "solp_fft_tdma2":
        # the last pivot of the (0,0) mode is zero in exact arithmetic and a
        # rounding residue in floating point, so that mode is excluded
        ni, nj, nk = (int(x) for x in args[2])
        m = np.zeros((nk, nj, ni), bool)
        m[nk // 2, :, ni // 2] = True
        singular = m.ravel()
    compare(routine, args, expect, GPU_TOL, singular, GPU_SCALAR_TOL)
