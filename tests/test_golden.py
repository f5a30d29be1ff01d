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
CASES = [os.path.splitext(os.path.basename(p))[0] for p in GOLDEN]

# relative tolerances (of each output's max) where bit equality is not expected
ORACLE_TOL = {"solp_fft_tdma1": 1e-13, "solp_fft_tdma3": 1e-13}
GPU_TOL = {"solp_fft_tdma1": 1e-13, "solp_fft_tdma3": 1e-13, "solphi_mthinc1": 1e-13,
           "gnbc": 1e-14}


def load(case):
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", case + ".npz"))
    spec = json.loads(str(z["spec"]))
    args, expect = [], {}
    for k, (kind, val) in enumerate(spec["args"]):
        if kind == "i":
            args.append(I(val))
        elif kind == "d":
            args.append(D(val))
        else:
            a = z[f"a{k}"]
            args.append(a.astype(np.int32 if kind == "ia" else np.float64).copy())
            if f"x{k}" in z:
                expect[k] = z[f"x{k}"]
    return spec["routine"], args, expect


def compare(routine, args, expect, tol, singular_mode=None):
    assert expect, "a golden case must change at least one array"
    for k, want in expect.items():
        got = args[k]
        if singular_mode is not None:
            got, want = got.copy(), want.copy()
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
    H.call_oracle(routine, *args)
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
    compare(routine, args, expect, GPU_TOL, singular)
