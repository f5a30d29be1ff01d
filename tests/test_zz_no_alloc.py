"""The allocation-free viscous solve (felgpu_solu_sor4_noalloc_) against the
drop-in solu_sor4_ it varies.

solu_sor4_ takes host arrays and the 45 coefficient arrays of
cal_arith_coef_vis_ and keeps its work memory (face densities, residual
partials, slip-wall coefficients) in library-allocated buffers; the variant
takes device arrays and the viscosity, forms coefficients and face densities
with the same expressions and keeps its work memory in a static device
array.  The velocities must be bit-identical, with the same sweep index and
err/err0, on two calls in a row, the second on a smaller grid.  A grid whose
work memory exceeds the static array is rejected before any launch.
Collected last (its file name): it must pass alone in a fresh process and
after every other test.
"""
import ctypes

import numpy as np
import pytest

import felharness as H
from felharness import D, I, Grid

pytestmark = pytest.mark.gpu

NID = np.full(6, -1, dtype=np.int32)
NOALLOC_BYTES = 16 << 20


def work_bytes(ni, nj, nk):
    """sor_work_bytes<double> of csrc/k_sor.hip (block partials of the 12
    launches of a sweep, 256 chunk sums, the residual sum; slip-wall
    coefficients of both walls)."""
    gx = ((ni + 1) // 2 + 63) // 64
    gy = ((nj + 1) // 2 + 3) // 4
    nb = gx * gy * nk
    a = 8 * (nb * 12 + 256 + 1)
    return (a + 255) // 256 * 256 + 8 * 4 * (ni + 6) * (nk + 6)


def inputs(g, seed):
    r = np.random.default_rng(seed)
    return dict(rho=r.uniform(0.5, 2.0, g.shape), rhon=r.uniform(0.5, 2.0, g.shape),
                rmun=r.uniform(0.05, 1.0, g.shape),
                src=[r.standard_normal(g.shape) for _ in range(3)],
                vel=[H.smooth_vel(g, r) for _ in range(3)],
                l1=r.uniform(0.01, 0.1, g.shape), l2=r.uniform(0.01, 0.1, g.shape),
                phi=r.uniform(0.0, 1.0, g.shape))


def with_drop_in(g, x, dt, uwall):
    """solu_sor4_ on host arrays with the 45 coefficient arrays."""
    cf = [g.zeros() for _ in range(45)]
    H.call_gpu("cal_arith_coef_vis", I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy),
               D(1 / g.dz), x["rmun"].copy(), *cf)
    vel = [v.copy() for v in x["vel"]]
    n3 = g.n3
    H.call_gpu("solu_sor4", I(0), I(0), NID.copy(), I(1), I(g.ni), I(g.nj), I(g.nk),
               np.zeros(2, dtype=np.int32), np.zeros(n3), np.zeros(n3), I(1), I(0), D(dt),
               x["rho"].copy(), x["rhon"].copy(), *cf, *[s.copy() for s in x["src"]], *vel,
               D(uwall), D(g.dy), x["l1"].copy(), x["l2"].copy(), x["phi"].copy())
    lp, ratio = ctypes.c_int(0), ctypes.c_double(0)
    H.gpu().felgpu_sor_last_(ctypes.byref(lp), ctypes.byref(ratio))
    return vel, lp.value, ratio.value


def hip():
    """The HIP runtime the library is linked with (loaded with it)."""
    H.gpu()
    lib = ctypes.CDLL("libamdhip64.so.7")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipDeviceSynchronize.argtypes = []
    for f in (lib.hipMalloc, lib.hipMemcpy, lib.hipFree, lib.hipDeviceSynchronize):
        f.restype = ctypes.c_int
    return lib


H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def noalloc(g, x, dt, uwall):
    """The allocation-free variant on device copies (the test allocates
    them with hipMalloc)."""
    rt = hip()
    host = [x["rho"], x["rhon"], x["rmun"], *x["src"], *[v.copy() for v in x["vel"]], x["l1"],
            x["l2"], x["phi"]]
    dev = []
    for a in host:
        d = ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(d), a.nbytes) == 0
        assert rt.hipMemcpy(d, a.ctypes.data, a.nbytes, H2D) == 0
        dev.append(d)
    p = dev
    lp, ratio = ctypes.c_int(-1), ctypes.c_double(0)
    fn = H.gpu().felgpu_solu_sor4_noalloc_
    fn.restype = ctypes.c_int
    rc = fn(NID.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ctypes.c_int(g.ni)),
            ctypes.byref(ctypes.c_int(g.nj)), ctypes.byref(ctypes.c_int(g.nk)),
            ctypes.byref(ctypes.c_double(dt)), ctypes.byref(ctypes.c_double(1 / g.dx)),
            ctypes.byref(ctypes.c_double(1 / g.dy)), ctypes.byref(ctypes.c_double(1 / g.dz)),
            *p[0:9], ctypes.byref(ctypes.c_double(uwall)), ctypes.byref(ctypes.c_double(g.dy)),
            *p[9:12], ctypes.byref(lp), ctypes.byref(ratio))
    H.gpu().felgpu_sync_()
    assert rt.hipDeviceSynchronize() == 0
    out = []
    for d, a in zip(dev[6:9], host[6:9]):
        assert rt.hipMemcpy(a.ctypes.data, d, a.nbytes, D2H) == 0
        out.append(a)
    for d in dev:
        assert rt.hipFree(d) == 0
    return rc, out, lp.value, ratio.value


@pytest.mark.parametrize("dt", [1e-3, 0.05], ids=["converged", "capped"])
def test_noalloc_equals_drop_in_on_two_calls(dt):
    for g, seed in ((Grid(32, 16, 32, 1 / 32, 1 / 16, 1 / 32), 51),
                    (Grid(16, 8, 16, 1 / 16, 1 / 8, 1 / 16), 52)):
        print(f"grid {g.ni}x{g.nj}x{g.nk} dt {dt}", flush=True)
        assert work_bytes(g.ni, g.nj, g.nk) <= NOALLOC_BYTES
        x = inputs(g, seed)
        want, lp0, r0 = with_drop_in(g, x, dt, 0.25)
        rc, got, lp1, r1 = noalloc(g, x, dt, 0.25)
        assert rc == 0
        # the same sweeps; err/err0 is the same sum of squares grouped by the
        # two paths' launch blocks (the drop-in runs pairs 1 and 2 as one
        # launch), so it agrees to rounding
        assert lp1 == lp0
        assert abs(r1 - r0) <= 1e-13 * abs(r0), (r1, r0)
        for a, b in zip(got, want):
            assert np.array_equal(a, b), np.max(np.abs(a - b))


def test_noalloc_rejects_grids_beyond_its_work_array():
    """A grid whose work memory exceeds the static array returns 1 before
    any launch (the pointers are never used)."""
    ni, nj, nk = 1024, 256, 1024
    assert work_bytes(ni, nj, nk) > NOALLOC_BYTES
    fn = H.gpu().felgpu_solu_sor4_noalloc_
    fn.restype = ctypes.c_int
    z = ctypes.c_void_p(0)
    lp, ratio = ctypes.c_int(-1), ctypes.c_double(0)
    d = lambda v: ctypes.byref(ctypes.c_double(v))  # noqa: E731
    rc = fn(NID.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ctypes.c_int(ni)),
            ctypes.byref(ctypes.c_int(nj)), ctypes.byref(ctypes.c_int(nk)), d(1e-3), d(1.0),
            d(1.0), d(1.0), *[z] * 9, d(0.0), d(1.0), *[z] * 3, ctypes.byref(lp),
            ctypes.byref(ratio))
    assert rc == 1 and lp.value == -1
