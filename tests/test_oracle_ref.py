"""Pins the CPU oracle against the reference's own Fortran sources.

Every routine whose reference source compiles on its own (oracle/Makefile,
target `ref`) is run through the compiled reference (oracle/_ref) and through
the oracle on the same seeded inputs; the outputs must be bit-identical,
except the FFT stages, whose reference (srcs/fourn.f) uses a trigonometric
recurrence: those agree to 1e-13 relative.  Skipped where the reference tree
is absent (the GPU box).
"""
import numpy as np
import pytest

import felharness as H
from felharness import D, I, Grid

pytestmark = pytest.mark.skipif(not H.ref_available(), reason="compiled reference not built")

# powers of two in every direction, as the reference requires (srcs/main.f90:117-126)
GRIDS = [Grid(16, 8, 32, 1 / 16, 1 / 8, 1 / 32), Grid(8, 16, 8, 0.1, 0.05, 0.1)]


def both(name, args_factory):
    """Run oracle and reference on deep copies of the same arguments."""
    a = args_factory()
    b = [x.copy() if isinstance(x, np.ndarray) else x for x in a]
    H.call_oracle(name, *a)
    H.call_ref(name, *b)
    return a, b


def assert_same(a, b):
    for x, y in zip(a, b):
        if isinstance(x, np.ndarray):
            assert np.array_equal(x, y, equal_nan=True), np.max(np.abs(x - y))


@pytest.fixture(params=range(len(GRIDS)))
def g(request):
    return GRIDS[request.param]


def test_cal_grad_p2a(g):
    rng = np.random.default_rng(1)
    phi = H.smooth_phi(g, rng)
    dyinv = np.full(g.nj + 6, 1 / g.dy)
    def mk():
        return [I(0), I(g.nj), I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), dyinv.copy(), D(1 / g.dz),
                phi.copy(), g.zeros(), g.zeros(), g.zeros()]
    assert_same(*both("cal_grad_p2a", mk))


def test_solphi_mthinc1(g):
    rng = np.random.default_rng(2)
    phi = H.smooth_phi(g, rng)
    gx, gy, gz = g.zeros(), g.zeros(), g.zeros()
    H.call_oracle("cal_grad_p2a", I(0), I(g.nj), I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx),
                  np.full(g.nj + 6, 1 / g.dy), D(1 / g.dz), phi, gx, gy, gz)
    u, v, w = (H.random_vel(g, rng, 0.5) for _ in range(3))
    u[5, 5, 5] = 0.0  # exercise the u == 0 faces
    def mk():
        return [I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), D(0.01), D(1.0),
                phi.copy(), gx.copy(), gy.copy(), gz.copy(), u.copy(), v.copy(), w.copy(),
                g.zeros(), g.zeros(), g.zeros()]
    assert_same(*both("solphi_mthinc1", mk))


def test_solphi_mthinc2(g):
    rng = np.random.default_rng(3)
    def mk():
        r = np.random.default_rng(3)
        return [I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), D(0.01),
                r.standard_normal(g.shape), r.standard_normal(g.shape), r.standard_normal(g.shape),
                0.01 * r.standard_normal(g.shape), 0.01 * r.standard_normal(g.shape),
                0.01 * r.standard_normal(g.shape), H.smooth_phi(g, r), g.zeros()]
    del rng
    assert_same(*both("solphi_mthinc2", mk))


def test_summation(g):
    def mk():
        r = np.random.default_rng(4)
        return [I(g.ni), I(g.nj), I(g.nk), r.uniform(-0.2, 1.2, (2,) + g.shape), I(1)]
    assert_same(*both("summation", mk))


@pytest.mark.parametrize("kind", ["random", "smooth"])
def test_cal_advu_weno5(g, kind):
    def mk():
        r = np.random.default_rng(5)
        f = H.random_vel if kind == "random" else H.smooth_vel
        return [I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz),
                f(g, r), f(g, r), f(g, r), g.zeros(), g.zeros(), g.zeros()]
    assert_same(*both("cal_advu_weno5", mk))


def test_cal_advu_weno5_tiny_differences(g):
    """Differences near 1e-30 drive the weights through the 1e-80 floor and
    into subnormals; results must still agree bit for bit."""
    def mk():
        r = np.random.default_rng(6)
        return [I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz),
                1e-12 * r.standard_normal(g.shape), 1e-20 * r.standard_normal(g.shape),
                1e-30 * r.standard_normal(g.shape), g.zeros(), g.zeros(), g.zeros()]
    assert_same(*both("cal_advu_weno5", mk))


def test_cal_arith_mean(g):
    def mk():
        r = np.random.default_rng(7)
        return [I(g.ni), I(g.nj), I(g.nk), r.uniform(0, 1, g.shape), D(0.81), D(1e-3), g.zeros()]
    assert_same(*both("cal_arith_mean", mk))


def test_calphat_grad_prs(g):
    def mk():
        r = np.random.default_rng(8)
        return [I(g.ni), I(g.nj), I(g.nk), r.standard_normal(g.shape), r.standard_normal(g.shape),
                r.standard_normal(g.shape)]
    assert_same(*both("calphat", mk))
    def mk2():
        r = np.random.default_rng(9)
        return [I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz),
                r.standard_normal(g.shape), g.zeros(), g.zeros(), g.zeros()]
    assert_same(*both("cal_grad_prs", mk2))


def test_calsij_tau_divtensor(g):
    def mk():
        r = np.random.default_rng(10)
        return [I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz),
                r.standard_normal(g.shape), r.standard_normal(g.shape), r.standard_normal(g.shape),
                g.zeros(6)]
    assert_same(*both("calsij", mk))
    def mk2():
        r = np.random.default_rng(11)
        return [I(g.ni), I(g.nj), I(g.nk), r.standard_normal((6,) + g.shape),
                r.uniform(0.1, 2, g.shape), g.zeros(6)]
    assert_same(*both("cal_arith_tau", mk2))
    def mk3():
        r = np.random.default_rng(12)
        return [I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz),
                r.standard_normal((6,) + g.shape), g.zeros(), g.zeros(), g.zeros()]
    assert_same(*both("cal_div_tensor", mk3))


def test_csf_chain(g):
    dyinv = np.full(g.nj + 6, 1 / g.dy)
    def mk():
        r = np.random.default_rng(13)
        return [I(0), I(g.nj), I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), dyinv.copy(), D(1 / g.dz),
                H.smooth_phi(g, r), g.zeros(), g.zeros(), g.zeros()]
    assert_same(*both("cal_grad_p2uvw", mk))
    for ip, jp, kp in [(1, 1, 1), (0, 0, 1), (0, 1, 0), (1, 0, 0)]:
        def mk2():
            r = np.random.default_rng(14)
            return [I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), I(ip), I(jp),
                    I(kp), r.standard_normal(g.shape), r.standard_normal(g.shape),
                    r.standard_normal(g.shape), g.zeros(3)]
        assert_same(*both("calrn", mk2))
    def mk3():
        r = np.random.default_rng(15)
        return [I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz)] + \
            [r.standard_normal((3,) + g.shape) for _ in range(4)] + [g.zeros(3)]
    assert_same(*both("calrkap", mk3))
    def mk4():
        r = np.random.default_rng(16)
        return [I(g.ni), I(g.nj), I(g.nk), D(0.81), D(8.1e-4), D(5.5)] + \
            [r.standard_normal(g.shape) for _ in range(3)] + [r.standard_normal((3,) + g.shape)] + \
            [g.zeros() for _ in range(3)]
    assert_same(*both("calfst", mk4))
    def mk5():
        r = np.random.default_rng(17)
        return [I(g.ni), I(g.nj), I(g.nk)] + [r.standard_normal(g.shape) for _ in range(6)]
    assert_same(*both("sum_fst", mk5))
    def mk6():
        r = np.random.default_rng(18)
        return [I(g.ni), I(g.nj), I(g.nk)] + [r.standard_normal(g.shape) for _ in range(3)]
    assert_same(*both("init_q", mk6))


def test_cal_srcu(g):
    nid = np.full(6, -1, dtype=np.int32)
    def mk():
        r = np.random.default_rng(19)
        arr = [r.standard_normal(g.shape) for _ in range(3)] + \
            [r.uniform(0.5, 2.0, g.shape) for _ in range(2)] + \
            [r.standard_normal(g.shape) for _ in range(18)] + [g.zeros() for _ in range(3)]
        return [nid.copy(), I(g.ni), I(g.nj), I(g.nk), D(0.9), D(0.003), D(0.1), D(-0.2)] + arr
    assert_same(*both("cal_srcu", mk))


def test_cal_arith_coef_vis(g):
    def mk():
        r = np.random.default_rng(20)
        return [I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz),
                r.uniform(0.01, 2.0, g.shape)] + [r.standard_normal(g.shape) for _ in range(45)]
    assert_same(*both("cal_arith_coef_vis", mk))


def test_bnd_routines(g):
    nid = np.full(6, -1, dtype=np.int32)
    def mk():
        r = np.random.default_rng(21)
        return [I(g.ni), I(g.nj), I(g.nk), r.standard_normal(g.shape)]
    assert_same(*both("bnd_periodic", mk))
    def mk2():
        r = np.random.default_rng(22)
        return [nid.copy(), I(g.ni), I(g.nj), I(g.nk), r.standard_normal(g.shape)]
    assert_same(*both("bnd_neumann", mk2))
    def mk3():
        r = np.random.default_rng(23)
        return [nid.copy(), I(g.ni), I(g.nj), I(g.nk), r.standard_normal(g.shape),
                r.standard_normal(g.shape), r.standard_normal(g.shape), D(0.25), D(g.dy),
                r.uniform(0.01, 0.1, g.shape), r.uniform(0.01, 0.1, g.shape),
                r.uniform(0, 1, g.shape)]
    assert_same(*both("bndu", mk3))
    def mk4():
        r = np.random.default_rng(24)
        return [I(g.ni), I(g.nj), I(g.nk), r.standard_normal(g.shape), g.zeros()]
    assert_same(*both("cpy", mk4))


def test_gnbc(g):
    nid = np.full(6, -1, dtype=np.int32)
    def mk():
        r = np.random.default_rng(25)
        return [nid.copy(), I(g.ni), I(g.nj), I(g.nk), r.standard_normal(g.shape),
                r.standard_normal(g.shape), D(0.25), r.uniform(20, 160, g.shape), D(5.5),
                r.uniform(0, 3, g.shape), g.zeros()]
    assert_same(*both("gnbc", mk))


def test_pressure_coefficients(g):
    n1 = g.nj + 6
    def mk():
        return [I(g.nj), D(0.81), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz)] + [np.zeros(n1) for _ in range(7)]
    a, b = both("cal_coef_prs", mk)
    assert_same(a, b)
    def mk2():
        return [I(g.nj), a[7].copy(), a[8].copy(), np.zeros(n1), np.zeros(n1), np.zeros(n1)]
    assert_same(*both("cal_coef_solp_fft", mk2))


def test_fft_stages(g):
    r = np.random.default_rng(26)
    div = r.standard_normal(g.shape)
    nm = g.ni * g.nj * g.nk
    fo = [I(g.ni), I(g.nj), I(g.nk), np.zeros(2 * nm), div.copy(), np.zeros(nm), np.zeros(nm)]
    fr = [I(g.ni), I(g.nj), I(g.nk), np.zeros(2 * nm), div.copy(), np.zeros(nm), np.zeros(nm)]
    H.call_oracle("solp_fft_tdma1", *fo)
    H.call_ref("solp_fft_tdma1", *fr)
    scale = np.max(np.abs(fr[5])) + np.max(np.abs(fr[6]))
    assert np.max(np.abs(fo[5] - fr[5])) <= 1e-13 * scale
    assert np.max(np.abs(fo[6] - fr[6])) <= 1e-13 * scale
    # inverse
    dpo, dpr = g.zeros(), g.zeros()
    H.call_oracle("solp_fft_tdma3", I(g.ni), I(g.nj), I(g.nk), np.zeros(2 * nm), fr[5], fr[6], dpo)
    H.call_ref("solp_fft_tdma3", I(g.ni), I(g.nj), I(g.nk), np.zeros(2 * nm), fr[5].copy(),
               fr[6].copy(), dpr)
    assert np.max(np.abs(dpo - dpr)) <= 1e-13 * np.max(np.abs(dpr))
    # tdma2 is bitwise (same recursion)
    n1 = max(g.nj, g.nk) + 6
    pc = [np.zeros(g.nj + 6) for _ in range(7)]
    H.call_oracle("cal_coef_prs", I(g.nj), D(0.81), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), *pc)
    fft = [np.zeros(g.nj + 6) for _ in range(3)]
    H.call_oracle("cal_coef_solp_fft", I(g.nj), pc[2], pc[3], *fft)
    svall = np.array([g.ni, g.nj, g.nk], dtype=np.int32)
    def mk():
        return [I(0), I(1), svall.copy(), D(0.81), D(1 / g.dx), D(1 / g.dz), fft[0].copy(),
                fft[1].copy(), fft[2].copy(), np.zeros(n1), np.zeros(n1), np.zeros(n1),
                fr[5].copy(), fr[6].copy()]
    assert_same(*both("solp_fft_tdma2", mk))


def test_corunp_explicit(g):
    nid = np.full(6, -1, dtype=np.int32)
    def mk():
        r = np.random.default_rng(27)
        return [nid.copy(), I(g.ni), I(g.nj), I(g.nk), D(0.81), D(1 / g.dx), D(1 / g.dy),
                D(1 / g.dz), D(0.003)] + [r.standard_normal(g.shape) for _ in range(6)]
    assert_same(*both("corunp_explicit", mk))


# ---------------------------------------------------------------------------
# Routines that include mpif.h (compiled against the image's MPICH; at one
# rank they make no MPI call) and bnd_contact_angle.f90 (compiled from a copy
# with the one dble() cast amdflang needs, oracle/Makefile).
# ---------------------------------------------------------------------------

import ctypes  # noqa: E402

NID_WALLS = np.full(6, -1, dtype=np.int32)


def twice(factory):
    """Two independent argument lists (arrays and boxed scalars) from a
    deterministic factory, one for the oracle and one for the reference."""
    return factory(), factory()


def run_pair(name, factory):
    a, b = twice(factory)
    H.call_oracle(name, *a)
    H.call_ref(name, *b)
    return a, b


def assert_same_all(a, b):
    for k, (x, y) in enumerate(zip(a, b)):
        if isinstance(x, np.ndarray):
            assert np.array_equal(x, y, equal_nan=True), (k, np.max(np.abs(x - y)))
        elif isinstance(x, H.Ref):
            assert x.value == y.value, (k, x.value, y.value)


def flush_both():
    H.ref().felref_flush()
    ctypes.CDLL(None).fflush(None)


def test_solphi_mthinc3(g):
    dyinv = np.full(g.nj + 6, 1 / g.dy)

    def mk():
        r = np.random.default_rng(31)
        phi = H.smooth_phi(g, r)
        phin = np.clip(phi + 0.02 * r.standard_normal(g.shape), -0.05, 1.05)
        phin[phin < 1e-9] = 1e-12  # cells that the truncation zeroes
        gx, gy, gz = g.zeros(), g.zeros(), g.zeros()
        H.call_oracle("cal_grad_p2a", I(0), I(g.nj), I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx),
                      dyinv.copy(), D(1 / g.dz), phin, gx, gy, gz)
        return [I(0), I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), D(2.0),
                gx, gy, gz, phi, phin]
    assert_same_all(*run_pair("solphi_mthinc3", mk))


def sor_args(g, seed, dt):
    r = np.random.default_rng(seed)
    rmu = r.uniform(0.05, 1.0, g.shape)
    cf = [g.zeros() for _ in range(45)]
    H.call_ref("cal_arith_coef_vis", I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy),
               D(1 / g.dz), rmu, *cf)
    n3 = g.n3
    return ([I(0), I(0), NID_WALLS.copy(), I(1), I(g.ni), I(g.nj), I(g.nk),
             np.zeros(2, dtype=np.int32), np.zeros(n3), np.zeros(n3), I(1), I(1), D(dt),
             r.uniform(0.5, 2.0, g.shape), r.uniform(0.5, 2.0, g.shape)] + cf +
            [r.standard_normal(g.shape) for _ in range(3)] +
            [H.smooth_vel(g, r) for _ in range(3)] +
            [D(0.25), D(g.dy), r.uniform(0.01, 0.1, g.shape), r.uniform(0.01, 0.1, g.shape),
             r.uniform(0.0, 1.0, g.shape)])


@pytest.mark.parametrize("dt", [1e-3, 0.05])
def test_solu_sor4(g, dt, capfd):
    """Colour order, refreshes and stop rule of the 8-colour Gauss-Seidel:
    the velocities bit for bit, and the monitor line's sweep count and
    err/err0 (srcs/solu_sor4.f:298-305) equal."""
    a, b = twice(lambda: sor_args(g, 32, dt))
    flush_both()
    capfd.readouterr()
    H.call_ref("solu_sor4", *b)
    flush_both()
    ref_line = [l for l in capfd.readouterr().out.splitlines() if l.startswith("Err_solu")]
    H.call_oracle("solu_sor4", *a)
    flush_both()
    capfd.readouterr()
    assert_same_all(a, b)
    assert len(ref_line) == 1
    t = ref_line[0].split()
    lp = ctypes.c_int.in_dll(H.oracle(), "orc_sor_last_lp").value
    ratio = ctypes.c_double.in_dll(H.oracle(), "orc_sor_last_ratio").value
    assert lp == int(t[2])
    assert abs(ratio / float(t[3]) - 1.0) <= 1e-9


def test_caldiv_cal_av_bnd_avzero(g):
    def mk():
        r = np.random.default_rng(33)
        return [I(0), I(1), I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz),
                D(0.003), H.smooth_vel(g, r), H.smooth_vel(g, r), H.smooth_vel(g, r), g.zeros(),
                D(0.0)]
    assert_same_all(*run_pair("caldiv", mk))

    def mk2():
        r = np.random.default_rng(34)
        return [I(0), I(1), I(g.ni), I(g.nj), I(g.nk), r.standard_normal(g.shape) + 0.3, D(0.0)]
    assert_same_all(*run_pair("cal_av", mk2))

    def mk3():
        r = np.random.default_rng(35)
        return [I(0), I(1), I(g.ni), I(g.nj), I(g.nk), r.standard_normal(g.shape) + 0.3]
    assert_same_all(*run_pair("bnd_avzero", mk3))


def test_solp_fft_tdma4_monitor(g, capfd):
    """The Poisson residual monitor (srcs/solp_fft_tdma4.f:28-95): its printed
    err/err0, err and err0 equal orc_poisson_residual's to the 10 digits
    printed."""
    r = np.random.default_rng(36)
    div, dp = r.standard_normal(g.shape), r.standard_normal(g.shape)
    flush_both()
    capfd.readouterr()
    H.call_ref("solp_fft_tdma4", I(0), I(0), I(1), I(g.ni), I(g.nj), I(g.nk), I(3), I(3),
               D(0.81), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), div, dp)
    flush_both()
    lines = [l for l in capfd.readouterr().out.splitlines() if l.startswith("Err_solp")]
    assert len(lines) == 1
    ratio, err, err0 = (float(x) for x in lines[0].split()[2:5])
    lib = H.oracle()
    lib.orc_poisson_residual.restype = ctypes.c_double
    lib.orc_poisson_residual.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 4 + \
        [ctypes.c_void_p] * 3
    eabs = ctypes.c_double(0.0)
    rel = lib.orc_poisson_residual(g.ni, g.nj, g.nk, 0.81, 1 / g.dx, 1 / g.dy, 1 / g.dz,
                                   div.ctypes.data, dp.ctypes.data, ctypes.addressof(eabs))
    assert abs(rel / ratio - 1) <= 1e-9 and abs(eabs.value / err - 1) <= 1e-9
    assert abs(eabs.value / rel / err0 - 1) <= 1e-9
    # not printed between monitor steps
    H.call_ref("solp_fft_tdma4", I(0), I(0), I(1), I(g.ni), I(g.nj), I(g.nk), I(4), I(3),
               D(0.81), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), div, dp)
    flush_both()
    assert "Err_solp" not in capfd.readouterr().out


def test_trans_r2w_w2r(g):
    nm = g.ni * g.nj * g.nk
    svall = np.array([g.ni, g.nj, g.nk], dtype=np.int32)
    for name in ("trans_r2w", "trans_w2r"):
        def mk():
            r = np.random.default_rng(37)
            return [I(0), I(0), I(1), svall.copy(), np.zeros(2, dtype=np.int32), np.zeros(nm),
                    np.zeros(nm), r.standard_normal(nm), np.zeros(nm)]
        assert_same_all(*run_pair(name, mk))


def test_cal_center_cal_vel(g):
    for area, lap in ((1, 0), (3, 2), (2, 1)):
        def mk():
            r = np.random.default_rng(38)
            return [I(0), I(0), I(g.nj), I(g.ni), I(g.nj), I(g.nk), D(g.ni * g.dx),
                    D(g.nj * g.dy), D(g.nk * g.dz), D(0.003), D(8.1e-4), H.smooth_phi(g, r),
                    np.zeros(3), D(0.4), D(0.3), I(area), I(3), I(lap), D(0.0)]
        assert_same_all(*run_pair("cal_center", mk))

    def mk2():
        r = np.random.default_rng(39)
        return [I(0), I(g.ni), I(g.nj), I(g.nk), D(5e-4), D(0.81), D(1.95)] + \
            [r.standard_normal(g.shape) for _ in range(3)] + [H.smooth_phi(g, r), np.zeros(3),
                                                              np.zeros(3)]
    assert_same_all(*run_pair("cal_vel", mk2))


def test_caldt(g, capfd):
    """The adaptive step (srcs/caldt.f:27-116) for each of its three limits."""
    for scale, rmu, sig in ((1.0, 1.95, 5.5), (1e-3, 50.0, 5.5), (1e-3, 1e-3, 5e4)):
        def mk():
            r = np.random.default_rng(40)
            return [I(0), NID_WALLS.copy(), I(0), I(1), I(g.ni), I(g.nj), I(g.nk), I(2), I(1),
                    D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), D(0.05), D(0.81), D(8.1e-4), D(rmu),
                    D(rmu / 10), D(sig)] + [scale * r.standard_normal(g.shape) for _ in range(3)] + \
                [D(0.0), D(1.5)]
        a, b = twice(mk)
        H.call_ref("caldt", *b)
        lim = np.zeros(3)
        H.call_oracle("caldt", *a, lim)
        assert_same_all(a, b)
    flush_both()
    capfd.readouterr()


def test_bnd_contact_angle(g):
    """The contact-angle ghosts of both walls.  The reference reads one cell
    outside its row at i = -2, ni+3 and k = -2, nk+3 (:42-44,75-77), where
    the oracle reads 0; those edge columns, which the following periodic fill
    rewrites (srcs/main.f90:500,515), are left out of the comparison.  The
    reference's array sits inside a zero-padded buffer so that its reads
    before the first and after the last plane stay in memory the test owns."""
    plane = (g.ni + 6) * (g.nj + 6)
    r = np.random.default_rng(41)
    q0 = H.smooth_phi(g, r)
    th = r.uniform(20.0, 160.0, g.shape)
    th[:, 3, :][r.random(th[:, 3, :].shape) < 0.1] = 90.0
    big = np.zeros(g.n3 + 2 * plane)
    qr = big[plane:plane + g.n3].reshape(g.shape)
    qr[...] = q0
    qo = q0.copy()
    args = lambda q: [NID_WALLS.copy(), I(g.ni), I(g.nj), I(g.nk), q, th.copy(), D(g.dx),  # noqa: E731
                      D(g.dy), D(g.dz)]
    H.call_ref("bnd_contact_angle", *args(qr))
    H.call_oracle("bnd_contact_angle", *args(qo))
    keep = np.ones(g.shape, bool)
    keep[:, :, [0, -1]] = False
    keep[[0, -1], :, :] = False
    assert np.array_equal(qo[keep], qr[keep]), np.max(np.abs(qo[keep] - qr[keep]))
    assert not np.array_equal(qo, q0)
