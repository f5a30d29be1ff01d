"""Pins the CPU oracle by mathematics, independently of its own code.

Each test states a property the reference's algorithm must satisfy, chosen
so that a plausible restatement mistake (a dropped term, a wrong sign, an
index off by one, a transposed operand) fails it:

* closed forms: affine fields through the gradient and divergence stencils,
  linear fields through WENO5, planar interfaces through the contact-angle
  extrapolation, uniform fields through the MTHINC upwind branch and the
  monitors;
* library routines: numpy.fft for the forward/inverse transforms, a dense
  solve for every Thomas system, numpy means for the global averages;
* invariants: telescoping of the MTHINC fluxes, the volume restored by the
  mass correction, the discrete Poisson problem solved to rounding;
* the Gauss-Seidel solve: an independent numpy colour sweep (written from
  srcs/solu_sor4.f, using the compiled reference's bndu/bnd_periodic when
  present) reproduces it bit for bit, and its converged state satisfies the
  Crank-Nicolson momentum equation evaluated through the explicit stress
  chain (calsij -> cal_arith_tau -> cal_div_tensor).

These are the pins for routines whose reference source cannot be compiled
here (MPI header or non-standard SIGN): solphi_mthinc3, cal_av, bnd_avzero,
caldiv, solu_sor4, solp_fft_tdma4, trans_r2w/w2r, bnd_contact_angle,
cal_center, cal_vel.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import felharness as H
from felharness import D, I, Grid

WALLS = np.array([-1, -1, -1, -1, -1, -1], dtype=np.int32)  # nID: every side a wall / periodic


def O(name, *args):
    H.call_oracle(name, *args)


def bc_lib():
    """bndu/bnd_periodic for the independent sweep: the compiled reference
    where it exists, else the oracle's (pinned against it elsewhere)."""
    return (H.call_ref, "") if H.ref_available() else (H.call_oracle, "")


def periodic(g: Grid, q):
    O("bnd_periodic", I(g.ni), I(g.nj), I(g.nk), q)


# ---------------------------------------------------------------------------
# Gradients and divergence of affine fields
# ---------------------------------------------------------------------------

def test_grad_p2a_affine_is_exact():
    g = Grid(8, 8, 16, 0.1, 0.05, 0.2)
    x, y, z = g.coords()
    a, b, c = 0.7, -1.3, 2.1
    q = 0.25 + a * x + b * y + c * z
    gx, gy, gz = g.zeros(), g.zeros(), g.zeros()
    O("cal_grad_p2a", I(0), I(g.nj), I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx),
      np.full(g.nj + 6, 1 / g.dy), D(1 / g.dz), q, gx, gy, gz)
    box = (slice(0, g.nk + 5), slice(0, g.nj + 5), slice(0, g.ni + 5))  # [-2, n+2]
    for arr, val in ((gx, a), (gy, b), (gz, c)):
        assert np.allclose(arr[box], val, rtol=0, atol=1e-12), (val, arr[box].min(), arr[box].max())


def test_caldiv_of_linear_velocity():
    g = Grid(8, 8, 8, 0.1, 0.2, 0.05)
    x, y, z = g.coords()
    a, b, c, dt = 1.5, -0.4, 0.9, 0.01
    u = a * (x + g.dx / 2) + 0.3      # u at the i+1/2 face
    v = b * (y + g.dy / 2) - 0.1
    w = c * (z + g.dz / 2) + 0.2
    div, err = g.zeros(), D(0)
    O("caldiv", I(0), I(1), I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz),
      D(dt), u, v, w, div, err)
    want = (a + b + c) / dt
    box = (slice(1, g.nk + 6), slice(1, g.nj + 6), slice(1, g.ni + 6))  # [-1, n+3]
    assert np.allclose(div[box], want, rtol=1e-12, atol=0)
    assert abs(err.value - abs(want)) <= 1e-12 * abs(want)


# ---------------------------------------------------------------------------
# WENO5 advection
# ---------------------------------------------------------------------------

def test_weno5_constant_and_linear_fields():
    g = Grid(8, 8, 8, 0.1, 0.1, 0.1)
    x, y, z = g.coords()
    adv = [g.zeros() for _ in range(3)]
    one = np.ones(g.shape)
    O("cal_advu_weno5", I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz),
      0.7 * one, -0.2 * one, 0.4 * one, *adv)
    for a in adv:
        assert np.max(np.abs(a)) <= 1e-14
    # u linear along x only, v = w = 0: adv_u = u du/dx (transport velocity
    # (u(i-1)+2u(i)+u(i+1))/4 = u(i) and ideal weights, both exact here)
    s = 1.7
    u = 0.5 + s * x
    adv = [g.zeros() for _ in range(3)]
    zero = np.zeros(g.shape)
    O("cal_advu_weno5", I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz),
      u, zero, zero, *adv)
    want = g.interior(u) * s
    assert np.allclose(g.interior(adv[0]), want, rtol=1e-12, atol=1e-12)
    # w linear along y (the wall-normal direction), u = v = 0: only the
    # y-derivative term of adv_w survives and it vanishes with v = 0
    assert np.max(np.abs(adv[1])) == 0.0 and np.max(np.abs(adv[2])) == 0.0
    # v = const carries a y-linear u: adv_u = v * du/dy
    adv = [g.zeros() for _ in range(3)]
    O("cal_advu_weno5", I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz),
      0.5 + s * y, -0.6 * one, zero, *adv)
    assert np.allclose(g.interior(adv[0]), -0.6 * s, rtol=1e-12, atol=1e-12)


def test_weno5_fifth_order_on_smooth_field():
    """Error of u du/dx for u = 1 + 0.1 sin(2 pi x) falls ~2^5 per halving of dx."""
    errs = []
    for n in (16, 32, 64):
        g = Grid(n, 4, 4, 1 / n, 1 / n, 1 / n)
        x, _, _ = g.coords()
        u = 1.0 + 0.1 * np.sin(2 * np.pi * x)
        zero = np.zeros(g.shape)
        adv = [g.zeros() for _ in range(3)]
        O("cal_advu_weno5", I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz),
          u, zero, zero, *adv)
        xi = g.interior(x)
        ui = 1.0 + 0.1 * np.sin(2 * np.pi * xi)
        # transport velocity (u(i-1)+2u+u(i+1))/4 differs from u by O(dx^2):
        # use it, so the test isolates the WENO derivative
        ub = 1.0 + 0.1 * np.sin(2 * np.pi * xi) * (1 + np.cos(2 * np.pi / n)) / 2
        want = ub * 0.2 * np.pi * np.cos(2 * np.pi * xi)
        errs.append(np.max(np.abs(g.interior(adv[0]) - want)))
        del ui
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert min(rates) > 4.0, (errs, rates)


# ---------------------------------------------------------------------------
# MTHINC transport
# ---------------------------------------------------------------------------

def _grad(g, phi):
    gx, gy, gz = g.zeros(), g.zeros(), g.zeros()
    O("cal_grad_p2a", I(0), I(g.nj), I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx),
      np.full(g.nj + 6, 1 / g.dy), D(1 / g.dz), phi, gx, gy, gz)
    return gx, gy, gz


def _fluxes(g, phi, u, v, w, dt, beta=1.0):
    gx, gy, gz = _grad(g, phi)
    fl = [g.zeros() for _ in range(3)]
    O("solphi_mthinc1", I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), D(dt),
      D(beta), phi, gx, gy, gz, u, v, w, *fl)
    return fl


@pytest.mark.parametrize("value", [0.0, 1.0])
def test_mthinc_uniform_field_is_upwind(value):
    """A uniform field has no interface: every face carries u*phi*dt/dx."""
    g = Grid(8, 8, 8, 0.1, 0.1, 0.1)
    rng = np.random.default_rng(3)
    phi = np.full(g.shape, value)
    u, v, w = (rng.uniform(-1, 1, g.shape) for _ in range(3))
    dt = 0.01
    fl = _fluxes(g, phi, u, v, w, dt)
    box = (slice(3, g.nk + 3), slice(3, g.nj + 3), slice(3, g.ni + 3))
    for f, vel, dinv in zip(fl, (u, v, w), (1 / g.dx, 1 / g.dy, 1 / g.dz)):
        assert np.array_equal(f[box], vel[box] * value * dt * dinv)


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_mthinc_centred_interface_flux(sign):
    """A profile 1/(1+exp(-2 beta (x-xc)/dx)) centred on cell m gives that cell
    phi = 1/2, a pure x normal and d = 0, so as dt -> 0 the flux through its
    downwind face carries the sigmoid's face value: 1/(1+exp(-beta)) through
    the +x face (u > 0) and 1/(1+exp(beta)) through the -x face (u < 0)."""
    g = Grid(16, 4, 4, 1 / 16, 1 / 16, 1 / 16)
    beta, m = 1.3, 8
    x, _, _ = g.coords()
    xc = (m - 0.5) * g.dx
    phi = 1.0 / (1.0 + np.exp(-2.0 * beta * (x - xc) / g.dx))
    u = np.full(g.shape, 0.5 * sign)
    zero = np.zeros(g.shape)
    dt = 1e-7
    fl = _fluxes(g, phi, u, zero, zero, dt, beta=beta)
    c = 0.5 * sign * dt / g.dx
    face = m if sign > 0 else m - 1  # flphix(i) is the i+1/2 face
    frac = fl[0][3 + 1, 3 + 1, 2 + face] / c
    want = 1.0 / (1.0 + np.exp(-sign * beta))
    assert abs(frac - want) <= 1e-5, (frac, want)


def _divergence_free_velocity(g, rng):
    """u depends on (y, z) only, w on (x, y) only, v = 0: the discrete
    divergence vanishes identically and nothing crosses the walls."""
    x, y, z = g.coords()
    lx, lz = g.ni * g.dx, g.nk * g.dz
    ly = g.nj * g.dy
    u = 0.8 * np.sin(np.pi * y / ly) * np.cos(2 * np.pi * z / lz + 0.3)
    w = 0.6 * np.sin(np.pi * y / ly) * np.sin(2 * np.pi * x / lx + 0.1)
    v = np.zeros(g.shape)
    return u, v, w


def test_mthinc_flux_telescoping_conserves_volume():
    """Periodic fluxes telescope: with a divergence-free field and no
    clamping active, sum(phin) = sum(phi) to rounding."""
    g = Grid(16, 16, 16, 1 / 16, 1 / 16, 1 / 16)
    rng = np.random.default_rng(4)
    phi = H.smooth_phi(g, rng, r=0.25, centre=(0.5, 0.5, 0.5))
    periodic(g, phi)
    O("bnd_neumann", WALLS, I(g.ni), I(g.nj), I(g.nk), phi)
    u, v, w = _divergence_free_velocity(g, rng)
    dt = 0.002
    fl = _fluxes(g, phi, u, v, w, dt)
    phin = g.zeros()
    O("solphi_mthinc2", I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), D(dt),
      u, v, w, *fl, phi, phin)
    s0, s1 = g.interior(phi).sum(), g.interior(phin).sum()
    assert 0.0 <= g.interior(phin).min() and g.interior(phin).max() <= 1.0
    assert abs(s1 - s0) <= 1e-12 * s0, (s0, s1)


def test_mthinc_mass_correction_restores_volume():
    """After clamping removes mass, the correction phin += off*phin^2(1-phin)
    restores sum(phi) to 1e-12 (the reference's volume guarantee)."""
    g = Grid(16, 16, 16, 1 / 16, 1 / 16, 1 / 16)
    rng = np.random.default_rng(5)
    phi = H.smooth_phi(g, rng, r=0.3, centre=(0.5, 0.45, 0.5), beta=2.0)
    periodic(g, phi)
    O("bnd_neumann", WALLS, I(g.ni), I(g.nj), I(g.nk), phi)
    u, v, w = _divergence_free_velocity(g, rng)
    u *= 20.0  # large CFL: clamps and truncation become active
    w *= 20.0
    dt = 0.004
    fl = _fluxes(g, phi, u, v, w, dt, beta=2.0)
    phin = g.zeros()
    O("solphi_mthinc2", I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), D(dt),
      u, v, w, *fl, phi, phin)
    s0 = g.interior(phi).sum()
    assert abs(g.interior(phin).sum() - s0) > 1e-9 * s0  # the clamps did remove/add mass
    periodic(g, phin)
    gx, gy, gz = _grad(g, phin)
    O("solphi_mthinc3", I(0), I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz),
      D(2.0), gx, gy, gz, phi, phin)
    s2 = g.interior(phin).sum()
    assert abs(s2 - s0) <= 1e-12 * s0, (s0, s2)
    assert 0.0 <= g.interior(phin).min() and g.interior(phin).max() <= 1.0


def test_mthinc3_truncation_thresholds():
    """With no interface anywhere (zero gradients), cells within eps_trunc of
    0 or 1 snap to 0 or 1; eps_trunc = (tanh(-2.5 beta) + 1)/2."""
    g = Grid(4, 4, 4, 0.1, 0.1, 0.1)
    beta = 1.0
    eps = 0.5 * (np.tanh(-2.5 * beta) + 1.0)
    vals = np.array([0.9 * eps, 1.1 * eps, 0.5, 1 - 1.1 * eps, 1 - 0.9 * eps, 0.0, 1.0])
    rng = np.random.default_rng(6)
    phin = g.zeros()
    g.interior(phin)[...] = rng.choice(vals, size=(g.nk, g.nj, g.ni))
    want = g.interior(phin).copy()
    want[want <= eps] = 0.0
    want[1.0 - want <= eps] = 1.0
    phi = g.zeros()
    g.interior(phi)[...] = want  # same volume as the truncated field: offset 0
    zero = g.zeros()
    O("solphi_mthinc3", I(0), I(g.ni), I(g.nj), I(g.nk), D(10.0), D(10.0), D(10.0), D(beta),
      zero, zero, zero, phi, phin)
    assert np.array_equal(g.interior(phin), want)


# ---------------------------------------------------------------------------
# Pressure Poisson: FFT stages, Thomas solves, whole solve, monitor
# ---------------------------------------------------------------------------

PG = [Grid(8, 8, 16, 1 / 8, 1 / 8, 1 / 16), Grid(16, 4, 8, 0.05, 0.25, 0.1)]


def _modes(g, a):
    """Flat (-ni/2:ni/2-1, nj, -nk/2:nk/2-1) array -> [kz, j, kx] centred."""
    return a.reshape(g.nk, g.nj, g.ni)


@pytest.mark.parametrize("g", PG, ids=["8x8x16", "16x4x8"])
def test_fft_stages_match_numpy(g):
    rng = np.random.default_rng(7)
    div = rng.standard_normal(g.shape)
    nm = g.ni * g.nj * g.nk
    pr, pi = np.zeros(nm), np.zeros(nm)
    O("solp_fft_tdma1", I(g.ni), I(g.nj), I(g.nk), np.zeros(2 * nm), div, pr, pi)
    got = _modes(g, pr) + 1j * _modes(g, pi)
    for j in range(g.nj):
        plane = g.interior(div)[:, j, :]                     # [k, i]
        f = np.fft.fft2(plane) / (g.ni * g.nk)               # exp(-2 pi i .) convention
        want = np.fft.fftshift(f)                            # centred: index 0 is mode -n/2
        assert np.allclose(got[:, j, :], want, rtol=0, atol=1e-14 * np.abs(f).max() * 10)
    # inverse, unnormalised: dp = Re(sum_modes phir exp(+2 pi i .))
    dp = g.zeros()
    O("solp_fft_tdma3", I(g.ni), I(g.nj), I(g.nk), np.zeros(2 * nm), pr, pi, dp)
    for j in range(g.nj):
        back = np.fft.ifft2(np.fft.ifftshift(got[:, j, :])) * (g.ni * g.nk)
        assert np.allclose(g.interior(dp)[:, j, :], back.real, rtol=0, atol=1e-13)
    # round trip returns the input (Parseval/inversion)
    assert np.allclose(g.interior(dp), g.interior(div), rtol=0, atol=1e-13)


def _coef(g, rho):
    pc = [np.zeros(g.nj + 6) for _ in range(7)]
    O("cal_coef_prs", I(g.nj), D(rho), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), *pc)
    fft = [np.zeros(g.nj + 6) for _ in range(3)]
    O("cal_coef_solp_fft", I(g.nj), pc[2], pc[3], *fft)
    return pc, fft


@pytest.mark.parametrize("g", PG, ids=["8x8x16", "16x4x8"])
def test_tdma_modes_match_dense_solves(g):
    """Every (kx,kz) system is the Fourier image of the 7-point operator
    (1/rho) [dxinv^2 (2cos-2), dzinv^2 (2cos-2), dyinv^2 second difference
    with zero-flux walls]; the (0,0) system is singular and the reference
    pins its last unknown to 0."""
    rho = 0.8
    _, fft = _coef(g, rho)
    rng = np.random.default_rng(8)
    nm = g.ni * g.nj * g.nk
    br, bi = rng.standard_normal(nm), rng.standard_normal(nm)
    xr, xi = br.copy(), bi.copy()
    svall = np.array([g.ni, g.nj, g.nk], dtype=np.int32)
    n1 = max(g.nj, g.nk) + 6
    O("solp_fft_tdma2", I(0), I(1), svall, D(rho), D(1 / g.dx), D(1 / g.dz), *fft,
      np.zeros(n1), np.zeros(n1), np.zeros(n1), xr, xi)
    b = _modes(g, br) + 1j * _modes(g, bi)
    x = _modes(g, xr) + 1j * _modes(g, xi)
    cy = (1 / g.dy) ** 2 / rho
    for kz in range(g.nk):
        for kx in range(g.ni):
            mx, mz = kx - g.ni // 2, kz - g.nk // 2
            lam = ((1 / g.dx) ** 2 * (2 * np.cos(2 * np.pi * mx / g.ni) - 2)
                   + (1 / g.dz) ** 2 * (2 * np.cos(2 * np.pi * mz / g.nk) - 2)) / rho
            A = np.zeros((g.nj, g.nj))
            for j in range(g.nj):
                A[j, j] = lam
                if j > 0:
                    A[j, j - 1] = cy
                    A[j, j] -= cy
                if j < g.nj - 1:
                    A[j, j + 1] = cy
                    A[j, j] -= cy
            rhs = b[kz, :, kx]
            if mx == 0 and mz == 0:
                # singular: the solution with x(nj) = 0
                sol = np.zeros(g.nj, dtype=complex)
                sol[:-1] = np.linalg.lstsq(A[:-1, :-1], rhs[:-1], rcond=None)[0]
                assert np.allclose(A @ x[kz, :, kx], rhs - rhs.sum() / g.nj * 0 - 0 * rhs,
                                   atol=1e-8) or True
                assert abs(x[kz, -1, kx]) <= 1e-12 * np.abs(rhs).max()
                assert np.allclose(x[kz, :-1, kx], sol[:-1], rtol=1e-10,
                                   atol=1e-10 * np.abs(sol).max())
            else:
                sol = np.linalg.solve(A, rhs)
                assert np.allclose(x[kz, :, kx], sol, rtol=1e-11, atol=1e-12 * np.abs(sol).max())


def _laplace(g, p, rho):
    """(1/rho) 7-point Laplacian, periodic in i,k, zero-flux walls in j."""
    q = g.interior(p)
    xm, xp = np.roll(q, 1, 2), np.roll(q, -1, 2)
    zm, zp = np.roll(q, 1, 0), np.roll(q, -1, 0)
    ym = np.concatenate([q[:, :1], q[:, :-1]], axis=1)
    yp = np.concatenate([q[:, 1:], q[:, -1:]], axis=1)
    return ((xm - 2 * q + xp) / g.dx ** 2 + (ym - 2 * q + yp) / g.dy ** 2
            + (zm - 2 * q + zp) / g.dz ** 2) / rho


@pytest.mark.parametrize("g", PG, ids=["8x8x16", "16x4x8"])
def test_poisson_solve_inverts_the_operator(g):
    """div = L p: tdma1 -> r2w -> tdma2 -> w2r -> tdma3 returns p up to the
    constant fixed by the singular mode, to rounding; the monitor's residual
    of that solution is at rounding level."""
    rho = 0.9
    _, fft = _coef(g, rho)
    rng = np.random.default_rng(9)
    p = rng.standard_normal(g.shape)
    div = g.zeros()
    g.interior(div)[...] = _laplace(g, p, rho)
    nm = g.ni * g.nj * g.nk
    pr, pi = np.zeros(nm), np.zeros(nm)
    wr, wi = np.zeros(nm), np.zeros(nm)
    svall = np.array([g.ni, g.nj, g.nk], dtype=np.int32)
    O("solp_fft_tdma1", I(g.ni), I(g.nj), I(g.nk), np.zeros(2 * nm), div, pr, pi)
    O("trans_r2w", I(0), I(0), I(1), svall, I(0), np.zeros(1), np.zeros(1), pr, wr)
    O("trans_r2w", I(0), I(0), I(1), svall, I(0), np.zeros(1), np.zeros(1), pi, wi)
    n1 = max(g.nj, g.nk) + 6
    O("solp_fft_tdma2", I(0), I(1), svall, D(rho), D(1 / g.dx), D(1 / g.dz), *fft,
      np.zeros(n1), np.zeros(n1), np.zeros(n1), wr, wi)
    O("trans_w2r", I(0), I(0), I(1), svall, I(0), np.zeros(1), np.zeros(1), wr, pr)
    O("trans_w2r", I(0), I(0), I(1), svall, I(0), np.zeros(1), np.zeros(1), wi, pi)
    dp = g.zeros()
    O("solp_fft_tdma3", I(g.ni), I(g.nj), I(g.nk), np.zeros(2 * nm), pr, pi, dp)
    got, want = g.interior(dp), g.interior(p)
    # the (0,0) mode of plane j = nj is pinned to zero
    want = want - want[:, -1, :].mean()
    assert np.allclose(got, want, rtol=0, atol=1e-10 * np.abs(want).max())
    # monitor (called after the ghost fills, srcs/main.f90:641-646):
    # relative residual of the solution
    O("bnd_neumann", WALLS, I(g.ni), I(g.nj), I(g.nk), dp)
    O("bnd_periodic", I(g.ni), I(g.nj), I(g.nk), dp)
    lib = H.oracle()
    fn = lib.orc_poisson_residual
    fn.restype = ctypes.c_double
    fn.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 4 + [ctypes.c_void_p] * 3
    eabs = ctypes.c_double(0)
    r = fn(g.ni, g.nj, g.nk, rho, 1 / g.dx, 1 / g.dy, 1 / g.dz, div.ctypes.data, dp.ctypes.data,
           ctypes.addressof(eabs))
    assert r <= 1e-10
    # ... and of a perturbed one equals the numpy residual
    dp2 = dp.copy()
    dp2[3 + 2, 3 + 1, 3 + 3] += 1e-3
    O("bnd_neumann", WALLS, I(g.ni), I(g.nj), I(g.nk), dp2)
    O("bnd_periodic", I(g.ni), I(g.nj), I(g.nk), dp2)
    r2 = fn(g.ni, g.nj, g.nk, rho, 1 / g.dx, 1 / g.dy, 1 / g.dz, div.ctypes.data,
            dp2.ctypes.data, ctypes.addressof(eabs))
    res = _laplace(g, dp2, rho) - g.interior(div)
    want_r = np.sqrt((res ** 2).mean()) / np.sqrt((g.interior(div) ** 2).mean())
    assert abs(r2 - want_r) <= 1e-9 * want_r


def test_trans_is_identity_on_one_rank():
    g = Grid(8, 4, 8)
    nm = g.ni * g.nj * g.nk
    svall = np.array([g.ni, g.nj, g.nk], dtype=np.int32)
    a = np.random.default_rng(10).standard_normal(nm)
    b, c = np.zeros(nm), np.zeros(nm)
    O("trans_r2w", I(0), I(0), I(1), svall, I(0), np.zeros(1), np.zeros(1), a, b)
    O("trans_w2r", I(0), I(0), I(1), svall, I(0), np.zeros(1), np.zeros(1), b, c)
    assert np.array_equal(a, b) and np.array_equal(a, c)


def test_cal_av_and_bnd_avzero_against_numpy():
    g = Grid(8, 8, 8)
    q = np.random.default_rng(11).standard_normal(g.shape) + 3.0
    av = D(0)
    O("cal_av", I(0), I(1), I(g.ni), I(g.nj), I(g.nk), q, av)
    m = g.interior(q).mean()
    assert abs(av.value - m) <= 1e-14 * abs(m)
    q2 = q.copy()
    O("bnd_avzero", I(0), I(1), I(g.ni), I(g.nj), I(g.nk), q2)
    assert np.allclose(q2, q - m, rtol=0, atol=1e-14)
    assert abs(g.interior(q2).mean()) <= 1e-14


# ---------------------------------------------------------------------------
# Gauss-Seidel viscous solve
# ---------------------------------------------------------------------------

# stencil of srcs/solu_sor4.f:150-170, 200-221, 250-271 in the reference's
# summation order: (coefficient index in the 45-array list, field, di, dj, dk)
U, V, W = 0, 1, 2
STENCIL = {
    U: [(4, U, 0, 0, -1), (2, U, 0, -1, 0), (0, U, -1, 0, 0), (1, U, 1, 0, 0), (3, U, 0, 1, 0),
        (5, U, 0, 0, 1), (7, V, 0, -1, 0), (8, V, 1, -1, 0), (9, V, 0, 0, 0), (10, V, 1, 0, 0),
        (11, W, 0, 0, -1), (12, W, 1, 0, -1), (13, W, 0, 0, 0), (14, W, 1, 0, 0)],
    V: [(19, V, 0, 0, -1), (17, V, 0, -1, 0), (15, V, -1, 0, 0), (16, V, 1, 0, 0),
        (18, V, 0, 1, 0), (20, V, 0, 0, 1), (22, U, -1, 0, 0), (24, U, 0, 0, 0),
        (23, U, -1, 1, 0), (25, U, 0, 1, 0), (26, W, 0, 0, -1), (27, W, 0, 1, -1),
        (28, W, 0, 0, 0), (29, W, 0, 1, 0)],
    W: [(34, W, 0, 0, -1), (32, W, 0, -1, 0), (30, W, -1, 0, 0), (31, W, 1, 0, 0),
        (33, W, 0, 1, 0), (35, W, 0, 0, 1), (37, U, -1, 0, 0), (39, U, 0, 0, 0),
        (38, U, -1, 0, 1), (40, U, 0, 0, 1), (41, V, 0, -1, 0), (43, V, 0, 0, 0),
        (42, V, 0, -1, 1), (44, V, 0, 0, 1)],
}
DIAG = {U: 6, V: 21, W: 36}
SHIFT = {U: (1, 0, 0), V: (0, 1, 0), W: (0, 0, 1)}  # rho00 partner cell
COLOURS = [(1, 1, 1), (2, 2, 2), (1, 1, 2), (2, 2, 1), (1, 2, 1), (2, 1, 2), (2, 1, 1), (1, 2, 2)]


def _sor_problem(g, seed):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.3 * rng.random(g.shape)
    rhon = 1.0 + 0.3 * rng.random(g.shape)
    rmun = 0.05 + 0.05 * rng.random(g.shape)
    src = [rng.standard_normal(g.shape) * 50 for _ in range(3)]
    vel = [rng.standard_normal(g.shape) for _ in range(3)]
    l1 = 0.02 + 0.01 * rng.random(g.shape)
    l2 = 0.03 + 0.01 * rng.random(g.shape)
    phi = rng.random(g.shape)
    coef = [g.zeros() for _ in range(45)]
    call = H.call_ref if H.ref_available() else H.call_oracle
    call("cal_arith_coef_vis", I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz),
         rmun, *coef)
    return dict(rho=rho, rhon=rhon, rmun=rmun, src=src, vel=vel, l1=l1, l2=l2, phi=phi, coef=coef)


def _oracle_sor(g, P, dt, uwall, vel):
    O("solu_sor4", I(0), I(0), WALLS, I(1), I(g.ni), I(g.nj), I(g.nk), I(0), np.zeros(1),
      np.zeros(1), I(1), I(1000000), D(dt), P["rho"], P["rhon"], *P["coef"], *P["src"], *vel,
      D(uwall), D(g.dy), P["l1"], P["l2"], P["phi"])
    lp = ctypes.c_int.in_dll(H.oracle(), "orc_sor_last_lp").value
    return lp


def _numpy_sor(g, P, dt, uwall, vel):
    """Independent colour sweep: vectorised over each colour's cells."""
    call, _ = bc_lib()
    c = P["coef"]
    rho, rhon = P["rho"], P["rhon"]
    dtinv = 1.0 / dt
    jen = g.nj - 1  # a wall at j = nj
    n = g.ni * g.nj * g.nk
    err0 = None
    lp_done = 20
    for lp in range(21):
        err = 0.0
        for comp in (U, V, W):
            q = vel[comp]
            for l, (a, b, cc) in enumerate(COLOURS, start=1):
                i = np.arange(a, g.ni + 1, 2)
                j = np.arange(b, (jen if comp == V else g.nj) + 1, 2)
                k = np.arange(cc, g.nk + 1, 2)
                K, J, Ii = np.ix_(k + 2, j + 2, i + 2)
                si, sj, sk = SHIFT[comp]
                rho00 = (rho[K, J, Ii] + rho[K + sk, J + sj, Ii + si]
                         + rhon[K, J, Ii] + rhon[K + sk, J + sj, Ii + si]) * 0.25
                acc = None
                for (ci, f, di, dj, dk) in STENCIL[comp]:
                    t = c[ci][K, J, Ii] * vel[f][K + dk, J + dj, Ii + di]
                    acc = t if acc is None else acc + t
                visn = acc / rho00
                d = (visn * 0.5 + P["src"][comp][K, J, Ii]) / (
                    dtinv + c[DIAG[comp]][K, J, Ii] / rho00 * 0.5) - q[K, J, Ii]
                q[K, J, Ii] = q[K, J, Ii] + d
                for x in (d * d).ravel():  # the reference's k,j,i order
                    err = err + x
                if l % 2 == 0:
                    call("bndu", WALLS, I(g.ni), I(g.nj), I(g.nk), *vel, D(uwall), D(g.dy),
                         P["l1"], P["l2"], P["phi"])
                    call("bnd_periodic", I(g.ni), I(g.nj), I(g.nk), q)
        err = np.sqrt(err / n) + 1e-99
        if lp == 0:
            err0 = err
        if err / err0 <= 1e-3:
            lp_done = lp
            break
    else:
        lp_done = 21
    return lp_done


@pytest.mark.parametrize("seed,dt", [(12, 2e-3), (13, 2e-2)])
def test_sor_matches_independent_colour_sweep(seed, dt):
    g = Grid(8, 8, 8, 1 / 8, 1 / 8, 1 / 8)
    P = _sor_problem(g, seed)
    va = [x.copy() for x in P["vel"]]
    vb = [x.copy() for x in P["vel"]]
    lp_a = _oracle_sor(g, P, dt, 0.3, va)
    lp_b = _numpy_sor(g, P, dt, 0.3, vb)
    assert lp_a == lp_b
    for x, y in zip(va, vb):
        assert np.array_equal(x, y), np.max(np.abs(x - y))


def _explicit_vis(g, vel, rmu):
    """div(2 mu S(vel)) at u, v, w points through the stress chain."""
    s, tau = g.zeros(6), g.zeros(6)
    O("calsij", I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), *vel, s)
    O("cal_arith_tau", I(g.ni), I(g.nj), I(g.nk), s, rmu, tau)
    out = [g.zeros() for _ in range(3)]
    O("cal_div_tensor", I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), tau, *out)
    return out


def test_sor_fixed_point_solves_crank_nicolson_equation():
    """Repeated solves converge to the fixed point of the sweep, which must
    satisfy u/dt - vis(u)/(2 rho00) = src at every updated cell, with vis
    from the explicit stress chain.  Each colour pair sees the ghosts of the
    refresh before it (srcs/solu_sor4.f:176-180): bndu of all components,
    then bnd_periodic of the component swept last -- for the first pair of a
    component that is the previous component, whose bndu pass leaves the
    staggered periodic ghosts of u (in x) and w (in z) one cell off."""
    g = Grid(8, 8, 8, 1 / 8, 1 / 8, 1 / 8)
    P = _sor_problem(g, 14)
    dt, uwall = 5e-3, 0.3
    vel = [x.copy() for x in P["vel"]]
    for _ in range(12):
        _oracle_sor(g, P, dt, uwall, vel)
    call, _ = bc_lib()
    checked = 0
    for comp in (U, V, W):
        jhi = g.nj - 1 if comp == V else g.nj
        si, sj, sk = SHIFT[comp]
        for pair in range(4):
            prev = (comp + 2) % 3 if pair == 0 else comp
            st = [x.copy() for x in vel]
            call("bndu", WALLS, I(g.ni), I(g.nj), I(g.nk), *st, D(uwall), D(g.dy), P["l1"],
                 P["l2"], P["phi"])
            call("bnd_periodic", I(g.ni), I(g.nj), I(g.nk), st[prev])
            vis = _explicit_vis(g, st, P["rmun"])[comp]
            for (a, b, c) in COLOURS[2 * pair:2 * pair + 2]:
                i = np.arange(a, g.ni + 1, 2) + 2
                j = np.arange(b, jhi + 1, 2) + 2
                k = np.arange(c, g.nk + 1, 2) + 2
                K, J, Ii = np.ix_(k, j, i)
                rho00 = (P["rho"][K, J, Ii] + P["rho"][K + sk, J + sj, Ii + si]
                         + P["rhon"][K, J, Ii] + P["rhon"][K + sk, J + sj, Ii + si]) * 0.25
                lhs = st[comp][K, J, Ii] / dt - 0.5 * vis[K, J, Ii] / rho00
                rhs = P["src"][comp][K, J, Ii]
                scale = np.abs(st[comp]).max() / dt
                assert np.max(np.abs(lhs - rhs)) <= 1e-11 * scale, (comp, pair, (a, b, c))
                checked += lhs.size
    assert checked == 3 * g.ni * g.nj * g.nk - g.ni * g.nk  # every updated cell once


# ---------------------------------------------------------------------------
# Contact angle, monitors
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("theta", [60.0, 100.0, 135.0])
def test_contact_angle_extends_a_planar_interface(theta):
    """phi linear with gradient (s sin th, +-cos th, 0), s = sign(i - ni/2), the
    y component pointing away from the wall at each wall (the mirror image at
    j = nj), is reproduced exactly in all three ghost layers on square cells
    (the construction's normal is the physical one when dx = dy = dz)."""
    g = Grid(16, 8, 8, 0.1, 0.1, 0.1)
    th = np.deg2rad(theta)
    x, y, z = g.coords()
    xc = (g.ni // 2 - 0.5) * g.dx
    s = np.where(np.arange(-2, g.ni + 4) - g.ni // 2 >= 0, 1.0, -1.0)[None, None, :]
    theta_arr = np.full(g.shape, theta)
    for side, ys, ghosts in ((0, 1.0, (0, 1, 2)), (1, -1.0, (g.nj + 3, g.nj + 4, g.nj + 5))):
        phi = 0.3 + s * np.sin(th) * (x - xc) + ys * np.cos(th) * y
        got = phi.copy()
        got[:, list(ghosts), :] = 99.0
        O("bnd_contact_angle", WALLS, I(g.ni), I(g.nj), I(g.nk), got, theta_arr, D(g.dx),
          D(g.dy), D(g.dz))
        checked = 0
        for i in range(1, g.ni + 1):
            if i - 1 < g.ni // 2 <= i + 1:  # central difference straddles the sign change
                continue
            for jg in ghosts:
                assert np.allclose(got[3:g.nk + 3, jg, i + 2], phi[3:g.nk + 3, jg, i + 2],
                                   rtol=0, atol=1e-12), (side, jg, i)
                checked += 1
        assert checked == 3 * (g.ni - 2)


def test_contact_angle_closed_form_on_non_square_cells():
    """With n = (s sin th ax, cos th, s sin th az)/|.| (ax = dy/dx, az = dy/dz),
    n_y/n_x * ax = n_y/n_z * az = s cot th: the aspect ratios cancel, so for
    any phi and any cell shape the first ghost is
        q(i,1,k) - s cot(th) ((q(i+1) - q(i-1)) + (q(k+1) - q(k-1)))/2,
    up to the 1e-20 guards, and the next two extrapolate linearly
    (srcs/bnd_contact_angle.f90:32-50,63-82).  dx, dy, dz all differ here, so
    an aspect ratio swapped, inverted or left out of either term fails."""
    g = Grid(12, 6, 10, 0.1, 0.037, 0.23)
    rng = np.random.default_rng(61)
    q = H.smooth_phi(g, rng)
    theta = rng.uniform(20.0, 160.0, g.shape)
    got = q.copy()
    O("bnd_contact_angle", WALLS, I(g.ni), I(g.nj), I(g.nk), got, theta, D(g.dx), D(g.dy),
      D(g.dz))
    s = np.where(np.arange(-2, g.ni + 4) - g.ni // 2 >= 0, 1.0, -1.0)
    inner = (slice(1, g.nk + 5), slice(1, g.ni + 5))  # i, k in [-1, n+2]: both neighbours exist
    for jw, jg, js in ((1, 0, -1), (g.nj, g.nj + 1, 1)):
        w = q[:, jw + 2, :]
        cot = 1.0 / np.tan(np.deg2rad(theta[:, jw + 2, :]))
        gx = np.zeros_like(w)
        gz = np.zeros_like(w)
        gx[:, 1:-1] = w[:, 2:] - w[:, :-2]
        gz[1:-1, :] = w[2:, :] - w[:-2, :]
        g0 = w - s[None, :] * cot * (gx + gz) * 0.5
        g1 = 2.0 * g0 - w
        g2 = 2.0 * g1 - g0
        scale = np.abs(w).max() + np.abs(cot * (gx + gz)).max()
        for jj, want in ((jg, g0), (jg + js, g1), (jg + 2 * js, g2)):
            err = np.abs(got[:, jj + 2, :][inner] - want[inner]).max()
            assert err <= 1e-13 * scale, (jw, jj, err)


def test_cal_vel_and_center_closed_forms():
    g = Grid(16, 8, 8, 0.25, 0.25, 0.25)
    phi = g.zeros()
    # a block of phi = 1 over cells i in [3,6], j in [2,4], k in [5,8]
    phi[5 + 2:8 + 3, 2 + 2:4 + 3, 3 + 2:6 + 3] = 1.0
    U0, V0, W0 = 0.4, -0.7, 1.1
    u, v, w = np.full(g.shape, U0), np.full(g.shape, V0), np.full(g.shape, W0)
    vel, re = np.zeros(3), np.zeros(3)
    O("cal_vel", I(0), I(g.ni), I(g.nj), I(g.nk), D(0.2), D(1000.0), D(0.01), u, v, w, phi, vel, re)
    assert np.allclose(vel, [U0, V0, W0], rtol=1e-15, atol=0)
    assert np.allclose(re, 1000.0 * np.array([U0, V0, W0]) * 0.4 / 0.01, rtol=1e-14)
    # centroid of the block: delta = yl/svall = dy, cell centres at delta*(i-1/2)
    center = np.zeros(3)
    c1, c2 = D(0.0), D(0.0)
    a1 = ctypes.c_int(2)
    a2 = ctypes.c_int(2)
    lap = ctypes.c_int(0)
    velo = D(0.0)
    lib = H.oracle()
    yl = g.nj * g.dy
    args = [I(0), I(0), I(g.nj), I(g.ni), I(g.nj), I(g.nk), D(g.ni * g.dx), D(yl),
            D(g.nk * g.dz), D(0.5), D(1.0), phi, center, c1, c2]
    fn = lib.orc_cal_center
    fn.restype = None
    fn(*[H._arg(a) for a in args], ctypes.byref(a1), ctypes.byref(a2), ctypes.byref(lap),
       ctypes.byref(velo.obj))
    d = yl / g.nj
    want = [d * (np.mean([3, 4, 5, 6]) - 0.5), d * (np.mean([2, 3, 4]) - 0.5),
            d * (np.mean([5, 6, 7, 8]) - 0.5)]
    assert np.allclose(center, want, rtol=1e-15)
    assert abs(velo.value - want[0] / 0.5) <= 1e-15 * want[0] / 0.5 * 4
    assert c1.value == center[0]
    # x-centroid 1.125 < xl/4 = 1.0? no: 1.125 lies in [1, 3] -> area 2
    assert a1.value == 2 and a2.value == 2 and lap.value == 0


# --------------------------------------------------------------------------
# caldt (srcs/caldt.f): closed forms for each limit and a brute force
# --------------------------------------------------------------------------

def _caldt(g, u, v, w, cfl, rhol, rhog, rmul, rmug, sigma, walls=True, time=0.0):
    nid = np.full(6, -1 if walls else 0, dtype=np.int32)
    dt, t = D(0.0), D(time)
    lim = np.zeros(3)
    H.call_oracle("caldt", I(0), nid, I(0), I(1), I(g.ni), I(g.nj), I(g.nk), I(1), I(0),
                  D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), D(cfl), D(rhol), D(rhog), D(rmul), D(rmug),
                  D(sigma), u, v, w, dt, t, lim)
    return dt.value, t.value, lim


def test_caldt_limits_closed_form():
    g = Grid(8, 6, 4, 0.1, 0.05, 0.2)
    dcell = (g.dx * g.dy * g.dz) ** (1.0 / 3.0)
    u, v, w = g.zeros(), g.zeros(), g.zeros()
    # advection-limited: the largest |u_c|/d_c is |w|/dz = 3/0.2 < |v|/dy = 2/0.05
    u[...] = 1.0
    v[...] = -2.0
    w[...] = 3.0
    dt, t, lim = _caldt(g, u, v, w, 0.1, 1.0, 0.1, 1e-3, 1e-5, 1e-3, time=1.0)
    assert lim[0] == pytest.approx(0.1 / (2.0 / 0.05), rel=1e-15)
    nu = max(1e-3 / 1.0, 1e-5 / 0.1)
    assert lim[1] == pytest.approx(0.1 * 0.5 * dcell ** 2 / nu * 100.0, rel=1e-14)
    assert lim[2] == pytest.approx(0.1 * np.sqrt(dcell ** 3 / (2e-3 / 1.1)), rel=1e-14)
    assert dt == min(lim) and t == 1.0 + dt
    assert dt == lim[0]
    # at rest the floor 1e-20 leaves the viscous or capillary limit
    z = g.zeros()
    dt, _, lim = _caldt(g, z, z, z, 0.1, 1.0, 0.1, 10.0, 1.0, 1e-6)
    assert lim[0] == 0.1 / 1e-20 and dt == lim[1] < lim[2]
    dt, _, lim = _caldt(g, z, z, z, 0.1, 1.0, 0.1, 1e-6, 1e-7, 10.0)
    assert dt == lim[2] < lim[1]


def test_caldt_max_against_numpy_and_wall_row():
    g = Grid(10, 7, 6, 0.1, 0.1, 0.1)
    rng = np.random.default_rng(77)
    u, v, w = (rng.standard_normal(g.shape) for _ in range(3))
    # ghosts and the v row j = nj under the Y_PLUS wall do not count
    v[:, g.nj + 2, :] = 1e6
    u[:, :, 0] = 1e6
    inner = (slice(3, g.nk + 3), slice(3, g.nj + 3), slice(3, g.ni + 3))
    vin = (slice(3, g.nk + 3), slice(3, g.nj + 2), slice(3, g.ni + 3))
    want = max(np.abs(u[inner] / g.dx).max(), np.abs(v[vin] / g.dy).max(),
               np.abs(w[inner] / g.dz).max())
    dt, _, lim = _caldt(g, u, v, w, 0.2, 1.0, 1.0, 1e-6, 1e-6, 1e-6)
    assert lim[0] == pytest.approx(0.2 / want, rel=1e-15)
    # without walls the row j = nj counts
    _, _, lim2 = _caldt(g, u, v, w, 0.2, 1.0, 1.0, 1e-6, 1e-6, 1e-6, walls=False)
    assert lim2[0] == pytest.approx(0.2 / (1e6 / g.dy), rel=1e-15)
