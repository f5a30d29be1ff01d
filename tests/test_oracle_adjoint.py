"""Pins the oracle's adjoint routines (oracle/fel_oracle_adj.c).

The reference has no adjoint, so the pin is the defining identity of a
transpose: for the linear map F of each forward routine (pinned to the
compiled reference elsewhere) and seeded random x, y,

    <F x, y> = <x, F^T y>

to rounding.  A dropped term, a wrong index or sign, a coefficient taken
from the wrong cell or a reversed order of the forward's assignments each
break it by O(1) (the last test shows it for the one place where the
forward and its transpose differ only by the order of the coefficients).
"""
from __future__ import annotations

import numpy as np
import pytest

import adjoint_cases as A
import felharness as H
from felharness import Grid

GRIDS = [Grid(8, 6, 8, 1 / 8, 1 / 6, 1 / 8), Grid(16, 8, 4, 0.1, 0.05, 0.2)]
GRID_VOF = [Grid(16, 16, 16, 1 / 16, 1 / 16, 1 / 16), Grid(16, 8, 16, 0.05, 0.1, 0.05)]


def dot_check(case, call=H.call_oracle, rtol=1e-12):
    x = case.inputs()
    y = case.out_adjoints()
    fx = case.forward(call, x)
    by = case.adjoint(call, y)
    lhs, rhs = A.dot(fx, y), A.dot(x, by)
    sc = max(A.scale(fx, y), A.scale(x, by))
    assert abs(lhs - rhs) <= rtol * sc, (case.name, lhs, rhs, sc)
    return lhs, rhs, sc


@pytest.mark.parametrize("gi", range(len(GRIDS)))
@pytest.mark.parametrize("cls", [A.Caldiv, A.Tdma1, A.Tdma3, A.Periodic, A.Neumann, A.Avzero,
                                 A.Corunp, A.Bndu], ids=lambda c: c.name)
def test_dot_product(cls, gi):
    dot_check(cls(GRIDS[gi], 11 + gi))


@pytest.mark.parametrize("gi", range(len(GRIDS)))
@pytest.mark.parametrize("cls", [A.Calphat, A.GradPrs, A.Calsij, A.ArithTau, A.DivTensor, A.Srcu],
                         ids=lambda c: c.name)
def test_dot_product_explicit_terms(cls, gi):
    dot_check(cls(GRIDS[gi], 111 + gi))


@pytest.mark.parametrize("gi", range(len(GRIDS)))
def test_weno5_adjoint_against_differences(gi):
    """The WENO5 adjoint at x0 against Richardson central differences of the
    forward along random directions: <(F(x0+hv)-F(x0-hv))/2h, y> = <v, J^T y>
    to the differences' accuracy (the identity fails by O(1) for a wrong
    weight derivative)."""
    lhs, rhs, sc = dot_check(A.Weno(GRIDS[gi], 121 + gi), rtol=1e-8)


@pytest.mark.parametrize("gi", range(len(GRIDS)))
def test_dot_product_grad_p2a(gi):
    dot_check(A.GradP2a(GRIDS[gi], 181 + gi))


@pytest.mark.parametrize("gi", range(len(GRIDS)))
@pytest.mark.parametrize("cls", [A.Mthinc1, A.Mthinc2, A.Mthinc3, A.Summation, A.Gnbc,
                                 A.ContactAngle], ids=lambda c: c.name)
def test_vof_adjoints_against_differences(cls, gi):
    """The MTHINC chain's adjoints at x0 against Richardson central
    differences along random directions (branches, upwind directions and
    clamps fixed at x0; the Newton offset by the implicit-function theorem)."""
    dot_check(cls(GRID_VOF[gi], 191 + gi), rtol=1e-7)


@pytest.mark.parametrize("skew", [False, True], ids=["reference_coefficients", "skewed"])
def test_dot_product_tdma2(skew):
    dot_check(A.Tdma2(GRIDS[0], 21, skew=skew))


@pytest.mark.parametrize("gi", range(len(GRIDS)))
@pytest.mark.parametrize("smooth", [True, False], ids=["smooth", "random"])
def test_dot_product_sor(gi, smooth):
    case = A.Sor(GRIDS[gi], 31 + gi, smooth=smooth)
    dot_check(case, rtol=1e-11)
    assert 0 <= case.lp <= 21


@pytest.mark.parametrize("consistent", [True, False], ids=["consistent_ghosts", "random_ghosts"])
def test_projection_chain(consistent):
    """caldiv .. corunp_explicit (srcs/main.f90:634-650) forward, and the
    adjoints in reverse order."""
    g = Grid(16, 8, 8, 1 / 16, 1 / 8, 1 / 8)
    case = A.Tdma2(g, 93)
    rng = np.random.default_rng(94)
    x = A.consistent_velocity(g, rng) if consistent else \
        {k: rng.standard_normal(g.shape) for k in ("un", "vn", "wn", "phat")}
    y = {k: rng.standard_normal(g.shape) for k in ("un", "vn", "wn", "pn")}
    fx = A.projection_forward(H.call_oracle, g, case.fft, x)
    bx = A.projection_adjoint(H.call_oracle, g, case.fft, y)
    lhs, rhs = A.dot(fx, y), A.dot(x, bx)
    assert abs(lhs - rhs) <= 1e-12 * max(A.scale(fx, y), A.scale(x, bx)), (lhs, rhs)


def momentum_cfg():
    """A 16x8x16 droplet case with a moving wall and a step long enough for
    the Gauss-Seidel solve to take several sweeps."""
    import dataclasses
    from felmpf import config
    return dataclasses.replace(config.sessile_droplet().scaled(16, 8, 16), dt=1e-3, uwall=0.05)


@pytest.mark.parametrize("steps", [1, 2])
def test_momentum_window_against_differences(steps):
    """The reverse of `steps` velocity-pressure updates at a fixed phase
    (felmpf.adjoint.MomentumWindow through the oracle's routines) against
    Richardson central differences of the same forward along a random
    direction of the state; the Gauss-Seidel sweep counts of the perturbed
    runs equal the unperturbed ones (the adjoint freezes them)."""
    from felmpf import adjoint
    cfg = momentum_cfg()
    fixed, x0 = A.momentum_inputs(cfg, 151)
    step = adjoint.MomentumStep(cfg, fixed, call=H.call_oracle, lp_reader=A.sor_lp_oracle)
    win = adjoint.MomentumWindow(step, steps)
    win.forward(x0)
    lps = []

    def fwd(x):
        out = win.forward(x)
        lps.append(step.lp)
        return out

    rng = np.random.default_rng(152)
    v = A.random_like(x0, rng)
    y = A.random_like(x0, rng)
    jv = A.richardson(fwd, x0, v, 1e-5)
    win.forward(x0)
    base_lp = step.lp
    assert all(lp == base_lp for lp in lps), (lps, base_lp)
    bx = win.reverse(y)
    lhs, rhs = A.dot(jv, y), A.dot(v, {k: bx[k] for k in v})
    assert abs(lhs - rhs) <= 1e-7 * A.scale(jv, y), (lhs, rhs)


def test_momentum_uwall_adjoint_against_differences():
    """d(outputs)/d(uwall) . y against the reverse's uwall adjoint."""
    import dataclasses
    from felmpf import adjoint
    cfg = momentum_cfg()
    fixed, x0 = A.momentum_inputs(cfg, 153)
    y = A.random_like({k: x0["u"] for k in ("un", "vn", "wn", "pn")}, np.random.default_rng(154))

    def run(uw):
        c = dataclasses.replace(cfg, uwall=cfg.uwall + uw)
        st = adjoint.MomentumStep(c, fixed, call=H.call_oracle, lp_reader=A.sor_lp_oracle)
        return st, st.forward(x0)

    h = 1e-4
    outs = {s: run(s * h)[1] for s in (-2, -1, 1, 2)}
    d1 = {k: (outs[1][k] - outs[-1][k]) / (2 * h) for k in y}
    d2 = {k: (outs[2][k] - outs[-2][k]) / (4 * h) for k in y}
    jd = {k: (4 * d1[k] - d2[k]) / 3 for k in y}
    st, _ = run(0.0)
    assert st.lp >= 1
    bx = st.reverse(y)
    lhs, rhs = A.dot(jd, y), float(bx["uwall"][0])
    assert abs(lhs - rhs) <= 1e-7 * max(abs(lhs), 1e-30), (lhs, rhs)


def phase_cfg():
    import dataclasses
    from felmpf import config
    return dataclasses.replace(config.sessile_droplet().scaled(16, 8, 16), dt=2e-3, uwall=0.05)


@pytest.mark.parametrize("steps", [1, 2])
def test_phase_window_against_differences(steps):
    """The reverse of `steps` volume-fraction updates at a fixed velocity
    (felmpf.adjoint.PhaseWindow through the oracle's routines) against
    Richardson central differences of the same forward along a random
    direction of (phi, u, v, w, theta_0, zeta)."""
    from felmpf import adjoint
    cfg = phase_cfg()
    phi0, params = A.phase_inputs(cfg, 221)
    rng = np.random.default_rng(222)
    keys = ("phi",) + adjoint.PHASE_PARAMS
    x0 = {"phi": phi0, **params}
    v = {k: rng.standard_normal(phi0.shape) for k in keys}
    y = rng.standard_normal((2,) + phi0.shape)

    def fwd(x):
        st = adjoint.PhaseStep(cfg, {k: x[k] for k in adjoint.PHASE_PARAMS}, call=H.call_oracle)
        return {"phin": adjoint.PhaseWindow(st, steps).forward(x["phi"])}

    jv = A.richardson(fwd, x0, v, 1e-5)
    st = adjoint.PhaseStep(cfg, params, call=H.call_oracle)
    win = adjoint.PhaseWindow(st, steps)
    win.forward(phi0)
    bx = win.reverse(y)
    lhs = float(np.sum(jv["phin"] * y))
    rhs = sum(float(np.sum(v[k] * bx[k])) for k in keys)
    assert abs(lhs - rhs) <= 1e-7 * abs(lhs), (lhs, rhs)


def test_planted_transpose_mistake_is_caught():
    """The skewed Thomas systems are not symmetric: using the forward solve
    as its own adjoint (a plausible mistake) violates the identity by O(1)."""
    case = A.Tdma2(GRIDS[0], 22, skew=True)
    x, y = case.inputs(), case.out_adjoints()
    fx = case.forward(H.call_oracle, x)
    wrong = case.forward(H.call_oracle, y)  # F used in place of F^T
    lhs, rhs = A.dot(fx, y), A.dot(x, wrong)
    assert abs(lhs - rhs) > 1e-6 * A.scale(fx, y)


@pytest.mark.parametrize("gi", range(len(GRIDS)))
@pytest.mark.parametrize("cls", [A.CoefVis, A.Calrkap, A.GradP2uvw, A.SumFst], ids=lambda c: c.name)
def test_dot_product_properties_and_surface_tension(cls, gi):
    dot_check(cls(GRIDS[gi], 261 + gi))


@pytest.mark.parametrize("gi", range(len(GRIDS)))
@pytest.mark.parametrize("cls", [A.ArithMean, A.ArithTauP, A.SrcuP, A.BnduP, A.Calfst],
                         ids=lambda c: c.name)
def test_parameter_adjoints_against_differences(cls, gi):
    """All-argument adjoints (_bp) and the nonlinear property and
    surface-tension adjoints against Richardson central differences."""
    dot_check(cls(GRIDS[gi], 271 + gi), rtol=1e-8)


@pytest.mark.parametrize("gi", range(len(GRIDS)))
@pytest.mark.parametrize("stag", [(1, 1, 1), (0, 0, 1), (0, 1, 0), (1, 0, 0)], ids=str)
def test_calrn_adjoint_against_differences(gi, stag):
    dot_check(A.Calrn(GRIDS[gi], 281 + gi, stag), rtol=1e-8)


@pytest.mark.parametrize("gi", range(len(GRIDS)))
def test_sor_parameter_adjoint_against_differences(gi):
    """solu_sor4_bp (densities, coefficients, slip lengths, phi, sources,
    state, uwall) against Richardson central differences of the solve at a
    frozen sweep count (asserted equal in every perturbed solve)."""
    case = A.SorP(GRIDS[gi], 291 + gi)
    dot_check(case, rtol=1e-8)
    assert len(set(case.lps)) == 1


def test_sor_parameter_adjoint_reduces_to_state_adjoint():
    """With the parameter adjoints ignored, solu_sor4_bp gives what
    solu_sor4_b gives (same sweeps reversed, same order)."""
    g = GRIDS[0]
    case = A.Sor(g, 301)
    x = case.inputs()
    case.forward(H.call_oracle, x)
    y = case.out_adjoints()
    b1 = case.adjoint(H.call_oracle, y)
    p = A.SorP(g, 302)
    p.x0.update({k: x[k] for k in Sor_keys()})
    p.x0["uwall"] = x["uwall"]
    p.x0["rho"], p.x0["rhon"] = case.rho, case.rhon
    p.x0["l1"], p.x0["l2"], p.x0["phi"] = case.l1, case.l1, case.phi
    for i in range(45):
        p.x0[f"c{i}"] = case.coef[i]
    b2 = p.adjoint(H.call_oracle, y)
    for k in Sor_keys() + ("uwall",):
        assert np.array_equal(b1[k], b2[k]), k


def Sor_keys():
    return ("src_u", "src_v", "src_w", "un", "vn", "wn")


@pytest.mark.parametrize("steps", [1, 2])
def test_full_window_against_differences(steps):
    """The reverse of `steps` whole timesteps (felmpf.adjoint.FullWindow
    through the oracle's routines: phase update, contact-angle fill of
    phi(.,0), material properties, surface tension, velocity-pressure
    update with the adjoints of the densities, viscosities, forces, phi and
    slip lengths) against Richardson central differences of the same
    forward along a random direction of the whole state and the wall maps
    theta_0, zeta, l1, l2.  The Gauss-Seidel sweep counts of the perturbed
    runs equal the unperturbed ones."""
    from felmpf import adjoint
    cfg = phase_cfg()
    state, params = A.full_inputs(cfg, 331)
    rng = np.random.default_rng(332)
    y = {k: rng.standard_normal(cfg.shape) for k in adjoint.FULL_STATE}
    lps = []

    def fwd(x):
        st = adjoint.FullStep(cfg, {k: x[k] for k in adjoint.FULL_PARAMS}, call=H.call_oracle,
                              lp_reader=A.sor_lp_oracle)
        win = adjoint.FullWindow(st, steps)
        out = win.forward({k: x[k] for k in adjoint.FULL_STATE})
        lps.append(st.lp)
        return out

    x0 = {**state, **params}
    v = {k: rng.standard_normal(cfg.shape) for k in x0}
    # the WENO5 transport velocities are means of several cells and come
    # within 1e-6 of 0 here (a kink of the upwind weights); 2e-7 steps stay
    # clear of it
    jv = A.richardson(fwd, x0, v, 2e-7)
    st = adjoint.FullStep(cfg, params, call=H.call_oracle, lp_reader=A.sor_lp_oracle)
    win = adjoint.FullWindow(st, steps)
    win.forward(state)
    bx = win.reverse(y)
    assert len(set(lps)) == 1, lps
    lhs = sum(float(np.sum(jv[k] * y[k])) for k in adjoint.FULL_STATE)
    rhs = sum(float(np.sum(v[k] * bx[k])) for k in x0)
    assert abs(lhs - rhs) <= 1e-7 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("every", [1, 2, 3])
def test_full_window_checkpointing_gives_the_same_adjoint(every):
    """FullWindow with checkpoints every K steps (the segment's states
    recomputed in the reverse) gives exactly the adjoint of the window that
    keeps every state, and runs the forward steps + recomputations it
    should."""
    from felmpf import adjoint
    cfg = phase_cfg()
    state, params = A.full_inputs(cfg, 361)
    rng = np.random.default_rng(362)
    y = {k: rng.standard_normal(cfg.shape) for k in adjoint.FULL_STATE}
    steps = 4
    out = {}
    for K in (None, every):
        st = adjoint.FullStep(cfg, params, call=H.call_oracle, lp_reader=A.sor_lp_oracle)
        win = adjoint.FullWindow(st, steps, checkpoint_every=K)
        fwd = win.forward(state)
        out[K] = (fwd, win.reverse(y), win.forward_steps)
    f0, b0, n0 = out[None]
    f1, b1, n1 = out[every]
    for k in f0:
        assert np.array_equal(f0[k], f1[k]), k
    for k in b0:
        assert np.array_equal(b0[k], b1[k]), k
    # forward: steps; reverse: one recompute per reversed step plus the
    # segments' recomputed states (steps - number of checkpoints)
    ncp = -(-steps // every)
    assert n0 == 2 * steps
    assert n1 == 2 * steps + (steps - ncp)


@pytest.mark.parametrize("gi", range(len(GRIDS)))
def test_sor_viscosity_adjoint_against_differences(gi):
    """solu_sor4_bpv (the solve with its coefficients from the viscosity)
    against Richardson central differences of cal_arith_coef_vis then
    solu_sor4, at a frozen sweep count."""
    case = A.SorPV(GRIDS[gi], 371 + gi)
    dot_check(case, rtol=1e-8)
    assert len(set(case.lps)) == 1


def test_wall_gradients_on_a_developed_state_against_differences():
    """The checkered wall of the patterned-wall case (32x16x32), advanced 2
    steps by the oracle's own timestep: the gradients of a random function
    of the state after 3 more steps with respect to each contact-angle
    patch group and the slip length (l1 = l2 = ls with zeta following it),
    from one reverse sweep (FullWindow through the oracle's routines),
    against Richardson central differences of the same window (h = 0.005
    deg and 1% of the slip length, as for the tangent mode)."""
    from felmpf import adjoint
    from felmpf import config as C
    from felmpf import sensitivity as S
    from oracle_run import OracleRun
    cfg = C.patterned_wall().scaled(32, 16, 32)
    o = OracleRun(cfg)
    o.steps(2)
    state = {k: o.field(k) for k in adjoint.FULL_STATE}
    maps = {"theta0": o.field("theta_0_array"), "zeta": o.field("zeta_array"),
            "l1": o.field("l1_array"), "l2": o.field("l2_array")}
    n = 3
    rng = np.random.default_rng(381)
    y = {k: rng.standard_normal(cfg.shape) for k in adjoint.FULL_STATE}
    step = adjoint.FullStep(cfg, maps, call=H.call_oracle, lp_reader=A.sor_lp_oracle)
    win = adjoint.FullWindow(step, n)
    win.forward(state)
    bx = win.reverse(y)
    names = {"theta_0_array": "theta0", "zeta_array": "zeta", "l1_array": "l1", "l2_array": "l2"}
    seeds = S.parameter_seeds(cfg)
    lps = []

    def f(seed, t):
        p = {k: v.copy() for k, v in maps.items()}
        for name, dirn in seed.items():
            p[names[name]] = p[names[name]] + t * dirn
        st = adjoint.FullStep(cfg, p, call=H.call_oracle, lp_reader=A.sor_lp_oracle)
        out = adjoint.FullWindow(st, n).forward(state)
        lps.append(st.lp)
        return sum(float(np.sum(out[k] * y[k])) for k in adjoint.FULL_STATE)

    for seed in seeds:
        h = 0.005 if "theta_0_array" in seed else 0.01 * cfg.bottom.slip
        d1 = (f(seed, h) - f(seed, -h)) / (2 * h)
        d2 = (f(seed, 2 * h) - f(seed, -2 * h)) / (4 * h)
        fd = (4 * d1 - d2) / 3
        g = sum(float(np.sum(dirn * bx[names[name]])) for name, dirn in seed.items())
        assert abs(fd - g) <= 1e-6 * abs(fd), (list(seed), fd, g)
    assert len(set(lps)) == 1, lps


@pytest.mark.parametrize("which", ["area", "height"])
def test_observable_weights_are_gradients(which):
    """felmpf.sensitivity.observable_weights: the value equals the droplet
    observables' (wetted area, centroid height), and the weights are its
    gradient (central differences along a random direction; the area is
    linear, the height a ratio)."""
    torch = pytest.importorskip("torch")
    from felmpf import config as C
    from felmpf import sensitivity as S
    cfg = C.patterned_wall().scaled(16, 8, 16)
    rng = np.random.default_rng(391)
    phi = rng.uniform(0.0, 1.0, cfg.shape)
    val, w = S.observable_weights(cfg, phi, which)
    obs = S.observables(cfg, torch.from_numpy(np.ascontiguousarray(
        phi[3:cfg.nk + 3, 3:cfg.nj + 3, 3:cfg.ni + 3]))).numpy()
    assert abs(val - obs[3 if which == "area" else 1]) <= 1e-13 * abs(val)
    v = rng.standard_normal(cfg.shape)
    h = 1e-6
    fd = (S.observable_weights(cfg, phi + h * v, which)[0] -
          S.observable_weights(cfg, phi - h * v, which)[0]) / (2 * h)
    assert abs(fd - float(np.sum(w * v))) <= 1e-7 * abs(fd)


@pytest.mark.parametrize("every", [None, 2])
def test_device_window_equals_host_window(every):
    """DeviceFullWindow (states, checkpoints and adjoints moved by cpy_ and
    summed by cpy_b_, here on host arrays through the oracle's routines)
    gives FullWindow's forward and adjoint bit for bit, with a start-up
    phin(.,1) and with checkpoints."""
    from felmpf import adjoint
    cfg = phase_cfg()
    state, params = A.full_inputs(cfg, 401)
    state["phin1"] = state["phi1"] + 0.01 * np.random.default_rng(402).standard_normal(cfg.shape)
    rng = np.random.default_rng(403)
    y = {k: rng.standard_normal(cfg.shape) for k in adjoint.FULL_STATE}
    res = {}
    for cls in (adjoint.FullWindow, adjoint.DeviceFullWindow):
        st = adjoint.FullStep(cfg, params, call=H.call_oracle, lp_reader=A.sor_lp_oracle)
        win = cls(st, 3, checkpoint_every=every)
        f = win.forward(state)
        b = win.reverse(y)
        res[cls.__name__] = (f, b, win.forward_steps)
    f0, b0, n0 = res["FullWindow"]
    f1, b1, n1 = res["DeviceFullWindow"]
    for k in f0:
        assert np.array_equal(f0[k], f1[k]), k
    for k in b0:
        assert np.array_equal(b0[k], b1[k]), k
    assert n0 == n1
