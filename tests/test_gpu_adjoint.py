"""Parity of the GPU adjoint routines <name>_b_ with the oracle's.

For every routine, on the same seeded inputs: the GPU adjoint equals the
oracle adjoint (itself pinned by the dot-product identity against the
forward oracle, tests/test_oracle_adjoint.py) to rounding, and the GPU
adjoint is the transpose of the GPU forward (<F x, y> = <x, F^T y>).  The
Gauss-Seidel adjoint is checked with coefficient arrays and with the
coefficients formed on the fly from the viscosity (registered arrays,
cal_arith_coef_vis_ deferred), at the sweep count of the GPU forward call.
The last test runs the whole pressure projection (caldiv_ .. corunp_explicit_,
srcs/main.f90:634-650) forward and reversed through the C ABI.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import adjoint_cases as A
import felharness as H
from felharness import D, I, Grid

pytestmark = pytest.mark.gpu

GRIDS = [Grid(8, 6, 8, 1 / 8, 1 / 6, 1 / 8), Grid(32, 16, 8, 0.1, 0.05, 0.2)]


def rel_diff(a: dict, b: dict) -> float:
    worst = 0.0
    for k in a:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        sc = max(float(np.max(np.abs(x))), 1e-300)
        worst = max(worst, float(np.max(np.abs(x - y))) / sc)
    return worst


def check_case(case, rtol=1e-13):
    print(f"adjoint case {case.name} grid {case.g.ni}x{case.g.nj}x{case.g.nk}", flush=True)
    x = case.inputs()
    y = case.out_adjoints()
    bo = case.adjoint(H.call_oracle, {k: v.copy() for k, v in y.items()})
    bg = case.adjoint(H.call_gpu, {k: v.copy() for k, v in y.items()})
    assert rel_diff(bo, bg) <= rtol, (case.name, rel_diff(bo, bg))
    fx = case.forward(H.call_gpu, x)
    lhs, rhs = A.dot(fx, y), A.dot(x, bg)
    assert abs(lhs - rhs) <= 1e-12 * max(A.scale(fx, y), A.scale(x, bg)), (case.name, lhs, rhs)


@pytest.mark.parametrize("gi", range(len(GRIDS)))
@pytest.mark.parametrize("cls", [A.Caldiv, A.Tdma1, A.Tdma3, A.Periodic, A.Neumann, A.Avzero,
                                 A.Corunp, A.Bndu], ids=lambda c: c.name)
def test_adjoint_matches_oracle(cls, gi):
    check_case(cls(GRIDS[gi], 51 + gi))


@pytest.mark.parametrize("gi", range(len(GRIDS)))
@pytest.mark.parametrize("cls", [A.Calphat, A.GradPrs, A.Calsij, A.ArithTau, A.DivTensor, A.Srcu],
                         ids=lambda c: c.name)
def test_explicit_term_adjoint_matches_oracle(cls, gi):
    check_case(cls(GRIDS[gi], 131 + gi))


@pytest.mark.parametrize("gi", range(len(GRIDS)))
def test_weno5_adjoint_matches_oracle(gi):
    """Nonlinear: the GPU adjoint at x0 equals the oracle's to rounding, and
    both match central differences of the GPU forward along v."""
    case = A.Weno(GRIDS[gi], 141 + gi)
    print(f"adjoint case cal_advu_weno5 grid {case.g.ni}x{case.g.nj}x{case.g.nk}", flush=True)
    x, y = case.inputs(), case.out_adjoints()
    bo = case.adjoint(H.call_oracle, {k: v.copy() for k, v in y.items()})
    bg = case.adjoint(H.call_gpu, {k: v.copy() for k, v in y.items()})
    assert rel_diff(bo, bg) <= 1e-12, rel_diff(bo, bg)
    fx = case.forward(H.call_gpu, x)
    lhs, rhs = A.dot(fx, y), A.dot(x, bg)
    assert abs(lhs - rhs) <= 1e-8 * max(A.scale(fx, y), A.scale(x, bg)), (lhs, rhs)


GRID_VOF = [Grid(16, 16, 16, 1 / 16, 1 / 16, 1 / 16), Grid(32, 16, 16, 0.05, 0.1, 0.05)]


@pytest.mark.parametrize("gi", range(len(GRIDS)))
def test_grad_p2a_adjoint(gi):
    check_case(A.GradP2a(GRIDS[gi], 201 + gi))


@pytest.mark.parametrize("gi", range(len(GRID_VOF)))
@pytest.mark.parametrize("cls", [A.Mthinc1, A.Mthinc2, A.Mthinc3, A.Summation, A.Gnbc,
                                 A.ContactAngle], ids=lambda c: c.name)
def test_vof_adjoint_matches_oracle(cls, gi):
    """Nonlinear: the GPU adjoint at x0 equals the oracle's to rounding
    (transcendentals: 1e-12), and matches central differences of the GPU
    forward along random directions."""
    case = cls(GRID_VOF[gi], 211 + gi)
    print(f"adjoint case {case.name} grid {case.g.ni}x{case.g.nj}x{case.g.nk}", flush=True)
    x, y = case.inputs(), case.out_adjoints()
    bo = case.adjoint(H.call_oracle, {k: v.copy() for k, v in y.items()})
    bg = case.adjoint(H.call_gpu, {k: v.copy() for k, v in y.items()})
    assert rel_diff(bo, bg) <= 1e-12, rel_diff(bo, bg)
    fx = case.forward(H.call_gpu, x)
    lhs, rhs = A.dot(fx, y), A.dot(x, bg)
    assert abs(lhs - rhs) <= 1e-7 * max(A.scale(fx, y), A.scale(x, bg)), (lhs, rhs)


@pytest.mark.parametrize("skew", [False, True], ids=["reference_coefficients", "skewed"])
def test_tdma2_adjoint(skew):
    check_case(A.Tdma2(GRIDS[1], 61, skew=skew))


def test_trans_adjoint():
    g = GRIDS[0]
    n = g.ni * g.nj * g.nk
    r = np.random.default_rng(3)
    svall = np.array([g.ni, g.nj, g.nk], dtype=np.int32)
    for name in ("trans_r2w_b", "trans_w2r_b"):
        a, b = r.standard_normal(n), r.standard_normal(n)
        ao, bo = a.copy(), b.copy()
        H.call_oracle(name, I(0), I(0), I(1), svall, np.zeros(1, np.int32), np.zeros(1), np.zeros(1), ao, bo)
        H.call_gpu(name, I(0), I(0), I(1), svall, np.zeros(1, np.int32), np.zeros(1), np.zeros(1), a, b)
        assert np.array_equal(a, ao) and np.array_equal(b, bo)


@pytest.mark.parametrize("gi", range(len(GRIDS)))
def test_sor_adjoint_coefficient_arrays(gi):
    g = GRIDS[gi]
    case = A.Sor(g, 71 + gi)
    print(f"adjoint case solu_sor4 arrays grid {g.ni}x{g.nj}x{g.nk}", flush=True)
    x = case.inputs()
    y = case.out_adjoints()
    fx = case.forward(H.call_gpu, x, lp_reader=A.sor_lp_gpu)
    bg = case.adjoint(H.call_gpu, {k: v.copy() for k, v in y.items()})
    assert H.gpu().felgpu_sor_b_onfly_() == 0
    bo = case.adjoint(H.call_oracle, {k: v.copy() for k, v in y.items()})
    assert rel_diff(bo, bg) <= 1e-12, rel_diff(bo, bg)
    lhs, rhs = A.dot(fx, y), A.dot(x, bg)
    assert abs(lhs - rhs) <= 1e-11 * max(A.scale(fx, y), A.scale(x, bg)), (lhs, rhs)


class Registered:
    def __init__(self):
        self.arrays = []

    def add(self, a):
        n, up = ctypes.c_int64(a.nbytes), ctypes.c_int(1)
        H.gpu().felgpu_register_(ctypes.c_void_p(a.ctypes.data), ctypes.byref(n), ctypes.byref(up))
        self.arrays.append(a)
        return a

    def download(self, a):
        H.gpu().felgpu_download_(ctypes.c_void_p(a.ctypes.data))
        return a

    def close(self):
        for a in self.arrays:
            H.gpu().felgpu_unregister_(ctypes.c_void_p(a.ctypes.data))
        self.arrays = []


def test_sor_adjoint_on_the_fly_coefficients():
    """Registered arrays: cal_arith_coef_vis_ defers the 45 coefficients,
    solu_sor4_ and solu_sor4_b_ form them from the viscosity."""
    g = GRIDS[1]
    case = A.Sor(g, 81)
    print(f"adjoint case solu_sor4 on-the-fly grid {g.ni}x{g.nj}x{g.nk}", flush=True)
    x = case.inputs()
    y = case.out_adjoints()
    reg = Registered()
    try:
        rmu = reg.add(case.rmu.copy())
        coef = [reg.add(g.zeros()) for _ in range(45)]
        H.call_gpu("cal_arith_coef_vis", I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy),
                   D(1 / g.dz), rmu, *coef)
        rho, rhon = reg.add(case.rho.copy()), reg.add(case.rhon.copy())
        l1, phi = reg.add(case.l1.copy()), reg.add(case.phi.copy())
        a = {k: reg.add(x[k].copy()) for k in case.keys}
        head = [I(0), I(0), A.WALLS.copy(), I(1), I(g.ni), I(g.nj), I(g.nk), np.zeros(8, np.int32),
                np.zeros(8), np.zeros(8), I(1), I(0), D(case.dt), rho, rhon] + coef
        H.call_gpu("solu_sor4", *head, *[a[k] for k in case.keys], D(float(x["uwall"][0])), D(g.dy),
                   l1, l1, phi)
        lp = A.sor_lp_gpu()
        fx = {k: reg.download(a[k]).copy() for k in ("un", "vn", "wn")}
        b = {k: g.zeros() for k in case.keys[:3]}
        b.update({k: y[k].copy() for k in ("un", "vn", "wn")})
        ub = D(0.0)
        H.call_gpu("solu_sor4_b", *head, *[b[k] for k in case.keys], ub, D(g.dy), l1, l1, phi, I(lp))
        assert H.gpu().felgpu_sor_b_onfly_() == 1
        b["uwall"] = np.array([ub.value])
    finally:
        reg.close()
    # the same adjoint from the oracle with the coefficient arrays
    case.lp = lp
    bo = case.adjoint(H.call_oracle, {k: v.copy() for k, v in y.items()})
    assert rel_diff(bo, b) <= 1e-12, rel_diff(bo, b)
    lhs, rhs = A.dot(fx, y), A.dot(x, b)
    assert abs(lhs - rhs) <= 1e-11 * max(A.scale(fx, y), A.scale(x, b)), (lhs, rhs)


def test_projection_chain_transpose():
    g = Grid(32, 16, 16, 1 / 32, 1 / 16, 1 / 16)
    print("adjoint case projection chain 32x16x16", flush=True)
    case = A.Tdma2(g, 91)
    rng = np.random.default_rng(92)
    x = A.consistent_velocity(g, rng)
    y = {k: rng.standard_normal(g.shape) for k in ("un", "vn", "wn", "pn")}
    fx = A.projection_forward(H.call_gpu, g, case.fft, x)
    bg = A.projection_adjoint(H.call_gpu, g, case.fft, y)
    bo = A.projection_adjoint(H.call_oracle, g, case.fft, y)
    assert rel_diff(bo, bg) <= 1e-11, rel_diff(bo, bg)
    lhs, rhs = A.dot(fx, y), A.dot(x, bg)
    assert abs(lhs - rhs) <= 1e-11 * max(A.scale(fx, y), A.scale(x, bg)), (lhs, rhs)


def momentum_cfg():
    import dataclasses
    from felmpf import config
    return dataclasses.replace(config.sessile_droplet().scaled(32, 16, 16), dt=1e-3, uwall=0.05)


@pytest.mark.parametrize("resident", [False, True], ids=["host_arrays", "registered"])
def test_momentum_step_reverse_matches_oracle(resident):
    """felmpf.adjoint.MomentumStep on the GPU (host arrays: coefficient
    arrays; registered: device-resident, coefficients on the fly) against
    the same reverse sweep through the oracle's routines."""
    from felmpf import adjoint
    cfg = momentum_cfg()
    print(f"adjoint case momentum step resident={resident} 32x16x16", flush=True)
    fixed, x0 = A.momentum_inputs(cfg, 161)
    y = A.random_like({k: x0["u"] for k in adjoint.OUT}, np.random.default_rng(162))
    so = adjoint.MomentumStep(cfg, fixed, call=H.call_oracle, lp_reader=A.sor_lp_oracle)
    fo = so.forward(x0)
    bo = so.reverse(y)
    sg = adjoint.MomentumStep(cfg, fixed, resident=resident)
    try:
        fg = sg.forward(x0)
        assert sg.lp == so.lp
        bg = sg.reverse(y)
        if resident:
            assert H.gpu().felgpu_sor_b_onfly_() == 1
    finally:
        sg.close()
    assert rel_diff(fo, fg) <= 1e-12, rel_diff(fo, fg)
    assert rel_diff(bo, bg) <= 1e-11, rel_diff(bo, bg)


def test_momentum_window_reverse_against_differences():
    """Three steps with the state rotation, reversed on the GPU, against
    Richardson central differences of the GPU forward."""
    from felmpf import adjoint
    cfg = momentum_cfg()
    print("adjoint case momentum window 3 steps 32x16x16", flush=True)
    fixed, x0 = A.momentum_inputs(cfg, 171)
    step = adjoint.MomentumStep(cfg, fixed)
    win = adjoint.MomentumWindow(step, 3)
    lps = []

    def fwd(x):
        out = win.forward(x)
        lps.append(step.lp)
        return out

    rng = np.random.default_rng(172)
    v, y = A.random_like(x0, rng), A.random_like(x0, rng)
    jv = A.richardson(fwd, x0, v, 1e-5)
    win.forward(x0)
    assert all(lp == step.lp for lp in lps), lps
    bx = win.reverse(y)
    lhs, rhs = A.dot(jv, y), A.dot(v, {k: bx[k] for k in v})
    assert abs(lhs - rhs) <= 1e-7 * A.scale(jv, y), (lhs, rhs)


def phase_cfg():
    import dataclasses
    from felmpf import config
    return dataclasses.replace(config.sessile_droplet().scaled(32, 16, 16), dt=2e-3, uwall=0.05)


@pytest.mark.parametrize("resident", [False, True], ids=["host_arrays", "registered"])
def test_phase_window_reverse_matches_oracle(resident):
    """felmpf.adjoint.PhaseWindow (2 steps at a fixed velocity) on the GPU
    against the same reverse through the oracle's routines, and the GPU
    reverse against central differences of the GPU forward."""
    from felmpf import adjoint
    cfg = phase_cfg()
    print(f"adjoint case phase window resident={resident} 32x16x16", flush=True)
    phi0, params = A.phase_inputs(cfg, 231)
    rng = np.random.default_rng(232)
    y = rng.standard_normal((2,) + phi0.shape)
    wo = adjoint.PhaseWindow(adjoint.PhaseStep(cfg, params, call=H.call_oracle), 2)
    fo = wo.forward(phi0)
    bo = wo.reverse(y)
    sg = adjoint.PhaseStep(cfg, params, resident=resident)
    try:
        wg = adjoint.PhaseWindow(sg, 2)
        fg = wg.forward(phi0)
        bg = wg.reverse(y)
    finally:
        sg.close()
    assert np.max(np.abs(fo - fg)) <= 1e-12 * np.max(np.abs(fo))
    for k in ("phi",) + adjoint.PHASE_PARAMS:
        scale = max(float(np.max(np.abs(bo[k]))), 1e-300)
        assert float(np.max(np.abs(bo[k] - bg[k]))) <= 1e-10 * scale, k
    assert abs(float(bo["uwall"][0]) - float(bg["uwall"][0])) <= 1e-10 * abs(float(bo["uwall"][0]))
    if resident:
        return
    keys = ("phi",) + adjoint.PHASE_PARAMS
    x0 = {"phi": phi0, **params}
    v = {k: rng.standard_normal(phi0.shape) for k in keys}

    def fwd(x):
        st = adjoint.PhaseStep(cfg, {k: x[k] for k in adjoint.PHASE_PARAMS})
        return {"phin": adjoint.PhaseWindow(st, 2).forward(x["phi"])}

    # a step of 1e-5 along this direction moves some cell across a switch
    # of the interface reconstruction within two steps; 1e-6 stays clear
    jv = A.richardson(fwd, x0, v, 1e-6)
    lhs = float(np.sum(jv["phin"] * y))
    rhs = sum(float(np.sum(v[k] * bg[k])) for k in keys)
    assert abs(lhs - rhs) <= 1e-7 * abs(lhs), (lhs, rhs)


# ---------------------------------------------------------------------------
# Reverse mode against the library's tangent mode: <J v, y> = <v, J^T y>
# with J v from forward-mode arithmetic (every device array a value and its
# derivative) through the same drop-in routines.  The two derivatives are
# written independently (the tangent mode differentiates the forward
# kernels' arithmetic, including the ten Newton iterations of the MTHINC
# offset, the reverse mode transposes each routine by hand), so agreement to
# rounding pins both.
# ---------------------------------------------------------------------------

def _tangent_run(make_step, state, seeds):
    """Forward of the step made by make_step() in the tangent mode, with
    the derivative directions seeds {array: direction}; returns the step's
    registered arrays' derivatives through fetch(array)."""
    import ctypes
    from felmpf import _lib
    gpu = _lib.gpu()
    gpu.felgpu_set_tangent_(ctypes.byref(ctypes.c_int(1)))
    fetched = {}
    try:
        st = make_step()
        try:
            st.set_state(state)
            for arr, dirn in seeds(st):
                full = np.ascontiguousarray(dirn, dtype=np.float64)
                assert full.size == arr.size
                gpu.felgpu_tangent_seed_(ctypes.c_void_p(arr.ctypes.data),
                                         full.ctypes.data_as(ctypes.c_void_p))
            st.forward_set()
            for name, arr in st.tangent_outputs().items():
                out = np.zeros(arr.shape)
                gpu.felgpu_tangent_fetch_(ctypes.c_void_p(arr.ctypes.data),
                                          out.ctypes.data_as(ctypes.c_void_p))
                fetched[name] = out
        finally:
            st.close()
    finally:
        gpu.felgpu_set_tangent_(ctypes.byref(ctypes.c_int(0)))
    return fetched


def test_phase_step_reverse_against_tangent_mode():
    from felmpf import adjoint
    cfg = phase_cfg()
    print("adjoint case phase step vs tangent mode 32x16x16", flush=True)
    phi0, params = A.phase_inputs(cfg, 241)
    rng = np.random.default_rng(242)
    keys = ("phi",) + adjoint.PHASE_PARAMS
    v = {k: rng.standard_normal(phi0.shape) for k in keys}
    y = rng.standard_normal((2,) + phi0.shape)

    class Step(adjoint.PhaseStep):
        def tangent_outputs(self):
            return {"phin": self.a["phin"]}

    jv = _tangent_run(lambda: Step(cfg, params, resident=True), {"phi": phi0},
                      lambda st: [(st.a["phi"], v["phi"])] +
                      [(st.par[k], v[k]) for k in adjoint.PHASE_PARAMS])
    st = adjoint.PhaseStep(cfg, params, resident=True)
    try:
        st.forward({"phi": phi0}, keep=False)
        bx = st.reverse({"phin": y})
    finally:
        st.close()
    lhs = float(np.sum(jv["phin"] * y))
    rhs = sum(float(np.sum(v[k] * bx[k])) for k in keys)
    print("phase <Jv,y>", lhs, "<v,J^T y>", rhs, flush=True)
    assert abs(lhs - rhs) <= 1e-10 * float(np.sqrt(np.sum(jv["phin"] ** 2) * np.sum(y ** 2))), \
        (lhs, rhs)


def test_momentum_step_reverse_against_tangent_mode():
    from felmpf import adjoint
    cfg = momentum_cfg()
    print("adjoint case momentum step vs tangent mode", cfg.shape, flush=True)
    fixed, state = A.momentum_inputs(cfg, 251)
    rng = np.random.default_rng(252)
    fst = ("fst_u", "fst_v", "fst_w", "fst_un", "fst_vn", "fst_wn")
    v = {k: rng.standard_normal(cfg.shape) for k in adjoint.STATE + fst}
    y = {k: rng.standard_normal(cfg.shape) for k in adjoint.OUT}

    class Step(adjoint.MomentumStep):
        def tangent_outputs(self):
            return {k: self.a[k] for k in adjoint.OUT}

    jv = _tangent_run(lambda: Step(cfg, fixed, resident=True), state,
                      lambda st: [(st.a[k], v[k]) for k in adjoint.STATE] +
                      [(st.fx[k], v[k]) for k in fst])
    st = adjoint.MomentumStep(cfg, fixed, resident=True)
    try:
        st.forward(state, keep=False)
        bx = st.reverse(y)
        fb = {k: st._get(st.b[k]) for k in fst}
    finally:
        st.close()
    lhs = sum(float(np.sum(jv[k] * y[k])) for k in adjoint.OUT)
    rhs = sum(float(np.sum(v[k] * bx[k])) for k in adjoint.STATE) + \
        sum(float(np.sum(v[k] * fb[k])) for k in fst)
    nj = np.sqrt(sum(float(np.sum(jv[k] ** 2)) for k in adjoint.OUT))
    ny = np.sqrt(sum(float(np.sum(y[k] ** 2)) for k in adjoint.OUT))
    print("momentum <Jv,y>", lhs, "<v,J^T y>", rhs, flush=True)
    assert abs(lhs - rhs) <= 1e-11 * nj * ny, (lhs, rhs)


# ---------------------------------------------------------------------------
# Material properties, surface tension and the parameter (all-argument)
# adjoints of the momentum routines.
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("gi", range(len(GRIDS)))
@pytest.mark.parametrize("cls", [A.CoefVis, A.Calrkap, A.GradP2uvw, A.SumFst], ids=lambda c: c.name)
def test_property_and_surface_tension_adjoint_matches_oracle(cls, gi):
    check_case(cls(GRIDS[gi], 311 + gi))


PARAM_CASES = [A.ArithMean, A.ArithTauP, A.SrcuP, A.BnduP, A.Calfst, A.SorP, A.SorPV,
               lambda g, s: A.Calrn(g, s, (1, 1, 1)), lambda g, s: A.Calrn(g, s, (0, 1, 0))]


@pytest.mark.parametrize("gi", range(len(GRIDS)))
@pytest.mark.parametrize("ci", range(len(PARAM_CASES)),
                         ids=["cal_arith_mean", "cal_arith_tau_bp", "cal_srcu_bp", "bndu_bp", "calfst",
                              "solu_sor4_bp", "solu_sor4_bpv", "calrn_111", "calrn_010"])
def test_parameter_adjoint_matches_oracle(ci, gi):
    """The GPU adjoint at x0 equals the oracle's to rounding, and matches
    Richardson central differences of the GPU forward along random
    directions (the solve at the frozen sweep count of the GPU forward)."""
    case = PARAM_CASES[ci](GRIDS[gi], 321 + gi)
    print(f"adjoint case {case.name} grid {case.g.ni}x{case.g.nj}x{case.g.nk}", flush=True)
    x, y = case.inputs(), case.out_adjoints()
    bo = case.adjoint(H.call_oracle, {k: v.copy() for k, v in y.items()})
    bg = case.adjoint(H.call_gpu, {k: v.copy() for k, v in y.items()})
    tol = 1e-11 if case.name.startswith("solu_sor4") else 1e-12
    assert rel_diff({k: bo[k] for k in x}, {k: bg[k] for k in x}) <= tol
    fx = case.forward(H.call_gpu, x)
    lhs, rhs = A.dot(fx, y), A.dot(x, bg)
    assert abs(lhs - rhs) <= 1e-7 * max(A.scale(fx, y), A.scale(x, bg)), (lhs, rhs)
    if case.name.startswith("solu_sor4"):
        assert len(set(case.lps)) == 1, case.lps


# ---------------------------------------------------------------------------
# The whole step
# ---------------------------------------------------------------------------

def _patterned():
    from felmpf import config as C
    return C.patterned_wall().scaled(32, 16, 32)


def _driver_state(cfg, warm):
    """State and wall maps of the drop-in host driver after `warm` steps,
    and its fields after one more step."""
    from felmpf.driver import GpuRun
    run = GpuRun(cfg)
    try:
        run.steps(warm)
        state = {k: run.field(k) for k in ("phi0", "phi1", "u", "v", "w", "uo", "vo", "wo", "p", "po")}
        params = {"theta0": run.field("theta_0_array"), "zeta": run.field("zeta_array"),
                  "l1": run.field("l1_array"), "l2": run.field("l2_array")}
        assert np.array_equal(run.field("phin1"), state["phi1"])
        run.steps(1)
        after = {k: run.field(k) for k in ("phi0", "phi1", "u", "v", "w", "p")}
    finally:
        run.close()
    return state, params, after


def test_full_step_is_the_driver_step_and_reverse_matches_oracle():
    """felmpf.adjoint.FullStep's forward through the drop-in routines gives
    the host driver's step (host/fel_driver.f90:step1) from the driver's own
    state, and its reverse on the GPU equals the reverse through the
    oracle's routines."""
    from felmpf import adjoint
    cfg = _patterned()
    print("adjoint case full step vs driver 32x16x32", flush=True)
    state, params, after = _driver_state(cfg, 3)
    st = adjoint.FullStep(cfg, params, resident=True)
    try:
        out = st.forward(state)
        y = {k: np.random.default_rng(341).standard_normal(cfg.shape) for k in OUT_FULL}
        bg = st.reverse(_out_b(y))
        lp = st.lp
    finally:
        st.close()
    for k, o in (("u", out["un"]), ("v", out["vn"]), ("w", out["wn"]), ("p", out["pn"]),
                 ("phi0", out["phin"][0]), ("phi1", out["phin"][1])):
        sc = max(float(np.max(np.abs(after[k]))), 1e-300)
        assert float(np.max(np.abs(o - after[k]))) <= 1e-12 * sc, k
    so = adjoint.FullStep(cfg, params, call=H.call_oracle, lp_reader=A.sor_lp_oracle)
    so.forward(state, keep=False)
    assert so.lp == lp
    bo = so.reverse(_out_b(y))
    for k in bo:
        sc = max(float(np.max(np.abs(bo[k]))), 1e-300)
        assert float(np.max(np.abs(bo[k] - bg[k]))) <= 1e-10 * sc, k


OUT_FULL = ("phi0", "phi1", "un", "vn", "wn", "pn")


def _out_b(y):
    return {"un": y["un"], "vn": y["vn"], "wn": y["wn"], "pn": y["pn"],
            "phin": np.stack([y["phi0"], y["phi1"]])}


@pytest.mark.parametrize("which", ["angle_patch0", "angle_patch1", "slip"])
def test_full_window_gradients_against_tangent_mode(which):
    """Wall-parameter gradients of the state after a 3-step window on the
    checkered wall, by the reverse sweep of the whole step (FullWindow),
    against the library's tangent mode run through the host driver along
    the same parameter direction: <J v, y> = <v, J^T y> for a contact-angle
    patch group or the slip length (l1 = l2 = ls with the friction zeta
    following it)."""
    from felmpf import adjoint
    from felmpf import sensitivity as S
    from felmpf.driver import GpuRun
    cfg = _patterned()
    print(f"adjoint case full window vs tangent {which} 32x16x32", flush=True)
    warm, nsteps = 2, 3
    seeds = S.parameter_seeds(cfg)
    seed = {"angle_patch0": seeds[0], "angle_patch1": seeds[1], "slip": seeds[-1]}[which]
    run = GpuRun(cfg, tangent=True)
    try:
        run.steps(warm)
        state = {k: run.field(k) for k in adjoint.FULL_STATE}
        params = {"theta0": run.field("theta_0_array"), "zeta": run.field("zeta_array"),
                  "l1": run.field("l1_array"), "l2": run.field("l2_array")}
        for name, dirn in seed.items():
            run.seed(name, dirn)
        run.steps(nsteps)
        jv = {k: run.tangent(k) for k in adjoint.FULL_STATE}
    finally:
        run.close()
    y = {k: np.random.default_rng(351).standard_normal(cfg.shape) for k in adjoint.FULL_STATE}
    st = adjoint.FullStep(cfg, params, resident=True)
    try:
        win = adjoint.FullWindow(st, nsteps)
        win.forward(state)
        bx = win.reverse(y)
    finally:
        st.close()
    names = {"theta_0_array": "theta0", "zeta_array": "zeta", "l1_array": "l1", "l2_array": "l2"}
    lhs = sum(float(np.sum(jv[k] * y[k])) for k in adjoint.FULL_STATE)
    rhs = sum(float(np.sum(dirn * bx[names[n]])) for n, dirn in seed.items())
    print(which, "<Jv,y>", lhs, "<v,J^T y>", rhs, flush=True)
    assert abs(lhs) > 0.0
    assert abs(lhs - rhs) <= 1e-8 * abs(lhs), (lhs, rhs)


def test_observable_gradients_reverse_equals_tangent():
    """Gradients of the droplet observables (liquid centroid, wetted area
    and its second moment) after 5 steps on the checkered wall, with respect
    to every patch's contact angle and the slip length: one reverse sweep
    (felmpf.sensitivity.AdjointWindow) against one tangent run per
    parameter (StepWindow, itself checked against central differences)."""
    torch = pytest.importorskip("torch")
    from felmpf import sensitivity as S
    cfg = _patterned()
    print("adjoint case observable gradients, all patches + slip, 32x16x32", flush=True)
    seeds = S.parameter_seeds(cfg, per_patch=True)
    assert len(seeds) >= 5
    nsteps = 5
    grads, vals = {}, {}
    for name, fn in (("tangent", S.step_window()), ("reverse", S.adjoint_window())):
        params = torch.zeros(len(seeds), dtype=torch.float64, requires_grad=True)
        phi = fn.apply(params, cfg, nsteps, seeds)
        obs = S.observables(cfg, phi)
        rows = []
        for m in range(obs.numel()):
            (gm,) = torch.autograd.grad(obs[m], params, retain_graph=True)
            rows.append(gm.numpy())
        grads[name] = np.array(rows)
        vals[name] = obs.detach().numpy()
    # the host driver and the routine-by-routine forward agree to rounding
    assert np.allclose(vals["tangent"], vals["reverse"], rtol=1e-12, atol=0.0)
    t, r = grads["tangent"], grads["reverse"]
    for m in range(t.shape[0]):
        sc = float(np.max(np.abs(t[m])))
        # the reverse sweep takes the MTHINC Newton offset's derivative by
        # the implicit-function theorem at the tenth iterate, the tangent
        # mode differentiates the ten iterations: where an iteration has not
        # converged to rounding (phi near 0 or 1) they differ slightly.  The
        # bar is the one the tangent mode meets against differences (1e-5),
        # with a floor of 1e-12 of the observable (these 5-step gradients
        # from rest are 1e-9..1e-7 of the observables' values)
        floor = 1e-12 * abs(float(vals["tangent"][m]))
        err = float(np.max(np.abs(t[m] - r[m])))
        print("observable", m, "max |grad|", sc, "max diff", err, flush=True)
        assert err <= 1e-5 * sc + floor, (m, t[m], r[m])


def test_device_window_equals_host_window_on_the_gpu():
    """DeviceFullWindow (every state, checkpoint and adjoint in device
    memory) against FullWindow (states through host memory), both
    device-resident on the checkered wall from the driver's start state:
    the same forward and adjoint, bit for bit."""
    from felmpf import adjoint
    from felmpf import sensitivity as S
    cfg = _patterned()
    print("adjoint case device window vs host window 32x16x32", flush=True)
    st, maps = S.driver_start(cfg)
    rng = np.random.default_rng(411)
    y = {k: rng.standard_normal(cfg.shape) for k in adjoint.FULL_STATE}
    res = {}
    for cls in (adjoint.FullWindow, adjoint.DeviceFullWindow):
        step = adjoint.FullStep(cfg, maps, resident=True)
        try:
            win = cls(step, 4, checkpoint_every=3)
            f = win.forward(st)
            b = win.reverse(y)
        finally:
            step.close()
        res[cls.__name__] = (f, b)
    f0, b0 = res["FullWindow"]
    f1, b1 = res["DeviceFullWindow"]
    for k in f0:
        assert np.array_equal(f0[k], f1[k]), k
    for k in b0:
        assert np.array_equal(b0[k], b1[k]), k
