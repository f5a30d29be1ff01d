"""Reverse mode (adjoint) of the timestep, through the library's drop-in
routines and their adjoint entry points (<name>_b_, <name>_bp_).

* MomentumStep: the velocity-pressure update (srcs/main.f90:526-667) with
  the phase fields fixed -- WENO5 advection of u and uo, the pressure
  predictor, the explicit viscous stress, the momentum source, the
  Crank-Nicolson Gauss-Seidel solve, the projection and the ghost
  refreshes.  params=True also gives the adjoints of the densities,
  viscosities, surface-tension forces, phi and slip lengths.
* PhaseStep: the volume-fraction update (srcs/main.f90:489-517) at a fixed
  velocity -- vertex gradients, MTHINC fluxes, update, truncation and mass
  correction, contact angle and its ghosts, summation.
* FullStep: the whole step (srcs/main.f90:489-694): PhaseStep, the
  contact-angle fill of phi(.,0), the material properties, the surface
  tension and MomentumStep(params=True); its reverse gives the adjoints of
  the state and of the wall maps theta_0, zeta, l1, l2 and uwall.

Each runs the reference's call sequence forward (forward(state), which
records what the reverse needs: primal velocities, sweep counts, the
refreshes' inputs) and reverses it (reverse(out_b)).  The *Window classes
chain steps with the reference's state rotation (srcs/main.f90:687-698)
and reverse a window, recomputing each step from its stored input state
(checkpoint and recompute); DeviceFullWindow keeps every state in device
memory.

The routines are FP64 (the library must be in its FP64 mode).  `call` is
the function that runs a drop-in routine by name -- by default the GPU
library's <name>_ entry point; arrays are numpy arrays in the reference
layout.  With resident=True every array is registered with the library, so
the sequence runs without host transfers.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .config import RunConfig

WALLS = np.array([-1, -1, -1, -1, -1, -1], dtype=np.int32)
STATE = ("u", "v", "w", "uo", "vo", "wo", "p", "po")
OUT = ("un", "vn", "wn", "pn")
FIXED = ("rho", "rhon", "rmu", "rmun", "phi", "l1", "l2", "fst_u", "fst_v", "fst_w", "fst_un",
         "fst_vn", "fst_wn")
# fixed fields whose adjoints MomentumStep(params=True) forms (the
# surface-tension forces' adjoints are always formed)
PARAMS = ("rho", "rhon", "rmu", "rmun", "phi", "l1", "l2")


class _Ref:
    def __init__(self, value, ctype):
        self.obj = ctype(value)

    @property
    def value(self):
        return self.obj.value


def I(v):  # noqa: E743 - Fortran INTEGER by reference
    return _Ref(int(v), ctypes.c_int)


def D(v):
    return _Ref(float(v), ctypes.c_double)


def _arg(a):
    if isinstance(a, _Ref):
        return ctypes.byref(a.obj)
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays must be contiguous")
        ct = ctypes.c_int if a.dtype == np.int32 else ctypes.c_double
        if a.dtype not in (np.int32, np.float64):
            raise TypeError(a.dtype)
        return a.ctypes.data_as(ctypes.POINTER(ct))
    raise TypeError(type(a))


def gpu_call(name: str, *args):
    """Runs the GPU library's drop-in routine <name>_."""
    fn = getattr(_lib.gpu(), name + "_")
    fn.restype = None
    fn(*[_arg(a) for a in args])


def sor_last_lp() -> int:
    lp, ratio = ctypes.c_int(0), ctypes.c_double(0.0)
    _lib.gpu().felgpu_sor_last_(ctypes.byref(lp), ctypes.byref(ratio))
    return lp.value


class MomentumStep:
    """The velocity-pressure update of one step, forward and reversed, for
    the fixed fields `fixed` (FIXED: densities rho, rhon, viscosities rmu,
    rmun, phi and the slip lengths l1, l2 of bndu, the surface-tension
    forces of both time levels)."""

    def __init__(self, cfg: RunConfig, fixed: dict, call=None, lp_reader=None,
                 resident: bool = False, params: bool = False):
        self.cfg = cfg
        self.call = call or gpu_call
        self.lp_reader = lp_reader or sor_last_lp
        self.resident = resident
        self.params = params
        self._registered = []
        g = cfg
        self.nm = g.ni * g.nj * g.nk
        self.n1 = max(g.nj, g.nk) + 6
        z = lambda: self.array(np.zeros(g.shape))  # noqa: E731
        self.fx = {k: self.array(np.array(fixed[k], dtype=np.float64)) for k in FIXED}  # own copies
        # pressure matrix and its modal form (srcs/main.f90:232-236, set-up)
        pc = [np.zeros(g.nj + 6) for _ in range(7)]
        self.call("cal_coef_prs", I(g.nj), D(g.rhog), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), *pc)
        self.fft = [np.zeros(g.nj + 6) for _ in range(3)]
        self.call("cal_coef_solp_fft", I(g.nj), pc[2], pc[3], *self.fft)
        # primal work arrays
        self.a = {k: z() for k in STATE + OUT + (
            "adv_u", "adv_v", "adv_w", "adv_uo", "adv_vo", "adv_wo", "phat", "prs_u", "prs_v",
            "prs_w", "vis_u", "vis_v", "vis_w", "src_u", "src_v", "src_w", "div", "dp")}
        self.a["s"] = self.array(np.zeros((6,) + g.shape))
        self.a["tau"] = self.array(np.zeros((6,) + g.shape))
        self.a["pr"] = self.array(np.zeros(self.nm))
        self.a["pi"] = self.array(np.zeros(self.nm))
        self.coef = [z() for _ in range(45)]
        # adjoints
        self.b = {k: z() for k in STATE + OUT + (
            "adv_u", "adv_v", "adv_w", "adv_uo", "adv_vo", "adv_wo", "phat", "prs_u", "prs_v",
            "prs_w", "vis_u", "vis_v", "vis_w", "src_u", "src_v", "src_w", "div", "dp",
            "fst_u", "fst_v", "fst_w", "fst_un", "fst_vn", "fst_wn")}
        self.b["s"] = self.array(np.zeros((6,) + g.shape))
        self.b["tau"] = self.array(np.zeros((6,) + g.shape))
        self.b["pr"] = self.array(np.zeros(self.nm))
        self.b["pi"] = self.array(np.zeros(self.nm))
        if params:
            # adjoints of the fixed fields (PARAMS); the refreshes' input u
            # and w (bndu_bp_ reads their wall rows)
            for k in PARAMS:
                self.b[k] = z()
            for k in ("ur1", "wr1", "ur2", "wr2"):
                self.a[k] = z()
        self.lp = None

    # -- arrays ---------------------------------------------------------
    def array(self, a: np.ndarray, upload: bool = True) -> np.ndarray:
        """Returns a (registered with the library when resident; upload=False
        leaves the device copy zero-filled instead of copying a's values)."""
        if self.resident:
            lib = _lib.gpu()
            n, up = ctypes.c_int64(a.nbytes), ctypes.c_int(1 if upload else 0)
            lib.felgpu_register_(ctypes.c_void_p(a.ctypes.data), ctypes.byref(n), ctypes.byref(up))
            self._registered.append(a)
        return a

    def _set(self, dst: np.ndarray, src: np.ndarray):
        dst[...] = src
        if self.resident:
            _lib.gpu().felgpu_upload_(ctypes.c_void_p(dst.ctypes.data))

    def _get(self, a: np.ndarray) -> np.ndarray:
        if self.resident:
            _lib.gpu().felgpu_download_(ctypes.c_void_p(a.ctypes.data))
        return a.copy()

    def close(self):
        lib = _lib.gpu()
        for a in self._registered:
            lib.felgpu_unregister_(ctypes.c_void_p(a.ctypes.data))
        self._registered = []

    # -- argument lists -------------------------------------------------
    def _g(self):
        g = self.cfg
        return [I(g.ni), I(g.nj), I(g.nk)]

    def _d(self):
        g = self.cfg
        return [D(1 / g.dx), D(1 / g.dy), D(1 / g.dz)]

    def _svall(self):
        g = self.cfg
        return np.array([g.ni, g.nj, g.nk], dtype=np.int32)

    def _tdma2_head(self):
        g = self.cfg
        return [I(0), I(1), self._svall(), D(g.rhog), D(1 / g.dx), D(1 / g.dz), *self.fft,
                np.zeros(self.n1), np.zeros(self.n1), np.zeros(self.n1)]

    def _sor_head(self):
        g = self.cfg
        return [I(0), I(0), WALLS, I(1), *self._g(), np.zeros(8, np.int32), np.zeros(8), np.zeros(8),
                I(1), I(0), D(g.dt), self.fx["rho"], self.fx["rhon"]] + self.coef

    def _bc_tail(self):
        g = self.cfg
        return [D(g.dy), self.fx["l1"], self.fx["l2"], self.fx["phi"]]

    # -- forward --------------------------------------------------------
    def forward(self, state: dict, keep: bool = True) -> dict:
        """srcs/main.f90:526-667 with the phase fixed; returns copies of
        (un, vn, wn, pn)."""
        self.set_state(state)
        self.forward_set()
        if not keep:
            return {}
        return {k: self._get(self.a[k]) for k in OUT}

    def set_state(self, state: dict):
        for k in STATE:
            self._set(self.a[k], state[k])

    def forward_set(self):
        """forward() on the state already in place (self.a[k], k in STATE);
        the outputs are left in self.a."""
        c, a, f, g = self.call, self.a, self.fx, self.cfg
        gh, dd = self._g(), self._d()
        c("cal_advu_weno5", *gh, *dd, a["u"], a["v"], a["w"], a["adv_u"], a["adv_v"], a["adv_w"])
        c("cal_advu_weno5", *gh, *dd, a["uo"], a["vo"], a["wo"], a["adv_uo"], a["adv_vo"], a["adv_wo"])
        c("calphat", *gh, a["p"], a["po"], a["phat"])
        c("cal_grad_prs", *gh, *dd, a["phat"], a["prs_u"], a["prs_v"], a["prs_w"])
        c("calsij", *gh, *dd, a["u"], a["v"], a["w"], a["s"])
        c("cal_arith_tau", *gh, a["s"], f["rmu"], a["tau"])
        c("cal_div_tensor", *gh, *dd, a["tau"], a["vis_u"], a["vis_v"], a["vis_w"])
        c("cal_srcu", WALLS, *gh, D(0.0), D(g.dt), D(0.0), D(g.grvp), a["u"], a["v"], a["w"],
          f["rho"], f["rhon"], a["adv_u"], a["adv_v"], a["adv_w"], a["adv_uo"], a["adv_vo"],
          a["adv_wo"], a["prs_u"], a["prs_v"], a["prs_w"], a["vis_u"], a["vis_v"], a["vis_w"],
          f["fst_u"], f["fst_v"], f["fst_w"], f["fst_un"], f["fst_vn"], f["fst_wn"], a["src_u"],
          a["src_v"], a["src_w"])
        c("cal_arith_coef_vis", *gh, *dd, f["rmun"], *self.coef)
        for s, t in (("u", "un"), ("v", "vn"), ("w", "wn")):
            c("cpy", *gh, a[s], a[t])
        c("solu_sor4", *self._sor_head(), a["src_u"], a["src_v"], a["src_w"], a["un"], a["vn"],
          a["wn"], D(g.uwall), *self._bc_tail())
        self.lp = self.lp_reader()
        self._keep_refresh_input(1)
        self._refresh(a)
        nm = self.nm
        c("caldiv", I(0), I(1), *gh, *dd, D(g.dt), a["un"], a["vn"], a["wn"], a["div"], D(0.0))
        c("solp_fft_tdma1", *gh, np.zeros(2 * nm), a["div"], a["pr"], a["pi"])
        c("solp_fft_tdma2", *self._tdma2_head(), a["pr"], a["pi"])
        c("solp_fft_tdma3", *gh, np.zeros(2 * nm), a["pr"], a["pi"], a["dp"])
        c("bnd_avzero", I(0), I(1), *gh, a["dp"])
        c("bnd_neumann", WALLS, *gh, a["dp"])
        c("bnd_periodic", *gh, a["dp"])
        c("corunp_explicit", WALLS, *gh, D(g.rhog), *dd, D(g.dt), a["dp"], a["phat"], a["un"],
          a["vn"], a["wn"], a["pn"])
        self._keep_refresh_input(2)
        self._refresh(a)
        c("bnd_avzero", I(0), I(1), *gh, a["pn"])
        c("bnd_neumann", WALLS, *gh, a["pn"])
        c("bnd_periodic", *gh, a["pn"])

    def _refresh(self, a):
        """bndu + bnd_periodic of un, vn, wn (srcs/main.f90:619-622, 652-655)."""
        c, g, gh = self.call, self.cfg, self._g()
        c("bndu", WALLS, *gh, a["un"], a["vn"], a["wn"], D(g.uwall), *self._bc_tail())
        for k in ("un", "vn", "wn"):
            c("bnd_periodic", *gh, a[k])

    def _keep_refresh_input(self, n: int):
        """Copies of un, wn before refresh n (bndu_bp_ reads their wall rows)."""
        if self.params:
            gh, a = self._g(), self.a
            self.call("cpy", *gh, a["un"], a[f"ur{n}"])
            self.call("cpy", *gh, a["wn"], a[f"wr{n}"])

    def _refresh_b(self, b, uwallb, n: int = 0):
        c, gh, g = self.call, self._g(), self.cfg
        for k in ("wn", "vn", "un"):
            c("bnd_periodic_b", *gh, b[k])
        if self.params:
            a, f = self.a, self.fx
            c("bndu_bp", WALLS, *gh, a[f"ur{n}"], b["un"], b["vn"], a[f"wr{n}"], b["wn"],
              D(g.uwall), uwallb, D(g.dy), f["l1"], b["l1"], f["l2"], b["l2"], f["phi"], b["phi"])
        else:
            c("bndu_b", WALLS, *gh, b["un"], b["vn"], b["wn"], uwallb, *self._bc_tail())

    # -- reverse --------------------------------------------------------
    def reverse(self, out_b: dict) -> dict:
        """Adjoints of (u, v, w, uo, vo, wo, p, po) and uwall from those of
        (un, vn, wn, pn), at the state of the last forward call."""
        for k in OUT:
            self._set(self.b[k], out_b[k])
        uwallb = self.reverse_set()
        out = {k: self._get(self.b[k]) for k in STATE}
        out["uwall"] = np.array([uwallb])
        if self.params:
            for k in FIXED:
                out[k] = self._get(self.b[k])
        return out

    def reverse_set(self) -> float:
        """reverse() on the output adjoints already in place (the arrays
        self.b[k], k in OUT); the state adjoints are left in self.b and the
        uwall adjoint returned.  Every other adjoint array the sweep
        consumes is zero again when it ends, so only the state adjoints
        are cleared here."""
        if self.lp is None:
            raise RuntimeError("reverse() needs a forward() of the same step first")
        c, a, b, f, g = self.call, self.a, self.b, self.fx, self.cfg
        gh, dd, nm = self._g(), self._d(), self.nm
        c("init_q", *gh, b["u"], b["v"], b["w"])
        c("init_q", *gh, b["uo"], b["vo"], b["wo"])
        c("init_q", *gh, b["p"], b["po"], b["div"])
        if self.params:
            for k in PARAMS:
                c("init_q", *gh, b[k], b["div"], b["div"])
            c("init_q", *gh, b["fst_u"], b["fst_v"], b["fst_w"])
            c("init_q", *gh, b["fst_un"], b["fst_vn"], b["fst_wn"])
        uwallb = D(0.0)
        c("bnd_periodic_b", *gh, b["pn"])
        c("bnd_neumann_b", WALLS, *gh, b["pn"])
        c("bnd_avzero_b", I(0), I(1), *gh, b["pn"])
        self._refresh_b(b, uwallb, 2)
        c("corunp_explicit_b", WALLS, *gh, D(g.rhog), *dd, D(g.dt), b["dp"], b["phat"], b["un"],
          b["vn"], b["wn"], b["pn"])
        c("bnd_periodic_b", *gh, b["dp"])
        c("bnd_neumann_b", WALLS, *gh, b["dp"])
        c("bnd_avzero_b", I(0), I(1), *gh, b["dp"])
        c("solp_fft_tdma3_b", *gh, np.zeros(2 * nm), b["pr"], b["pi"], b["dp"])
        c("solp_fft_tdma2_b", *self._tdma2_head(), b["pr"], b["pi"])
        c("solp_fft_tdma1_b", *gh, np.zeros(2 * nm), b["div"], b["pr"], b["pi"])
        c("caldiv_b", I(0), I(1), *gh, *dd, D(g.dt), b["un"], b["vn"], b["wn"], b["div"])
        self._refresh_b(b, uwallb, 1)
        if self.params:
            pairs = []
            for k in ("src_u", "src_v", "src_w"):
                pairs += [a[k], b[k]]
            for s_, t in (("u", "un"), ("v", "vn"), ("w", "wn")):
                pairs += [a[s_], b[t]]  # the solve's input state is the copy of u, v, w
            # the solve and cal_arith_coef_vis reversed together: the
            # coefficients are those of the viscosity rmun
            head = self._sor_head()
            c("solu_sor4_bpv", *head[:13], f["rho"], b["rho"], f["rhon"], b["rhon"], *dd,
              f["rmun"], b["rmun"], *pairs, D(g.uwall), uwallb, D(g.dy), f["l1"], b["l1"],
              f["l2"], b["l2"], f["phi"], b["phi"], I(self.lp))
        else:
            c("solu_sor4_b", *self._sor_head(), b["src_u"], b["src_v"], b["src_w"], b["un"],
              b["vn"], b["wn"], uwallb, *self._bc_tail(), I(self.lp))
        for s_, t in (("u", "un"), ("v", "vn"), ("w", "wn")):
            c("cpy_b", *gh, b[s_], b[t])
        if self.params:
            c("cal_srcu_bp", WALLS, *gh, D(0.0), D(g.dt), D(0.0), D(g.grvp), b["u"], b["v"],
              b["w"], f["rho"], b["rho"], f["rhon"], b["rhon"], b["adv_u"], b["adv_v"],
              b["adv_w"], b["adv_uo"], b["adv_vo"], b["adv_wo"], a["prs_u"], b["prs_u"],
              a["prs_v"], b["prs_v"], a["prs_w"], b["prs_w"], a["vis_u"], b["vis_u"], a["vis_v"],
              b["vis_v"], a["vis_w"], b["vis_w"], b["fst_u"], b["fst_v"], b["fst_w"],
              b["fst_un"], b["fst_vn"], b["fst_wn"], b["src_u"], b["src_v"], b["src_w"])
        else:
            c("cal_srcu_b", WALLS, *gh, D(0.0), D(g.dt), D(0.0), D(g.grvp), b["u"], b["v"],
              b["w"], f["rho"], f["rhon"], b["adv_u"], b["adv_v"], b["adv_w"], b["adv_uo"],
              b["adv_vo"], b["adv_wo"], b["prs_u"], b["prs_v"], b["prs_w"], b["vis_u"],
              b["vis_v"], b["vis_w"], b["fst_u"], b["fst_v"], b["fst_w"], b["fst_un"],
              b["fst_vn"], b["fst_wn"], b["src_u"], b["src_v"], b["src_w"])
        c("cal_div_tensor_b", *gh, *dd, b["tau"], b["vis_u"], b["vis_v"], b["vis_w"])
        if self.params:
            c("cal_arith_tau_bp", *gh, a["s"], b["s"], f["rmu"], b["rmu"], b["tau"])
        else:
            c("cal_arith_tau_b", *gh, b["s"], f["rmu"], b["tau"])
        c("calsij_b", *gh, *dd, b["u"], b["v"], b["w"], b["s"])
        c("cal_grad_prs_b", *gh, *dd, b["phat"], b["prs_u"], b["prs_v"], b["prs_w"])
        c("calphat_b", *gh, b["p"], b["po"], b["phat"])
        c("cal_advu_weno5_b", *gh, *dd, a["uo"], b["uo"], a["vo"], b["vo"], a["wo"], b["wo"],
          b["adv_uo"], b["adv_vo"], b["adv_wo"])
        c("cal_advu_weno5_b", *gh, *dd, a["u"], b["u"], a["v"], b["v"], a["w"], b["w"],
          b["adv_u"], b["adv_v"], b["adv_w"])
        return uwallb.value


class MomentumWindow:
    """`steps` momentum updates with the state rotation between them, and
    their reverse: the input state of every step is kept (host copies) and
    each step is recomputed before it is reversed, so its primal WENO5
    inputs and Gauss-Seidel sweep count are those of the forward pass."""

    def __init__(self, step: MomentumStep, steps: int):
        self.step = step
        self.steps = steps
        self.states = []

    def forward(self, state: dict) -> dict:
        st = {k: np.array(state[k], dtype=np.float64) for k in STATE}
        self.states = []
        for _ in range(self.steps):
            self.states.append({k: v.copy() for k, v in st.items()})
            out = self.step.forward(st)
            st = {"uo": st["u"], "vo": st["v"], "wo": st["w"], "po": st["p"],
                  "u": out["un"], "v": out["vn"], "w": out["wn"], "p": out["pn"]}
        return st

    def reverse(self, state_b: dict) -> dict:
        """Adjoints of the window's input state (and uwall) from those of
        its output state (u, v, w, uo, vo, wo, p, po after the last step)."""
        sb = {k: np.array(state_b[k], dtype=np.float64) for k in STATE}
        uw = 0.0
        for n in reversed(range(self.steps)):
            # rotation uo' = u, u' = un, po' = p, p' = pn reversed
            out_b = {"un": sb["u"], "vn": sb["v"], "wn": sb["w"], "pn": sb["p"]}
            self.step.forward(self.states[n], keep=False)
            ib = self.step.reverse(out_b)
            uw += float(ib["uwall"][0])
            nb = {k: ib[k] for k in STATE}
            for s, o in (("u", "uo"), ("v", "vo"), ("w", "wo"), ("p", "po")):
                nb[s] = nb[s] + sb[o]
            sb = nb
        sb["uwall"] = np.array([uw])
        return sb


# ---------------------------------------------------------------------------
# The volume-fraction transport at a fixed velocity
# ---------------------------------------------------------------------------

PHASE_PARAMS = ("u", "v", "w", "theta0", "zeta")


class PhaseStep(MomentumStep):
    """The phase half of one step (srcs/main.f90:489-517, one phase slice):
    vertex gradients, MTHINC fluxes, update, truncation and mass
    correction, contact angle (gnbc) and its ghosts, periodic fills,
    summation into slice 0 and its fills -- forward and reversed, with the
    velocity held fixed.  The step maps phi(.,1) (the whole padded array;
    the forward also reads the ghost layer of phin(.,1), which holds the
    previous step's copy of phi(.,1)) to phin(.,0:1); its adjoint gives the
    adjoints of phi(.,1) and of the parameters u, v, w, theta_0_array,
    zeta_array and uwall.  The third call site's contact-angle fill of
    phi(.,0) (srcs/main.f90:509) feeds only the material properties, which
    are held fixed, and is not part of the map."""

    def __init__(self, cfg: RunConfig, params: dict, call=None, resident: bool = False):
        self.cfg = cfg
        self.call = call or gpu_call
        self.resident = resident
        self._registered = []
        g = cfg
        z = lambda: self.array(np.zeros(g.shape))  # noqa: E731
        self.par = {k: self.array(np.array(params[k], dtype=np.float64))  # own copies
                    for k in PHASE_PARAMS}
        self.dyinv = np.full(g.nj + 6, 1.0 / g.dy)
        self.a = {k: z() for k in ("phi", "gx", "gy", "gz", "hx", "hy", "hz", "fx", "fy", "fz",
                                   "theta", "pre3", "preca", "phin1_in")}
        self.a["phin"] = self.array(np.zeros((2,) + g.shape))
        self.b = {k: z() for k in ("phi", "gx", "gy", "gz", "fx", "fy", "fz", "theta",
                                   "phin1_in") + PHASE_PARAMS}
        self.b["phin"] = self.array(np.zeros((2,) + g.shape))
        self.own_phin = False

    def _slice(self, arr, l):
        return arr[l]

    def set_state(self, state: dict):
        """state["phi"]: phi(.,1); state["phin1"] (optional): the contents
        of phin(.,1) before the step, of which the update keeps the part it
        does not write -- after a step of the reference's loop it is the
        copy of phi(.,1) (the default), before the first step whatever the
        start-up left there."""
        self._set(self.a["phi"], state["phi"])
        self.own_phin = "phin1" in state
        if self.own_phin:
            self._set(self.a["phin1_in"], state["phin1"])

    def forward_set(self):
        c, a, p, g = self.call, self.a, self.par, self.cfg
        gh, dd = self._g(), self._d()
        grad_head = [I(0), I(g.nj), *gh, D(1 / g.dx), self.dyinv, D(1 / g.dz)]
        pn1 = a["phin"][1]
        # the previous step's copy phi = phin, or the given contents
        c("cpy", *gh, a["phin1_in"] if self.own_phin else a["phi"], pn1)
        c("cal_grad_p2a", *grad_head, a["phi"], a["gx"], a["gy"], a["gz"])
        c("solphi_mthinc1", *gh, *dd, D(g.dt), D(g.bet), a["phi"], a["gx"], a["gy"], a["gz"],
          p["u"], p["v"], p["w"], a["fx"], a["fy"], a["fz"])
        c("solphi_mthinc2", *gh, *dd, D(g.dt), p["u"], p["v"], p["w"], a["fx"], a["fy"], a["fz"],
          a["phi"], pn1)
        c("cal_grad_p2a", *grad_head, pn1, a["hx"], a["hy"], a["hz"])
        c("cpy", *gh, pn1, a["pre3"])
        c("solphi_mthinc3", I(0), *gh, *dd, D(g.bet), a["hx"], a["hy"], a["hz"], a["phi"], pn1)
        c("gnbc", WALLS, *gh, p["u"], p["w"], D(g.uwall), p["theta0"], D(g.sigma), p["zeta"],
          a["theta"])
        c("cpy", *gh, pn1, a["preca"])
        c("bnd_contact_angle", WALLS, *gh, pn1, a["theta"], D(g.dx), D(g.dy), D(g.dz))
        c("bnd_periodic", *gh, pn1)
        c("summation", *gh, a["phin"], I(1))
        c("bnd_neumann", WALLS, *gh, a["phin"][0])
        c("bnd_periodic", *gh, a["phin"][0])
        self.lp = 0

    def forward(self, state: dict, keep: bool = True) -> dict:
        self.set_state(state)
        self.forward_set()
        if not keep:
            return {}
        return {"phin": self._get(self.a["phin"])}

    def reverse(self, out_b: dict) -> dict:
        """Adjoints of phi(.,1) and of the parameters from those of
        phin(.,0:1), at the state of the last forward call."""
        self._set(self.b["phin"], out_b["phin"])
        uwallb = self.reverse_set()
        out = {"phi": self._get(self.b["phi"])}
        for k in PHASE_PARAMS:
            out[k] = self._get(self.b[k])
        out["uwall"] = np.array([uwallb])
        return out

    def reverse_set(self, theta_b_hook=None) -> float:
        """reverse() on the output adjoint already in place (b["phin"]);
        theta_b_hook(), if given, runs once the adjoints are cleared and
        before the contact angle is reversed: it may add to b["theta"] (the
        contact-angle fill of phi(.,0) in the full step)."""
        c, a, b, p, g = self.call, self.a, self.b, self.par, self.cfg
        gh, dd = self._g(), self._d()
        grad_head = [I(0), I(g.nj), *gh, D(1 / g.dx), self.dyinv, D(1 / g.dz)]
        for k in ("phi", "theta") + PHASE_PARAMS:
            c("init_q", *gh, b[k], b["gx"], b["gy"])
        c("init_q", *gh, b["gz"], b["fx"], b["fy"])
        c("init_q", *gh, b["fz"], b["gx"], b["gy"])
        if theta_b_hook is not None:
            theta_b_hook()
        uwallb = D(0.0)
        pn0b, pn1b = b["phin"][0], b["phin"][1]
        c("bnd_periodic_b", *gh, pn0b)
        c("bnd_neumann_b", WALLS, *gh, pn0b)
        c("summation_b", *gh, a["phin"], b["phin"], I(1))
        c("bnd_periodic_b", *gh, pn1b)
        c("bnd_contact_angle_b", WALLS, *gh, a["preca"], pn1b, a["theta"], b["theta"], D(g.dx),
          D(g.dy), D(g.dz))
        c("gnbc_b", WALLS, *gh, p["u"], b["u"], p["w"], b["w"], D(g.uwall), uwallb, p["theta0"],
          b["theta0"], D(g.sigma), p["zeta"], b["zeta"], b["theta"])
        c("solphi_mthinc3_b", I(0), *gh, *dd, D(g.bet), a["hx"], a["hy"], a["hz"], a["phi"],
          b["phi"], a["pre3"], pn1b)
        c("solphi_mthinc2_b", *gh, *dd, D(g.dt), p["u"], b["u"], p["v"], b["v"], p["w"], b["w"],
          a["fx"], b["fx"], a["fy"], b["fy"], a["fz"], b["fz"], a["phi"], b["phi"], pn1b)
        c("solphi_mthinc1_b", *gh, *dd, D(g.dt), D(g.bet), a["phi"], b["phi"], a["gx"], b["gx"],
          a["gy"], b["gy"], a["gz"], b["gz"], p["u"], b["u"], p["v"], b["v"], p["w"], b["w"],
          b["fx"], b["fy"], b["fz"])
        c("cal_grad_p2a_b", *grad_head, b["phi"], b["gx"], b["gy"], b["gz"])
        # phin(.,1) started as the copy of phi (or the given contents): its
        # unwritten part
        if self.own_phin:
            c("init_q", *gh, b["phin1_in"], b["gx"], b["gy"])
            c("cpy_b", *gh, b["phin1_in"], pn1b)
        else:
            c("cpy_b", *gh, b["phi"], pn1b)
        return uwallb.value


class PhaseWindow:
    """`steps` phase updates at a fixed velocity (phi(.,1) <- phin(.,1)
    between steps) and their reverse, recomputing each step from its stored
    input.  reverse() takes the adjoint of the last step's phin(.,0:1) and
    returns those of the first step's phi(.,1) and of the parameters,
    accumulated over the window."""

    def __init__(self, step: PhaseStep, steps: int):
        self.step = step
        self.steps = steps
        self.states = []

    def forward(self, phi1: np.ndarray) -> np.ndarray:
        st = np.array(phi1, dtype=np.float64)
        self.states = []
        out = None
        for _ in range(self.steps):
            self.states.append(st.copy())
            out = self.step.forward({"phi": st})["phin"]
            st = out[1].copy()
        return out

    def reverse(self, phin_b: np.ndarray) -> dict:
        ob = np.array(phin_b, dtype=np.float64)
        acc = {k: 0.0 for k in PHASE_PARAMS + ("uwall",)}
        res = None
        for n in reversed(range(self.steps)):
            self.step.forward({"phi": self.states[n]}, keep=False)
            res = self.step.reverse({"phin": ob})
            for k in acc:
                acc[k] = acc[k] + res[k]
            ob = np.zeros_like(ob)
            ob[1] = res["phi"]
        out = {"phi": res["phi"]}
        out.update(acc)
        return out


# ---------------------------------------------------------------------------
# The whole step: phase transport, material properties, surface tension and
# the velocity-pressure update
# ---------------------------------------------------------------------------

FULL_STATE = ("phi0", "phi1") + STATE
FULL_PARAMS = ("theta0", "zeta", "l1", "l2")


class FullStep:
    """One whole timestep of the reference (srcs/main.f90:489-694 for one
    phase slice, the order of host/fel_driver.f90:step1) through the
    drop-in routines, and its reverse.  The step maps the state (phi(.,0),
    phi(.,1), u, v, w, uo, vo, wo, p, po) to (phin(.,0:1), un, vn, wn, pn)
    for the wall maps theta_0_array, zeta_array, l1_array, l2_array and
    uwall; reverse() gives the adjoints of the state and of all five.

    Forward: the phase update (PhaseStep, at the state's u, v, w); the
    contact-angle fill of phi(.,0) with the same contact angle
    (srcs/main.f90:509, gnbc computed once: both of its calls read the
    same u, w); densities and viscosities of phi(.,0) and phin(.,0)
    (cal_arith_mean x4); the surface-tension forces of phi(.,1) and
    phin(.,1) (vertex gradients, normals on four staggerings, curvature,
    face gradients, calfst, sum_fst); then the velocity-pressure update
    (MomentumStep with parameter adjoints) on those fields, with bndu's
    phi = phi(.,1).  Reverse: the momentum update with the adjoints of the
    densities, viscosities, forces, phi, l1 and l2; the surface-tension
    chain; the material properties; the contact-angle fill of phi(.,0)
    (its theta adjoint joining the phase reverse's before gnbc); the
    phase update."""

    def __init__(self, cfg: RunConfig, params: dict, call=None, lp_reader=None,
                 resident: bool = False):
        self.cfg = cfg
        self.call = call or gpu_call
        g = cfg
        zero = np.zeros(g.shape)
        self.ph = PhaseStep(cfg, {"u": zero, "v": zero, "w": zero, "theta0": params["theta0"],
                                  "zeta": params["zeta"]}, call=call, resident=resident)
        fixed = {k: zero for k in FIXED}
        fixed["l1"], fixed["l2"] = params["l1"], params["l2"]
        self.mo = MomentumStep(cfg, fixed, call=call, lp_reader=lp_reader, resident=resident,
                               params=True)
        z = lambda n=None: self.mo.array(np.zeros(g.shape if n is None else (n,) + g.shape))  # noqa: E731
        # primal arrays of the full step beyond the two parts'
        self.a = {k: z() for k in ("phi0", "phi0c", "gx", "gy", "gz", "hx", "hy", "hz", "fgx",
                                   "fgy", "fgz", "fhx", "fhy", "fhz", "fst_u", "fst_v", "fst_w",
                                   "fst_un", "fst_vn", "fst_wn")}
        for k in ("rn_pp", "rn_12", "rn_13", "rn_23", "rnn_pp", "rnn_12", "rnn_13", "rnn_23",
                  "rkap", "rkapn"):
            self.a[k] = z(3)
        self.b = {k: z() for k in ("phi0c", "phi1", "gx", "gy", "gz", "fgx", "fgy", "fgz")}
        for k in ("rn_pp", "rn_12", "rn_13", "rn_23", "rkap"):
            self.b[k] = z(3)
        self.dyinv = np.full(g.nj + 6, 1.0 / g.dy)
        self.lp = None

    def close(self):
        self.ph.close()
        self.mo.close()

    # -- pieces -----------------------------------------------------------
    def _grad_head(self):
        g = self.cfg
        return [I(0), I(g.nj), *self.mo._g(), D(1 / g.dx), self.dyinv, D(1 / g.dz)]

    _STAG = (("pp", (1, 1, 1)), ("12", (0, 0, 1)), ("13", (0, 1, 0)), ("23", (1, 0, 0)))

    def _csf(self, q, gv, rn, rkap, fg, fst):
        """Surface-tension force of the volume fraction q (srcs/main.f90:
        547-568): vertex gradients gv, normals rn[...], curvature rkap,
        face gradients fg, force fst."""
        c, a, g = self.call, self.a, self.cfg
        gh, dd = self.mo._g(), self.mo._d()
        c("cal_grad_p2a", *self._grad_head(), q, *[a[k] for k in gv])
        for tag, st in self._STAG:
            c("calrn", *gh, *dd, *[I(v) for v in st], *[a[k] for k in gv], a[rn + tag])
        c("calrkap", *gh, *dd, *[a[rn + tag] for tag, _ in self._STAG], a[rkap])
        c("cal_grad_p2uvw", *self._grad_head(), q, *[a[k] for k in fg])
        c("calfst", *gh, D(g.rhol), D(g.rhog), D(g.sigma), *[a[k] for k in fg], a[rkap],
          *[a[k] for k in fst])

    def _csf_b(self, q, qb, gv, rn, rkap, fg, fst_b):
        """Reverse of _csf for q (primal) into qb, from the force adjoints
        fst_b (consumed)."""
        c, a, b, g = self.call, self.a, self.b, self.cfg
        gh, dd = self.mo._g(), self.mo._d()
        # the gradient adjoints are shared by both levels: cleared first (the
        # parts the forward leaves unwritten carry no derivative)
        c("init_q", *gh, b["gx"], b["gy"], b["gz"])
        c("init_q", *gh, b["fgx"], b["fgy"], b["fgz"])
        c("calfst_b", *gh, D(g.rhol), D(g.rhog), D(g.sigma), a[fg[0]], b["fgx"], a[fg[1]],
          b["fgy"], a[fg[2]], b["fgz"], a[rkap], b["rkap"], *fst_b)
        c("cal_grad_p2uvw_b", *self._grad_head(), qb, b["fgx"], b["fgy"], b["fgz"])
        c("calrkap_b", *gh, *dd, *[b["rn_" + tag] for tag, _ in self._STAG], b["rkap"])
        for tag, st in reversed(self._STAG):
            c("calrn_b", *gh, *dd, *[I(v) for v in st], a[gv[0]], b["gx"], a[gv[1]], b["gy"],
              a[gv[2]], b["gz"], b["rn_" + tag])
        c("cal_grad_p2a_b", *self._grad_head(), qb, b["gx"], b["gy"], b["gz"])

    # -- forward ----------------------------------------------------------
    def set_state(self, state: dict):
        ph, mo = self.ph, self.mo
        for k in ("u", "v", "w"):
            ph._set(ph.par[k], state[k])
        ps = {"phi": state["phi1"]}
        if "phin1" in state:
            ps["phin1"] = state["phin1"]
        ph.set_state(ps)
        mo._set(self.a["phi0"], state["phi0"])
        mo.set_state(state)

    def forward_set(self):
        c, a, g = self.call, self.a, self.cfg
        ph, mo = self.ph, self.mo
        gh = mo._g()
        f = mo.fx
        ph.forward_set()
        pn0, pn1 = ph.a["phin"][0], ph.a["phin"][1]
        # contact-angle fill of phi(.,0) (srcs/main.f90:509)
        c("cpy", *gh, a["phi0"], a["phi0c"])
        c("bnd_contact_angle", WALLS, *gh, a["phi0c"], ph.a["theta"], D(g.dx), D(g.dy), D(g.dz))
        # material properties (srcs/main.f90:526-529)
        c("cal_arith_mean", *gh, a["phi0c"], D(g.rhol), D(g.rhog), f["rho"])
        c("cal_arith_mean", *gh, pn0, D(g.rhol), D(g.rhog), f["rhon"])
        c("cal_arith_mean", *gh, a["phi0c"], D(g.rmul), D(g.rmug), f["rmu"])
        c("cal_arith_mean", *gh, pn0, D(g.rmul), D(g.rmug), f["rmun"])
        # surface tension (srcs/main.f90:545-573)
        c("init_q", *gh, f["fst_u"], f["fst_v"], f["fst_w"])
        c("init_q", *gh, f["fst_un"], f["fst_vn"], f["fst_wn"])
        self._csf(ph.a["phi"], ("gx", "gy", "gz"), "rn_", "rkap", ("fgx", "fgy", "fgz"),
                  ("fst_u", "fst_v", "fst_w"))
        self._csf(pn1, ("hx", "hy", "hz"), "rnn_", "rkapn", ("fhx", "fhy", "fhz"),
                  ("fst_un", "fst_vn", "fst_wn"))
        c("sum_fst", *gh, a["fst_u"], a["fst_v"], a["fst_w"], f["fst_u"], f["fst_v"], f["fst_w"])
        c("sum_fst", *gh, a["fst_un"], a["fst_vn"], a["fst_wn"], f["fst_un"], f["fst_vn"],
          f["fst_wn"])
        # velocity-pressure update, bndu's phi = phi(.,1)
        c("cpy", *gh, ph.a["phi"], f["phi"])
        mo.forward_set()
        self.lp = mo.lp

    def forward(self, state: dict, keep: bool = True) -> dict:
        self.set_state(state)
        self.forward_set()
        if not keep:
            return {}
        out = {k: self.mo._get(self.mo.a[k]) for k in OUT}
        out["phin"] = self.ph._get(self.ph.a["phin"])
        return out

    # -- reverse ----------------------------------------------------------
    def reverse(self, out_b: dict) -> dict:
        """Adjoints of the state and of the parameters from those of
        (phin(.,0:1), un, vn, wn, pn), at the state of the last forward."""
        ph, mo, b = self.ph, self.mo, self.b
        for k in OUT:
            mo._set(mo.b[k], out_b[k])
        ph._set(ph.b["phin"], out_b["phin"])
        uw = self.reverse_set()
        mb = mo.b
        out = {"phi0": mo._get(b["phi0c"])}
        out["phi1"] = ph._get(ph.b["phi"]) + mo._get(b["phi1"])
        for k in ("u", "v", "w"):
            out[k] = mo._get(mb[k]) + ph._get(ph.b[k])
        for k in ("uo", "vo", "wo", "p", "po"):
            out[k] = mo._get(mb[k])
        out["theta0"] = ph._get(ph.b["theta0"])
        out["zeta"] = ph._get(ph.b["zeta"])
        out["l1"] = mo._get(mb["l1"])
        out["l2"] = mo._get(mb["l2"])
        out["uwall"] = np.array([uw])
        return out

    def reverse_set(self) -> float:
        """reverse() on the output adjoints already in place (mo.b[un, vn,
        wn, pn], ph.b["phin"]), without transfers: the adjoints are left in
        the arrays reverse() reads; returns uwall's."""
        c, a, b, g = self.call, self.a, self.b, self.cfg
        ph, mo = self.ph, self.mo
        gh = mo._g()
        if self.lp is None:
            raise RuntimeError("reverse() needs a forward() of the same step first")
        pn0b, pn1b = ph.b["phin"][0], ph.b["phin"][1]
        uw = mo.reverse_set()
        mb = mo.b
        c("init_q", *gh, b["phi0c"], b["phi1"], b["gx"])
        # bndu's phi is phi(.,1); its copy into the fixed field reversed
        c("cpy_b", *gh, b["phi1"], mb["phi"])
        # surface tension: the sums (zeroed, then one force added: sum_fst
        # reversed gives the force the sum's adjoint, init_q ends it) hand
        # their adjoints to the forces, consumed by each level's chain
        fb_n = [mb["fst_un"], mb["fst_vn"], mb["fst_wn"]]
        fb = [mb["fst_u"], mb["fst_v"], mb["fst_w"]]
        self._csf_b(ph.a["phin"][1], pn1b, ("hx", "hy", "hz"), "rnn_", "rkapn",
                    ("fhx", "fhy", "fhz"), fb_n)
        self._csf_b(ph.a["phi"], b["phi1"], ("gx", "gy", "gz"), "rn_", "rkap",
                    ("fgx", "fgy", "fgz"), fb)
        # material properties
        c("cal_arith_mean_b", *gh, pn0b, D(g.rmul), D(g.rmug), mb["rmun"])
        c("cal_arith_mean_b", *gh, b["phi0c"], D(g.rmul), D(g.rmug), mb["rmu"])
        c("cal_arith_mean_b", *gh, pn0b, D(g.rhol), D(g.rhog), mb["rhon"])
        c("cal_arith_mean_b", *gh, b["phi0c"], D(g.rhol), D(g.rhog), mb["rho"])

        # the contact-angle fill of phi(.,0): its theta adjoint joins the
        # phase reverse's before gnbc is reversed
        def theta_hook():
            c("bnd_contact_angle_b", WALLS, *gh, a["phi0"], b["phi0c"], ph.a["theta"],
              ph.b["theta"], D(g.dx), D(g.dy), D(g.dz))
        uw += ph.reverse_set(theta_hook)
        return uw


class FullWindow:
    """`steps` whole timesteps with the reference's state rotation
    (srcs/main.f90:687-698: uo = u, u = un, po = p, p = pn, phi = phin) and
    their reverse, each step recomputed from its input state before it is
    reversed.  checkpoint_every = K keeps the input state of every K-th step
    only (host memory for ceil(steps/K) + K states instead of `steps`): the
    reverse recomputes each segment of K steps from its checkpoint, keeping
    its K input states, and reverses them, so the forward runs about twice
    in all.  K = None (or 1) keeps every input state."""

    def __init__(self, step: FullStep, steps: int, checkpoint_every: int | None = None,
                 on_step=None):
        self.step = step
        self.steps = steps
        self.every = checkpoint_every if checkpoint_every else 1
        if self.every < 1:
            raise ValueError("checkpoint_every must be >= 1")
        self.checkpoints = {}
        self.forward_steps = 0  # steps run forward, recomputation included
        self.on_step = on_step  # on_step("forward" | "reverse", n) after each step

    @staticmethod
    def _rotate(st: dict, out: dict) -> dict:
        return {"uo": st["u"], "vo": st["v"], "wo": st["w"], "po": st["p"],
                "u": out["un"], "v": out["vn"], "w": out["wn"], "p": out["pn"],
                "phi0": out["phin"][0].copy(), "phi1": out["phin"][1].copy()}

    def _advance(self, st: dict) -> dict:
        self.forward_steps += 1
        return self._rotate(st, self.step.forward(st))

    def forward(self, state: dict) -> dict:
        """state: FULL_STATE, and optionally "phin1", the contents of
        phin(.,1) before the first step (PhaseStep.set_state)."""
        st = {k: np.array(state[k], dtype=np.float64) for k in FULL_STATE}
        if "phin1" in state:
            st["phin1"] = np.array(state["phin1"], dtype=np.float64)
        self.checkpoints = {}
        self.forward_steps = 0
        for n in range(self.steps):
            if n % self.every == 0:
                self.checkpoints[n] = {k: v.copy() for k, v in st.items()}
            st = self._advance(st)
            if self.on_step:
                self.on_step("forward", n)
        return st

    def reverse(self, state_b: dict) -> dict:
        """Adjoints of the window's input state and of the parameters
        (accumulated) from those of its output state."""
        sb = {k: np.array(state_b[k], dtype=np.float64) for k in FULL_STATE}
        acc = {k: 0.0 for k in FULL_PARAMS + ("uwall",)}
        for c0 in sorted(self.checkpoints, reverse=True):
            # the segment's input states, recomputed from its checkpoint
            seg = [self.checkpoints[c0]]
            for _ in range(c0 + 1, min(c0 + self.every, self.steps)):
                seg.append(self._advance(seg[-1]))
            for m, st in reversed(list(enumerate(seg))):
                out_b = {"un": sb["u"], "vn": sb["v"], "wn": sb["w"], "pn": sb["p"],
                         "phin": np.stack([sb["phi0"], sb["phi1"]])}
                self.step.forward(st, keep=False)
                self.forward_steps += 1
                ib = self.step.reverse(out_b)
                for k in acc:
                    acc[k] = acc[k] + ib[k]
                nb = {k: ib[k] for k in FULL_STATE}
                for s_, o in (("u", "uo"), ("v", "vo"), ("w", "wo"), ("p", "po")):
                    nb[s_] = nb[s_] + sb[o]
                sb = nb
                if self.on_step:
                    self.on_step("reverse", c0 + m)
        sb.update(acc)
        return sb


class DeviceFullWindow:
    """FullWindow with every state, checkpoint and adjoint in device memory
    (with FullStep(resident=True); with host arrays it runs the same
    sequence on them): states move between steps by the library's cpy_
    (device copies), adjoints are summed by cpy_b_ (q1b += q2b), and only
    the window's input state, the output adjoint and the results cross to
    the host.  checkpoint_every = K keeps every K-th input state and
    recomputes the segments in the reverse, as FullWindow does; the slots
    (one state each: ten padded arrays) are reused, ceil(steps/K) + K + 3
    at most.  Results are those of FullWindow, bit for bit."""

    def __init__(self, step: FullStep, steps: int, checkpoint_every: int | None = None,
                 on_step=None):
        self.step = step
        self.steps = steps
        self.every = checkpoint_every if checkpoint_every else 1
        if self.every < 1:
            raise ValueError("checkpoint_every must be >= 1")
        self.on_step = on_step
        self.pool = []
        self.slots = 0
        self.checkpoints = {}
        self.forward_steps = 0
        g = step.cfg
        self.g3 = [I(g.ni), I(g.nj), I(g.nk)]
        self.phin1 = None

    # -- slots ----------------------------------------------------------------
    def _take(self) -> dict:
        if self.pool:
            return self.pool.pop()
        self.slots += 1
        return {k: self.step.mo.array(np.zeros(self.step.cfg.shape), upload=False)
                for k in FULL_STATE}

    def _give(self, slot: dict):
        self.pool.append(slot)

    def _cp(self, src, dst):
        self.step.call("cpy", *self.g3, src, dst)

    def _add(self, dst, src):
        """dst += src (src consumed)."""
        self.step.call("cpy_b", *self.g3, dst, src)

    # -- forward ----------------------------------------------------------------
    def _load(self, slot: dict, first: bool):
        st, ph, mo = self.step, self.step.ph, self.step.mo
        for k in ("u", "v", "w"):
            self._cp(slot[k], ph.par[k])
        self._cp(slot["phi1"], ph.a["phi"])
        ph.own_phin = first and self.phin1 is not None
        if ph.own_phin:
            self._cp(self.phin1, ph.a["phin1_in"])
        self._cp(slot["phi0"], st.a["phi0"])
        for k in STATE:
            self._cp(slot[k], mo.a[k])

    def _advance(self, slot: dict, first: bool) -> dict:
        """The step from the state in `slot`; returns a new slot with the
        next state (the reference's rotation)."""
        st, ph, mo = self.step, self.step.ph, self.step.mo
        self._load(slot, first)
        st.forward_set()
        self.forward_steps += 1
        nxt = self._take()
        for s_, o in (("u", "uo"), ("v", "vo"), ("w", "wo"), ("p", "po")):
            self._cp(slot[s_], nxt[o])
        for s_, o in (("un", "u"), ("vn", "v"), ("wn", "w"), ("pn", "p")):
            self._cp(mo.a[s_], nxt[o])
        self._cp(ph.a["phin"][0], nxt["phi0"])
        self._cp(ph.a["phin"][1], nxt["phi1"])
        return nxt

    def forward(self, state: dict) -> dict:
        """state: FULL_STATE, and optionally "phin1" (PhaseStep.set_state);
        returns the window's output state (host copies)."""
        mo = self.step.mo
        for c in self.checkpoints.values():
            self._give(c)
        self.checkpoints = {}
        self.forward_steps = 0
        cur = self._take()
        for k in FULL_STATE:
            mo._set(cur[k], np.asarray(state[k], dtype=np.float64))
        self.phin1 = None
        if "phin1" in state:
            self.phin1 = mo.array(np.array(state["phin1"], dtype=np.float64))
        for n in range(self.steps):
            if n % self.every == 0:
                self.checkpoints[n] = cur
            nxt = self._advance(cur, n == 0)
            if n % self.every != 0:
                self._give(cur)
            cur = nxt
            if self.on_step:
                self.on_step("forward", n)
        out = {k: mo._get(cur[k]) for k in FULL_STATE}
        self._give(cur)
        return out

    # -- reverse ----------------------------------------------------------------
    def reverse(self, state_b: dict) -> dict:
        """Adjoints of the window's input state and of the wall maps
        (accumulated) from those of its output state (host arrays in, host
        arrays out)."""
        st, ph, mo = self.step, self.step.ph, self.step.mo
        sb = self._take()
        for k in FULL_STATE:
            mo._set(sb[k], np.asarray(state_b[k], dtype=np.float64))
        acc = self._take()  # theta0, zeta, l1, l2 accumulate in four of its arrays
        for k in FULL_STATE:
            mo._set(acc[k], np.zeros(st.cfg.shape))
        slots = {"theta0": "phi0", "zeta": "phi1", "l1": "u", "l2": "v"}
        uw = 0.0
        for c0 in sorted(self.checkpoints, reverse=True):
            seg = [self.checkpoints[c0]]
            for m in range(c0 + 1, min(c0 + self.every, self.steps)):
                seg.append(self._advance(seg[-1], m - 1 == 0))
            for m in reversed(range(len(seg))):
                n = c0 + m
                self._load(seg[m], n == 0)
                st.forward_set()
                self.forward_steps += 1
                for s_, o in (("u", "un"), ("v", "vn"), ("w", "wn"), ("p", "pn")):
                    self._cp(sb[s_], mo.b[o])
                self._cp(sb["phi0"], ph.b["phin"][0])
                self._cp(sb["phi1"], ph.b["phin"][1])
                uw += st.reverse_set()
                nb = self._take()
                self._cp(st.b["phi0c"], nb["phi0"])
                self._cp(ph.b["phi"], nb["phi1"])
                self._add(nb["phi1"], st.b["phi1"])
                for k in ("u", "v", "w"):
                    self._cp(mo.b[k], nb[k])
                    self._add(nb[k], ph.b[k])
                for k in ("uo", "vo", "wo", "p", "po"):
                    self._cp(mo.b[k], nb[k])
                for s_, o in (("u", "uo"), ("v", "vo"), ("w", "wo"), ("p", "po")):
                    self._add(nb[s_], sb[o])
                self._add(acc[slots["theta0"]], ph.b["theta0"])
                self._add(acc[slots["zeta"]], ph.b["zeta"])
                self._add(acc[slots["l1"]], mo.b["l1"])
                self._add(acc[slots["l2"]], mo.b["l2"])
                self._give(sb)
                sb = nb
                if m > 0:
                    self._give(seg[m])
                if self.on_step:
                    self.on_step("reverse", n)
        out = {k: mo._get(sb[k]) for k in FULL_STATE}
        for k, a in slots.items():
            out[k] = mo._get(acc[a])
        out["uwall"] = np.array([uw])
        self._give(sb)
        self._give(acc)
        return out
