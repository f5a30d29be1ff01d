"""Reverse mode (adjoint) of the velocity-pressure update at a fixed phase field.

One timestep of the reference (srcs/main.f90:526-667) maps the flow state
(u, v, w, uo, vo, wo, p, po) to (un, vn, wn, pn) through WENO5 advection of
u and uo, the pressure predictor, the explicit viscous stress, the momentum
source, the Crank-Nicolson Gauss-Seidel solve, the projection and the ghost
refreshes; the interface enters through the densities, viscosities, the
surface-tension forces, phi and the slip lengths.  With those held fixed
(the phase field frozen), MomentumStep runs that sequence through the
library's drop-in routines and reverses it through their adjoint
entry points (<name>_b_), in the reverse order:

    forward(state)   -> outputs, recording the primal velocities the WENO5
                        adjoint needs and the Gauss-Seidel sweep count
    reverse(out_b)   -> adjoints of the state and of uwall

MomentumWindow chains steps with the reference's state rotation
(srcs/main.f90:687-694: uo = u, u = un, po = p, p = pn) and reverses a
window of them, storing each step's input state (the checkpoints).

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
                 resident: bool = False):
        self.cfg = cfg
        self.call = call or gpu_call
        self.lp_reader = lp_reader or sor_last_lp
        self.resident = resident
        self._registered = []
        g = cfg
        self.nm = g.ni * g.nj * g.nk
        self.n1 = max(g.nj, g.nk) + 6
        z = lambda: self.array(np.zeros(g.shape))  # noqa: E731
        self.fx = {k: self.array(np.ascontiguousarray(fixed[k], dtype=np.float64)) for k in FIXED}
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
        self.lp = None

    # -- arrays ---------------------------------------------------------
    def array(self, a: np.ndarray) -> np.ndarray:
        """Returns a (registered with the library when resident)."""
        if self.resident:
            lib = _lib.gpu()
            n, up = ctypes.c_int64(a.nbytes), ctypes.c_int(1)
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

    def _refresh_b(self, b, uwallb):
        c, gh = self.call, self._g()
        for k in ("wn", "vn", "un"):
            c("bnd_periodic_b", *gh, b[k])
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
        uwallb = D(0.0)
        c("bnd_periodic_b", *gh, b["pn"])
        c("bnd_neumann_b", WALLS, *gh, b["pn"])
        c("bnd_avzero_b", I(0), I(1), *gh, b["pn"])
        self._refresh_b(b, uwallb)
        c("corunp_explicit_b", WALLS, *gh, D(g.rhog), *dd, D(g.dt), b["dp"], b["phat"], b["un"],
          b["vn"], b["wn"], b["pn"])
        c("bnd_periodic_b", *gh, b["dp"])
        c("bnd_neumann_b", WALLS, *gh, b["dp"])
        c("bnd_avzero_b", I(0), I(1), *gh, b["dp"])
        c("solp_fft_tdma3_b", *gh, np.zeros(2 * nm), b["pr"], b["pi"], b["dp"])
        c("solp_fft_tdma2_b", *self._tdma2_head(), b["pr"], b["pi"])
        c("solp_fft_tdma1_b", *gh, np.zeros(2 * nm), b["div"], b["pr"], b["pi"])
        c("caldiv_b", I(0), I(1), *gh, *dd, D(g.dt), b["un"], b["vn"], b["wn"], b["div"])
        self._refresh_b(b, uwallb)
        c("solu_sor4_b", *self._sor_head(), b["src_u"], b["src_v"], b["src_w"], b["un"], b["vn"],
          b["wn"], uwallb, *self._bc_tail(), I(self.lp))
        for s, t in (("u", "un"), ("v", "vn"), ("w", "wn")):
            c("cpy_b", *gh, b[s], b[t])
        c("cal_srcu_b", WALLS, *gh, D(0.0), D(g.dt), D(0.0), D(g.grvp), b["u"], b["v"], b["w"],
          f["rho"], f["rhon"], b["adv_u"], b["adv_v"], b["adv_w"], b["adv_uo"], b["adv_vo"],
          b["adv_wo"], b["prs_u"], b["prs_v"], b["prs_w"], b["vis_u"], b["vis_v"], b["vis_w"],
          b["fst_u"], b["fst_v"], b["fst_w"], b["fst_un"], b["fst_vn"], b["fst_wn"], b["src_u"],
          b["src_v"], b["src_w"])
        c("cal_div_tensor_b", *gh, *dd, b["tau"], b["vis_u"], b["vis_v"], b["vis_w"])
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
        self.par = {k: self.array(np.ascontiguousarray(params[k], dtype=np.float64))
                    for k in PHASE_PARAMS}
        self.dyinv = np.full(g.nj + 6, 1.0 / g.dy)
        self.a = {k: z() for k in ("phi", "gx", "gy", "gz", "hx", "hy", "hz", "fx", "fy", "fz",
                                   "theta", "pre3", "preca")}
        self.a["phin"] = self.array(np.zeros((2,) + g.shape))
        self.b = {k: z() for k in ("phi", "gx", "gy", "gz", "fx", "fy", "fz", "theta") + PHASE_PARAMS}
        self.b["phin"] = self.array(np.zeros((2,) + g.shape))

    def _slice(self, arr, l):
        return arr[l]

    def set_state(self, state: dict):
        self._set(self.a["phi"], state["phi"])

    def forward_set(self):
        c, a, p, g = self.call, self.a, self.par, self.cfg
        gh, dd = self._g(), self._d()
        grad_head = [I(0), I(g.nj), *gh, D(1 / g.dx), self.dyinv, D(1 / g.dz)]
        pn1 = a["phin"][1]
        c("cpy", *gh, a["phi"], pn1)  # the previous step's copy phi = phin
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

    def reverse_set(self) -> float:
        c, a, b, p, g = self.call, self.a, self.b, self.par, self.cfg
        gh, dd = self._g(), self._d()
        grad_head = [I(0), I(g.nj), *gh, D(1 / g.dx), self.dyinv, D(1 / g.dz)]
        for k in ("phi", "theta") + PHASE_PARAMS:
            c("init_q", *gh, b[k], b["gx"], b["gy"])
        c("init_q", *gh, b["gz"], b["fx"], b["fy"])
        c("init_q", *gh, b["fz"], b["gx"], b["gy"])
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
        # phin(.,1) started as the copy of phi: its unwritten (ghost) part
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
