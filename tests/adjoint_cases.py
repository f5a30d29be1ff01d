"""Seeded cases for the adjoint (reverse-mode) routines <name>_b.

Each case is a linear map F of the forward routine <name> (its active
arrays in, its active arrays out, passive data fixed) and the matching
adjoint routine B.  ``forward(call, x)`` and ``adjoint(call, y)`` run them
through a caller -- the oracle (``H.call_oracle``) or the GPU library
(``H.call_gpu``) -- on dicts of arrays, so the same case serves the
dot-product identity <F x, y> = <x, B y> and the GPU/oracle comparison.
"""
from __future__ import annotations

import ctypes

import numpy as np

import felharness as H
from felharness import D, I, Grid

WALLS = np.array([-1, -1, -1, -1, -1, -1], dtype=np.int32)


def dot(a: dict, b: dict) -> float:
    return float(sum(np.dot(np.ravel(a[k]), np.ravel(b[k])) for k in a))


def scale(a: dict, b: dict) -> float:
    """|a| |b| over the entries that take part in the dot product (keys
    starting with "_" hold adjoints passed through unchanged)."""
    na = np.sqrt(sum(float(np.sum(np.square(a[k]))) for k in a if not k.startswith("_")))
    nb = np.sqrt(sum(float(np.sum(np.square(b[k]))) for k in b if not k.startswith("_")))
    return na * nb


def box_mask(g: Grid, lo: int, hi_off: int):
    """Mask of the padded cells in [lo, n+hi_off]^3."""
    m = np.zeros(g.shape, bool)
    m[lo + 2:g.nk + hi_off + 3, lo + 2:g.nj + hi_off + 3, lo + 2:g.ni + hi_off + 3] = True
    return m


def interior_mask(g: Grid):
    return box_mask(g, 1, 0)


class Case:
    name = ""

    def __init__(self, g: Grid, seed: int):
        self.g = g
        self.rng = np.random.default_rng(seed)

    def inputs(self) -> dict:
        raise NotImplementedError

    def out_adjoints(self) -> dict:
        raise NotImplementedError

    def forward(self, call, x: dict) -> dict:
        raise NotImplementedError

    def adjoint(self, call, y: dict) -> dict:
        raise NotImplementedError


class Caldiv(Case):
    name = "caldiv"
    dt = 0.003

    def inputs(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("u", "v", "w")}

    def out_adjoints(self):
        return {"div": np.where(box_mask(self.g, -1, 3), self.rng.standard_normal(self.g.shape), 0.0)}

    def head(self):
        g = self.g
        return [I(0), I(1), I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), D(self.dt)]

    def forward(self, call, x):
        div = self.g.zeros()
        call("caldiv", *self.head(), x["u"].copy(), x["v"].copy(), x["w"].copy(), div, D(0.0))
        return {"div": np.where(box_mask(self.g, -1, 3), div, 0.0)}

    def adjoint(self, call, y):
        ub, vb, wb = self.g.zeros(), self.g.zeros(), self.g.zeros()
        divb = y["div"].copy()
        call("caldiv_b", *self.head(), ub, vb, wb, divb)
        assert not np.any(divb[box_mask(self.g, -1, 3)]), "div_b must be consumed"
        return {"u": ub, "v": vb, "w": wb}


class Tdma1(Case):
    name = "solp_fft_tdma1"

    def nm(self):
        g = self.g
        return g.ni * g.nj * g.nk

    def inputs(self):
        return {"div": np.where(interior_mask(self.g), self.rng.standard_normal(self.g.shape), 0.0)}

    def out_adjoints(self):
        return {"pr": self.rng.standard_normal(self.nm()), "pi": self.rng.standard_normal(self.nm())}

    def forward(self, call, x):
        g = self.g
        pr, pi = np.zeros(self.nm()), np.zeros(self.nm())
        call("solp_fft_tdma1", I(g.ni), I(g.nj), I(g.nk), np.zeros(2 * self.nm()), x["div"].copy(), pr, pi)
        return {"pr": pr, "pi": pi}

    def adjoint(self, call, y):
        g = self.g
        divb = g.zeros()
        pr, pi = y["pr"].copy(), y["pi"].copy()
        call("solp_fft_tdma1_b", I(g.ni), I(g.nj), I(g.nk), np.zeros(2 * self.nm()), divb, pr, pi)
        assert not np.any(pr) and not np.any(pi)
        return {"div": np.where(interior_mask(g), divb, 0.0)}


class Tdma3(Tdma1):
    name = "solp_fft_tdma3"

    def inputs(self):
        return {"pr": self.rng.standard_normal(self.nm()), "pi": self.rng.standard_normal(self.nm())}

    def out_adjoints(self):
        return {"dp": np.where(interior_mask(self.g), self.rng.standard_normal(self.g.shape), 0.0)}

    def forward(self, call, x):
        g = self.g
        dp = g.zeros()
        call("solp_fft_tdma3", I(g.ni), I(g.nj), I(g.nk), np.zeros(2 * self.nm()), x["pr"].copy(),
             x["pi"].copy(), dp)
        return {"dp": np.where(interior_mask(g), dp, 0.0)}

    def adjoint(self, call, y):
        g = self.g
        pr, pi = np.zeros(self.nm()), np.zeros(self.nm())
        dpb = y["dp"].copy()
        call("solp_fft_tdma3_b", I(g.ni), I(g.nj), I(g.nk), np.zeros(2 * self.nm()), pr, pi, dpb)
        assert not np.any(dpb[interior_mask(g)])
        return {"pr": pr, "pi": pi}


class Tdma2(Tdma1):
    """Thomas solves per mode.  `skew` makes the j coefficients
    non-symmetric (as(j+1) != an(j)), so that the transpose differs from the
    forward solve; the singular (0,0) mode's entries are zero in both the
    inputs and the output adjoints (its solve is defined up to a constant)."""
    name = "solp_fft_tdma2"
    rho = 0.81

    def __init__(self, g, seed, skew=False):
        super().__init__(g, seed)
        pc = [np.zeros(g.nj + 6) for _ in range(7)]
        H.call_oracle("cal_coef_prs", I(g.nj), D(self.rho), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), *pc)
        fft = [np.zeros(g.nj + 6) for _ in range(3)]
        H.call_oracle("cal_coef_solp_fft", I(g.nj), pc[2], pc[3], *fft)
        if skew:
            r = np.random.default_rng(seed + 1000)
            fft[0][3:] *= r.uniform(0.5, 1.5, g.nj + 3)
            fft[1][:-3] *= r.uniform(0.5, 1.5, g.nj + 3)
            fft[2][:] = fft[0] + fft[1]
            fft[0][3] = 0.0  # Neumann rows keep as(1) = an(nj) = 0
            fft[1][g.nj + 2] = 0.0
        self.fft = fft

    def singular_zero(self, a):
        g = self.g
        a = a.reshape(g.nk, g.nj, g.ni).copy()
        a[g.nk // 2, :, g.ni // 2] = 0.0
        return a.ravel()

    def inputs(self):
        return {"pr": self.singular_zero(self.rng.standard_normal(self.nm())),
                "pi": self.singular_zero(self.rng.standard_normal(self.nm()))}

    def out_adjoints(self):
        return self.inputs()

    def head(self):
        g = self.g
        svall = np.array([g.ni, g.nj, g.nk], dtype=np.int32)
        n1 = max(g.nj, g.nk) + 6
        return [I(0), I(1), svall, D(self.rho), D(1 / g.dx), D(1 / g.dz), self.fft[0].copy(),
                self.fft[1].copy(), self.fft[2].copy(), np.zeros(n1), np.zeros(n1), np.zeros(n1)]

    def forward(self, call, x):
        pr, pi = x["pr"].copy(), x["pi"].copy()
        call("solp_fft_tdma2", *self.head(), pr, pi)
        return {"pr": pr, "pi": pi}

    def adjoint(self, call, y):
        pr, pi = y["pr"].copy(), y["pi"].copy()
        call("solp_fft_tdma2_b", *self.head(), pr, pi)
        return {"pr": pr, "pi": pi}


class InPlace(Case):
    """Routines that map one padded array to itself in place."""

    def inputs(self):
        return {"q": self.rng.standard_normal(self.g.shape)}

    def out_adjoints(self):
        return {"q": self.rng.standard_normal(self.g.shape)}

    def args(self, q):
        g = self.g
        return [I(g.ni), I(g.nj), I(g.nk), q]

    def forward(self, call, x):
        q = x["q"].copy()
        call(self.name, *self.args(q))
        return {"q": q}

    def adjoint(self, call, y):
        q = y["q"].copy()
        call(self.name + "_b", *self.args(q))
        return {"q": q}


class Periodic(InPlace):
    name = "bnd_periodic"


class Neumann(InPlace):
    name = "bnd_neumann"

    def args(self, q):
        g = self.g
        return [WALLS.copy(), I(g.ni), I(g.nj), I(g.nk), q]


class Avzero(InPlace):
    name = "bnd_avzero"

    def args(self, q):
        g = self.g
        return [I(0), I(1), I(g.ni), I(g.nj), I(g.nk), q]


class Corunp(Case):
    name = "corunp_explicit"
    rho, dt = 0.81, 0.003
    keys = ("dp", "phat", "un", "vn", "wn", "pn")

    def inputs(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in self.keys}

    def out_adjoints(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("un", "vn", "wn", "pn")}

    def head(self):
        g = self.g
        return [WALLS.copy(), I(g.ni), I(g.nj), I(g.nk), D(self.rho), D(1 / g.dx), D(1 / g.dy),
                D(1 / g.dz), D(self.dt)]

    def forward(self, call, x):
        a = {k: x[k].copy() for k in self.keys}
        call("corunp_explicit", *self.head(), *[a[k] for k in self.keys])
        return {k: a[k] for k in ("un", "vn", "wn", "pn")}

    def adjoint(self, call, y):
        g = self.g
        b = {"dp": g.zeros(), "phat": g.zeros()}
        b.update({k: y[k].copy() for k in ("un", "vn", "wn", "pn")})
        call("corunp_explicit_b", *self.head(), *[b[k] for k in self.keys])
        return b


class Bndu(Case):
    name = "bndu"
    keys = ("u", "v", "w")

    def __init__(self, g, seed):
        super().__init__(g, seed)
        r = np.random.default_rng(seed + 500)
        self.l1 = r.uniform(0.01, 0.1, g.shape)
        self.l2 = r.uniform(0.01, 0.1, g.shape)
        self.phi = r.uniform(0, 1, g.shape)

    def inputs(self):
        x = {k: self.rng.standard_normal(self.g.shape) for k in self.keys}
        x["uwall"] = np.array([self.rng.standard_normal()])
        return x

    def out_adjoints(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in self.keys}

    def forward(self, call, x):
        g = self.g
        a = [x[k].copy() for k in self.keys]
        call("bndu", WALLS.copy(), I(g.ni), I(g.nj), I(g.nk), *a, D(float(x["uwall"][0])), D(g.dy),
             self.l1, self.l2, self.phi)
        return dict(zip(self.keys, a))

    def adjoint(self, call, y):
        g = self.g
        a = [y[k].copy() for k in self.keys]
        ub = D(0.0)
        call("bndu_b", WALLS.copy(), I(g.ni), I(g.nj), I(g.nk), *a, ub, D(g.dy), self.l1, self.l2,
             self.phi)
        out = dict(zip(self.keys, a))
        out["uwall"] = np.array([ub.value])
        return out


def sor_lp_oracle() -> int:
    return ctypes.c_int.in_dll(H.oracle(), "orc_sor_last_lp").value


def sor_lp_gpu() -> int:
    lp, ratio = ctypes.c_int(0), ctypes.c_double(0.0)
    H.gpu().felgpu_sor_last_(ctypes.byref(lp), ctypes.byref(ratio))
    return lp.value


class Sor(Case):
    """solu_sor4 with its sweep count frozen at the forward call's: linear in
    (src_u, src_v, src_w, un, vn, wn, uwall) for fixed densities and
    coefficients."""
    name = "solu_sor4"
    keys = ("src_u", "src_v", "src_w", "un", "vn", "wn")
    dt = 0.01

    def __init__(self, g, seed, smooth=True):
        super().__init__(g, seed)
        r = np.random.default_rng(seed + 700)
        self.rmu = r.uniform(0.01, 0.2, g.shape)
        self.coef = [g.zeros() for _ in range(45)]
        H.call_oracle("cal_arith_coef_vis", I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy),
                      D(1 / g.dz), self.rmu, *self.coef)
        self.rho = r.uniform(0.5, 1.5, g.shape)
        self.rhon = r.uniform(0.5, 1.5, g.shape)
        self.l1 = np.full(g.shape, 0.05)
        self.phi = r.uniform(0, 1, g.shape)
        self.smooth = smooth
        self.lp = None

    def inputs(self):
        g, r = self.g, self.rng
        x = {k: r.standard_normal(g.shape) for k in self.keys[:3]}
        for k in self.keys[3:]:
            x[k] = H.smooth_vel(g, r) if self.smooth else r.standard_normal(g.shape)
        x["uwall"] = np.array([r.standard_normal()])
        return x

    def out_adjoints(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("un", "vn", "wn")}

    def head(self):
        g = self.g
        return [I(0), I(0), WALLS.copy(), I(1), I(g.ni), I(g.nj), I(g.nk), np.zeros(8, np.int32),
                np.zeros(8), np.zeros(8), I(1), I(0), D(self.dt), self.rho, self.rhon] + \
            [c.copy() for c in self.coef]

    def forward(self, call, x, lp_reader=None):
        g = self.g
        a = {k: x[k].copy() for k in self.keys}
        call("solu_sor4", *self.head(), *[a[k] for k in self.keys], D(float(x["uwall"][0])), D(g.dy),
             self.l1, self.l1, self.phi)
        self.lp = (lp_reader or sor_lp_oracle)()
        return {k: a[k] for k in ("un", "vn", "wn")}

    def adjoint(self, call, y):
        g = self.g
        assert self.lp is not None, "run the forward first (its sweep count is frozen)"
        b = {k: g.zeros() for k in self.keys[:3]}
        b.update({k: y[k].copy() for k in ("un", "vn", "wn")})
        ub = D(0.0)
        call("solu_sor4_b", *self.head(), *[b[k] for k in self.keys], ub, D(g.dy), self.l1, self.l1,
             self.phi, I(self.lp))
        b["uwall"] = np.array([ub.value])
        return b


def projection_forward(call, g, fft, x):
    """caldiv_ .. corunp_explicit_ (srcs/main.f90:634-650) on (un, vn, wn,
    phat) -> (un, vn, wn, pn); the divergence, modes and dp are work arrays."""
    rho, dt = 0.81, 0.003
    nm = g.ni * g.nj * g.nk
    svall = np.array([g.ni, g.nj, g.nk], dtype=np.int32)
    u, v, w, phat = (x[k].copy() for k in ("un", "vn", "wn", "phat"))
    div, dp, pn = g.zeros(), g.zeros(), g.zeros()
    pr, pi = np.zeros(nm), np.zeros(nm)
    n1 = max(g.nj, g.nk) + 6
    call("caldiv", I(0), I(1), I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), D(dt),
         u, v, w, div, D(0.0))
    call("solp_fft_tdma1", I(g.ni), I(g.nj), I(g.nk), np.zeros(2 * nm), div, pr, pi)
    call("solp_fft_tdma2", I(0), I(1), svall, D(rho), D(1 / g.dx), D(1 / g.dz), fft[0], fft[1], fft[2],
         np.zeros(n1), np.zeros(n1), np.zeros(n1), pr, pi)
    call("solp_fft_tdma3", I(g.ni), I(g.nj), I(g.nk), np.zeros(2 * nm), pr, pi, dp)
    call("bnd_avzero", I(0), I(1), I(g.ni), I(g.nj), I(g.nk), dp)
    call("bnd_neumann", WALLS.copy(), I(g.ni), I(g.nj), I(g.nk), dp)
    call("bnd_periodic", I(g.ni), I(g.nj), I(g.nk), dp)
    call("corunp_explicit", WALLS.copy(), I(g.ni), I(g.nj), I(g.nk), D(rho), D(1 / g.dx), D(1 / g.dy),
         D(1 / g.dz), D(dt), dp, phat, u, v, w, pn)
    return {"un": u, "vn": v, "wn": w, "pn": np.where(interior_mask(g), pn, 0.0)}


def projection_adjoint(call, g, fft, y):
    rho, dt = 0.81, 0.003
    nm = g.ni * g.nj * g.nk
    svall = np.array([g.ni, g.nj, g.nk], dtype=np.int32)
    n1 = max(g.nj, g.nk) + 6
    ub, vb, wb = (y[k].copy() for k in ("un", "vn", "wn"))
    pnb = np.where(interior_mask(g), y["pn"], 0.0)
    dpb, phatb, divb = g.zeros(), g.zeros(), g.zeros()
    prb, pib = np.zeros(nm), np.zeros(nm)
    call("corunp_explicit_b", WALLS.copy(), I(g.ni), I(g.nj), I(g.nk), D(rho), D(1 / g.dx), D(1 / g.dy),
         D(1 / g.dz), D(dt), dpb, phatb, ub, vb, wb, pnb)
    call("bnd_periodic_b", I(g.ni), I(g.nj), I(g.nk), dpb)
    call("bnd_neumann_b", WALLS.copy(), I(g.ni), I(g.nj), I(g.nk), dpb)
    call("bnd_avzero_b", I(0), I(1), I(g.ni), I(g.nj), I(g.nk), dpb)
    call("solp_fft_tdma3_b", I(g.ni), I(g.nj), I(g.nk), np.zeros(2 * nm), prb, pib, dpb)
    call("solp_fft_tdma2_b", I(0), I(1), svall, D(rho), D(1 / g.dx), D(1 / g.dz), fft[0], fft[1], fft[2],
         np.zeros(n1), np.zeros(n1), np.zeros(n1), prb, pib)
    call("solp_fft_tdma1_b", I(g.ni), I(g.nj), I(g.nk), np.zeros(2 * nm), divb, prb, pib)
    call("caldiv_b", I(0), I(1), I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), D(dt),
         ub, vb, wb, divb)
    return {"un": ub, "vn": vb, "wn": wb, "phat": phatb}


def consistent_velocity(g, rng):
    """Random interior with the ghost images the divergence reads made
    consistent (x and z periodic, v zero on both wall faces), so that every
    plane's divergence sums to zero and the singular pressure mode's
    right-hand side is compatible, as it is in a run."""
    out = {}
    for k in ("un", "vn", "wn"):
        q = rng.standard_normal(g.shape)
        if k == "vn":
            q[:, 2, :] = 0.0          # j = 0
            q[:, g.nj + 2, :] = 0.0   # j = nj
        H.call_oracle("bnd_periodic", I(g.ni), I(g.nj), I(g.nk), q)
        out[k] = q
    out["phat"] = rng.standard_normal(g.shape)
    return out


# --------------------------------------------------------------------------
# Explicit momentum terms
# --------------------------------------------------------------------------

def grid_head(g: Grid):
    return [I(g.ni), I(g.nj), I(g.nk)]


def dxyz(g: Grid):
    return [D(1 / g.dx), D(1 / g.dy), D(1 / g.dz)]


class Calphat(Case):
    name = "calphat"

    def inputs(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("p", "po")}

    def out_adjoints(self):
        return {"phat": self.rng.standard_normal(self.g.shape)}

    def forward(self, call, x):
        phat = self.g.zeros()
        call("calphat", *grid_head(self.g), x["p"].copy(), x["po"].copy(), phat)
        return {"phat": phat}

    def adjoint(self, call, y):
        g = self.g
        pb, pob, phb = g.zeros(), g.zeros(), y["phat"].copy()
        call("calphat_b", *grid_head(g), pb, pob, phb)
        # the forward leaves i = -2 of phat unwritten: its adjoint passes through
        return {"p": pb, "po": pob, "_phat_kept": phb}


class GradPrs(Case):
    name = "cal_grad_prs"

    def inputs(self):
        return {"q": self.rng.standard_normal(self.g.shape)}

    def out_adjoints(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("qx", "qy", "qz")}

    def forward(self, call, x):
        g = self.g
        o = [g.zeros() for _ in range(3)]
        call("cal_grad_prs", *grid_head(g), *dxyz(g), x["q"].copy(), *o)
        return dict(zip(("qx", "qy", "qz"), o))

    def adjoint(self, call, y):
        g = self.g
        qb = g.zeros()
        o = [y[k].copy() for k in ("qx", "qy", "qz")]
        call("cal_grad_prs_b", *grid_head(g), *dxyz(g), qb, *o)
        return {"q": qb, "_kept": o}


class Calsij(Case):
    name = "calsij"

    def inputs(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("u", "v", "w")}

    def out_adjoints(self):
        return {"s": self.rng.standard_normal((6,) + self.g.shape)}

    def forward(self, call, x):
        g = self.g
        s = g.zeros(6)
        call("calsij", *grid_head(g), *dxyz(g), x["u"].copy(), x["v"].copy(), x["w"].copy(), s)
        return {"s": s}

    def adjoint(self, call, y):
        g = self.g
        b = [g.zeros() for _ in range(3)]
        s = y["s"].copy()
        call("calsij_b", *grid_head(g), *dxyz(g), *b, s)
        return {"u": b[0], "v": b[1], "w": b[2], "_kept": s}


class ArithTau(Case):
    name = "cal_arith_tau"

    def __init__(self, g, seed):
        super().__init__(g, seed)
        self.rmu = np.random.default_rng(seed + 300).uniform(0.01, 2.0, g.shape)

    def inputs(self):
        return {"s": self.rng.standard_normal((6,) + self.g.shape)}

    def out_adjoints(self):
        return {"tau": self.rng.standard_normal((6,) + self.g.shape)}

    def forward(self, call, x):
        g = self.g
        tau = g.zeros(6)
        call("cal_arith_tau", *grid_head(g), x["s"].copy(), self.rmu, tau)
        return {"tau": tau}

    def adjoint(self, call, y):
        g = self.g
        sb = g.zeros(6)
        tb = y["tau"].copy()
        call("cal_arith_tau_b", *grid_head(g), sb, self.rmu, tb)
        return {"s": sb, "_kept": tb}


class DivTensor(Case):
    name = "cal_div_tensor"

    def inputs(self):
        return {"tau": self.rng.standard_normal((6,) + self.g.shape)}

    def out_adjoints(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("u", "v", "w")}

    def forward(self, call, x):
        g = self.g
        o = [g.zeros() for _ in range(3)]
        call("cal_div_tensor", *grid_head(g), *dxyz(g), x["tau"].copy(), *o)
        return dict(zip(("u", "v", "w"), o))

    def adjoint(self, call, y):
        g = self.g
        tb = g.zeros(6)
        o = [y[k].copy() for k in ("u", "v", "w")]
        call("cal_div_tensor_b", *grid_head(g), *dxyz(g), tb, *o)
        assert not any(np.any(a) for a in o), "vis adjoints are consumed"
        return {"tau": tb}


SRCU_IN = ("u", "v", "w", "adv_u", "adv_v", "adv_w", "adv_uo", "adv_vo", "adv_wo", "prs_u", "prs_v",
           "prs_w", "vis_u", "vis_v", "vis_w", "fst_u", "fst_v", "fst_w", "fst_un", "fst_vn", "fst_wn")


class Srcu(Case):
    name = "cal_srcu"
    dt = 0.003

    def __init__(self, g, seed):
        super().__init__(g, seed)
        r = np.random.default_rng(seed + 400)
        self.rho = r.uniform(0.5, 2.0, g.shape)
        self.rhon = r.uniform(0.5, 2.0, g.shape)

    def inputs(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in SRCU_IN}

    def out_adjoints(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("src_u", "src_v", "src_w")}

    def head(self):
        g = self.g
        return [WALLS.copy(), *grid_head(g), D(0.9), D(self.dt), D(0.0), D(0.0)]

    def forward(self, call, x):
        g = self.g
        a = {k: x[k].copy() for k in SRCU_IN}
        src = [g.zeros() for _ in range(3)]
        call("cal_srcu", *self.head(), a["u"], a["v"], a["w"], self.rho, self.rhon,
             *[a[k] for k in SRCU_IN[3:]], *src)
        return dict(zip(("src_u", "src_v", "src_w"), src))

    def adjoint(self, call, y):
        g = self.g
        b = {k: g.zeros() for k in SRCU_IN}
        src = [y[k].copy() for k in ("src_u", "src_v", "src_w")]
        call("cal_srcu_b", *self.head(), b["u"], b["v"], b["w"], self.rho, self.rhon,
             *[b[k] for k in SRCU_IN[3:]], *src)
        b["_kept"] = src
        return b


class Weno(Case):
    """cal_advu_weno5 is not linear: its Jacobian at x0 applied to v comes
    from Richardson-extrapolated central differences of the forward."""
    name = "cal_advu_weno5"

    def __init__(self, g, seed):
        super().__init__(g, seed)
        r = np.random.default_rng(seed + 600)
        self.x0 = {k: H.smooth_vel(g, r) + 0.05 * r.standard_normal(g.shape) for k in ("u", "v", "w")}

    def inputs(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("u", "v", "w")}

    def out_adjoints(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("adv_u", "adv_v", "adv_w")}

    def value(self, call, x):
        g = self.g
        o = [g.zeros() for _ in range(3)]
        call("cal_advu_weno5", *grid_head(g), *dxyz(g), x["u"].copy(), x["v"].copy(), x["w"].copy(), *o)
        return dict(zip(("adv_u", "adv_v", "adv_w"), o))

    def forward(self, call, x, h=1e-5):
        def central(hh):
            p = self.value(call, {k: self.x0[k] + hh * x[k] for k in x})
            m = self.value(call, {k: self.x0[k] - hh * x[k] for k in x})
            return {k: (p[k] - m[k]) / (2 * hh) for k in p}
        d1, d2 = central(h), central(2 * h)
        return {k: (4 * d1[k] - d2[k]) / 3 for k in d1}

    def adjoint(self, call, y):
        g = self.g
        b = {k: g.zeros() for k in ("u", "v", "w")}
        a = [y[k].copy() for k in ("adv_u", "adv_v", "adv_w")]
        call("cal_advu_weno5_b", *grid_head(g), *dxyz(g), self.x0["u"], b["u"], self.x0["v"], b["v"],
             self.x0["w"], b["w"], *a)
        assert not any(np.any(q) for q in a), "adv adjoints are consumed"
        return b


# --------------------------------------------------------------------------
# The velocity-pressure update at a fixed phase field (felmpf.adjoint)
# --------------------------------------------------------------------------

def momentum_inputs(cfg, seed: int):
    """Fixed fields (densities and viscosities of a droplet's two time
    levels, phi, slip lengths, surface-tension forces) and a flow state with
    generic (nonzero) transport velocities."""
    r = np.random.default_rng(seed)
    grid = Grid(cfg.ni, cfg.nj, cfg.nk, cfg.dx, cfg.dy, cfg.dz)
    phi = H.smooth_phi(grid, r)
    phin = np.clip(phi + 0.02 * r.standard_normal(phi.shape), 0.0, 1.0)
    fixed = {
        "rho": cfg.rhol * (1 - phi) + cfg.rhog * phi, "rhon": cfg.rhol * (1 - phin) + cfg.rhog * phin,
        "rmu": cfg.rmul * (1 - phi) + cfg.rmug * phi, "rmun": cfg.rmul * (1 - phin) + cfg.rmug * phin,
        "phi": phi, "l1": np.full(grid.shape, 0.3 * cfg.dy), "l2": np.full(grid.shape, 0.5 * cfg.dy),
    }
    for k in ("fst_u", "fst_v", "fst_w", "fst_un", "fst_vn", "fst_wn"):
        fixed[k] = 0.1 * r.standard_normal(grid.shape)
    state = {}
    for k in ("u", "v", "w", "uo", "vo", "wo"):
        state[k] = 0.1 * H.smooth_vel(grid, r) + 0.02 * r.standard_normal(grid.shape)
    for k in ("p", "po"):
        state[k] = 0.1 * r.standard_normal(grid.shape)
    return fixed, state


def random_like(d: dict, rng) -> dict:
    return {k: rng.standard_normal(np.shape(v)) for k, v in d.items()}


def richardson(fwd, x0: dict, v: dict, h: float):
    """(4 D(h) - D(2h)) / 3 of the central differences D of fwd along v."""
    def central(hh):
        p = fwd({k: x0[k] + hh * v[k] for k in x0})
        m = fwd({k: x0[k] - hh * v[k] for k in x0})
        return {k: (p[k] - m[k]) / (2 * hh) for k in p}
    d1, d2 = central(h), central(2 * h)
    return {k: (4 * d1[k] - d2[k]) / 3 for k in d1}


# --------------------------------------------------------------------------
# Volume-fraction transport (MTHINC)
# --------------------------------------------------------------------------

class FDCase(Case):
    """A nonlinear routine: forward(call, v) is its Jacobian at self.x0
    applied to v, by Richardson-extrapolated central differences."""
    h = 1e-6

    def value(self, call, x):
        raise NotImplementedError

    def forward(self, call, x):
        def central(hh):
            p = self.value(call, {k: self.x0[k] + hh * x[k] for k in x})
            m = self.value(call, {k: self.x0[k] - hh * x[k] for k in x})
            return {k: (p[k] - m[k]) / (2 * hh) for k in p}
        d1, d2 = central(self.h), central(2 * self.h)
        return {k: (4 * d1[k] - d2[k]) / 3 for k in d1}

    def inputs(self):
        return {k: self.rng.standard_normal(np.shape(v)) for k, v in self.x0.items()}


def vof_grad(g: Grid, phi, call=H.call_oracle):
    o = [g.zeros() for _ in range(3)]
    call("cal_grad_p2a", I(0), I(g.nj), I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx),
         np.full(g.nj + 6, 1 / g.dy), D(1 / g.dz), phi, *o)
    return o


def away_from_zero(r, shape, scale):
    """Velocities bounded away from 0, so that no upwind direction flips
    under a small perturbation."""
    return np.where(r.uniform(size=shape) < 0.5, -1.0, 1.0) * (0.2 + r.uniform(size=shape)) * scale


class GradP2a(Case):
    name = "cal_grad_p2a"

    def inputs(self):
        return {"q": self.rng.standard_normal(self.g.shape)}

    def out_adjoints(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("qx", "qy", "qz")}

    def head(self):
        g = self.g
        return [I(0), I(g.nj), I(g.ni), I(g.nj), I(g.nk), D(1 / g.dx), np.full(g.nj + 6, 1 / g.dy),
                D(1 / g.dz)]

    def forward(self, call, x):
        g = self.g
        o = [g.zeros() for _ in range(3)]
        call("cal_grad_p2a", *self.head(), x["q"].copy(), *o)
        return dict(zip(("qx", "qy", "qz"), o))

    def adjoint(self, call, y):
        g = self.g
        qb = g.zeros()
        o = [y[k].copy() for k in ("qx", "qy", "qz")]
        call("cal_grad_p2a_b", *self.head(), qb, *o)
        return {"q": qb, "_kept": o}


class Mthinc1(FDCase):
    name = "solphi_mthinc1"
    bet = 2.0

    def __init__(self, g, seed):
        super().__init__(g, seed)
        r = np.random.default_rng(seed + 800)
        phi = H.smooth_phi(g, r, beta=1.0)
        gx, gy, gz = vof_grad(g, phi)
        vel = {k: away_from_zero(r, g.shape, 0.5) for k in ("u", "v", "w")}
        self.dt = 0.2 * min(g.dx, g.dy, g.dz) / 1.2
        self.x0 = {"phi": phi, "phix": gx, "phiy": gy, "phiz": gz, **vel}

    def inputs(self):
        """Random directions, without phi components at cells within 1e-5 of
        the truncation thresholds (the differences would cross them)."""
        v = super().inputs()
        et = 0.5 * (np.tanh(-self.bet * 2.5) + 1.0)
        phi = self.x0["phi"]
        near = (np.abs(phi - et) < 1e-5) | (np.abs(1.0 - phi - et) < 1e-5)
        v["phi"] = np.where(near, 0.0, v["phi"])
        return v

    def out_adjoints(self):
        m = box_mask(self.g, -1, 1)
        return {k: np.where(m, self.rng.standard_normal(self.g.shape), 0.0) for k in ("fx", "fy", "fz")}

    def head(self):
        g = self.g
        return [I(g.ni), I(g.nj), I(g.nk), *dxyz(g), D(self.dt), D(self.bet)]

    def value(self, call, x):
        g = self.g
        o = [g.zeros() for _ in range(3)]
        call("solphi_mthinc1", *self.head(), x["phi"].copy(), x["phix"].copy(), x["phiy"].copy(),
             x["phiz"].copy(), x["u"].copy(), x["v"].copy(), x["w"].copy(), *o)
        return dict(zip(("fx", "fy", "fz"), o))

    def adjoint(self, call, y):
        g = self.g
        b = {k: g.zeros() for k in self.x0}
        o = [y[k].copy() for k in ("fx", "fy", "fz")]
        x0 = self.x0
        call("solphi_mthinc1_b", *self.head(), x0["phi"], b["phi"], x0["phix"], b["phix"],
             x0["phiy"], b["phiy"], x0["phiz"], b["phiz"], x0["u"], b["u"], x0["v"], b["v"],
             x0["w"], b["w"], *o)
        assert not any(np.any(q[box_mask(g, -1, 1)]) for q in o), "flux adjoints are consumed"
        return b


class Mthinc2(FDCase):
    name = "solphi_mthinc2"

    def __init__(self, g, seed):
        super().__init__(g, seed)
        r = np.random.default_rng(seed + 900)
        self.dt = 0.05 * g.dx
        self.x0 = {"u": r.standard_normal(g.shape), "v": r.standard_normal(g.shape),
                   "w": r.standard_normal(g.shape), "fx": 0.02 * r.standard_normal(g.shape),
                   "fy": 0.02 * r.standard_normal(g.shape), "fz": 0.02 * r.standard_normal(g.shape),
                   "phi": r.uniform(0.1, 0.9, g.shape)}
        # a few cells clamped at each bound
        self.x0["phi"][5, 4, 4] = 1.3
        self.x0["phi"][6, 5, 7] = -0.4

    def out_adjoints(self):
        return {"phin": np.where(interior_mask(self.g), self.rng.standard_normal(self.g.shape), 0.0)}

    def head(self):
        g = self.g
        return [I(g.ni), I(g.nj), I(g.nk), *dxyz(g), D(self.dt)]

    def value(self, call, x):
        g = self.g
        phin = g.zeros()
        call("solphi_mthinc2", *self.head(), x["u"].copy(), x["v"].copy(), x["w"].copy(),
             x["fx"].copy(), x["fy"].copy(), x["fz"].copy(), x["phi"].copy(), phin)
        return {"phin": np.where(interior_mask(g), phin, 0.0)}

    def adjoint(self, call, y):
        g = self.g
        b = {k: g.zeros() for k in self.x0}
        pb = y["phin"].copy()
        x0 = self.x0
        call("solphi_mthinc2_b", *self.head(), x0["u"], b["u"], x0["v"], b["v"], x0["w"], b["w"],
             x0["fx"], b["fx"], x0["fy"], b["fy"], x0["fz"], b["fz"], x0["phi"], b["phi"], pb)
        return b


class Mthinc3(FDCase):
    name = "solphi_mthinc3"
    bet = 2.0

    def __init__(self, g, seed):
        super().__init__(g, seed)
        r = np.random.default_rng(seed + 1000)
        phi = H.smooth_phi(g, r, beta=1.0)
        phin = np.clip(phi + 0.02 * r.standard_normal(g.shape), 0.0, 1.0)
        self.grad = vof_grad(g, phin)
        self.x0 = {"phi": phi, "phin": phin}

    def inputs(self):
        """Random directions, without phin components at cells within 1e-5 of
        the truncation thresholds or the clamp bounds."""
        v = super().inputs()
        et = 0.5 * (np.tanh(-self.bet * 2.5) + 1.0)
        p = self.x0["phin"]
        near = (np.abs(p - et) < 1e-5) | (np.abs(1.0 - p - et) < 1e-5) | (p < 1e-5) | (p > 1 - 1e-5)
        v["phin"] = np.where(near, 0.0, v["phin"])
        return v

    def out_adjoints(self):
        return {"phin": self.rng.standard_normal(self.g.shape)}

    def head(self):
        g = self.g
        return [I(0), I(g.ni), I(g.nj), I(g.nk), *dxyz(g), D(self.bet)]

    def value(self, call, x):
        g = self.g
        phin = x["phin"].copy()
        call("solphi_mthinc3", *self.head(), *self.grad, x["phi"].copy(), phin)
        return {"phin": phin}

    def adjoint(self, call, y):
        g = self.g
        b = {"phi": g.zeros(), "phin": y["phin"].copy()}
        x0 = self.x0
        call("solphi_mthinc3_b", *self.head(), *self.grad, x0["phi"], b["phi"], x0["phin"], b["phin"])
        return b


class Summation(FDCase):
    name = "summation"

    def __init__(self, g, seed):
        super().__init__(g, seed)
        r = np.random.default_rng(seed + 1100)
        q = np.zeros((2,) + g.shape)
        q[1] = r.uniform(0.1, 0.9, g.shape)
        q[1][5, 4, 4] = 1.2
        self.x0 = {"q": q}

    def inputs(self):
        v = self.rng.standard_normal((2,) + self.g.shape)
        v[0] = 0.0  # slice 0 is overwritten: not an input
        return {"q": v}

    def out_adjoints(self):
        y = np.zeros((2,) + self.g.shape)
        y[0] = self.rng.standard_normal(self.g.shape)
        return {"q": y}

    def value(self, call, x):
        g = self.g
        q = x["q"].copy()
        call("summation", I(g.ni), I(g.nj), I(g.nk), q, I(1))
        out = np.zeros_like(q)
        out[0] = q[0]
        return {"q": out}

    def adjoint(self, call, y):
        g = self.g
        qb = y["q"].copy()
        call("summation_b", I(g.ni), I(g.nj), I(g.nk), self.x0["q"], qb, I(1))
        assert not np.any(qb[0])
        return {"q": qb}


# --------------------------------------------------------------------------
# Contact line
# --------------------------------------------------------------------------

def wall_rows_mask(g: Grid):
    m = np.zeros(g.shape, bool)
    m[:, 3, :] = True          # j = 1
    m[:, g.nj + 2, :] = True   # j = nj
    return m


class Gnbc(FDCase):
    name = "gnbc"
    sigma = 5.5

    def __init__(self, g, seed):
        super().__init__(g, seed)
        r = np.random.default_rng(seed + 1200)
        self.x0 = {"u": away_from_zero(r, g.shape, 0.5), "w": r.standard_normal(g.shape),
                   "uwall": np.array([0.13]), "theta0": r.uniform(40.0, 140.0, g.shape),
                   "zeta": r.uniform(0.5, 2.0, g.shape)}

    def out_adjoints(self):
        return {"theta": np.where(wall_rows_mask(self.g), self.rng.standard_normal(self.g.shape), 0.0)}

    def value(self, call, x):
        g = self.g
        th = g.zeros()
        call("gnbc", WALLS.copy(), I(g.ni), I(g.nj), I(g.nk), x["u"].copy(), x["w"].copy(),
             D(float(x["uwall"][0])), x["theta0"].copy(), D(self.sigma), x["zeta"].copy(), th)
        return {"theta": np.where(wall_rows_mask(g), th, 0.0)}

    def adjoint(self, call, y):
        g = self.g
        b = {k: g.zeros() for k in ("u", "w", "theta0", "zeta")}
        ub = D(0.0)
        tb = y["theta"].copy()
        x0 = self.x0
        call("gnbc_b", WALLS.copy(), I(g.ni), I(g.nj), I(g.nk), x0["u"], b["u"], x0["w"], b["w"],
             D(float(x0["uwall"][0])), ub, x0["theta0"], b["theta0"], D(self.sigma), x0["zeta"],
             b["zeta"], tb)
        assert not np.any(tb[wall_rows_mask(g)])
        b["uwall"] = np.array([ub.value])
        return b


class ContactAngle(FDCase):
    name = "bnd_contact_angle"

    def __init__(self, g, seed):
        super().__init__(g, seed)
        r = np.random.default_rng(seed + 1300)
        self.x0 = {"q": H.smooth_phi(g, r), "theta": r.uniform(30.0, 150.0, g.shape)}

    def out_adjoints(self):
        return {"q": self.rng.standard_normal(self.g.shape)}

    def value(self, call, x):
        g = self.g
        q = x["q"].copy()
        call("bnd_contact_angle", WALLS.copy(), I(g.ni), I(g.nj), I(g.nk), q, x["theta"].copy(),
             D(g.dx), D(g.dy), D(g.dz))
        return {"q": q}

    def adjoint(self, call, y):
        g = self.g
        qb, tb = y["q"].copy(), g.zeros()
        x0 = self.x0
        call("bnd_contact_angle_b", WALLS.copy(), I(g.ni), I(g.nj), I(g.nk), x0["q"], qb,
             x0["theta"], tb, D(g.dx), D(g.dy), D(g.dz))
        return {"q": qb, "theta": tb}


def phase_inputs(cfg, seed: int):
    """A smooth volume fraction strictly inside (0.15, 0.85) (no truncation
    threshold, clamp bound or flat-cell switch near it), velocities bounded
    away from 0, contact angles and friction of the wall rows."""
    r = np.random.default_rng(seed)
    grid = Grid(cfg.ni, cfg.nj, cfg.nk, cfg.dx, cfg.dy, cfg.dz)
    x, y, z = grid.coords()
    a = r.uniform(0, 2 * np.pi, 4)
    phi = 0.5 + 0.35 * np.sin(2 * np.pi * x / (cfg.ni * cfg.dx) + a[0]) * \
        np.cos(np.pi * y / (cfg.nj * cfg.dy) + a[1]) * np.cos(2 * np.pi * z / (cfg.nk * cfg.dz) + a[2])
    params = {"u": away_from_zero(r, grid.shape, 0.3), "v": away_from_zero(r, grid.shape, 0.3),
              "w": away_from_zero(r, grid.shape, 0.3), "theta0": r.uniform(60.0, 120.0, grid.shape),
              "zeta": r.uniform(0.001, 0.01, grid.shape)}
    return phi, params


# --------------------------------------------------------------------------
# Material properties, surface tension and the parameter adjoints (_bp) of
# the momentum routines
# --------------------------------------------------------------------------

class ArithMean(FDCase):
    name = "cal_arith_mean"
    ql, qg = 1.3, 0.07

    def __init__(self, g, seed):
        super().__init__(g, seed)
        self.x0 = {"phi": np.random.default_rng(seed + 1400).uniform(0, 1, g.shape)}

    def out_adjoints(self):
        return {"q": self.rng.standard_normal(self.g.shape)}

    def value(self, call, x):
        q = self.g.zeros()
        call("cal_arith_mean", *grid_head(self.g), x["phi"].copy(), D(self.ql), D(self.qg), q)
        return {"q": q}

    def adjoint(self, call, y):
        pb, qb = self.g.zeros(), y["q"].copy()
        call("cal_arith_mean_b", *grid_head(self.g), pb, D(self.ql), D(self.qg), qb)
        assert not np.any(qb)
        return {"phi": pb}


class ArithTauP(FDCase):
    """cal_arith_tau_bp: the stress's adjoint with respect to the strain
    rates and the viscosity."""
    name = "cal_arith_tau_bp"

    def __init__(self, g, seed):
        super().__init__(g, seed)
        r = np.random.default_rng(seed + 1500)
        self.x0 = {"s": r.standard_normal((6,) + g.shape), "rmu": r.uniform(0.01, 2.0, g.shape)}

    def out_adjoints(self):
        return {"tau": self.rng.standard_normal((6,) + self.g.shape)}

    def value(self, call, x):
        tau = self.g.zeros(6)
        call("cal_arith_tau", *grid_head(self.g), x["s"].copy(), x["rmu"].copy(), tau)
        return {"tau": tau}

    def adjoint(self, call, y):
        g, x0 = self.g, self.x0
        sb, mb, tb = g.zeros(6), g.zeros(), y["tau"].copy()
        call("cal_arith_tau_bp", *grid_head(g), x0["s"], sb, x0["rmu"], mb, tb)
        return {"s": sb, "rmu": mb}


class SrcuP(FDCase):
    """cal_srcu_bp: the source's adjoint including the densities."""
    name = "cal_srcu_bp"
    dt = 0.003

    def __init__(self, g, seed):
        super().__init__(g, seed)
        r = np.random.default_rng(seed + 1600)
        self.x0 = {k: r.standard_normal(g.shape) for k in SRCU_IN}
        self.x0["rho"] = r.uniform(0.5, 2.0, g.shape)
        self.x0["rhon"] = r.uniform(0.5, 2.0, g.shape)

    def out_adjoints(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("src_u", "src_v", "src_w")}

    def head(self):
        return [WALLS.copy(), *grid_head(self.g), D(0.9), D(self.dt), D(0.0), D(0.0)]

    def value(self, call, x):
        g = self.g
        src = [g.zeros() for _ in range(3)]
        call("cal_srcu", *self.head(), x["u"].copy(), x["v"].copy(), x["w"].copy(), x["rho"].copy(),
             x["rhon"].copy(), *[x[k].copy() for k in SRCU_IN[3:]], *src)
        return dict(zip(("src_u", "src_v", "src_w"), src))

    def adjoint(self, call, y):
        g, x0 = self.g, self.x0
        b = {k: g.zeros() for k in SRCU_IN + ("rho", "rhon")}
        src = [y[k].copy() for k in ("src_u", "src_v", "src_w")]
        pv = []
        for k in ("prs_u", "prs_v", "prs_w", "vis_u", "vis_v", "vis_w"):
            pv += [x0[k], b[k]]
        call("cal_srcu_bp", *self.head(), b["u"], b["v"], b["w"], x0["rho"], b["rho"], x0["rhon"],
             b["rhon"], *[b[k] for k in ("adv_u", "adv_v", "adv_w", "adv_uo", "adv_vo", "adv_wo")],
             *pv, *[b[k] for k in SRCU_IN[15:]], *src)
        b["_kept"] = src
        return b


class CoefVis(Case):
    """cal_arith_coef_vis_b: linear in the viscosity."""
    name = "cal_arith_coef_vis"

    def inputs(self):
        return {"rmu": self.rng.standard_normal(self.g.shape)}

    def out_adjoints(self):
        return {f"c{i}": self.rng.standard_normal(self.g.shape) for i in range(45)}

    def head(self):
        g = self.g
        return [*grid_head(g), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz)]

    def forward(self, call, x):
        c = [self.g.zeros() for _ in range(45)]
        call("cal_arith_coef_vis", *self.head(), x["rmu"].copy(), *c)
        return {f"c{i}": c[i] for i in range(45)}

    def adjoint(self, call, y):
        mb = self.g.zeros()
        cb = [y[f"c{i}"].copy() for i in range(45)]
        call("cal_arith_coef_vis_b", *self.head(), mb, *cb)
        assert not any(np.any(c) for c in cb)
        return {"rmu": mb}


class BnduP(FDCase):
    """bndu_bp: the wall refresh's adjoint including the slip lengths and
    phi."""
    name = "bndu_bp"

    def __init__(self, g, seed):
        super().__init__(g, seed)
        r = np.random.default_rng(seed + 1700)
        self.x0 = {k: r.standard_normal(g.shape) for k in ("u", "v", "w")}
        self.x0.update({"uwall": np.array([0.37]), "l1": r.uniform(0.01, 0.1, g.shape),
                        "l2": r.uniform(0.01, 0.1, g.shape), "phi": r.uniform(0.05, 0.95, g.shape)})

    def out_adjoints(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("u", "v", "w")}

    def value(self, call, x):
        g = self.g
        a = [x[k].copy() for k in ("u", "v", "w")]
        call("bndu", WALLS.copy(), *grid_head(g), *a, D(float(x["uwall"][0])), D(g.dy),
             x["l1"].copy(), x["l2"].copy(), x["phi"].copy())
        return dict(zip(("u", "v", "w"), a))

    def adjoint(self, call, y):
        g, x0 = self.g, self.x0
        a = [y[k].copy() for k in ("u", "v", "w")]
        pb = {k: g.zeros() for k in ("l1", "l2", "phi")}
        ub = D(0.0)
        call("bndu_bp", WALLS.copy(), *grid_head(g), x0["u"], a[0], a[1], x0["w"], a[2],
             D(float(x0["uwall"][0])), ub, D(g.dy), x0["l1"], pb["l1"], x0["l2"], pb["l2"],
             x0["phi"], pb["phi"])
        out = dict(zip(("u", "v", "w"), a))
        out.update(pb)
        out["uwall"] = np.array([ub.value])
        return out


class SorP(FDCase):
    """solu_sor4_bp with the sweep count frozen at the unperturbed call's:
    the adjoint with respect to the sources, the initial state, uwall, the
    densities, the 45 coefficients, the slip lengths and phi."""
    name = "solu_sor4_bp"
    dt = 0.01
    h = 1e-6

    def __init__(self, g, seed):
        super().__init__(g, seed)
        r = np.random.default_rng(seed + 1800)
        rmu = r.uniform(0.01, 0.2, g.shape)
        coef = [g.zeros() for _ in range(45)]
        H.call_oracle("cal_arith_coef_vis", *grid_head(g), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz),
                      rmu, *coef)
        self.x0 = {f"c{i}": coef[i] for i in range(45)}
        for k in ("src_u", "src_v", "src_w"):
            self.x0[k] = r.standard_normal(g.shape)
        for k in ("un", "vn", "wn"):
            self.x0[k] = H.smooth_vel(g, r)
        self.x0.update({"uwall": np.array([0.21]), "rho": r.uniform(0.5, 1.5, g.shape),
                        "rhon": r.uniform(0.5, 1.5, g.shape), "l1": r.uniform(0.02, 0.1, g.shape),
                        "l2": r.uniform(0.02, 0.1, g.shape), "phi": r.uniform(0.05, 0.95, g.shape)})
        self.lps = []
        self.lp = None

    def head(self, x):
        g = self.g
        return [I(0), I(0), WALLS.copy(), I(1), *grid_head(g), np.zeros(8, np.int32), np.zeros(8),
                np.zeros(8), I(1), I(0), D(self.dt)]

    def out_adjoints(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("un", "vn", "wn")}

    def value(self, call, x, lp_reader=None):
        g = self.g
        a = {k: x[k].copy() for k in ("un", "vn", "wn")}
        call("solu_sor4", *self.head(x), x["rho"].copy(), x["rhon"].copy(),
             *[x[f"c{i}"].copy() for i in range(45)],
             *[x[k].copy() for k in ("src_u", "src_v", "src_w")], a["un"], a["vn"], a["wn"],
             D(float(x["uwall"][0])), D(g.dy), x["l1"].copy(), x["l2"].copy(), x["phi"].copy())
        reader = lp_reader or (sor_lp_gpu if call is H.call_gpu else sor_lp_oracle)
        self.lps.append(reader())
        return a

    def adjoint(self, call, y, lp=None):
        g, x0 = self.g, self.x0
        if lp is None:
            self.value(H.call_oracle, x0)
            lp = self.lps[-1]
        self.lp = lp
        assert all(v == lp for v in self.lps), self.lps
        b = {k: g.zeros() for k in x0 if k not in ("un", "vn", "wn", "uwall")}
        b.update({k: y[k].copy() for k in ("un", "vn", "wn")})
        ub = D(0.0)
        cargs = [x0[f"c{i}"] for i in range(45)] + [b[f"c{i}"] for i in range(45)]
        pairs = []
        for k in ("src_u", "src_v", "src_w", "un", "vn", "wn"):
            pairs += [x0[k], b[k]]
        call("solu_sor4_bp", *self.head(x0), x0["rho"], b["rho"], x0["rhon"], b["rhon"], *cargs,
             *pairs, D(float(x0["uwall"][0])), ub, D(g.dy), x0["l1"], b["l1"], x0["l2"], b["l2"],
             x0["phi"], b["phi"], I(lp))
        b["uwall"] = np.array([ub.value])
        return b


def vertex_gradients(g: Grid, r):
    """Vertex gradients of a smooth volume fraction plus noise (no zero
    mean normal on the staggered boxes)."""
    phi = H.smooth_phi(g, r)
    o = vof_grad(g, phi)
    return [a + 0.3 * r.standard_normal(g.shape) for a in o]


class Calrn(FDCase):
    name = "calrn"

    def __init__(self, g, seed, stag=(1, 1, 1)):
        super().__init__(g, seed)
        r = np.random.default_rng(seed + 1900)
        self.stag = stag
        gx, gy, gz = vertex_gradients(g, r)
        self.x0 = {"phix": gx, "phiy": gy, "phiz": gz}

    def out_adjoints(self):
        return {"rn": self.rng.standard_normal((3,) + self.g.shape)}

    def head(self):
        g = self.g
        return [*grid_head(g), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz), *[I(v) for v in self.stag]]

    def value(self, call, x):
        rn = self.g.zeros(3)
        call("calrn", *self.head(), x["phix"].copy(), x["phiy"].copy(), x["phiz"].copy(), rn)
        return {"rn": rn}

    def adjoint(self, call, y):
        g, x0 = self.g, self.x0
        b = {k: g.zeros() for k in ("phix", "phiy", "phiz")}
        rb = y["rn"].copy()
        call("calrn_b", *self.head(), x0["phix"], b["phix"], x0["phiy"], b["phiy"], x0["phiz"],
             b["phiz"], rb)
        assert not np.any(rb)
        return b


class Calrkap(Case):
    name = "calrkap"
    keys = ("rn_pp", "rn_12", "rn_13", "rn_23")

    def inputs(self):
        return {k: self.rng.standard_normal((3,) + self.g.shape) for k in self.keys}

    def out_adjoints(self):
        return {"rkap": self.rng.standard_normal((3,) + self.g.shape)}

    def head(self):
        g = self.g
        return [*grid_head(g), D(1 / g.dx), D(1 / g.dy), D(1 / g.dz)]

    def forward(self, call, x):
        rk = self.g.zeros(3)
        call("calrkap", *self.head(), *[x[k].copy() for k in self.keys], rk)
        return {"rkap": rk}

    def adjoint(self, call, y):
        b = {k: self.g.zeros(3) for k in self.keys}
        rk = y["rkap"].copy()
        call("calrkap_b", *self.head(), *[b[k] for k in self.keys], rk)
        assert not np.any(rk)
        return b


class GradP2uvw(Case):
    name = "cal_grad_p2uvw"

    def inputs(self):
        return {"q": self.rng.standard_normal(self.g.shape)}

    def out_adjoints(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("qx", "qy", "qz")}

    def head(self):
        g = self.g
        return [I(0), I(g.nj), *grid_head(g), D(1 / g.dx), np.full(g.nj + 6, 1 / g.dy), D(1 / g.dz)]

    def forward(self, call, x):
        o = [self.g.zeros() for _ in range(3)]
        call("cal_grad_p2uvw", *self.head(), x["q"].copy(), *o)
        return dict(zip(("qx", "qy", "qz"), o))

    def adjoint(self, call, y):
        qb = self.g.zeros()
        o = [y[k].copy() for k in ("qx", "qy", "qz")]
        call("cal_grad_p2uvw_b", *self.head(), qb, *o)
        return {"q": qb, "_kept": o}


class Calfst(FDCase):
    name = "calfst"
    rhol, rhog, sigma = 1.0, 0.01, 0.7

    def __init__(self, g, seed):
        super().__init__(g, seed)
        r = np.random.default_rng(seed + 2000)
        self.x0 = {k: r.standard_normal(g.shape) for k in ("gx", "gy", "gz")}
        self.x0["rkap"] = r.standard_normal((3,) + g.shape)

    def out_adjoints(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("fu", "fv", "fw")}

    def consts(self):
        return [D(self.rhol), D(self.rhog), D(self.sigma)]

    def value(self, call, x):
        g = self.g
        f = [g.zeros() for _ in range(3)]
        call("calfst", *grid_head(g), *self.consts(), x["gx"].copy(), x["gy"].copy(), x["gz"].copy(),
             x["rkap"].copy(), *f)
        return dict(zip(("fu", "fv", "fw"), f))

    def adjoint(self, call, y):
        g, x0 = self.g, self.x0
        b = {k: g.zeros() for k in ("gx", "gy", "gz")}
        b["rkap"] = g.zeros(3)
        f = [y[k].copy() for k in ("fu", "fv", "fw")]
        call("calfst_b", *grid_head(g), *self.consts(), x0["gx"], b["gx"], x0["gy"], b["gy"],
             x0["gz"], b["gz"], x0["rkap"], b["rkap"], *f)
        assert not any(np.any(a) for a in f)
        return b


class SumFst(Case):
    name = "sum_fst"

    def inputs(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("fu", "fv", "fw", "su", "sv", "sw")}

    def out_adjoints(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("su", "sv", "sw")}

    def forward(self, call, x):
        s = [x[k].copy() for k in ("su", "sv", "sw")]
        call("sum_fst", *grid_head(self.g), *[x[k].copy() for k in ("fu", "fv", "fw")], *s)
        return dict(zip(("su", "sv", "sw"), s))

    def adjoint(self, call, y):
        f = [self.g.zeros() for _ in range(3)]
        s = [y[k].copy() for k in ("su", "sv", "sw")]
        call("sum_fst_b", *grid_head(self.g), *f, *s)
        out = dict(zip(("fu", "fv", "fw"), f))
        out.update(dict(zip(("su", "sv", "sw"), s)))
        return out


def full_inputs(cfg, seed: int):
    """A state and wall maps for the whole step: the phase inputs' volume
    fraction (phi(.,0) a small perturbation of it), velocities bounded away
    from 0 at both time levels, pressures, contact angles, friction and
    slip lengths of the wall rows."""
    phi1, par = phase_inputs(cfg, seed)
    r = np.random.default_rng(seed + 1)
    state = {"phi1": phi1, "phi0": phi1 + 0.01 * r.standard_normal(phi1.shape),
             "u": par["u"], "v": par["v"], "w": par["w"]}
    for k in ("uo", "vo", "wo"):
        state[k] = away_from_zero(r, phi1.shape, 0.3)
    for k in ("p", "po"):
        state[k] = 0.1 * r.standard_normal(phi1.shape)
    params = {"theta0": par["theta0"], "zeta": par["zeta"],
              "l1": r.uniform(0.3, 0.6, phi1.shape) * cfg.dy,
              "l2": r.uniform(0.3, 0.6, phi1.shape) * cfg.dy}
    return state, params


class SorPV(FDCase):
    """solu_sor4_bpv: the adjoint of cal_arith_coef_vis followed by
    solu_sor4 (coefficients from the viscosity), with respect to the
    sources, state, uwall, densities, viscosity, slip lengths and phi; sweep
    count frozen at the unperturbed call's."""
    name = "solu_sor4_bpv"
    dt = 0.01
    h = 1e-6

    def __init__(self, g, seed):
        super().__init__(g, seed)
        r = np.random.default_rng(seed + 2100)
        self.x0 = {"rmu": r.uniform(0.01, 0.2, g.shape)}
        for k in ("src_u", "src_v", "src_w"):
            self.x0[k] = r.standard_normal(g.shape)
        for k in ("un", "vn", "wn"):
            self.x0[k] = H.smooth_vel(g, r)
        self.x0.update({"uwall": np.array([0.21]), "rho": r.uniform(0.5, 1.5, g.shape),
                        "rhon": r.uniform(0.5, 1.5, g.shape), "l1": r.uniform(0.02, 0.1, g.shape),
                        "l2": r.uniform(0.02, 0.1, g.shape), "phi": r.uniform(0.05, 0.95, g.shape)})
        self.lps = []

    def head(self):
        g = self.g
        return [I(0), I(0), WALLS.copy(), I(1), *grid_head(g), np.zeros(8, np.int32), np.zeros(8),
                np.zeros(8), I(1), I(0), D(self.dt)]

    def dinv(self):
        g = self.g
        return [D(1 / g.dx), D(1 / g.dy), D(1 / g.dz)]

    def out_adjoints(self):
        return {k: self.rng.standard_normal(self.g.shape) for k in ("un", "vn", "wn")}

    def value(self, call, x):
        g = self.g
        coef = [g.zeros() for _ in range(45)]
        call("cal_arith_coef_vis", *grid_head(g), *self.dinv(), x["rmu"].copy(), *coef)
        a = {k: x[k].copy() for k in ("un", "vn", "wn")}
        call("solu_sor4", *self.head(), x["rho"].copy(), x["rhon"].copy(), *coef,
             *[x[k].copy() for k in ("src_u", "src_v", "src_w")], a["un"], a["vn"], a["wn"],
             D(float(x["uwall"][0])), D(g.dy), x["l1"].copy(), x["l2"].copy(), x["phi"].copy())
        reader = sor_lp_gpu if call is H.call_gpu else sor_lp_oracle
        self.lps.append(reader())
        return a

    def adjoint(self, call, y):
        g, x0 = self.g, self.x0
        self.value(H.call_oracle, x0)
        lp = self.lps[-1]
        assert all(v == lp for v in self.lps), self.lps
        b = {k: g.zeros() for k in x0 if k not in ("un", "vn", "wn", "uwall")}
        b.update({k: y[k].copy() for k in ("un", "vn", "wn")})
        ub = D(0.0)
        pairs = []
        for k in ("src_u", "src_v", "src_w", "un", "vn", "wn"):
            pairs += [x0[k], b[k]]
        call("solu_sor4_bpv", *self.head(), x0["rho"], b["rho"], x0["rhon"], b["rhon"], *self.dinv(),
             x0["rmu"], b["rmu"], *pairs, D(float(x0["uwall"][0])), ub, D(g.dy), x0["l1"], b["l1"],
             x0["l2"], b["l2"], x0["phi"], b["phi"], I(lp))
        b["uwall"] = np.array([ub.value])
        return b
