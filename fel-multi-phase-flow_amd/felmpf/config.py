"""Run configurations and initial states.

The reference hard-codes its parameters (srcs/main.f90:91-249) and is
periodic in x (i) and z (k) with walls at the j faces (srcs/bndu.f:29,63;
srcs/bnd_periodic.f:16-43).  The benchmark configurations are stated with
walls at z, so a configuration's (x, y, z) maps to the reference's (i, k, j):
the wall-normal axis is j.  Phase convention of srcs/cal_arith_mean.f:18-26:
phi = 1 takes the "g" properties, so the ambient (light) phase is phi = 1 and
a droplet or liquid stripe is phi = 0.  Dimensionless groups are turned into
the reference's dimensional parameters with liquid scales rho_l = 1, U = 1
and L = the droplet diameter (stripe width for the Couette case):
mu_l = L/Re, sigma = L/We (or mu_l U / Ca), g = Bo sigma / L^2 acting along
-j.  zeta (the GNBC friction) follows srcs/main.f90:172.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field, replace

import numpy as np


class FelCfg(ctypes.Structure):
    """Mirror of fel_cfg (host/fel_driver.f90) and of the oracle's orc_cfg."""

    _fields_ = [
        ("ni", ctypes.c_int), ("nj", ctypes.c_int), ("nk", ctypes.c_int), ("imon_t", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double),
        ("bet_mthinc", ctypes.c_double), ("rhol", ctypes.c_double), ("rhog", ctypes.c_double),
        ("rmul", ctypes.c_double), ("rmug", ctypes.c_double), ("surface_tension", ctypes.c_double),
        ("uwall", ctypes.c_double), ("grvb", ctypes.c_double), ("grvp", ctypes.c_double),
        ("xl", ctypes.c_double), ("yl", ctypes.c_double), ("zl", ctypes.c_double),
        ("particle_radius", ctypes.c_double), ("cfl", ctypes.c_double),
    ]


@dataclass
class WallSide:
    theta0: float          # static angle [deg]
    slip: float            # slip length l1 = l2
    theta_alt: float = 0.0  # second angle of a checkerboard or stripe pattern
    patch: int = 0          # checkerboard patch size in cells (0: uniform wall)
    stripe: int = 0         # period in cells of x-stripes theta0 / theta_alt (0: none)


@dataclass
class RunConfig:
    name: str
    ni: int
    nj: int
    nk: int
    xl: float
    yl: float
    zl: float
    dt: float
    rhol: float
    rhog: float
    rmul: float
    rmug: float
    sigma: float
    bottom: WallSide
    top: WallSide
    uwall: float = 0.0
    grvp: float = 0.0
    bet: float = 2.0
    imon_t: int = 0
    steps: int = 100
    init: str = "droplet"
    params: dict = field(default_factory=dict)
    radius: float = 5e-4  # cal_vel's particle radius (diagnostic only)
    cfl: float = 0.0      # > 0: adaptive dt by caldt each step (srcs/caldt.f); 0: fixed dt

    @property
    def dx(self):
        return self.xl / self.ni

    @property
    def dy(self):
        return self.yl / self.nj

    @property
    def dz(self):
        return self.zl / self.nk

    @property
    def cells(self):
        return self.ni * self.nj * self.nk

    @property
    def shape(self):
        return (self.nk + 6, self.nj + 6, self.ni + 6)

    def cstruct(self) -> FelCfg:
        return FelCfg(self.ni, self.nj, self.nk, self.imon_t, self.dx, self.dy, self.dz, self.dt,
                      self.bet, self.rhol, self.rhog, self.rmul, self.rmug, self.sigma, self.uwall,
                      0.0, self.grvp, self.xl, self.yl, self.zl, self.radius, self.cfl)

    def scaled(self, ni, nj, nk, **kw) -> "RunConfig":
        """The same physics on another grid (same domain lengths)."""
        return replace(self, ni=ni, nj=nj, nk=nk, **kw)


def zeta_of(rmul, rmug, l1, l2):
    """GNBC friction coefficient, srcs/main.f90:172."""
    return 0.21 * 6.0 * (rmul / l1 + rmug / l2)


# ---------------------------------------------------------------------------
# The benchmark configurations, in reference orientation
# ---------------------------------------------------------------------------

def sessile_droplet() -> RunConfig:
    """64^3 unit cube, walls at the j faces with GNBC (slip 0.02), cap of
    radius 0.25 on the bottom wall at 90 deg, static angle 60 deg, density
    and viscosity ratios 10, Re 100, We 10, beta 1, dt 1e-4, 100 steps."""
    D = 0.5
    rmul = D / 100.0
    return RunConfig("sessile_droplet", 64, 64, 64, 1.0, 1.0, 1.0, dt=1e-4, rhol=1.0, rhog=0.1,
                     rmul=rmul, rmug=rmul / 10.0, sigma=D / 10.0,
                     bottom=WallSide(60.0, 0.02), top=WallSide(60.0, 0.02), bet=1.0, steps=100,
                     init="cap", params=dict(r=0.25, cx=0.5, cz=0.5))


def couette_stripe() -> RunConfig:
    """256x64x128 (Lx=4, Ly=1, Lz=2), walls moving at -1/+1 in x, a liquid
    stripe |x-2|<1 with a seeded sinusoidal interface perturbation of 0.5 dx,
    90 deg, ratios 1, Re 10, Ca 0.1, slip 0.05, dt 5e-4, 2000 steps."""
    L = 2.0
    rmu = L / 10.0
    return RunConfig("couette_stripe", 256, 128, 64, 4.0, 2.0, 1.0, dt=5e-4, rhol=1.0, rhog=1.0,
                     rmul=rmu, rmug=rmu, sigma=rmu * 1.0 / 0.1,
                     bottom=WallSide(90.0, 0.05), top=WallSide(90.0, 0.05), uwall=1.0, steps=2000,
                     init="stripe", params=dict(x0=1.0, x1=3.0, amp=0.5, seed=0))


def droplet_impact() -> RunConfig:
    """512x256x128 (Lx=4, Ly=2, Lz=1): droplet of radius 0.2 at (2,1,0.4)
    moving at w=-1 onto a GNBC wall at 100 deg (slip 0.02), free-slip top
    (infinite slip, 90 deg), density ratio 1000, viscosity ratio 100, Re 500,
    We 50, Bond 1.  dt is not given; 1e-4 (CFL 0.013 at the impact speed)."""
    D = 0.4
    rmul = D / 500.0
    sigma = D / 50.0
    g = 1.0 * sigma / D ** 2
    return RunConfig("droplet_impact", 512, 128, 256, 4.0, 1.0, 2.0, dt=1e-4, rhol=1.0, rhog=1e-3,
                     rmul=rmul, rmug=rmul / 100.0, sigma=sigma,
                     bottom=WallSide(100.0, 0.02), top=WallSide(90.0, 1e30), grvp=-g, steps=1000,
                     init="sphere", params=dict(r=0.2, cx=2.0, cy=0.4, cz=1.0, vy=-1.0))


def patterned_wall() -> RunConfig:
    """128x128x64 (Lx=Ly=2, Lz=1): a cap of radius 0.4 at (1,1,0) on a
    bottom wall checkered in 16-cell patches of 60/120 deg (64 patches),
    ratios 10, Re 50, We 5, dt 2e-4, 400 steps."""
    D = 0.8
    rmul = D / 50.0
    return RunConfig("patterned_wall", 128, 64, 128, 2.0, 1.0, 2.0, dt=2e-4, rhol=1.0, rhog=0.1,
                     rmul=rmul, rmug=rmul / 10.0, sigma=D / 5.0,
                     bottom=WallSide(60.0, 0.02, theta_alt=120.0, patch=16),
                     top=WallSide(90.0, 0.02), steps=400,
                     init="cap", params=dict(r=0.4, cx=1.0, cz=1.0))


def reference_walls(rmul: float, rmug: float, l1_a: float, theta_0_a: float, l1_b: float,
                    theta_0_b: float, period: int = 1, ratio_a: int = 1) -> dict:
    """The wall maps srcs/main.f90 builds with init_array_x_stripe
    (:169-186, :207-210): a uniform bottom wall (theta_0_a, l1_a = l2_a)
    and a top wall of x-stripes alternating the "b" side (theta_0_b, l1_b,
    l2_b = l1_b*3.67/1.625) and the "c" side (the mirrored angle and the
    swapped slip lengths), each with zeta = 0.21*6*(mul/l1 + mug/l2)."""
    l2_a = l1_a
    zeta_a = 0.21 * 6.0 * (rmul / l1_a + rmug / l2_a)
    l2_b = l1_b * 3.67 / 1.625
    zeta_b = 0.21 * 6.0 * (rmul / l1_b + rmug / l2_b)
    l1_c, l2_c = l2_b, l1_b
    theta_0_c = 180.0 - theta_0_b
    zeta_c = zeta_b
    return dict(theta=(theta_0_a, theta_0_b, theta_0_c), l1=(l1_a, l1_b, l1_c),
                l2=(l2_a, l2_b, l2_c), zeta=(zeta_a, zeta_b, zeta_c), period=period,
                ratio_a=ratio_a)


def with_reference_walls(cfg: RunConfig) -> RunConfig:
    """cfg with the wall maps the reference's program builds from its bottom
    and top wall parameters (reference_walls): what it runs when it restarts
    from a record of cfg's initial state."""
    walls = reference_walls(cfg.rmul, cfg.rmug, cfg.bottom.slip, cfg.bottom.theta0,
                            cfg.top.slip, cfg.top.theta0)
    return replace(cfg, params=dict(cfg.params, x_stripe=walls))


def reference_program(name: str = "reference_program", nsv: int = 8, svall=(32, 8, 32),
                      xl=17.0, yl=4.25, zl=17.0, rhol=8.1e-1, rhog=8.1e-1, rmul=1.95,
                      rmug=1.95, surface_tension=5.5, uwall=0.25, l1_a=2.165, theta_0_a=90.0,
                      l1_b=2.379, theta_0_b=64.0, period=1, ratio_a=1, grv=0.0, bet=2.0,
                      steps=100, imon_t=0) -> RunConfig:
    """The reference program's own configuration (srcs/main.f90:91-249):
    Couette flow between walls moving at -/+uwall past a stripe of the "g"
    phase (phi = 1 for nsv < i <= 3 nsv, srcs/init_phi_legendre.f90:15-28),
    a uniform bottom wall (theta_0_a, l1_a = l2_a) and a top wall of
    x-stripes alternating the "b" and "c" properties
    (srcs/init_array_x_stripe.f90:12-46).  Derived values are formed as
    main.f90 forms them: l2_b = l1_b*3.67/1.625, the "c" side swaps the
    slip lengths and mirrors the angle, zeta = 0.21*6*(mul/l1 + mug/l2)
    (:169-182), grvp = -grv cos(0) (:190-194) and dt = 0.32/nsv (:481).
    The defaults are the reference's physics on a 32x8x32 grid of cubic
    cells."""
    ni, nj, nk = svall
    stripes = reference_walls(rmul, rmug, l1_a, theta_0_a, l1_b, theta_0_b, period, ratio_a)
    dt = 32.0e-2 / nsv * 1
    return RunConfig(name, ni, nj, nk, xl, yl, zl, dt=dt, rhol=rhol, rhog=rhog, rmul=rmul,
                     rmug=rmug, sigma=surface_tension, bottom=WallSide(theta_0_a, l1_a),
                     top=WallSide(theta_0_b, l1_b), uwall=uwall, grvp=-grv * math.cos(0.0),
                     bet=bet, imon_t=imon_t, steps=steps, init="reference_stripe",
                     params=dict(nsv=nsv, x_stripe=stripes))


def seeded_droplets(n: int, seed: int, lx: float, lz: float, ly: float, rmin: float,
                    rmax: float, vstd: float, gap: float, max_tries: int = 100000) -> list:
    """n non-overlapping droplets in the lower half of the wall-normal
    extent (ref y in [0, ly/2]): radii uniform in [rmin, rmax], centres
    uniform with the sphere inside that half and `gap` clear of the wall,
    accepted when clear of every droplet placed before by `gap` (periodic
    distance in x and z); velocities normal with std vstd per component.
    Returns (x, y, z, r, u, v, w) tuples in reference axes."""
    rng = np.random.default_rng(seed)
    drops = []
    tries = 0
    while len(drops) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"placed {len(drops)} of {n} droplets in {max_tries} tries")
        r = rng.uniform(rmin, rmax)
        x, z = rng.uniform(0.0, lx), rng.uniform(0.0, lz)
        y = rng.uniform(r + gap, ly / 2 - r)
        ok = True
        for (x2, y2, z2, r2, *_) in drops:
            dx = abs(x - x2)
            dx = min(dx, lx - dx)
            dz = abs(z - z2)
            dz = min(dz, lz - dz)
            if dx * dx + (y - y2) ** 2 + dz * dz < (r + r2 + gap) ** 2:
                ok = False
                break
        if ok:
            u, v, w = rng.normal(0.0, vstd, 3)
            drops.append((x, y, z, r, u, v, w))
    return drops


def scale_run() -> RunConfig:
    """512x512x256 (Lx=Ly=4, Lz=2): 32 non-overlapping droplets (seed 0),
    radii uniform in [0.03, 0.08] Lx, centres uniform in the lower half with
    a rejection test, velocities normal with std 0.2 per component; bottom
    wall of x-stripes 60/100 deg with a 32-cell period (slip 0.02), free-slip
    top (infinite slip, 90 deg); density ratio 100, viscosity ratio 50,
    Re 1000, We 20 with L = the mean droplet diameter 0.44 and U = 1, no
    gravity; dt by the CFL condition of srcs/caldt.f (cfl 0.05), 500 steps."""
    lx, ly, lz = 4.0, 2.0, 4.0  # reference axes: x, wall-normal y, z
    D = 2.0 * 0.055 * lx
    rmul = D / 1000.0
    drops = seeded_droplets(32, 0, lx, lz, ly, 0.03 * lx, 0.08 * lx, 0.2, gap=4 * lx / 512)
    return RunConfig("scale_run", 512, 256, 512, lx, ly, lz, dt=0.0, rhol=1.0, rhog=0.01,
                     rmul=rmul, rmug=rmul / 50.0, sigma=D / 20.0,
                     bottom=WallSide(60.0, 0.02, theta_alt=100.0, stripe=32),
                     top=WallSide(90.0, 1e30), steps=500, init="droplets",
                     params=dict(droplets=drops), cfl=0.05)


CONFIGS = {
    "scale_run": scale_run,
    "sessile_droplet": sessile_droplet,
    "couette_stripe": couette_stripe,
    "droplet_impact": droplet_impact,
    "patterned_wall": patterned_wall,
    "reference_program": reference_program,
}


# ---------------------------------------------------------------------------
# Initial state
# ---------------------------------------------------------------------------

def _coords(cfg: RunConfig):
    i = np.arange(-2, cfg.ni + 4)
    j = np.arange(-2, cfg.nj + 4)
    k = np.arange(-2, cfg.nk + 4)
    xc, yc, zc = (i - 0.5) * cfg.dx, (j - 0.5) * cfg.dy, (k - 0.5) * cfg.dz
    return xc, yc, zc


def _tanh_profile(sdf, cfg: RunConfig):
    """phi = 1/(1+exp(-2 beta sdf/dcell)), sdf > 0 in the ambient phase,
    set to exactly 0 or 1 beyond |2 beta sdf/dcell| > 30 (the profile of
    srcs/init_phi.f with its truncation)."""
    dcell = (cfg.dx * cfg.dy * cfg.dz) ** (1.0 / 3.0)
    a = 2.0 * cfg.bet * sdf / dcell
    phi = 1.0 / (1.0 + np.exp(-np.clip(a, -60, 60)))
    phi[a > 30] = 1.0
    phi[a < -30] = 0.0
    return phi


def _wall_map(cfg: RunConfig, side: WallSide):
    """(nk+6, ni+6) static-angle map of one wall: uniform, or a checkerboard
    of side.patch-cell squares (periodic in the ghosts)."""
    if side.stripe > 0:
        # x-stripes: theta0 on the first half of each period, theta_alt on
        # the second (periodic in the ghosts)
        ii = ((np.arange(-2, cfg.ni + 4) - 1) % cfg.ni) % side.stripe < side.stripe // 2
        return np.broadcast_to(np.where(ii, side.theta0, side.theta_alt)[None, :],
                               (cfg.nk + 6, cfg.ni + 6)).copy()
    if side.patch <= 0:
        return np.full((cfg.nk + 6, cfg.ni + 6), float(side.theta0))
    ii = ((np.arange(-2, cfg.ni + 4) - 1) % cfg.ni) // side.patch
    kk = ((np.arange(-2, cfg.nk + 4) - 1) % cfg.nk) // side.patch
    return np.where(((kk[:, None] + ii[None, :]) % 2) == 0, side.theta0, side.theta_alt)


def x_stripe_map(cfg: RunConfig, qa: float, qb: float, qc: float, period: int,
                 ratio_a: int) -> np.ndarray:
    """srcs/init_array_x_stripe.f90:12-46 on the padded (nk+6, nj+6, ni+6)
    image: zero, q_a on the row j = 1, and on the row j = nj q_b where the
    parity is 0 and q_c elsewhere, the parity being mod(i+ni-1, 2) for
    period 1, else that of (ratio_a*kk + i+ni-1)/period with
    kk = (k+nk-1)/period (integer division; both numerators are positive)."""
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    i = np.arange(-2, ni + 4)[None, :]
    k = np.arange(-2, nk + 4)[:, None]
    if period == 1:
        parity = np.broadcast_to((i + ni - 1) % 2, (nk + 6, ni + 6))
    else:
        kk = (k + nk - 1) // period
        parity = ((ratio_a * kk + i + ni - 1) // period) % 2
    q = np.zeros(cfg.shape)
    q[:, 3, :] = qa
    q[:, nj + 2, :] = np.where(parity == 0, qb, qc)
    return q


def _droplets_state(cfg: RunConfig) -> dict[str, np.ndarray]:
    """Several droplets (phi = 0 inside, the tanh profile of _tanh_profile)
    with their own velocities on the faces whose liquid fraction exceeds one
    half; each droplet's profile is evaluated in its own padded box only
    (beyond 8 cells from its surface the profile is exactly 1)."""
    sh = cfg.shape
    ni, nj, nk = cfg.ni, cfg.nj, cfg.nk
    dcell = (cfg.dx * cfg.dy * cfg.dz) ** (1.0 / 3.0)
    phi = np.ones(sh)
    owner = np.full(sh, -1, dtype=np.int32)  # droplet of the cell (liquid fraction > 0)
    xi = (np.arange(-2, ni + 4) - 0.5) * cfg.dx
    yj = (np.arange(-2, nj + 4) - 0.5) * cfg.dy
    zk = (np.arange(-2, nk + 4) - 0.5) * cfg.dz
    drops = cfg.params["droplets"]
    for m, (x0, y0, z0, r, *_v) in enumerate(drops):
        reach = r + 8 * dcell
        # periodic images in x and z: distances measured to the nearest image
        def axis(c, lo_hi, n, L, d):
            idx = np.arange(n + 6)
            dist = np.abs(lo_hi - c)
            dist = np.minimum(dist, L - dist) if L else dist
            return idx[dist <= reach], dist[dist <= reach]
        ii, dxs = axis(x0, xi, ni, cfg.xl, cfg.dx)
        jj, dys = axis(y0, yj, nj, 0.0, cfg.dy)
        kk, dzs = axis(z0, zk, nk, cfg.zl, cfg.dz)
        sdf = np.sqrt(dzs[:, None, None] ** 2 + dys[None, :, None] ** 2 +
                      dxs[None, None, :] ** 2) - r
        a = 2.0 * cfg.bet * sdf / dcell
        prof = 1.0 / (1.0 + np.exp(-np.clip(a, -60, 60)))
        prof[a > 30] = 1.0
        prof[a < -30] = 0.0
        sub = np.ix_(kk, jj, ii)
        cur = phi[sub]
        mine = prof < cur
        phi[sub] = np.where(mine, prof, cur)
        o = owner[sub]
        owner[sub] = np.where(mine & (prof < 1.0), m, o)
    out_phi = np.zeros(sh)
    out_phi[:, 3:nj + 3, :] = phi[:, 3:nj + 3, :]
    u, v, w, p = (np.zeros(sh) for _ in range(4))
    vel = np.array([d[4:7] for d in drops] + [(0.0, 0.0, 0.0)])
    for comp, (arr, ax) in enumerate(((u, 2), (v, 1), (w, 0))):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        face = 0.5 * (out_phi[tuple(lo)] + out_phi[tuple(hi)])
        own = np.maximum(owner[tuple(lo)], owner[tuple(hi)])
        arr[tuple(lo)] = np.where(face < 0.5, vel[own, comp], 0.0)
    v[:, :3, :] = 0.0
    v[:, nj + 2:, :] = 0.0
    maps = {n: np.zeros(sh) for n in ("theta_0_array", "l1_array", "l2_array", "zeta_array")}
    for side, jrow in ((cfg.bottom, 1), (cfg.top, nj)):
        maps["theta_0_array"][:, jrow + 2, :] = _wall_map(cfg, side)
        maps["l1_array"][:, jrow + 2, :] = side.slip
        maps["l2_array"][:, jrow + 2, :] = side.slip
        maps["zeta_array"][:, jrow + 2, :] = zeta_of(cfg.rmul, cfg.rmug, side.slip, side.slip)
    out = dict(u=u, v=v, w=w, p=p, phi1=out_phi)
    out.update(maps)
    return out


def initial_state(cfg: RunConfig) -> dict[str, np.ndarray]:
    """Initial fields in the reference layout (nk+6, nj+6, ni+6): u, v, w,
    p, phi1 and the wall maps theta_0_array, l1_array, l2_array, zeta_array
    (set on the rows j=1 and j=nj only, as srcs/init_array_x_stripe.f90:12-46
    does).  phi1 is set on the rows j=1..nj over the whole padded i,k range
    (srcs/init_phi_legendre.f90:15-28)."""
    xc, yc, zc = _coords(cfg)
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    sh = cfg.shape
    u, v, w, p = (np.zeros(sh) for _ in range(4))
    phi = np.zeros(sh)
    q = cfg.params
    if cfg.init in ("sphere", "cap"):
        cy = q.get("cy", 0.0)
        d = np.sqrt((X - q["cx"]) ** 2 + (Y - cy) ** 2 + (Z - q["cz"]) ** 2) - q["r"]
        prof = _tanh_profile(d, cfg)
    elif cfg.init == "stripe":
        rng = np.random.default_rng(q.get("seed", 0))
        ph = rng.uniform(0, 2 * np.pi, 2)
        amp = q["amp"] * cfg.dx
        pert = amp * np.sin(2 * np.pi * Z / cfg.zl + ph[0]) * np.sin(np.pi * Y / cfg.yl + ph[1])
        d = np.maximum(q["x0"] - (X + pert), (X + pert) - q["x1"])
        prof = _tanh_profile(d, cfg)
    elif cfg.init == "droplets":
        return _droplets_state(cfg)
    elif cfg.init == "reference_stripe":
        # srcs/init_phi_legendre.f90:15-28: phi = 1 for nsv < i <= 3 nsv on
        # the rows j = 1..nj over the padded i and k ranges
        nsv = q["nsv"]
        i = np.arange(-2, cfg.ni + 4)
        prof = np.broadcast_to(np.where((i > nsv) & (i <= 3 * nsv), 1.0, 0.0), sh)
    else:
        raise ValueError(cfg.init)
    phi[:, 3:cfg.nj + 3, :] = prof[:, 3:cfg.nj + 3, :]
    if "vy" in q:
        # v lives on the j+1/2 faces: liquid fraction of the face
        face = 0.5 * (phi[:, :-1, :] + phi[:, 1:, :])
        v[:, :-1, :] = np.where(face < 0.5, q["vy"], 0.0)
        v[:, :3, :] = 0.0
        v[:, cfg.nj + 2:, :] = 0.0
    maps = {n: np.zeros(sh) for n in ("theta_0_array", "l1_array", "l2_array", "zeta_array")}
    if "x_stripe" in q:
        xs = q["x_stripe"]
        for n, key in (("theta_0_array", "theta"), ("l1_array", "l1"), ("l2_array", "l2"),
                       ("zeta_array", "zeta")):
            maps[n] = x_stripe_map(cfg, *xs[key], xs["period"], xs["ratio_a"])
        out = dict(u=u, v=v, w=w, p=p, phi1=phi)
        out.update(maps)
        return out
    for side, jrow in ((cfg.bottom, 1), (cfg.top, cfg.nj)):
        th = _wall_map(cfg, side)
        maps["theta_0_array"][:, jrow + 2, :] = th
        maps["l1_array"][:, jrow + 2, :] = side.slip
        maps["l2_array"][:, jrow + 2, :] = side.slip
        maps["zeta_array"][:, jrow + 2, :] = zeta_of(cfg.rmul, cfg.rmug, side.slip, side.slip)
    out = dict(u=u, v=v, w=w, p=p, phi1=phi)
    out.update(maps)
    return out
