"""Test harness shared by the parity tests.

Loads the three implementations that share the reference's Fortran argument
lists -- the CPU oracle (oracle/liboracle.so, names orc_<sub>), the compiled
reference sources (oracle/_ref/libfelref.so, names <sub>_, built only where
the reference tree exists) and the GPU library (names <sub>_) -- and calls
them with identical actual arguments: integers and doubles by reference,
arrays as base addresses of padded (-2:n+3) column-major images, which are
numpy arrays of shape (nk+6, nj+6, ni+6) in C order.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libfelref.so")
GPU_SO = os.path.join(REPO, "fel-multi-phase-flow_amd", "libfelgpu.so")

_libs: dict[str, ctypes.CDLL] = {}


def load(path: str) -> ctypes.CDLL:
    if path not in _libs:
        _libs[path] = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    return _libs[path]


def oracle() -> ctypes.CDLL:
    return load(ORACLE_SO)


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref() -> ctypes.CDLL:
    return load(REF_SO)


@dataclass
class Grid:
    ni: int
    nj: int
    nk: int
    dx: float = 1.0 / 16
    dy: float = 1.0 / 16
    dz: float = 1.0 / 16

    @property
    def shape(self):
        return (self.nk + 6, self.nj + 6, self.ni + 6)

    @property
    def n3(self):
        return (self.ni + 6) * (self.nj + 6) * (self.nk + 6)

    def zeros(self, m: int | None = None):
        if m is None:
            return np.zeros(self.shape)
        return np.zeros((m,) + self.shape)

    def interior(self, a):
        return a[..., 3:self.nk + 3, 3:self.nj + 3, 3:self.ni + 3]

    def coords(self):
        """Cell-centre coordinates of every padded cell (x along i)."""
        i = np.arange(-2, self.ni + 4)
        j = np.arange(-2, self.nj + 4)
        k = np.arange(-2, self.nk + 4)
        x = (i - 0.5) * self.dx
        y = (j - 0.5) * self.dy
        z = (k - 0.5) * self.dz
        return np.meshgrid(z, y, x, indexing="ij")[::-1]


class Ref:
    """Boxed scalar passed by reference."""

    def __init__(self, value, ctype):
        self.obj = ctype(value)

    @property
    def value(self):
        return self.obj.value


def I(v: int):  # noqa: E743 - Fortran INTEGER by reference
    return Ref(int(v), ctypes.c_int)


def D(v: float):
    return Ref(float(v), ctypes.c_double)


def _arg(a):
    if isinstance(a, Ref) or hasattr(a, "obj"):  # boxed scalars (also felmpf.adjoint's)
        return ctypes.byref(a.obj)
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"], "arrays must be contiguous"
        if a.dtype == np.int32:
            return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        assert a.dtype == np.float64, a.dtype
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    if isinstance(a, int):  # raw address
        return ctypes.c_void_p(a)
    raise TypeError(type(a))


def call(lib: ctypes.CDLL, name: str, *args):
    fn = getattr(lib, name)
    fn.restype = None
    fn(*[_arg(a) for a in args])


def gpu() -> ctypes.CDLL:
    if not os.path.exists(GPU_SO):
        raise RuntimeError(f"{GPU_SO} is not built (run __graft_entry__.build())")
    return load(GPU_SO)


def call_gpu(name, *args):
    call(gpu(), name + "_", *args)


def call_oracle(name, *args):
    call(oracle(), "orc_" + name, *args)


def call_ref(name, *args):
    call(ref(), name + "_", *args)


# --------------------------------------------------------------------------
# Seeded synthetic fields
# --------------------------------------------------------------------------

def smooth_phi(g: Grid, rng, r=None, centre=None, beta=1.0):
    """A tanh-profile droplet (phi=0 inside, the reference's liquid) with a
    seeded random perturbation, so that both MTHINC branches occur."""
    x, y, z = g.coords()
    lx, ly, lz = g.ni * g.dx, g.nj * g.dy, g.nk * g.dz
    c = centre if centre is not None else (lx * rng.uniform(0.35, 0.65), ly * rng.uniform(0.3, 0.6),
                                           lz * rng.uniform(0.35, 0.65))
    rad = r if r is not None else 0.3 * min(lx, ly, lz)
    d = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - rad
    d = d + 0.2 * g.dx * rng.standard_normal(d.shape)
    dcell = (g.dx * g.dy * g.dz) ** (1.0 / 3.0)
    return 1.0 / (1.0 + np.exp(-2.0 * beta * d / dcell))


def random_vel(g: Grid, rng, scale=1.0):
    return scale * rng.standard_normal(g.shape)


def smooth_vel(g: Grid, rng, scale=1.0):
    x, y, z = g.coords()
    lx, ly, lz = g.ni * g.dx, g.nj * g.dy, g.nk * g.dz
    a = rng.uniform(-1, 1, 6)
    return scale * (a[0] * np.sin(2 * np.pi * x / lx + a[1]) * np.cos(np.pi * y / ly + a[2])
                    + a[3] * np.cos(2 * np.pi * z / lz + a[4]) + a[5] * 0.1)
