"""The GPU timestep driven by the Fortran host (host/fel_driver.f90).

GpuRun(cfg) allocates the reference-layout host arrays in the Fortran
driver, lets the caller store the initial state, makes everything
device-resident and runs timesteps; fields are copied back on request.

precision=64 (default) is the FP64 path, whose every result equals the
reference's FP64 arithmetic; precision=32 is the library's FP32 mode (flow
fields, stencils, Gauss-Seidel and FFT+TDMA in FP32; volume-fraction
fluxes, mass sums, contact-line velocity and the pressure correction in
FP64 with iterative refinement).  The mode is process-wide while the run
is open and returns to FP64 when it closes.

tangent=True is the library's tangent mode (FP64 values and their
derivative along one direction of the inputs): after construction, seed()
sets the direction on input arrays (the wall maps theta_0_array, l1_array,
l2_array, zeta_array, or any field), steps() advances values and
derivatives together, and tangent() returns a field's derivative.  Values
are bitwise those of the FP64 run.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .config import RunConfig, initial_state


class GpuRun:
    # registered arrays that hold several slices: name -> (base name, slice)
    def __init__(self, cfg: RunConfig, state: dict[str, np.ndarray] | None = None,
                 precision: int = 64, tangent: bool = False, reuse: bool = True,
                 fp32_modes: bool = False):
        if precision not in (32, 64):
            raise ValueError(f"precision must be 32 or 64, not {precision}")
        if tangent and precision != 64:
            raise ValueError("the tangent mode is FP64")
        self.cfg = cfg
        self.precision = precision
        self.tangent_mode = bool(tangent)
        self.reuse = bool(reuse)
        self.drv = _lib.driver()
        self.gpu = _lib.gpu()
        self.drv.fel_drv_field.restype = ctypes.c_void_p
        self.gpu.felgpu_result_reuse_hits_.restype = ctypes.c_int64
        self.drv.fel_drv_field.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]
        self.drv.fel_drv_steps.argtypes = [ctypes.c_int]
        self._c = cfg.cstruct()
        self.drv.fel_drv_create(ctypes.byref(self._c))
        st = initial_state(cfg) if state is None else state
        for name, arr in st.items():
            self.view(name)[...] = arr
        self.gpu.felgpu_set_precision_(ctypes.byref(ctypes.c_int(precision)))
        self.gpu.felgpu_set_tangent_(ctypes.byref(ctypes.c_int(1 if tangent else 0)))
        # results already computed are taken over instead of recomputed
        # (bitwise the same either way); the library's setting before this
        # run is restored when it closes
        self.gpu.felgpu_result_reuse_.restype = ctypes.c_int
        self._reuse_before = int(self.gpu.felgpu_result_reuse_())
        self.gpu.felgpu_set_result_reuse_(ctypes.byref(ctypes.c_int(1 if reuse else 0)))
        # pressure-solve modes: FP64 (default), or the flow class (FP32 solve
        # with FP64 refinement in the FP32 mode)
        self.drv.fel_drv_set_modes_flow_class(ctypes.c_int(1 if fp32_modes else 0))
        self.drv.fel_drv_start()

    def view(self, name: str) -> np.ndarray:
        """Host array of the driver (reference layout, (nk+6, nj+6, ni+6))."""
        n = ctypes.c_int64(0)
        ptr = self.drv.fel_drv_field(name.encode(), ctypes.byref(n))
        if not ptr:
            raise KeyError(name)
        buf = (ctypes.c_double * n.value).from_address(ptr)
        a = np.ctypeslib.as_array(buf)
        per = int(np.prod(self.cfg.shape))
        if n.value == per:
            return a.reshape(self.cfg.shape)
        return a.reshape((n.value // per,) + tuple(self.cfg.shape))  # vector fields (rn, rkap)

    def field(self, name: str) -> np.ndarray:
        """Device copy of a field, as a new host array."""
        self.drv.fel_drv_download(name.encode())
        return self.view(name).copy()

    def _range(self, name: str):
        """(host pointer, element count) of the registered range `name`
        (each slice of phi and phin is a range of its own)."""
        v = self.view(name)
        return v.ctypes.data, v.size

    def upload(self, name: str):
        """Host -> device copy after the caller changed view(name)."""
        self.gpu.felgpu_upload_(ctypes.c_void_p(self._range(name)[0]))

    def seed(self, name: str, direction: np.ndarray):
        """Tangent mode: sets the derivative part of field `name` to
        `direction` (same shape as the field); the values are kept."""
        if not self.tangent_mode:
            raise RuntimeError("seed() needs tangent=True")
        ptr, n = self._range(name)
        full = np.ascontiguousarray(np.asarray(direction, dtype=np.float64).ravel())
        assert full.size == n, (name, full.size, n)
        self.gpu.felgpu_tangent_seed_(ctypes.c_void_p(ptr), full.ctypes.data_as(ctypes.c_void_p))

    def _tangent_full(self, ptr, n):
        out = np.empty(n)
        self.gpu.felgpu_tangent_fetch_(ctypes.c_void_p(ptr), out.ctypes.data_as(ctypes.c_void_p))
        return out

    def tangent(self, name: str) -> np.ndarray:
        """Tangent mode: the derivative part of field `name`."""
        if not self.tangent_mode:
            raise RuntimeError("tangent() needs tangent=True")
        ptr, n = self._range(name)
        return self._tangent_full(ptr, n).reshape(self.cfg.shape)

    def dataou(self, path: str):
        """Writes a checkpoint in the reference's restart format
        (srcs/dataou.f: u, v, w, p, uo, vo, wo, po, phi, time, nstep)."""
        b = path.encode()
        self.drv.fel_drv_dataou(ctypes.c_char_p(b), ctypes.c_int(len(b)))

    def datain(self, path: str):
        """Restarts this (started) run from a checkpoint file, with the
        reference's restart ghost fills (srcs/main.f90:375-404)."""
        b = path.encode()
        self.drv.fel_drv_datain(ctypes.c_char_p(b), ctypes.c_int(len(b)))

    def output(self, what: str):
        """One output cadence of srcs/main.f90:663-743 for the current step,
        written to the current directory: "vtk" (phi_<n>.vtk, p_<n>.vtk and
        the interface-position monitor lines), "uwphi" (the j-slice file
        uwphi<n>) or "center_vel" (a line of center_01.dat and
        velocity_01.dat)."""
        fn = {"vtk": self.drv.fel_drv_vtk, "uwphi": self.drv.fel_drv_uwphi,
              "center_vel": self.drv.fel_drv_center_vel}[what]
        fn()

    def steps(self, n: int):
        self.drv.fel_drv_steps(int(n))

    def sync(self):
        self.gpu.felgpu_sync_()

    def state(self) -> dict:
        out = (ctypes.c_double * 13)()
        self.drv.fel_drv_state(out)
        return dict(nstep=int(out[0]), time=out[1], sor_lp=int(out[2]), sor_ratio=out[3],
                    err_div0=out[4], center=tuple(out[5:8]), velocity=out[8], vel=tuple(out[9:12]),
                    dt=out[12])

    def close(self):
        if self.drv is not None:
            self.drv.fel_drv_destroy()
            self.drv = None
            if self.tangent_mode:
                self.gpu.felgpu_set_tangent_(ctypes.byref(ctypes.c_int(0)))
            if self.precision != 64:
                self.gpu.felgpu_set_precision_(ctypes.byref(ctypes.c_int(64)))
            # back to the library's setting before this run
            self.gpu.felgpu_set_result_reuse_(ctypes.byref(ctypes.c_int(self._reuse_before)))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
