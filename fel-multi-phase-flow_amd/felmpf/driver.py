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
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .config import RunConfig, initial_state


class GpuRun:
    def __init__(self, cfg: RunConfig, state: dict[str, np.ndarray] | None = None,
                 precision: int = 64):
        if precision not in (32, 64):
            raise ValueError(f"precision must be 32 or 64, not {precision}")
        self.cfg = cfg
        self.precision = precision
        self.drv = _lib.driver()
        self.gpu = _lib.gpu()
        self.drv.fel_drv_field.restype = ctypes.c_void_p
        self.drv.fel_drv_field.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]
        self.drv.fel_drv_steps.argtypes = [ctypes.c_int]
        self._c = cfg.cstruct()
        self.drv.fel_drv_create(ctypes.byref(self._c))
        st = initial_state(cfg) if state is None else state
        for name, arr in st.items():
            self.view(name)[...] = arr
        self.gpu.felgpu_set_precision_(ctypes.byref(ctypes.c_int(precision)))
        self.drv.fel_drv_start()

    def view(self, name: str) -> np.ndarray:
        """Host array of the driver (reference layout, (nk+6, nj+6, ni+6))."""
        n = ctypes.c_int64(0)
        ptr = self.drv.fel_drv_field(name.encode(), ctypes.byref(n))
        if not ptr:
            raise KeyError(name)
        buf = (ctypes.c_double * n.value).from_address(ptr)
        return np.ctypeslib.as_array(buf).reshape(self.cfg.shape)

    def field(self, name: str) -> np.ndarray:
        """Device copy of a field, as a new host array."""
        self.drv.fel_drv_download(name.encode())
        return self.view(name).copy()

    def steps(self, n: int):
        self.drv.fel_drv_steps(int(n))

    def sync(self):
        self.gpu.felgpu_sync_()

    def state(self) -> dict:
        out = (ctypes.c_double * 12)()
        self.drv.fel_drv_state(out)
        return dict(nstep=int(out[0]), time=out[1], sor_lp=int(out[2]), sor_ratio=out[3],
                    err_div0=out[4], center=tuple(out[5:8]), velocity=out[8], vel=tuple(out[9:12]))

    def close(self):
        if self.drv is not None:
            self.drv.fel_drv_destroy()
            self.drv = None
            if self.precision != 64:
                self.gpu.felgpu_set_precision_(ctypes.byref(ctypes.c_int(64)))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
