"""The oracle's timestep (oracle/fel_oracle_step.c) with the same interface
as felmpf.GpuRun, for whole-step comparisons.  Test infrastructure."""
from __future__ import annotations

import ctypes

import numpy as np

import felharness as H
from felmpf.config import RunConfig, initial_state


class OracleRun:
    def __init__(self, cfg: RunConfig, state=None):
        self.cfg = cfg
        self.lib = H.oracle()
        self.lib.orc_state_create.restype = ctypes.c_void_p
        self.lib.orc_state_field.restype = ctypes.c_void_p
        self.lib.orc_state_field.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        self.lib.orc_state_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        for f in ("orc_state_startup", "orc_state_destroy"):
            getattr(self.lib, f).argtypes = [ctypes.c_void_p]
        self.lib.orc_state_last_sor_lp.argtypes = [ctypes.c_void_p]
        self.lib.orc_state_last_sor_ratio.argtypes = [ctypes.c_void_p]
        self.lib.orc_state_last_sor_ratio.restype = ctypes.c_double
        self.lib.orc_state_diag.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        self.lib.orc_state_dt.argtypes = [ctypes.c_void_p]
        self.lib.orc_state_dt.restype = ctypes.c_double
        self._c = cfg.cstruct()
        self.h = self.lib.orc_state_create(ctypes.byref(self._c))
        st = initial_state(cfg) if state is None else state
        for name, arr in st.items():
            self.view(name)[...] = arr
        self.lib.orc_state_startup(self.h)

    def view(self, name):
        ptr = self.lib.orc_state_field(self.h, name.encode())
        if not ptr:
            raise KeyError(name)
        n = int(np.prod(self.cfg.shape))
        buf = (ctypes.c_double * n).from_address(ptr)
        return np.ctypeslib.as_array(buf).reshape(self.cfg.shape)

    def field(self, name):
        return self.view(name).copy()

    def steps(self, n):
        self.lib.orc_state_step(self.h, int(n))

    def datain(self, path: str):
        """Restart from a checkpoint file (srcs/main.f90:375-404)."""
        self.lib.orc_state_datain.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        if self.lib.orc_state_datain(self.h, path.encode()) != 0:
            raise ValueError(f"{path}: not a checkpoint of this grid")

    def state(self):
        d = (ctypes.c_double * 8)()
        self.lib.orc_state_diag(self.h, d)
        return dict(sor_lp=self.lib.orc_state_last_sor_lp(self.h),
                    sor_ratio=self.lib.orc_state_last_sor_ratio(self.h),
                    center=tuple(d[0:3]), velocity=d[3], vel=tuple(d[4:7]), time=d[7],
                    dt=self.lib.orc_state_dt(self.h))

    def close(self):
        if self.h:
            self.lib.orc_state_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rel_err(a, b):
    """max |a-b| / max |a| (max-norm relative error)."""
    s = np.max(np.abs(a))
    return float(np.max(np.abs(a - b)) / (s if s > 0 else 1.0))
