"""fel-multi-phase-flow on MI355X: the reference's per-timestep solver as
gfx950 kernels behind its own Fortran subroutine interface.

Product entry points:
  felmpf.GpuRun      -- Fortran host driver + GPU library, reference step order
  felmpf.config      -- run configurations and initial states
  felmpf._lib        -- the native libraries (C ABI of libfelgpu.so, Fortran driver)
  felmpf.sensitivity -- tangent-mode seeds, droplet observables, autograd window
  felmpf.longrun     -- long mixed-precision comparisons (python -m felmpf.longrun)
"""
from . import config
from .config import CONFIGS, RunConfig, initial_state
from .driver import GpuRun

__all__ = ["config", "CONFIGS", "RunConfig", "initial_state", "GpuRun"]
