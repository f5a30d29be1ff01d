"""fel-multi-phase-flow on MI355X: the reference's per-timestep solver as
gfx950 kernels behind its own Fortran subroutine interface.

Product entry points:
  felmpf.GpuRun      -- Fortran host driver + GPU library, reference step order
  felmpf.config      -- run configurations and initial states
  felmpf.abi         -- the C ABI (libfelgpu.so) for direct calls
"""
from . import config
from .config import CONFIGS, RunConfig, initial_state
from .driver import GpuRun

__all__ = ["config", "CONFIGS", "RunConfig", "initial_state", "GpuRun"]
