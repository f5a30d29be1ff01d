"""Loads the in-tree native libraries.  There is no fallback: if the HIP
library or the Fortran host driver has not been built, every entry point
raises."""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GPU_SO = os.path.join(PKG, "libfelgpu.so")
DRV_SO = os.path.join(PKG, "libfeldrv.so")

_cache: dict[str, ctypes.CDLL] = {}


class NativeLibraryMissing(RuntimeError):
    pass


def _load(path: str, mode=ctypes.RTLD_GLOBAL) -> ctypes.CDLL:
    if path not in _cache:
        if not os.path.exists(path):
            raise NativeLibraryMissing(
                f"{path} is not built; run __graft_entry__.build() (make -C fel-multi-phase-flow_amd)")
        _cache[path] = ctypes.CDLL(path, mode=mode)
    return _cache[path]


def gpu() -> ctypes.CDLL:
    return _load(GPU_SO)


def driver() -> ctypes.CDLL:
    gpu()
    return _load(DRV_SO)
