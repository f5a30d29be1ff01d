"""The drop-in boundary without a GPU: include/felgpu.h parses as C, the
built libfelgpu.so loads and exports exactly the entry points the header
declares, and each reference subroutine on the timestep path has its
drop-in.  No entry point is called (every one of them would touch the GPU).
"""
from __future__ import annotations

import ctypes
import os
import re
import shutil
import subprocess

import pytest

from felharness import GPU_SO, REPO

HEADER = os.path.join(REPO, "include", "felgpu.h")
DRV_SO = os.path.join(REPO, "fel-multi-phase-flow_amd", "libfeldrv.so")

# Reference subroutines called by the timestep (srcs/main.f90:467-748) and
# its start-up (srcs/main.f90:303-422); each one has a drop-in `name_`.
REFERENCE_PATH = """
cal_grad_p2a solphi_mthinc1 solphi_mthinc2 solphi_mthinc3 summation
cal_advu_weno5 cal_arith_mean calphat cal_grad_prs calsij cal_arith_tau
cal_div_tensor init_q cal_grad_p2uvw calrn calrkap calfst sum_fst cal_av
bnd_avzero cal_srcu cal_arith_coef_vis cpy bnd_periodic bnd_neumann bnd_comm
bndu solu_sor4 caldiv solp_fft_tdma1 solp_fft_tdma2 solp_fft_tdma3
solp_fft_tdma4 trans_r2w trans_w2r corunp_explicit cal_coef_prs
cal_coef_solp_fft gnbc bnd_contact_angle cal_center cal_vel
""".split()


def declared() -> list[str]:
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", text, flags=re.M)


def exported(path: str) -> set[str]:
    out = subprocess.run(["nm", "-D", "--defined-only", path], check=True, capture_output=True,
                         text=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def need(path: str):
    if not os.path.exists(path):
        pytest.skip(f"{os.path.basename(path)} is not built (run __graft_entry__.build())")


def test_header_declares_every_reference_routine():
    names = set(declared())
    missing = [r for r in REFERENCE_PATH if r + "_" not in names]
    assert not missing, f"no drop-in declared for {missing}"
    assert len(names) == len(declared()), "duplicate declaration"


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_header_is_valid_c_and_cxx(tmp_path):
    src = tmp_path / "h.c"
    src.write_text('#include "felgpu.h"\nint main(void) { return 0; }\n')
    inc = os.path.join(REPO, "include")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I", inc, str(src)],
                   check=True)
    subprocess.run(["g++", "-x", "c++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I",
                    inc, str(src)], check=True)


def test_library_loads_and_exports_exactly_the_header():
    need(GPU_SO)
    lib = ctypes.CDLL(GPU_SO, mode=ctypes.RTLD_LOCAL)
    names = declared()
    for n in names:
        assert hasattr(lib, n), f"libfelgpu.so does not export {n}"
    extra = {s for s in exported(GPU_SO) if not s.startswith("_Z")} - set(names)
    assert not extra, f"exported but not declared in include/felgpu.h: {sorted(extra)}"


def test_driver_library_exports_host_api():
    need(DRV_SO)
    syms = exported(DRV_SO)
    for n in ("fel_drv_create", "fel_drv_field", "fel_drv_start", "fel_drv_steps",
              "fel_drv_download", "fel_drv_state", "fel_drv_destroy"):
        assert n in syms, f"libfeldrv.so does not export {n}"


def test_monitor_lines_use_the_reference_edit_descriptor():
    """The drop-ins print their monitor lines (Err_solu, Err_solp, caldt's
    step line) with Fortran's Ew.d edit descriptor, as the reference's
    write statements do (srcs/solu_sor4.f:299, srcs/solp_fft_tdma4.f:88,
    srcs/caldt.f:111).  Expected strings are what amdflang prints for
    write(*,'(e20.10)') and '(e17.10)' of the same values."""
    lib = ctypes.CDLL(GPU_SO)
    fn = lib._ZN3fel9fortran_eEPcdii  # fel::fortran_e(char*, double, int, int), a host function
    fn.restype = ctypes.c_char_p
    fn.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    buf = ctypes.create_string_buffer(64)
    cases = [(0.0, 20, "    0.0000000000E+00"), (-0.0, 20, "   -0.0000000000E+00"),
             (3.535572739e-3, 20, "    0.3535572739E-02"),
             (-1.0e-101, 20, "   -0.1000000000-100"), (1.0e100, 20, "    0.1000000000+101"),
             (123456.789, 17, " 0.1234567890E+06"), (0.99999999999, 20, "    0.1000000000E+01")]
    for v, w, want in cases:
        assert fn(buf, v, w, 10).decode() == want, (v, want)
