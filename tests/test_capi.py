"""The C-ABI library loads and exports every entry point include/phasefield_amd.h declares, and
rejects bad descriptors before touching a device (CPU only)."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from phasefield_amd import _capi
from tests import helpers

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "phasefield_amd.h")


def _header_parts():
    """(C-linkage text, C++-linkage text) of the header: the extern "C" block ends at the
    `#ifdef __cplusplus` / `}` that closes it; what follows is declared with C++ linkage."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    cut = txt.rindex("#ifdef __cplusplus\n}")
    return txt[:cut], txt[cut:]


def _names(txt):
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", txt, re.M)
    return sorted({n for n in names if n not in ("if", "sizeof", "defined")})


def declared_functions():
    return _names(_header_parts()[0])


def declared_cxx_functions():
    return _names(_header_parts()[1])


def test_header_declares_expected_surface():
    names = declared_functions()
    for must in ("pf_create", "pf_step", "pf_upload_state", "pf_download_state", "pf_reduce",
                 "pf_adjoint_run", "pf_model_read", "pf_gpu_init", "pf_gpu_diffusion_solver",
                 "pf_gpu_cleanup"):
        assert must in names
    assert len(names) >= 40
    assert declared_cxx_functions() == ["GPU_Cleanup", "GPU_Diffusion_solver", "GPU_Init"]


def test_library_exports_every_declared_symbol(capi):
    missing = [n for n in declared_functions() if not hasattr(capi, n)]
    assert missing == []
    # and the ctypes table covers the header exactly
    assert set(declared_functions()) == set(_capi.SIGNATURES)


@pytest.mark.parametrize("field,value", [("nx", 0), ("depth", 5), ("nphase", 4), ("nsolute", 3),
                                         ("precision", 7), ("bc_z", 2)])
def test_create_rejects_bad_descriptors(capi, field, value):
    d = _capi.PfDesc(8, 8, 8, 4, 2, 1, 0, 0, 0)
    setattr(d, field, value)
    h = C.c_void_p()
    assert capi.pf_create(C.byref(d), C.byref(h)) == _capi.PF_ERR_INVALID
    assert not h.value
    assert capi.pf_last_error(None)


def test_create_rejects_1d_and_2d_noflux(capi):
    h = C.c_void_p()
    d = _capi.PfDesc(16, 1, 1, 4, 2, 1, 0, 0, 0)
    assert capi.pf_create(C.byref(d), C.byref(h)) == _capi.PF_ERR_INVALID
    d = _capi.PfDesc(16, 16, 1, 4, 2, 1, 0, _capi.PF_BC_NOFLUX, 0)
    assert capi.pf_create(C.byref(d), C.byref(h)) == _capi.PF_ERR_INVALID


def test_null_arguments(capi):
    assert capi.pf_create(None, None) == _capi.PF_ERR_INVALID
    assert capi.pf_step(None, 1) == _capi.PF_ERR_INVALID
    assert capi.pf_destroy(None) == 0


def test_library_exports_the_reference_cxx_names():
    """GPU_Init / GPU_Diffusion_solver / GPU_Cleanup with C++ linkage: the mangled names of the
    prototypes in the reference's gpu_solver.h:18-23 (real = double)."""
    out = subprocess.run(["nm", "-D", "--defined-only", helpers.LIB_SO], capture_output=True,
                         text=True, check=True).stdout
    syms = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    for m in ("_Z8GPU_Initiii", "_Z20GPU_Diffusion_solverPdS_S_S_S_iiidd", "_Z11GPU_Cleanupv"):
        assert m in syms, m


def test_reference_prototypes_link_and_fall_back_without_a_device(tmp_path):
    """A C++ translation unit that declares the three functions exactly as the reference's
    gpu_solver.h does (no extern "C"), as phase_field.cpp sees them, links against libpfamd.so
    with g++ and runs; without a GPU, GPU_Init returns 0 (the reference's CPU fallback).  A unit
    that includes both that header text and phasefield_amd.h compiles (no linkage conflict)."""
    exe = helpers.build_reference_shim_program(str(tmp_path))
    helpers.write_shim_input(str(tmp_path / "in.bin"), 8, 8, 1, 1e-3, 0.1, [0.5, 100.0, 0.0],
                             np.ones(9), np.zeros(9), np.tile([0.0, 1.0, 0.0], 64),
                             np.full(64, 0.02))
    res = subprocess.run([exe, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                         capture_output=True, text=True, timeout=120)
    if res.returncode == 0 and "GPU_Init=1" in res.stdout:
        pytest.skip("a GPU is present; the numeric run is tests/test_gpu_parity.py")
    assert res.returncode == 3, res.stdout + res.stderr
    assert "GPU_Init=0" in res.stdout
    both = tmp_path / "both.cpp"
    both.write_text(helpers.REFERENCE_GPU_SOLVER_PROTOTYPES + '#include "phasefield_amd.h"\n'
                    "int main() { return GPU_Init(1, 1, 1); }\n")
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", helpers.INCLUDE, str(both)],
                   check=True)
