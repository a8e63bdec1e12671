"""The C-ABI library loads and exports every entry point include/phasefield_amd.h declares, and
rejects bad descriptors before touching a device (CPU only)."""
from __future__ import annotations

import ctypes as C
import os
import re

import pytest

from phasefield_amd import _capi

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "phasefield_amd.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", txt, re.M)
    return sorted({n for n in names if n not in ("if", "sizeof", "defined")})


def test_header_declares_expected_surface():
    names = declared_functions()
    for must in ("pf_create", "pf_step", "pf_upload_state", "pf_download_state", "pf_reduce",
                 "pf_adjoint_run", "pf_model_read", "GPU_Init", "GPU_Diffusion_solver",
                 "GPU_Cleanup"):
        assert must in names
    assert len(names) >= 40


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
