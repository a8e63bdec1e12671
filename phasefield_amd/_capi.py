"""ctypes view of the C-ABI in include/phasefield_amd.h (libpfamd.so, built in-tree).

Loading fails loudly when the library has not been built: there is no Python or CPU fallback
for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpfamd.so")

PF_OK = 0
PF_ERR_INVALID = -1
PF_ERR_NODEVICE = -2
PF_ERR_OOM = -3
PF_ERR_HIP = -4
PF_ERR_NONFINITE = -5
PF_ERR_STATE = -6
PF_ERR_IO = -7
PF_ERR_RANGE = -8

PF_FP64 = 0
PF_MIXED = 1
PF_BC_PERIODIC = 0
PF_BC_NOFLUX = 1

_D = C.c_double
_I = C.c_int


class PfDesc(C.Structure):
    _fields_ = [("nx", _I), ("ny", _I), ("nz", _I), ("depth", _I), ("nphase", _I),
                ("nsolute", _I), ("precision", _I), ("bc_z", _I), ("device", _I)]


class PfParams(C.Structure):
    _fields_ = [
        ("dx", _D), ("eta", _D), ("cutoff", _D), ("dims", _I), ("reserved0", _I),
        ("mobility", _D * 3 * 3), ("sigma", _D * 3 * 3), ("entropy", _D * 3 * 3),
        ("eq_conc", _D * 3 * 3), ("eq_slope", _D * 3 * 3),
        ("part_k", _D * 3 * 3 * 2), ("part_c", _D * 3 * 3 * 2), ("diffusivity", _D * 3 * 2),
        ("pfm_dt", _D), ("dt_d", _D * 2), ("n_sub", _I * 2),
    ]


class PfHostState(C.Structure):
    _fields_ = [("index", C.POINTER(C.c_int32)), ("value", C.POINTER(_D)),
                ("phase", C.POINTER(_D)), ("conc", C.POINTER(_D)),
                ("conc_comp", C.POINTER(_D)), ("phase_conc", C.POINTER(_D))]


class PfStats(C.Structure):
    _fields_ = [("mass", _D * 2), ("phase_frac", _D * 3), ("max_conc", _D * 2),
                ("min_conc", _D * 2), ("max_sum_err", _D), ("n_interface", C.c_int64),
                ("n_nonfinite", C.c_int64)]


class PfAdjointSpec(C.Structure):
    _fields_ = [("n_steps", _I), ("checkpoint_every", _I), ("n_samples", _I),
                ("sample_steps", _I * 64), ("target_fraction", _D * 64),
                ("weight_conc", _D), ("target_conc", C.POINTER(_D))]


class PfAdjointResult(C.Structure):
    _fields_ = [("loss", _D), ("d_mobility", _D), ("d_sigma", _D), ("d_diff_gamma", _D),
                ("d_conc0", C.POINTER(_D)), ("fraction", _D * 64), ("forward_ms", _D),
                ("reverse_ms", _D)]


_H = C.c_void_p  # pf_handle*
_M = C.c_void_p  # pf_model*

# name -> (restype, argtypes); every symbol here is declared in include/phasefield_amd.h
SIGNATURES = {
    "pf_create": (_I, [C.POINTER(PfDesc), C.POINTER(_H)]),
    "pf_destroy": (_I, [_H]),
    "pf_last_error": (C.c_char_p, [_H]),
    "pf_set_params": (_I, [_H, C.POINTER(PfParams)]),
    "pf_upload_state": (_I, [_H, C.POINTER(PfHostState)]),
    "pf_download_state": (_I, [_H, C.POINTER(PfHostState)]),
    "pf_step": (_I, [_H, _I]),
    "pf_phase_field_sweep": (_I, [_H]),
    "pf_diffuse_substeps": (_I, [_H, _I, _I, _D]),
    "pf_reduce": (_I, [_H, C.POINTER(PfStats)]),
    "pf_synchronize": (_I, [_H]),
    "pf_dump_vti_async": (_I, [_H, C.c_char_p, _D]),
    "pf_dump_wait": (_I, [_H]),
    "pf_kernel_timing": (_I, [_H, _I, C.POINTER(_D), C.POINTER(C.c_int64)]),
    "pf_kernel_timing_reset": (_I, [_H, _I]),
    "pf_list_lengths": (_I, [_H, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "pf_march_shape": (_I, [_H, C.POINTER(_I), C.POINTER(_I)]),
    "pf_checkpoint_size": (_I, [_H, C.POINTER(C.c_size_t)]),
    "pf_checkpoint_save": (_I, [_H, C.c_void_p]),
    "pf_checkpoint_restore": (_I, [_H, C.c_void_p]),
    "pf_setup_construct": (_I, [_H, C.POINTER(_D), _I, _D, _I, C.POINTER(_D)]),
    "pf_setup_nucleate": (_I, [_H, _I, C.c_float, C.POINTER(_D), C.c_uint64]),
    "pf_setup_solute_noise": (_I, [_H, _I, _D, _D, C.c_uint64]),
    "pf_copy_state": (_I, [_H, _H]),
    "pf_rex_fraction": (_I, [_H, C.POINTER(_D)]),
    "pf_adjoint_run": (_I, [_H, C.POINTER(PfAdjointSpec), C.POINTER(PfAdjointResult)]),
    "pf_model_read": (_I, [C.c_char_p, C.POINTER(_M)]),
    "pf_model_free": (_I, [_M]),
    "pf_model_last_error": (C.c_char_p, []),
    "pf_model_get_int": (_I, [_M, C.c_char_p, C.POINTER(C.c_int64)]),
    "pf_model_get_real": (_I, [_M, C.c_char_p, C.POINTER(_D)]),
    "pf_model_get_string": (_I, [_M, C.c_char_p, C.c_char_p, C.c_size_t]),
    "pf_model_set_real": (_I, [_M, C.c_char_p, _D]),
    "pf_model_set_int": (_I, [_M, C.c_char_p, C.c_int64]),
    "pf_model_desc": (_I, [_M, C.POINTER(PfDesc)]),
    "pf_model_update_parameter": (_I, [_M, C.POINTER(PfParams)]),
    "pf_model_temperature": (_D, [_M, _D]),
    "pf_model_voronoi_seeds": (_I, [_M, C.POINTER(_D), _I]),
    "pf_model_output_due": (_I, [_M, _I, C.POINTER(_I)]),
    "pf_model_write_kinetics": (_I, [_M, C.POINTER(PfStats), _D, _I, C.POINTER(_I)]),
    "pf_write_vti": (_I, [C.c_char_p, _I, _I, _I, _I, _D, C.POINTER(C.c_int32), C.POINTER(_D)]),
    "pf_write_vti_raw": (_I, [C.c_char_p, _I, _I, _I, _D, C.POINTER(C.c_int32),
                              C.POINTER(C.c_float)]),
    "pf_write_state": (_I, [C.c_char_p, C.POINTER(PfDesc), _D, C.c_int64,
                            C.POINTER(PfHostState)]),
    "pf_read_state": (_I, [C.c_char_p, C.POINTER(PfDesc), C.POINTER(_D), C.POINTER(C.c_int64),
                           C.POINTER(PfHostState)]),
    "pf_gpu_init": (_I, [_I, _I, _I]),
    "pf_gpu_diffusion_solver": (_I, [C.POINTER(_D), C.POINTER(_D), C.POINTER(_D), C.POINTER(_D),
                                     C.POINTER(_D), _I, _I, _I, _D, _D]),
    "pf_gpu_cleanup": (None, []),
    "pf_gpu_diffusion_solver_noalloc": (_I, [C.POINTER(_D), C.POINTER(_D), C.POINTER(_D),
                                             C.c_void_p, C.c_void_p, _I, _I, _I, _D, _D,
                                             C.c_void_p]),
}

_lib = None


class PhaseFieldError(RuntimeError):
    """A C-ABI call returned a negative status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


def load() -> C.CDLL:
    """Load libpfamd.so (raises if it is missing: build it with phasefield_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m phasefield_amd.build` (gfx950 HIP library)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
