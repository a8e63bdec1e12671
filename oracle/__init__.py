"""CPU oracle — TEST INFRASTRUCTURE ONLY.

ctypes wrapper over liboracle.so (pf_oracle.cpp, a C++/OpenMP restatement of the reference's
step).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package,
as the checker; the product (phasefield_amd/) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ORC_FP64, ORC_MIXED, ORC_DUAL = 0, 1, 2
_D, _I = C.c_double, C.c_int


class OrcParams(C.Structure):
    """Field-for-field the layout of pf_params (include/phasefield_amd.h)."""
    _fields_ = [
        ("dx", _D), ("eta", _D), ("cutoff", _D), ("dims", _I), ("reserved0", _I),
        ("mobility", _D * 3 * 3), ("sigma", _D * 3 * 3), ("entropy", _D * 3 * 3),
        ("eq_conc", _D * 3 * 3), ("eq_slope", _D * 3 * 3),
        ("part_k", _D * 3 * 3 * 2), ("part_c", _D * 3 * 3 * 2), ("diffusivity", _D * 3 * 2),
        ("pfm_dt", _D), ("dt_d", _D * 2), ("n_sub", _I * 2),
    ]


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        sig = {
            "orc_create": (C.c_void_p, [_I, _I, _I, _I, _I, _I, _I]),
            "orc_destroy": (None, [C.c_void_p]),
            "orc_set_params": (None, [C.c_void_p, P(OrcParams)]),
            "orc_set_state": (None, [C.c_void_p, P(C.c_int32), P(_D), P(_D), P(_D)]),
            "orc_set_comp": (None, [C.c_void_p, P(_D)]),
            "orc_get_state": (None, [C.c_void_p, P(C.c_int32), P(_D), P(_D), P(_D), P(_D), P(_D)]),
            "orc_pf_sweep": (None, [C.c_void_p]),
            "orc_diffusion_substeps": (None, [C.c_void_p, _I, _I, _D]),
            "orc_diffusion_substeps_factor": (None, [C.c_void_p, _I, _I, _D]),
            "orc_step": (None, [C.c_void_p, _I]),
            "orc_reduce": (None, [C.c_void_p, P(_D), P(_D), P(_D), P(_D), P(_D), P(C.c_int64)]),
            "orc_num_threads": (_I, []),
            "orc_dual_seed": (None, [C.c_void_p, P(_D)]),
            "orc_dual_phase_fraction": (None, [C.c_void_p, P(_D), P(_D)]),
            "orc_dual_conc": (None, [C.c_void_p, P(_D), P(_D)]),
            "orc_tessellation": (None, [_I, _I, _I, P(_D), _I, _D, P(C.c_int32)]),
            "orc_construct": (None, [C.c_void_p, P(C.c_int32), _I, P(_D)]),
            "orc_nucleate": (None, [C.c_void_p, _I, C.c_float, P(_D), C.c_uint64]),
            "orc_solute_noise": (None, [C.c_void_p, _I, _D, _D, C.c_uint64]),
            "orc_update_parameter": (None, [_D, _D, _D, P(_D), P(_D), P(_D), P(_D), P(_D), P(_D),
                                            P(_D), P(OrcParams), P(_D), P(_D)]),
            "orc_mt19937_nth": (C.c_uint32, [C.c_uint32, _I]),
            "orc_splitmix64": (C.c_uint64, [C.c_uint64]),
            "orc_quantize_comp": (C.c_float, [C.c_float, C.c_float, C.c_uint32]),
            "orc_quantize_comp_mean": (C.c_double, [C.c_float, C.c_float]),
            "orc_mix32": (C.c_uint32, [C.c_uint32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int32))


def as_orc_params(p) -> OrcParams:
    """Copy a pf_params-layout ctypes structure into OrcParams."""
    o = OrcParams()
    assert C.sizeof(o) == C.sizeof(p)
    C.memmove(C.byref(o), C.byref(p), C.sizeof(o))
    return o


class OracleDomain:
    def __init__(self, nx, ny, nz, depth=4, nsolute=1, form=ORC_FP64, bc_z=0):
        self.nx, self.ny, self.nz, self.K, self.S = nx, ny, nz, depth, nsolute
        self.N = nx * ny * nz
        self._h = lib().orc_create(nx, ny, nz, depth, nsolute, form, bc_z)
        assert self._h, "orc_create failed"

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        self.params = as_orc_params(p)
        lib().orc_set_params(self._h, C.byref(self.params))

    def set_state(self, index, value, phase, conc):
        index = np.ascontiguousarray(index, np.int32).reshape(-1)
        value = np.ascontiguousarray(value, np.float64).reshape(-1)
        phase = np.ascontiguousarray(phase, np.float64).reshape(-1)
        conc = np.ascontiguousarray(conc, np.float64).reshape(-1)
        lib().orc_set_state(self._h, _ip(index), _dp(value), _dp(phase), _dp(conc))

    def set_comp(self, comp):
        comp = np.ascontiguousarray(comp, np.float64).reshape(-1)
        assert comp.size == self.N * self.S
        lib().orc_set_comp(self._h, _dp(comp))

    def get_state(self):
        N, K, S = self.N, self.K, self.S
        out = {"index": np.empty(N * K, np.int32), "value": np.empty(N * K),
               "phase": np.empty(N * 3), "conc": np.empty(N * S), "conc_comp": np.empty(N * S),
               "phase_conc": np.empty(N * S * 3)}
        lib().orc_get_state(self._h, _ip(out["index"]), _dp(out["value"]), _dp(out["phase"]),
                            _dp(out["conc"]), _dp(out["conc_comp"]), _dp(out["phase_conc"]))
        return out

    def sweep(self):
        lib().orc_pf_sweep(self._h)

    def diffuse(self, solute, n, dt_d):
        lib().orc_diffusion_substeps(self._h, solute, n, dt_d)

    def diffuse_factor(self, solute, n, factor):
        lib().orc_diffusion_substeps_factor(self._h, solute, n, factor)

    def step(self, n=1):
        lib().orc_step(self._h, n)

    def reduce(self):
        mass, pf = np.zeros(2), np.zeros(3)
        cmax, cmin, serr = np.zeros(2), np.zeros(2), np.zeros(1)
        nif = C.c_int64()
        lib().orc_reduce(self._h, _dp(mass), _dp(pf), _dp(cmax), _dp(cmin), _dp(serr),
                         C.byref(nif))
        return {"mass": mass[: self.S].tolist(), "phase_frac": pf.tolist(),
                "max_conc": cmax[: self.S].tolist(), "min_conc": cmin[: self.S].tolist(),
                "max_sum_err": float(serr[0]), "n_interface": nif.value}

    def construct(self, grain_index, n_iso, init_conc):
        g = np.ascontiguousarray(grain_index, np.int32)
        ic = np.zeros(6)
        ic[: np.asarray(init_conc).size] = np.asarray(init_conc, np.float64).reshape(-1)
        lib().orc_construct(self._h, _ip(g), int(n_iso), _dp(ic))

    def nucleate(self, n, radius, nucleus_conc=(-1.0, -1.0), seed=0):
        nc = np.asarray(list(nucleus_conc) + [-1.0] * (2 - len(nucleus_conc)), np.float64)
        lib().orc_nucleate(self._h, int(n), float(radius), _dp(nc), int(seed))

    def solute_noise(self, solute, mean, sd, seed):
        lib().orc_solute_noise(self._h, solute, mean, sd, int(seed))

    def dual_seed(self, dc0_dir=None):
        d = None if dc0_dir is None else np.ascontiguousarray(dc0_dir, np.float64)
        lib().orc_dual_seed(self._h, _dp(d))

    def dual_phase_fraction(self):
        v, d = np.zeros(1), np.zeros(4)
        lib().orc_dual_phase_fraction(self._h, _dp(v), _dp(d))
        return float(v[0]), d

    def dual_conc(self):
        c, dc = np.empty(self.N), np.empty(4 * self.N)
        lib().orc_dual_conc(self._h, _dp(c), _dp(dc))
        return c, dc.reshape(4, self.N)


def tessellation(nx, ny, nz, seeds, reduction=0.0):
    seeds = np.ascontiguousarray(seeds, np.float64)
    out = np.empty(nx * ny * nz, np.int32)
    lib().orc_tessellation(nx, ny, nz, _dp(seeds), seeds.size // 3, float(reduction), _ip(out))
    return out


def update_parameter(T, dx, dims, mob0, mobQ, sigma, diff0, diffQ, entropy_lines, para_lines):
    arr = [np.ascontiguousarray(a, np.float64).reshape(-1)
           for a in (mob0, mobQ, sigma, diff0, diffQ, entropy_lines, para_lines)]
    out = OrcParams()
    pdt, ddt = np.zeros(1), np.zeros(1)
    lib().orc_update_parameter(float(T), float(dx), float(dims), *[_dp(a) for a in arr],
                               C.byref(out), _dp(pdt), _dp(ddt))
    return out, float(pdt[0]), float(ddt[0])


def num_threads() -> int:
    return lib().orc_num_threads()
