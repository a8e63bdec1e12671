"""MI355X (gfx950) multi-phase-field + solute-diffusion step behind a C-ABI.

Host-side mirror of the reference's driver interface (zhubq/PhaseField, C++-openmp-cuda/src):

  reference                                 here
  phase_field::Read_data(file)              PhaseField.read_data(path)         -> Model
  phase_field::Construct(true, ..)          PhaseField.construct()             (device setup)
  nucleation::Nucleate1in2(F)               PhaseField.nucleate()
  phase_field::Update(false)                PhaseField.update()                (one GPU step)
  phase_field::Write_data(forced)           PhaseField.write_data(forced)
  GPU_Init / GPU_Diffusion_solver / ..      gpu_init / gpu_diffusion_solver / gpu_cleanup

`Domain` is the thin handle wrapper over pf_* (include/phasefield_amd.h); every compute call runs
the HIP kernels in libpfamd.so.  Arrays cross the boundary as numpy arrays in the reference layout
(cell n = i + nx*(j + ny*k), slots/phases fastest).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _capi
from ._capi import (PF_BC_NOFLUX, PF_BC_PERIODIC, PF_FP64, PF_MIXED, PfDesc, PfHostState,
                    PfParams, PfStats, PhaseFieldError)

__all__ = ["Domain", "Model", "PhaseField", "PfDesc", "PfParams", "PhaseFieldError",
           "PF_FP64", "PF_MIXED", "PF_BC_PERIODIC", "PF_BC_NOFLUX", "lib", "gpu_init",
           "gpu_diffusion_solver", "gpu_cleanup", "params_from_dict", "params_to_dict",
           "write_vti_raw", "read_vti_raw"]


def lib():
    return _capi.load()


def _dp(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int32))


def params_to_dict(p: PfParams) -> dict:
    return {
        "dx": p.dx, "eta": p.eta, "cutoff": p.cutoff, "dims": p.dims,
        "mobility": np.ctypeslib.as_array(p.mobility).copy(),
        "sigma": np.ctypeslib.as_array(p.sigma).copy(),
        "entropy": np.ctypeslib.as_array(p.entropy).copy(),
        "eq_conc": np.ctypeslib.as_array(p.eq_conc).copy(),
        "eq_slope": np.ctypeslib.as_array(p.eq_slope).copy(),
        "part_k": np.ctypeslib.as_array(p.part_k).copy(),
        "part_c": np.ctypeslib.as_array(p.part_c).copy(),
        "diffusivity": np.ctypeslib.as_array(p.diffusivity).copy(),
        "pfm_dt": p.pfm_dt, "dt_d": list(p.dt_d), "n_sub": list(p.n_sub),
    }


def params_from_dict(d: dict) -> PfParams:
    p = PfParams()
    p.dx, p.eta, p.cutoff, p.dims = d["dx"], d["eta"], d["cutoff"], int(d["dims"])
    for name in ("mobility", "sigma", "entropy", "eq_conc", "eq_slope", "part_k", "part_c",
                 "diffusivity"):
        np.ctypeslib.as_array(getattr(p, name))[...] = np.asarray(d[name], dtype=np.float64)
    p.pfm_dt = d["pfm_dt"]
    p.dt_d[0], p.dt_d[1] = d["dt_d"][0], d["dt_d"][1]
    p.n_sub[0], p.n_sub[1] = int(d["n_sub"][0]), int(d["n_sub"][1])
    return p


def _check(rc: int, h=None, what: str = "") -> None:
    if rc != 0:
        msg = lib().pf_last_error(h)
        raise PhaseFieldError(rc, f"{what}: {msg.decode() if msg else ''}")


class Domain:
    """One simulation domain on one GPU (pf_create ... pf_destroy)."""

    def __init__(self, nx: int, ny: int, nz: int, depth: int = 4, nphase: int = 2,
                 nsolute: int = 1, precision: str | int = "fp64", bc_z: str | int = "periodic",
                 device: int = 0):
        prec = precision if isinstance(precision, int) else (
            PF_MIXED if precision in ("mixed", "fp32") else PF_FP64)
        bc = bc_z if isinstance(bc_z, int) else (PF_BC_NOFLUX if bc_z == "noflux" else PF_BC_PERIODIC)
        self.desc = PfDesc(nx, ny, nz, depth, nphase, nsolute, prec, bc, device)
        self._h = C.c_void_p()
        _check(lib().pf_create(C.byref(self.desc), C.byref(self._h)), None, "pf_create")
        self.N = nx * ny * nz
        self.params: PfParams | None = None

    @classmethod
    def from_desc(cls, d: PfDesc) -> "Domain":
        return cls(d.nx, d.ny, d.nz, d.depth, d.nphase, d.nsolute, d.precision, d.bc_z, d.device)

    @property
    def handle(self):
        return self._h

    @property
    def mixed(self) -> bool:
        return self.desc.precision == PF_MIXED

    def close(self) -> None:
        if self._h:
            lib().pf_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- parameters / state ----------------------------------------------------------------
    def set_params(self, p: PfParams | dict) -> None:
        if isinstance(p, dict):
            p = params_from_dict(p)
        _check(lib().pf_set_params(self._h, C.byref(p)), self._h, "pf_set_params")
        self.params = p

    def upload(self, index, value, phase, conc, conc_comp=None) -> None:
        d = self.desc
        index = np.ascontiguousarray(index, dtype=np.int32).reshape(-1)
        value = np.ascontiguousarray(value, dtype=np.float64).reshape(-1)
        phase = np.ascontiguousarray(phase, dtype=np.float64).reshape(-1)
        conc = np.ascontiguousarray(conc, dtype=np.float64).reshape(-1)
        assert index.size == self.N * d.depth and value.size == self.N * d.depth
        assert phase.size == self.N * 3 and conc.size == self.N * d.nsolute
        comp = None if conc_comp is None else np.ascontiguousarray(conc_comp, np.float64).reshape(-1)
        s = PfHostState(_ip(index), _dp(value), _dp(phase), _dp(conc), _dp(comp), None)
        _check(lib().pf_upload_state(self._h, C.byref(s)), self._h, "pf_upload_state")

    def download(self, phase_conc: bool = False) -> dict:
        d = self.desc
        out = {
            "index": np.empty(self.N * d.depth, np.int32),
            "value": np.empty(self.N * d.depth, np.float64),
            "phase": np.empty(self.N * 3, np.float64),
            "conc": np.empty(self.N * d.nsolute, np.float64),
            "conc_comp": np.empty(self.N * d.nsolute, np.float64),
        }
        pc = np.empty(self.N * d.nsolute * 3, np.float64) if phase_conc else None
        s = PfHostState(_ip(out["index"]), _dp(out["value"]), _dp(out["phase"]), _dp(out["conc"]),
                        _dp(out["conc_comp"]), _dp(pc))
        _check(lib().pf_download_state(self._h, C.byref(s)), self._h, "pf_download_state")
        if pc is not None:
            out["phase_conc"] = pc
        return out

    def copy_from(self, src: "Domain") -> None:
        _check(lib().pf_copy_state(self._h, src._h), self._h, "pf_copy_state")

    # --- step --------------------------------------------------------------------------------
    def step(self, n: int = 1) -> None:
        _check(lib().pf_step(self._h, int(n)), self._h, "pf_step")

    def sweep(self) -> None:
        _check(lib().pf_phase_field_sweep(self._h), self._h, "pf_phase_field_sweep")

    def diffuse(self, solute: int, n: int, dt_d: float) -> None:
        _check(lib().pf_diffuse_substeps(self._h, solute, n, dt_d), self._h, "pf_diffuse_substeps")

    def synchronize(self) -> None:
        _check(lib().pf_synchronize(self._h), self._h, "pf_synchronize")

    def dump_vti_async(self, path: str, dx: float) -> None:
        """Binary VTI dump (korn, carbon) of the current state, written by a host thread while
        the caller keeps stepping; `dump_wait` returns when the file is complete."""
        _check(lib().pf_dump_vti_async(self._h, path.encode(), float(dx)), self._h,
               "pf_dump_vti_async")

    def dump_wait(self) -> None:
        _check(lib().pf_dump_wait(self._h), self._h, "pf_dump_wait")

    def reduce(self, check_finite: bool = True) -> dict:
        st = PfStats()
        rc = lib().pf_reduce(self._h, C.byref(st))
        if rc != 0 and (check_finite or rc != _capi.PF_ERR_NONFINITE):
            _check(rc, self._h, "pf_reduce")
        return {"mass": list(st.mass), "phase_frac": list(st.phase_frac),
                "max_conc": list(st.max_conc), "min_conc": list(st.min_conc),
                "max_sum_err": st.max_sum_err, "n_interface": st.n_interface,
                "n_nonfinite": st.n_nonfinite}

    # --- timing ------------------------------------------------------------------------------
    def timing_reset(self, enable: bool = True) -> None:
        _check(lib().pf_kernel_timing_reset(self._h, int(enable)), self._h, "timing")

    def timing(self, kind: int = 0) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_int64()
        _check(lib().pf_kernel_timing(self._h, kind, C.byref(ms), C.byref(n)), self._h, "timing")
        return ms.value, n.value

    def list_lengths(self) -> tuple[int, int]:
        """(cells the last sweep listed, cells the diffusion sub-steps list)."""
        a, b = C.c_int64(), C.c_int64()
        _check(lib().pf_list_lengths(self._h, C.byref(a), C.byref(b)), self._h, "list lengths")
        return a.value, b.value

    def march_shape(self) -> tuple[int, int]:
        """(tile rows, z-chunk) of the plane-marching diffusion kernel; (0, 0) before its first use."""
        a, b = C.c_int(), C.c_int()
        _check(lib().pf_march_shape(self._h, C.byref(a), C.byref(b)), self._h, "march shape")
        return a.value, b.value

    # --- checkpoints -------------------------------------------------------------------------
    def checkpoint(self) -> bytes:
        n = C.c_size_t()
        _check(lib().pf_checkpoint_size(self._h, C.byref(n)), self._h, "checkpoint")
        buf = C.create_string_buffer(n.value)
        _check(lib().pf_checkpoint_save(self._h, buf), self._h, "checkpoint")
        return buf.raw

    def restore(self, blob: bytes) -> None:
        n = C.c_size_t()
        _check(lib().pf_checkpoint_size(self._h, C.byref(n)), self._h, "restore")
        if len(blob) != n.value:
            raise PhaseFieldError(_capi.PF_ERR_INVALID,
                                  f"restore: checkpoint of {len(blob)} bytes, this domain's take "
                                  f"{n.value}")
        buf = C.create_string_buffer(blob, len(blob))
        _check(lib().pf_checkpoint_restore(self._h, buf), self._h, "restore")

    # --- seeded setup ------------------------------------------------------------------------
    def construct(self, seeds: np.ndarray, reduction: float, n_iso: int, init_conc) -> None:
        seeds = np.ascontiguousarray(seeds, np.float64)
        ic = np.zeros(6, np.float64)
        ic[: np.asarray(init_conc).size] = np.asarray(init_conc, np.float64).reshape(-1)
        _check(lib().pf_setup_construct(self._h, _dp(seeds), seeds.size // 3, float(reduction),
                                        int(n_iso), _dp(ic)), self._h, "pf_setup_construct")

    def nucleate(self, n: int, radius: float, nucleus_conc=(-1.0, -1.0), seed: int = 0) -> None:
        nc = np.asarray(list(nucleus_conc) + [-1.0] * (2 - len(nucleus_conc)), np.float64)
        _check(lib().pf_setup_nucleate(self._h, int(n), float(radius), _dp(nc), int(seed)),
               self._h, "pf_setup_nucleate")

    def solute_noise(self, solute: int, mean: float, sd: float, seed: int) -> None:
        _check(lib().pf_setup_solute_noise(self._h, solute, mean, sd, int(seed)), self._h,
               "pf_setup_solute_noise")

    def rex_fraction(self) -> float:
        v = C.c_double()
        _check(lib().pf_rex_fraction(self._h, C.byref(v)), self._h, "pf_rex_fraction")
        return v.value

    # --- discrete adjoint --------------------------------------------------------------------
    def adjoint(self, n_steps: int, sample_steps=(), target_fraction=(), weight_conc: float = 0.0,
                target_conc: np.ndarray | None = None, checkpoint_every: int = 1,
                want_conc0: bool = True) -> dict:
        """Forward `n_steps` steps from the current state with checkpoints, then the reverse sweep.

        Loss = sum_k (f_gamma(sample_steps[k]) - target_fraction[k])^2
               + weight_conc * ||c_final - target_conc||^2.
        Returns the loss, the forward f_gamma samples and dL/d(mobility scale),
        dL/d(interface-energy scale), dL/dD(phase 1, carbon) and dL/dc0 (carbon, per cell).
        The domain holds the final forward state afterwards.
        """
        samples = list(sample_steps)
        targets = list(target_fraction)
        if len(samples) != len(targets) or len(samples) > 64:
            raise ValueError("sample_steps and target_fraction must pair up (at most 64)")
        sp = _capi.PfAdjointSpec()
        sp.n_steps = int(n_steps)
        sp.checkpoint_every = int(checkpoint_every)
        sp.n_samples = len(samples)
        for q, (s, t) in enumerate(zip(samples, targets)):
            sp.sample_steps[q] = int(s)
            sp.target_fraction[q] = float(t)
        sp.weight_conc = float(weight_conc)
        tc = None
        if target_conc is not None:
            tc = np.ascontiguousarray(target_conc, np.float64).reshape(-1)
            if tc.size != self.N:
                raise ValueError("target_conc must hold one value per cell")
            sp.target_conc = _dp(tc)
        res = _capi.PfAdjointResult()
        g = np.zeros(self.N, np.float64) if want_conc0 else None
        res.d_conc0 = _dp(g)
        _check(lib().pf_adjoint_run(self._h, C.byref(sp), C.byref(res)), self._h, "pf_adjoint_run")
        return {"loss": res.loss, "fraction": list(res.fraction[: len(samples)]),
                "d_mobility": res.d_mobility, "d_sigma": res.d_sigma,
                "d_diff_gamma": res.d_diff_gamma, "d_conc0": g,
                "forward_ms": res.forward_ms, "reverse_ms": res.reverse_ms}


class Model:
    """Parsed input (pf_model): the reference's positional main file + thermodynamics file."""

    def __init__(self, path: str):
        self._m = C.c_void_p()
        rc = lib().pf_model_read(path.encode(), C.byref(self._m))
        if rc != 0:
            raise PhaseFieldError(rc, lib().pf_model_last_error().decode())
        self.path = path

    def close(self):
        if self._m:
            lib().pf_model_free(self._m)
            self._m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_int(self, key: str) -> int:
        v = C.c_int64()
        rc = lib().pf_model_get_int(self._m, key.encode(), C.byref(v))
        if rc != 0:
            raise PhaseFieldError(rc, lib().pf_model_last_error().decode())
        return v.value

    def get_real(self, key: str) -> float:
        v = C.c_double()
        rc = lib().pf_model_get_real(self._m, key.encode(), C.byref(v))
        if rc != 0:
            raise PhaseFieldError(rc, lib().pf_model_last_error().decode())
        return v.value

    def get_string(self, key: str) -> str:
        buf = C.create_string_buffer(4096)
        rc = lib().pf_model_get_string(self._m, key.encode(), buf, 4096)
        if rc != 0:
            raise PhaseFieldError(rc, lib().pf_model_last_error().decode())
        return buf.value.decode()

    def set_real(self, key: str, v: float) -> None:
        lib().pf_model_set_real(self._m, key.encode(), float(v))

    def set_int(self, key: str, v: int) -> None:
        lib().pf_model_set_int(self._m, key.encode(), int(v))

    def desc(self) -> PfDesc:
        d = PfDesc()
        lib().pf_model_desc(self._m, C.byref(d))
        return d

    def update_parameter(self) -> PfParams:
        """phase_field::Update_parameter: parameters at the current time, then time += pfm_dt."""
        p = PfParams()
        rc = lib().pf_model_update_parameter(self._m, C.byref(p))
        if rc != 0:
            raise PhaseFieldError(rc, "pf_model_update_parameter")
        return p

    def temperature(self, t: float) -> float:
        return lib().pf_model_temperature(self._m, float(t))

    def voronoi_seeds(self, n: int | None = None) -> np.ndarray:
        n = self.get_int("num_grain1") if n is None else n
        s = np.empty(3 * n, np.float64)
        _check(lib().pf_model_voronoi_seeds(self._m, _dp(s), n), None, "voronoi seeds")
        return s

    def output_due(self, forced: bool = False) -> bool:
        v = C.c_int()
        lib().pf_model_output_due(self._m, int(forced), C.byref(v))
        return bool(v.value)

    def write_kinetics(self, stats: dict, f_rex: float, forced: bool = False) -> bool:
        st = PfStats()
        for i in range(3):
            st.phase_frac[i] = stats["phase_frac"][i]
        w = C.c_int()
        rc = lib().pf_model_write_kinetics(self._m, C.byref(st), float(f_rex), int(forced),
                                           C.byref(w))
        if rc != 0:
            raise PhaseFieldError(rc, lib().pf_model_last_error().decode())
        return bool(w.value)


class PhaseField:
    """Mirror of the reference's phase_field driver object over one GPU domain."""

    def __init__(self):
        self.model: Model | None = None
        self.domain: Domain | None = None
        self.params: PfParams | None = None
        self.steps_done = 0

    def read_data(self, path: str) -> "PhaseField":
        self.model = Model(path)
        return self

    @property
    def time(self) -> float:
        return self.model.get_real("time")

    @property
    def end_time(self) -> float:
        return self.model.get_real("end_time")

    def construct(self) -> None:
        """Construct(true, ..): seeded Voronoi + Iso_growth + initial partition on the device.

        Setup runs on an FP64 domain; a mixed-precision input converts after nucleation.
        """
        m = self.model
        d = m.desc()
        self._target = PfDesc.from_buffer_copy(d)
        d.precision = PF_FP64
        self.domain = Domain.from_desc(d)
        self.params = m.update_parameter()
        self.domain.set_params(self.params)
        ic = [m.get_real(f"init_conc{i}") for i in range(3)]
        if d.nsolute > 1:
            ic += [m.get_real("mn_conc")] * 3
        self.domain.construct(m.voronoi_seeds(), m.get_real("reduction"),
                              m.get_int("num_init_iterations"), ic)
        if d.nsolute > 1 and m.get_real("mn_noise") > 0:
            self.domain.solute_noise(1, m.get_real("mn_conc"), m.get_real("mn_noise"),
                                     m.get_int("mn_seed"))

    def nucleate(self) -> None:
        m = self.model
        self.domain.nucleate(m.get_int("num_nuclei"), m.get_real("nucleus_radius"),
                             (m.get_real("nucleus_conc"), m.get_real("mn_nucleus_conc")),
                             m.get_int("nucleation_seed"))
        if self._target.precision == PF_MIXED:
            mixed = Domain.from_desc(self._target)
            mixed.set_params(self.params)
            mixed.copy_from(self.domain)
            self.domain.close()
            self.domain = mixed

    def update(self) -> None:
        """phase_field::Update(false): parameters, one phase-field sweep, the diffusion sub-cycle."""
        self.params = self.model.update_parameter()
        self.domain.set_params(self.params)
        self.domain.step(1)
        self.steps_done += 1

    def write_data(self, forced: bool = False) -> bool:
        if not self.model.output_due(forced):
            return False
        st = self.domain.reduce()
        return self.model.write_kinetics(st, self.domain.rex_fraction(), True)


# ---- binary field files --------------------------------------------------------------------
def write_vti_raw(path: str, shape, dx: float, korn: np.ndarray, conc: np.ndarray) -> None:
    """korn (int32) and conc (float32), storage order, as a binary VTI file (raw appended)."""
    nx, ny, nz = shape
    k = np.ascontiguousarray(korn, np.int32)
    c = np.ascontiguousarray(conc, np.float32)
    if k.size != nx * ny * nz or c.size != k.size:
        raise ValueError("korn/conc size does not match the shape")
    _check(lib().pf_write_vti_raw(path.encode(), nx, ny, nz, float(dx),
                                  k.ctypes.data_as(C.POINTER(C.c_int32)),
                                  c.ctypes.data_as(C.POINTER(C.c_float))), None, "pf_write_vti_raw")


def read_vti_raw(path: str) -> dict:
    """Parse a binary VTI file of write_vti_raw / Domain.dump_vti_async: extent, spacing and the
    korn / conc arrays (storage order)."""
    import re
    raw = open(path, "rb").read()
    head, _, rest = raw.partition(b"<AppendedData encoding=\"raw\">\n_")
    text = head.decode()
    ext = [int(v) for v in re.search(r'WholeExtent="([^"]+)"', text).group(1).split()]
    spacing = float(re.search(r'Spacing="([^ "]+)', text).group(1))
    n0 = int(np.frombuffer(rest[:8], np.uint64)[0])
    korn = np.frombuffer(rest[8:8 + n0], np.int32)
    n1 = int(np.frombuffer(rest[8 + n0:16 + n0], np.uint64)[0])
    conc = np.frombuffer(rest[16 + n0:16 + n0 + n1], np.float32)
    tail = rest[16 + n0 + n1:].decode()
    if "</AppendedData>" not in tail or "</VTKFile>" not in tail:
        raise ValueError("truncated VTI file")
    return {"shape": (ext[1], ext[3], ext[5]), "dx": spacing, "korn": korn, "conc": conc}


# ---- the reference's accelerator API (gpu_solver.h:18-23) -----------------------------------
def gpu_init(x: int, y: int, z: int) -> int:
    return lib().pf_gpu_init(x, y, z)


def gpu_diffusion_solver(diffusivity, partition_k, partition_c, value_phase, total_conc,
                         x: int, y: int, z: int, dt: float, dx: float) -> None:
    """In-place update of total_conc (float64 [N]) as GPU_Diffusion_solver does."""
    arrs = [np.ascontiguousarray(a, np.float64) for a in (diffusivity, partition_k, partition_c,
                                                          value_phase)]
    assert total_conc.dtype == np.float64 and total_conc.flags.c_contiguous
    rc = lib().pf_gpu_diffusion_solver(*[_dp(a) for a in arrs], _dp(total_conc), x, y, z,
                                       float(dt), float(dx))
    if rc != 0:
        raise PhaseFieldError(rc, "GPU_Diffusion_solver failed (see stderr)")


def gpu_cleanup() -> None:
    lib().pf_gpu_cleanup()
