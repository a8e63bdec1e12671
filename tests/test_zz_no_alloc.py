"""GPU: the allocation-free diffusion offload (pf_gpu_diffusion_solver_noalloc) against the entry
point it varies (pf_gpu_diffusion_solver, the reference's GPU_Diffusion_solver), bitwise, on two
calls in a row (the second smaller, and with the other parity of the sub-step count), and its
refusal of a domain larger than its scratch buffer.  Device buffers and the stream come from torch.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle
from phasefield_amd import gpu_cleanup, gpu_diffusion_solver, lib
from phasefield_amd._capi import PF_ERR_RANGE
from tests import helpers

pytestmark = pytest.mark.gpu


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _num_loop(dt, dx, D, dims):
    return int(np.ceil(dt / (dx * dx / D.max() / dims * float(np.float32(0.2)))))


def test_noalloc_matches_gpu_diffusion_solver(tmp_path):
    import torch

    m = helpers.small_model(str(tmp_path), "fec_3d_256", nx=16, ny=16, nz=16, grains=4, nuclei=2)
    p = helpers.model_params(m)
    D = np.array([p.diffusivity[0][i] for i in range(3)])
    pk = np.array([[p.part_k[0][i][j] for j in range(3)] for i in range(3)]).T.reshape(-1).copy()
    pc = np.array([[p.part_c[0][i][j] for j in range(3)] for i in range(3)]).T.reshape(-1).copy()
    rng = np.random.default_rng(4321)
    stream = torch.cuda.current_stream()
    parities = []
    # the second call is smaller (and 2D) and takes the other parity of the sub-step count, so the
    # result ends once in the caller's buffer and once in the scratch buffer (copied back)
    for (nx, ny, nz), parity in (((72, 40, 30), 0), ((40, 24, 1), 1)):
        dims = 3 if nz > 1 else 2
        dt = next(p.pfm_dt * mul for mul in (1.0, 0.997, 0.83, 0.71, 0.57, 0.49, 0.37)
                  if _num_loop(p.pfm_dt * mul, p.dx, D, dims) % 2 == parity)
        nloop = _num_loop(dt, p.dx, D, dims)
        parities.append(nloop % 2)
        print(f"[gpu] noalloc {nx}x{ny}x{nz}: {nloop} sub-steps", flush=True)
        ids, val, phase, conc = helpers.random_state(rng, nx, ny, nz, 4)
        ref = conc.copy()
        gpu_diffusion_solver(D, pk, pc, phase, ref, nx, ny, nz, dt, p.dx)
        dphase = torch.from_numpy(phase.copy()).to("cuda:0")
        dconc = torch.from_numpy(conc.copy()).to("cuda:0")
        torch.cuda.synchronize()
        rc = lib().pf_gpu_diffusion_solver_noalloc(_dp(D), _dp(pk), _dp(pc), dphase.data_ptr(),
                                                   dconc.data_ptr(), nx, ny, nz, dt, p.dx,
                                                   stream.cuda_stream)
        assert rc == 0, lib().pf_last_error(None)
        torch.cuda.synchronize()
        got = dconc.cpu().numpy()
        assert np.array_equal(got, ref), (nx, ny, nz)
        q = oracle.as_orc_params(p)
        q.dims = 3
        o = oracle.OracleDomain(nx, ny, nz, 4, 1)
        o.set_params(q)
        o.set_state(ids, val, phase, conc)
        o.diffuse_factor(0, nloop, dt / nloop / (p.dx * p.dx))
        assert np.array_equal(got, o.get_state()["conc"]), (nx, ny, nz)
    assert sorted(parities) == [0, 1]
    gpu_cleanup()
    # beyond the scratch buffer: refused before any device work
    big = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    rc = lib().pf_gpu_diffusion_solver_noalloc(_dp(D), _dp(pk), _dp(pc), big.data_ptr(),
                                               big.data_ptr(), 256, 128, 128, p.pfm_dt, p.dx, None)
    assert rc == PF_ERR_RANGE
    assert b"scratch" in lib().pf_last_error(None)
