"""Test helpers: small workloads, oracle-built initial states and random active-set states."""
from __future__ import annotations

import os
import tempfile

import numpy as np

import oracle
from phasefield_amd import Model, configs
from phasefield_amd._capi import PfParams


def small_model(tmpdir: str, workload: str = "fec_3d_256", **over) -> Model:
    """Parse a (shrunken) named workload with the library's own input parser."""
    path = configs.write_inputs(tmpdir, workload, **over)
    return Model(path)


def model_params(m: Model) -> PfParams:
    """Parameters after Construct's Update_parameter (as the reference computes them)."""
    return m.update_parameter()


def oracle_initial_state(m: Model, p: PfParams, form=oracle.ORC_FP64) -> oracle.OracleDomain:
    """Seeded setup on the CPU oracle: Tessellation + Iso_growth + partition + nucleation
    (+ manganese noise for two-solute workloads)."""
    d = m.desc()
    od = oracle.OracleDomain(d.nx, d.ny, d.nz, d.depth, d.nsolute, form, d.bc_z)
    od.set_params(p)
    seeds = m.voronoi_seeds()
    grains = oracle.tessellation(d.nx, d.ny, d.nz, seeds, m.get_real("reduction"))
    ic = [m.get_real(f"init_conc{i}") for i in range(3)]
    if d.nsolute > 1:
        ic += [m.get_real("mn_conc")] * 3
    od.construct(grains, m.get_int("num_init_iterations"), ic)
    if d.nsolute > 1 and m.get_real("mn_noise") > 0:
        od.solute_noise(1, m.get_real("mn_conc"), m.get_real("mn_noise"), m.get_int("mn_seed"))
    od.nucleate(m.get_int("num_nuclei"), m.get_real("nucleus_radius"),
                (m.get_real("nucleus_conc"), m.get_real("mn_nucleus_conc")),
                m.get_int("nucleation_seed"))
    return od


def random_state(rng, nx, ny, nz, K, n_grains=6, nphase=2, nsolute=1, max_slots=None):
    """Random valid active-set state: per cell 1..max_slots distinct grains with phi sorted
    descending and summing to one; value_phase consistent; random concentrations."""
    N = nx * ny * nz
    max_slots = max_slots or K
    ids = np.zeros((N, K), np.int32)
    val = np.zeros((N, K))
    pool = [200001 + g for g in range(n_grains)] + [100001 + g for g in range(max(2, n_grains // 2))]
    if nphase == 3:
        pool += [300001, 300002]
    ncell = rng.integers(1, max_slots + 1, size=N)
    for n in range(N):
        k = ncell[n]
        g = rng.choice(pool, size=k, replace=False)
        v = rng.random(k) + 0.05
        v /= v.sum()
        order = np.argsort(-v, kind="stable")
        ids[n, :k] = g[order]
        val[n, :k] = v[order]
    phase = np.zeros((N, 3))
    for s in range(K):
        ph = np.where(ids[:, s] >= 300000, 2, np.where(ids[:, s] >= 200000, 1, 0))
        live = ids[:, s] != 0
        np.add.at(phase, (np.nonzero(live)[0], ph[live]), val[live, s])
    conc = np.concatenate([0.02 + 0.8 * rng.random(N) for _ in range(nsolute)])
    return ids.reshape(-1), val.reshape(-1), phase.reshape(-1), conc


def tmpdir():
    return tempfile.mkdtemp(prefix="pfamd_")
