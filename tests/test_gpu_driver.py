"""The driver surface on the GPU: the `pfm` executable (the reference's main.cpp loop in C++) against
the Python PhaseField mirror on the same input files, and the 2D Fe-C workload (the reference's own
CPU-runnable case: 256 x 256, 12 grains, 20 nuclei, K = 6) over its full 1000 steps against the
oracle, bit for bit.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle
from phasefield_amd import Domain, Model, PhaseField, configs, lib
from phasefield_amd._capi import PfDesc, PfHostState

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PFM = os.path.join(ROOT, "phasefield_amd", "pfm")


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


def _read_dump(path: str, d: PfDesc) -> dict:
    N = d.nx * d.ny * d.nz
    out = {"index": np.zeros(N * d.depth, np.int32), "value": np.zeros(N * d.depth),
           "phase": np.zeros(N * 3), "conc": np.zeros(N * d.nsolute),
           "conc_comp": np.zeros(N * d.nsolute)}
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    s = PfHostState(out["index"].ctypes.data_as(C.POINTER(C.c_int32)), dp(out["value"]),
                    dp(out["phase"]), dp(out["conc"]), dp(out["conc_comp"]), None)
    t, st = C.c_double(), C.c_int64()
    assert lib().pf_read_state(path.encode(), C.byref(d), C.byref(t), C.byref(st), C.byref(s)) == 0
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pfm_matches_python_driver(tmp_path, precision):
    """Same inputs through `pfm` and through PhaseField.read_data/construct/nucleate/update/
    write_data: identical kinetics files and identical final states."""
    banner(f"pfm vs python driver ({precision})")
    assert os.path.exists(PFM), "phasefield_amd/pfm is not built"
    over = dict(nx=96, ny=80, nz=1, grains=6, nuclei=5, steps=30, output_dt=0.01,
                precision=precision)
    dump = str(tmp_path / "a" / "state.bin")
    a = configs.write_inputs(str(tmp_path / "a"), "fec_2d_256", extra={"dump_state": dump}, **over)
    r = subprocess.run([PFM, a], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "SIMULATION IS FINISHED." in r.stdout

    b = configs.write_inputs(str(tmp_path / "b"), "fec_2d_256", **over)
    pf = PhaseField().read_data(b)
    pf.construct()
    pf.write_data(forced=True)
    pf.nucleate()
    for _ in range(over["steps"]):
        pf.update()
        pf.write_data()
    ka = open(os.path.join(Model(a).get_string("directory"), "fraction_time.txt")).read()
    kb = open(os.path.join(pf.model.get_string("directory"), "fraction_time.txt")).read()
    assert ka.count("\n") >= 5 and ka == kb
    d = Model(a).desc()
    sa, sb = _read_dump(dump, d), pf.domain.download()
    for key in ("index", "value", "phase", "conc", "conc_comp"):
        assert np.array_equal(sa[key], sb[key]), key


@pytest.mark.slow
def test_fec_2d_256_1000_steps_bitwise(tmp_path):
    """The 2D Fe-C workload at full size and step count: device-built seeded state, then 1000 steps
    (1 sweep + 200 sub-steps each) on the GPU and in the oracle; identical maps and values."""
    banner("fec_2d_256: 1000 full steps vs oracle")
    m = Model(configs.write_inputs(str(tmp_path), "fec_2d_256"))
    p = m.update_parameter()
    d = m.desc()
    assert (d.nx, d.ny, d.nz, d.depth) == (256, 256, 1, 6)
    g = Domain.from_desc(d)
    g.set_params(p)
    g.construct(m.voronoi_seeds(), m.get_real("reduction"), m.get_int("num_init_iterations"),
                [m.get_real(f"init_conc{i}") for i in range(3)])
    g.nucleate(m.get_int("num_nuclei"), m.get_real("nucleus_radius"),
               (m.get_real("nucleus_conc"), -1.0), m.get_int("nucleation_seed"))
    s0 = g.download()
    o = oracle.OracleDomain(d.nx, d.ny, d.nz, d.depth, 1)
    o.set_params(oracle.as_orc_params(p))
    o.set_state(s0["index"], s0["value"], s0["phase"], s0["conc"])
    g.step(1000)
    o.step(1000)
    gs, os_ = g.download(), o.get_state()
    assert np.array_equal(gs["index"], os_["index"]), "active-set maps differ"
    for key in ("value", "phase", "conc"):
        assert np.array_equal(gs[key], os_[key]), key
    # the run did something: austenite grew from its nuclei
    assert gs["phase"].reshape(-1, 3)[:, 0].sum() > s0["phase"].reshape(-1, 3)[:, 0].sum()
