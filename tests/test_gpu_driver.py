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


@pytest.mark.slow
def test_fec_2d_256_1000_steps_mixed_bitwise(tmp_path):
    """The same workload and horizon in the mixed form against the oracle's mixed form: 200 000
    sub-steps of stochastically rounded compensation codes (keys counted per sub-step), list
    rebuilds after every sweep; identical maps, values, concentrations and codes."""
    banner("fec_2d_256 mixed: 1000 full steps vs oracle")
    m = Model(configs.write_inputs(str(tmp_path), "fec_2d_256"))
    p = m.update_parameter()
    d = m.desc()
    d.precision = 1
    g = Domain.from_desc(d)
    g.set_params(p)
    f = Domain.from_desc(m.desc())
    f.set_params(p)
    f.construct(m.voronoi_seeds(), m.get_real("reduction"), m.get_int("num_init_iterations"),
                [m.get_real(f"init_conc{i}") for i in range(3)])
    f.nucleate(m.get_int("num_nuclei"), m.get_real("nucleus_radius"),
               (m.get_real("nucleus_conc"), -1.0), m.get_int("nucleation_seed"))
    s0 = f.download()
    f.close()
    g.upload(s0["index"], s0["value"], s0["phase"], s0["conc"])
    o = oracle.OracleDomain(d.nx, d.ny, d.nz, d.depth, 1, oracle.ORC_MIXED)
    o.set_params(oracle.as_orc_params(p))
    o.set_state(s0["index"], s0["value"], s0["phase"], s0["conc"])
    g.step(1000)
    o.step(1000)
    gs, os_ = g.download(), o.get_state()
    assert np.array_equal(gs["index"], os_["index"]), "active-set maps differ"
    for key in ("value", "phase", "conc", "conc_comp"):
        assert np.array_equal(gs[key], os_[key]), key
    assert gs["phase"].reshape(-1, 3)[:, 0].sum() > s0["phase"].reshape(-1, 3)[:, 0].sum()
    g.close()


def _vti_fields(s: dict, d: PfDesc):
    """korn and float32 carbon of a downloaded state, as write_data.cpp:42-82 defines them."""
    idx = s["index"].reshape(-1, d.depth)
    g1, g2 = idx[:, 0], idx[:, 1]
    ph = np.where(g1 == 0, 0, np.where(g1 // 100000 > 0, g1 // 100000, 1))
    korn = np.where(g2 > 0, 0, ph).astype(np.int32)
    N = d.nx * d.ny * d.nz
    conc = (s["conc"][:N] - s["conc_comp"][:N]).astype(np.float32)
    return korn, conc


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_async_vti_dump(tmp_path, precision):
    """pf_dump_vti_async writes the state at the call although the domain keeps stepping while
    the copy and the file write run; a second dump first waits for the first."""
    from phasefield_amd import read_vti_raw
    banner(f"async VTI dump ({precision})")
    path = configs.write_inputs(str(tmp_path), "fec_3d_256", nx=64, ny=48, nz=40, grains=6,
                                nuclei=6, precision=precision)
    pf = PhaseField().read_data(path)
    pf.construct()
    pf.nucleate()
    g = pf.domain
    d = Model(path).desc()
    dx = Model(path).get_real("dx")
    g.step(2)
    f1, f2 = str(tmp_path / "a.vti"), str(tmp_path / "b.vti")
    g.dump_vti_async(f1, dx)
    s1 = g.download()
    g.step(3)
    g.dump_vti_async(f2, dx)
    s2 = g.download()
    g.step(1)
    g.dump_wait()
    for f, s in ((f1, s1), (f2, s2)):
        r = read_vti_raw(f)
        korn, conc = _vti_fields(s, d)
        assert r["shape"] == (d.nx, d.ny, d.nz) and r["dx"] == dx
        assert np.array_equal(r["korn"], korn), f
        assert np.array_equal(r["conc"], conc), f
    assert not np.array_equal(read_vti_raw(f1)["conc"], read_vti_raw(f2)["conc"])


def test_pfm_binary_vti_output(tmp_path):
    """`pfm` with micro output and vti_format=binary: one raw VTI per output time, written
    asynchronously; the last file holds the final state (also dumped in full precision)."""
    from phasefield_amd import read_vti_raw
    banner("pfm binary VTI output")
    assert os.path.exists(PFM), "phasefield_amd/pfm is not built"
    dump = str(tmp_path / "state.bin")
    path = configs.write_inputs(str(tmp_path), "fec_3d_256", nx=48, ny=40, nz=24, grains=4,
                                nuclei=3, steps=6, output_dt=0.001, micro_out=1,
                                extra={"vti_format": "binary", "dump_state": dump})
    r = subprocess.run([PFM, path], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = Model(path).get_string("directory")
    files = sorted(f for f in os.listdir(out) if f.endswith(".vti"))
    assert len(files) >= 2
    d = Model(path).desc()
    last = max(files, key=lambda f: int(f.rsplit("_", 1)[1][:-4]))
    v = read_vti_raw(os.path.join(out, last))
    korn, conc = _vti_fields(_read_dump(dump, d), d)
    assert np.array_equal(v["korn"], korn) and np.array_equal(v["conc"], conc)
