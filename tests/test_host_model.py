"""Host side of the drop-in boundary (CPU only): the reference's input dialect, the parameter
update, the seeded Voronoi draw and the writers."""
from __future__ import annotations

import os
import re
import shutil

import numpy as np
import pytest

import oracle
from phasefield_amd import Model, PhaseFieldError, configs, lib
from phasefield_amd._capi import PfDesc, PfHostState
from tests import helpers

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture
def ref_example(tmp_path):
    """The reference's example input pair (C++-openmp-cuda/build/log_file.txt, thermo_data.txt),
    kept verbatim as fixtures, under the names the main file expects."""
    shutil.copy(os.path.join(GOLDEN, "ref_example_log_file.txt"), tmp_path / "log_file.txt")
    shutil.copy(os.path.join(GOLDEN, "ref_example_thermo_data.txt"), tmp_path / "thermo_data.txt")
    os.chmod(tmp_path / "log_file.txt", 0o644)
    os.chmod(tmp_path / "thermo_data.txt", 0o644)
    return Model(str(tmp_path / "log_file.txt"))


def test_parse_reference_example(ref_example):
    m = ref_example
    assert m.get_string("directory") == "rex140k1_para1cs"
    assert (m.get_int("ix"), m.get_int("iy"), m.get_int("iz")) == (500, 500, 1)
    assert m.get_real("dx") == 0.15
    assert m.get_int("i_intf") == 5  # "5.00" read as an int
    assert m.get_real("cutoff") == 1e-5
    assert [m.get_int(f"num_grain{i}") for i in range(3)] == [8, 1, 0]
    assert [m.get_real(f"init_conc{i}") for i in range(3)] == [0.0063, 0.12, 0.1]
    assert m.get_int("n_temperature_points") == 2
    assert m.get_real("time") == 0.0 and m.get_real("end_time") == 4.6
    assert m.get_real("temperature") == 1003.0
    assert m.get_real("reduction") == 0.0
    assert m.get_int("num_init_iterations") == 10
    assert m.get_real("timestep_output") == 0.5
    assert m.get_int("is_micro_out") == 1
    assert m.get_real("num_nuclei") == 1.0
    assert m.get_real("mobility0_11") == 24e29 and m.get_real("mobilityQ_11") == 360000
    assert m.get_real("mobility0_12") == 200.e16 and m.get_real("mobility0_21") == 200.e16
    assert m.get_real("mobility0_33") == 0.0e-33 and m.get_real("mobilityQ_33") == 16000000
    assert m.get_real("sigma_11") == 0.76e-12 and m.get_real("sigma_23") == 0.94e-12
    assert m.get_real("sigma_33") == 0.3e-12
    assert (m.get_real("diff0_1"), m.get_real("diffQ_1")) == (0.1e8, 122500)
    assert (m.get_real("diff0_2"), m.get_real("diffQ_2")) == (0.15e8, 142000)
    assert m.get_string("thermo_file") == "thermo_data.txt"
    assert m.get_int("gpu_flag") == 1
    d = m.desc()
    assert (d.nx, d.ny, d.nz, d.depth) == (500, 500, 1, 6)


def test_reference_example_parameters_match_oracle(ref_example):
    """Update_parameter on the example files vs the oracle's restatement (bitwise)."""
    m = ref_example
    p = m.update_parameter()
    T = 1003.0
    mob0 = np.array([[m.get_real(f"mobility0_{i}{j}") for j in (1, 2, 3)] for i in (1, 2, 3)])
    mobQ = np.array([[m.get_real(f"mobilityQ_{i}{j}") for j in (1, 2, 3)] for i in (1, 2, 3)])
    sig = np.array([[m.get_real(f"sigma_{i}{j}") for j in (1, 2, 3)] for i in (1, 2, 3)])
    d0 = [m.get_real(f"diff0_{i}") for i in (1, 2, 3)]
    dq = [m.get_real(f"diffQ_{i}") for i in (1, 2, 3)]
    ent = np.array([[1, 2, 0., 0.2e-12], [1, 3, 0, 0.25e-12], [2, 3, 0, 1.0e-12]])
    para = np.array([[1, 2, 0., -1e-4, 0.11], [2, 1, 0., -5e-3, 5.395], [1, 3, 0, 1e-2, 0.001],
                     [3, 1, 0, 1e-3, 0.01], [2, 3, 0, -1e-3, 0.1], [3, 2, 0, -9e-3, 0.1]])
    o, pdt, ddt = oracle.update_parameter(T, 0.15, 2, mob0, mobQ, sig, d0, dq, ent, para)
    for name in ("mobility", "sigma", "entropy", "eq_conc", "eq_slope"):
        assert np.array_equal(np.ctypeslib.as_array(getattr(p, name)),
                              np.ctypeslib.as_array(getattr(o, name))), name
    assert np.array_equal(np.ctypeslib.as_array(p.part_k)[0], np.ctypeslib.as_array(o.part_k)[0])
    assert np.array_equal(np.ctypeslib.as_array(p.part_c)[0], np.ctypeslib.as_array(o.part_c)[0])
    assert np.array_equal(np.ctypeslib.as_array(p.diffusivity)[0],
                          np.ctypeslib.as_array(o.diffusivity)[0])
    assert p.pfm_dt == pdt
    assert p.n_sub[0] == max(int(np.ceil(pdt / ddt)), 1)
    assert p.dt_d[0] == pdt / p.n_sub[0]
    # time advanced by one phase-field step (phase_field.cpp:314)
    assert m.get_real("time") == pdt


def test_workload_parameters(tmp_path):
    """The named workloads: 200 carbon sub-steps per step in 2D and 3D, dt values of the
    physics (closed forms: pfm_dt = dx^2/dims*0.2/(M sigma))."""
    for wl, dims in (("fec_2d_256", 2), ("fec_3d_256", 3)):
        m = Model(configs.write_inputs(str(tmp_path / wl), wl))
        p = m.update_parameter()
        assert p.dims == dims
        assert p.n_sub[0] == 200
        assert p.pfm_dt == 0.1 * 0.1 / dims * 0.2 / (1e12 * 0.5e-12)
        assert p.dt_d[0] == p.pfm_dt / 200
        assert abs(p.part_k[0][0][1] - 0.04) < 1e-12
        assert abs(p.eq_conc[0][1] - 0.45) < 1e-12 and abs(p.eq_conc[1][0] - 0.018) < 1e-12
        # austenite with carbon above c_gamma_eq has a positive driving force against ferrite
        dg = p.entropy[0][1] * (0.8 - p.eq_conc[0][1]) / p.eq_slope[0][1]
        assert dg > 0
    m = Model(configs.write_inputs(str(tmp_path / "mn"), "fecmn_3d_512"))
    p = m.update_parameter()
    assert p.n_sub[1] == 1 and p.dt_d[1] == p.pfm_dt
    assert p.part_k[1][1][0] == 2.0 and p.part_k[1][0][1] == 0.5
    assert p.diffusivity[1][1] == 1e-4 and p.diffusivity[1][0] == 1e-6
    d = m.desc()
    assert (d.nsolute, d.precision, d.depth) == (2, 1, 4)


def test_temperature_path(tmp_path):
    path = configs.write_inputs(str(tmp_path), "fec_3d_256")
    txt = open(path).read().replace("2\n0 1013.0\n1e6 1013.0", "3\n0 1000\n10 1100\n20 1050")
    open(path, "w").write(txt)
    m = Model(path)
    assert m.temperature(0.0) == 1000.0
    assert m.temperature(5.0) == 1050.0
    assert m.temperature(15.0) == 1075.0
    assert m.temperature(25.0) == 1050.0  # held at the last point
    assert m.temperature(-1.0) == 1000.0


def test_voronoi_seeds(tmp_path):
    m = Model(configs.write_inputs(str(tmp_path), "fec_3d_256", nx=40, ny=30, nz=20, grains=7))
    s = m.voronoi_seeds().reshape(-1, 3)
    assert s.shape == (7, 3)
    for d, n in enumerate((40, 30, 20)):
        assert np.all(s[:, d] >= 1) and np.all(s[:, d] <= n)
    assert np.array_equal(s, m.voronoi_seeds().reshape(-1, 3))  # reproducible
    # the draw is std::mt19937 seeded with the grain count, mapped to [1, n]
    first = oracle.lib().orc_mt19937_nth(7, 1)
    assert s[0, 0] == first / 4294967295.0 * 39 + 1


def test_bad_input_reports_error(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("OUTPUT\nx\n")
    with pytest.raises(PhaseFieldError):
        Model(str(p))
    with pytest.raises(PhaseFieldError):
        Model(str(tmp_path / "missing.txt"))


def test_vti_writer(tmp_path):
    nx, ny, nz, K = 3, 2, 1, 4
    idx = np.zeros(nx * ny * nz * K, np.int32)
    idx[0::K] = [200001, 100002, 200001, 200002, 100001, 200003]
    idx[K + 1] = 200001  # cell 1 has a second slot -> korn 0
    conc = np.linspace(0.02, 0.8, nx * ny * nz)
    path = str(tmp_path / "f.vti")
    rc = lib().pf_write_vti(path.encode(), nx, ny, nz, K, 0.1,
                            idx.ctypes.data_as(np.ctypeslib.ctypes.POINTER(np.ctypeslib.ctypes.c_int32)),
                            conc.ctypes.data_as(np.ctypeslib.ctypes.POINTER(np.ctypeslib.ctypes.c_double)))
    assert rc == 0
    txt = open(path).read()
    assert 'WholeExtent="1 3 1 2 1 1"' in txt and 'Spacing="0.1 0.1 0.1"' in txt
    korn = re.search(r'Name="korn" format="ascii">\n(.*?)\n', txt).group(1).split()
    assert korn == ["2", "0", "2", "2", "1", "2"]
    vals = re.search(r'Name="conc" format="ascii">\n(.*?)</DataArray>', txt).group(1).split()
    assert [float(v) for v in vals] == [float(f"{c:.3g}") for c in conc]


def test_vti_raw_writer_roundtrip(tmp_path, rng):
    """Binary VTI (raw appended): header extent/spacing as the ASCII writer's, arrays back
    bit for bit, block sizes in the UInt64 headers, and a complete XML tail."""
    from phasefield_amd import read_vti_raw, write_vti_raw
    shape = (7, 5, 3)
    n = 7 * 5 * 3
    korn = rng.integers(0, 3, n).astype(np.int32)
    conc = rng.random(n).astype(np.float32)
    path = str(tmp_path / "f.vti")
    write_vti_raw(path, shape, 0.1, korn, conc)
    txt = open(path, "rb").read()
    assert b'WholeExtent="1 7 1 5 1 3"' in txt and b'header_type="UInt64"' in txt
    assert b'Name="conc" format="appended" offset="%d"' % (8 + 4 * n) in txt
    r = read_vti_raw(path)
    assert r["shape"] == shape and r["dx"] == 0.1
    assert np.array_equal(r["korn"], korn) and np.array_equal(r["conc"], conc)
    with pytest.raises(ValueError):
        write_vti_raw(path, (7, 5, 4), 0.1, korn, conc)


def test_state_dump_roundtrip(tmp_path, rng):
    import ctypes as C
    nx, ny, nz, K = 5, 4, 3, 4
    ids, val, phase, conc = helpers.random_state(rng, nx, ny, nz, K)
    d = PfDesc(nx, ny, nz, K, 2, 1, 0, 0, 0)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    s = PfHostState(ids.ctypes.data_as(C.POINTER(C.c_int32)), dp(val), dp(phase), dp(conc), None,
                    None)
    path = str(tmp_path / "state.bin")
    assert lib().pf_write_state(path.encode(), C.byref(d), 1.25, 17, C.byref(s)) == 0
    out = [np.zeros_like(ids), np.zeros_like(val), np.zeros_like(phase), np.zeros_like(conc)]
    s2 = PfHostState(out[0].ctypes.data_as(C.POINTER(C.c_int32)), dp(out[1]), dp(out[2]),
                     dp(out[3]), None, None)
    t, st = C.c_double(), C.c_int64()
    assert lib().pf_read_state(path.encode(), C.byref(d), C.byref(t), C.byref(st), C.byref(s2)) == 0
    assert t.value == 1.25 and st.value == 17
    for a, b in zip((ids, val, phase, conc), out):
        assert np.array_equal(a, b)
    d_bad = PfDesc(nx, ny, nz + 1, K, 2, 1, 0, 0, 0)
    assert lib().pf_read_state(path.encode(), C.byref(d_bad), C.byref(t), C.byref(st),
                               C.byref(s2)) != 0


def test_kinetics_writer(tmp_path):
    m = Model(configs.write_inputs(str(tmp_path), "fec_2d_256"))
    out = m.get_string("directory")
    assert m.output_due(True)
    assert m.write_kinetics({"phase_frac": [0.03, 0.97, 0.0]}, 0.0, True)
    rows = open(os.path.join(out, "fraction_time.txt")).read().split("\n")
    assert rows[0].split() == ["0", "1013", "0.03", "0.97", "0", "0"]
