"""GPU: state handling and the workloads' less common paths against the CPU oracle.

- nucleation on an uploaded state that already holds austenite grains (grain codes stay distinct);
- checkpoints restored into a second domain (grain-code tables travel with the blob);
- the 1024^3 Fe-C-Mn workload's combination (mixed form, two solutes, no-flux z) at a small size;
- a moving temperature with Arrhenius mobility and diffusivities (parameters change every step);
- the reference's NaN when every updated phi of a cell falls below the cutoff (solver.cpp:292-296),
  and the non-finite check of the reductions;
- late times of the 256^3 Fe-C workload (steps 1996-2000 of 2000).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from phasefield_amd import Domain, Model, PhaseFieldError, configs
from phasefield_amd._capi import PF_ERR_NONFINITE, PfParams
from tests import helpers
from tests.test_gpu_parity import assert_same, banner, run_pair

pytestmark = pytest.mark.gpu


def _oracle_state(m, p):
    od = helpers.oracle_initial_state(m, p)
    s = od.get_state()
    return s["index"], s["value"], s["phase"], s["conc"]


def test_nucleate_after_upload_keeps_phase1_grains(tmp_path):
    """Nucleus q takes the reference's id 100000 + q (nucleation.cpp:29).  Upload a state whose
    austenite grains are 100002, 100007 and 100009, nucleate three more (100001, 100002, 100003)
    on the device and on the oracle, then step: 100002 is one grain in both (equal ids), the others
    stay distinct, and index maps and fields agree bitwise."""
    banner("nucleate after upload")
    m = helpers.small_model(str(tmp_path), "fec_3d_256", nx=48, ny=40, nz=24, grains=6, nuclei=3)
    p = helpers.model_params(m)
    ids, val, phase, conc = _oracle_state(m, p)
    relabel = {100001: 100002, 100002: 100007, 100003: 100009}
    ids = np.array([relabel.get(int(g), int(g)) for g in ids], np.int32)
    assert {100002, 100007, 100009} <= set(np.unique(ids))
    g = Domain(48, 40, 24, depth=4)
    g.set_params(p)
    g.upload(ids, val, phase, conc)
    g.nucleate(3, 4.0, (0.8,), 99)
    o = oracle.OracleDomain(48, 40, 24, 4, 1)
    o.set_params(p)
    o.set_state(ids, val, phase, conc)
    o.nucleate(3, 4.0, (0.8,), 99)
    assert_same(g.download(), o.get_state(), 4, "after nucleation")
    assert {100001, 100003, 100007, 100009} <= set(np.unique(g.download()["index"]))
    g.step(2)
    o.step(2)
    assert_same(g.download(), o.get_state(), 4, "nucleation + 2 steps")
    g.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_checkpoint_restores_into_fresh_domain(tmp_path, precision):
    """A checkpoint carries the grain-code tables: restored into a second domain of the same shape
    (which never saw an upload), downloads, the rex fraction and continued steps match the first
    domain; a blob of the wrong length or from another shape is refused."""
    banner(f"checkpoint into a fresh domain ({precision})")
    m = helpers.small_model(str(tmp_path), "fecmn_3d_512", nx=40, ny=24, nz=20, grains=5,
                            nuclei=3, precision=precision)
    p = helpers.model_params(m)
    st = _oracle_state(m, p)
    a = Domain(40, 24, 20, depth=4, nsolute=2, precision=precision)
    a.set_params(p)
    a.upload(*st)
    a.step(1)
    blob = a.checkpoint()
    b = Domain(40, 24, 20, depth=4, nsolute=2, precision=precision)
    b.set_params(p)
    b.restore(blob)
    ra, rb = a.download(), b.download()
    assert_same(ra, rb, 4, "restored")
    assert np.array_equal(ra["conc_comp"], rb["conc_comp"])
    assert a.rex_fraction() == b.rex_fraction()
    a.step(2)
    b.step(2)
    ra, rb = a.download(), b.download()
    assert_same(ra, rb, 4, "restored + 2 steps")
    assert np.array_equal(ra["conc_comp"], rb["conc_comp"])
    with pytest.raises(PhaseFieldError):
        b.restore(blob[:-1])
    c = Domain(40, 24, 18, depth=4, nsolute=2, precision=precision)
    c.set_params(p)
    with pytest.raises(PhaseFieldError):
        c.restore(blob + bytes(len(blob) // 10))
    for d in (a, b, c):
        d.close()


@pytest.mark.parametrize("form", [oracle.ORC_FP64, oracle.ORC_MIXED])
def test_fecmn_1024_workload_small_bitwise(tmp_path, form):
    """The 1024^3 workload's combination at 64 x 48 x 40 (5 grains, 4 nuclei): Fe-C-Mn, no-flux z,
    K = 4, 3 full steps (sweep + 200 carbon sub-steps + 1 Mn update each) in the FP64 form and the
    mixed form, each against the oracle's matching form, bitwise (compensation codes included)."""
    banner(f"fecmn_3d_1024 small, form {form}")
    m = helpers.small_model(str(tmp_path), "fecmn_3d_1024", nx=64, ny=48, nz=40, grains=5,
                            nuclei=4)
    p = helpers.model_params(m)
    d = m.desc()
    assert (d.nsolute, d.bc_z, d.depth) == (2, 1, 4)
    assert p.n_sub[0] == 200 and p.n_sub[1] == 1
    st = _oracle_state(m, p)
    gs, os_ = run_pair((64, 48, 40), 4, 2, 2, 1, form, st, p,
                       lambda g, q: g.step(3), lambda o, q: o.step(3))
    assert_same(gs, os_, 4, "fecmn_3d_1024 small")


@pytest.mark.parametrize("form", [oracle.ORC_FP64, oracle.ORC_MIXED])
def test_temperature_path_with_arrhenius_bitwise(tmp_path, form):
    """Temperature 1013 K -> 995 K over the first 6 ms, then towards 1030 K, with Q != 0 for the
    mobility (140 kJ/mol) and both diffusivities (80 / 140 kJ/mol): every step takes the host
    model's Update_parameter (phase_field.cpp:232-242, 307-315), which matches the oracle's
    restatement bitwise, and the device and the oracle step with those values; the sub-step count
    and the partition tables change along the path, so the driving force and the first sub-step
    read partitions of the previous step's tables where the reference's stored value_conc1 still
    holds them.  Fields and maps agree bitwise after every step."""
    banner(f"temperature path + Arrhenius, form {form}")
    R, T0 = 8.314, 1013.0
    QM, QA, QG = 140e3, 80e3, 140e3
    wl = dict(nx=48, ny=32, nz=24, grains=5, nuclei=4,
              temperature=[(0.0, T0), (0.006, 995.0), (1.0, 1030.0)],
              mobility=(1e12 / np.exp(-QM / (R * T0)), QM),
              diff_alpha=(100.0 / np.exp(-QA / (R * T0)), QA),
              diff_gamma=(0.5 / np.exp(-QG / (R * T0)), QG))
    m = helpers.small_model(str(tmp_path), "fec_3d_256", **wl)
    p0 = m.update_parameter()
    st = _oracle_state(m, p0)
    mob0 = np.array([[m.get_real(f"mobility0_{i}{j}") for j in (1, 2, 3)] for i in (1, 2, 3)])
    mobQ = np.array([[m.get_real(f"mobilityQ_{i}{j}") for j in (1, 2, 3)] for i in (1, 2, 3)])
    sig = np.array([[m.get_real(f"sigma_{i}{j}") for j in (1, 2, 3)] for i in (1, 2, 3)])
    d0 = [m.get_real(f"diff0_{i}") for i in (1, 2, 3)]
    dq = [m.get_real(f"diffQ_{i}") for i in (1, 2, 3)]
    assert mobQ[0, 1] == QM and dq[0] == QG and dq[1] == QA
    ent = np.array([[1, 2, 0.0, -0.2e-12], [1, 3, 0.0, 0.0], [2, 3, 0.0, 0.0]])
    text = configs.thermo_text().split("#PARA wt%\n")[1]
    para = np.array([[float(x) for x in ln.split()] for ln in text.strip().splitlines()])
    # the state is partitioned with the tables in force when it is set (Construct's parameters)
    g = Domain(48, 32, 24, depth=4, precision="mixed" if form == oracle.ORC_MIXED else "fp64")
    g.set_params(p0)
    g.upload(*st)
    o = oracle.OracleDomain(48, 32, 24, 4, 1, form)
    o.set_params(p0)
    o.set_state(*st)
    n_subs, temps = set(), []
    for k in range(10):
        T = m.temperature(m.get_real("time"))
        p = m.update_parameter()
        q, pdt, ddt = oracle.update_parameter(T, p.dx, p.dims, mob0, mobQ, sig, d0, dq, ent, para)
        for name in ("mobility", "eq_conc", "eq_slope", "entropy"):
            assert np.array_equal(np.ctypeslib.as_array(getattr(p, name)),
                                  np.ctypeslib.as_array(getattr(q, name))), (k, name)
        assert np.array_equal(np.ctypeslib.as_array(p.diffusivity)[0],
                              np.ctypeslib.as_array(q.diffusivity)[0])
        assert p.pfm_dt == pdt and p.n_sub[0] == max(int(np.ceil(pdt / ddt)), 1)
        n_subs.add(p.n_sub[0])
        temps.append(T)
        g.set_params(p)
        g.step(1)
        o.set_params(p)
        o.step(1)
        gs, os_ = g.download(), o.get_state()
        assert_same(gs, os_, 4, f"T path step {k + 1} (T = {T:.2f} K)")
        assert np.array_equal(gs["conc_comp"], os_["conc_comp"])
    assert len(n_subs) >= 3, n_subs
    assert min(temps) < 1000.0 and temps[-1] > min(temps)
    g.close()


def _two_grain_uniform(N, K):
    ids = np.zeros((N, K), np.int32)
    val = np.zeros((N, K))
    ids[:, 0], val[:, 0] = 100001, 0.5
    ids[:, 1], val[:, 1] = 200001, 0.5
    phase = np.zeros((N, 3))
    phase[:, 0] = phase[:, 1] = 0.5
    return ids.reshape(-1), val.reshape(-1), phase.reshape(-1)


@pytest.mark.parametrize("precision,nphase", [("fp64", 3), ("mixed", 3), ("fp64", 2)])
def test_all_phi_below_cutoff_gives_the_reference_nan(base_params_any, precision, nphase):
    """Both driving forces negative: in every cell both grains fall below the cutoff, so the sum is
    0 and the reference's renormalisation `*= 1./sum` (solver.cpp:292-296) writes NaN.  The device
    reproduces that state exactly (NaN in the same places, empty maps), and the reductions refuse it
    with PF_ERR_NONFINITE.  A domain carrying two phases stores no phase-3 fraction (the
    reference's 0 * inf there is NaN too); its two stored phases are compared."""
    banner(f"sum = 0 -> NaN ({precision}, P = {nphase})")
    nx, ny, nz, K = 16, 12, 8, 4
    N = nx * ny * nz
    q = oracle.as_orc_params(base_params_any)
    for i in range(3):
        for j in range(3):
            q.eq_conc[i][j], q.eq_slope[i][j] = 0.0, 1.0
            for s in range(2):
                q.part_k[s][i][j], q.part_c[s][i][j] = 1.0, 0.0
    q.entropy[0][1] = q.entropy[1][0] = -1e-9
    ids, val, phase = _two_grain_uniform(N, K)
    form = oracle.ORC_MIXED if precision == "mixed" else oracle.ORC_FP64
    gp = PfParams.from_buffer_copy(q)
    g = Domain(nx, ny, nz, depth=K, nphase=nphase, precision=precision)
    g.set_params(gp)
    g.upload(ids, val, phase, np.full(N, 1.0))
    g.sweep()
    o = oracle.OracleDomain(nx, ny, nz, K, 1, form)
    o.set_params(q)
    o.set_state(ids, val, phase, np.full(N, 1.0))
    o.sweep()
    gs, os_ = g.download(), o.get_state()
    assert np.all(os_["index"] == 0) and np.all(np.isnan(os_["value"]))
    assert np.array_equal(gs["index"], os_["index"])
    assert np.array_equal(gs["value"], os_["value"], equal_nan=True)
    ph_g, ph_o = gs["phase"].reshape(N, 3), os_["phase"].reshape(N, 3)
    assert np.array_equal(ph_g[:, :nphase], ph_o[:, :nphase], equal_nan=True)
    assert np.all(np.isnan(ph_g[:, :nphase]))
    with pytest.raises(PhaseFieldError) as e:
        g.reduce()
    assert e.value.code == PF_ERR_NONFINITE
    g.close()


def test_unstable_diffusion_reaches_nonfinite(base_params_any, rng):
    """A diffusion sub-step 1000x the workload's (r = D dt/dx^2 ~ 67 in ferrite) grows the
    concentration's shortest wave by a factor ~800 per sub-step, past the FP64 range within about
    110 of the 200 sub-steps; pf_reduce then reports PF_ERR_NONFINITE instead of statistics."""
    banner("unstable diffusion -> PF_ERR_NONFINITE")
    nx, ny, nz, K = 24, 16, 12, 4
    ids, val, phase, conc = helpers.random_state(rng, nx, ny, nz, K)
    g = Domain(nx, ny, nz, depth=K)
    g.set_params(base_params_any)
    g.upload(ids, val, phase, conc)
    r = g.reduce()
    assert r["n_nonfinite"] == 0
    p = base_params_any
    g.diffuse(0, 200, p.dt_d[0] * 1000)
    with pytest.raises(PhaseFieldError) as e:
        g.reduce()
    assert e.value.code == PF_ERR_NONFINITE
    g.close()


@pytest.fixture(scope="module")
def base_params_any(tmp_path_factory):
    m = helpers.small_model(str(tmp_path_factory.mktemp("st")), "fec_3d_256", nx=16, ny=16, nz=16,
                            grains=4, nuclei=2)
    return helpers.model_params(m)


@pytest.mark.slow
def test_fec_3d_256_late_steps_bitwise(tmp_path):
    """The 256^3 Fe-C parity workload to its horizon: the device runs 1995 steps from the seeded
    state, its state is handed to the oracle, and both run steps 1996-2000; active-set maps, phi,
    value_phase and concentrations agree bitwise at step 2000 (grown austenite, long lists)."""
    banner("fec_3d_256: steps 1996-2000 vs oracle")
    m = Model(configs.write_inputs(str(tmp_path), "fec_3d_256"))
    p = m.update_parameter()
    d = m.desc()
    g = Domain.from_desc(d)
    g.set_params(p)
    g.construct(m.voronoi_seeds(), m.get_real("reduction"), m.get_int("num_init_iterations"),
                [m.get_real(f"init_conc{i}") for i in range(3)])
    g.nucleate(m.get_int("num_nuclei"), m.get_real("nucleus_radius"),
               (m.get_real("nucleus_conc"), -1.0), m.get_int("nucleation_seed"))
    f0 = g.reduce()["phase_frac"][0]
    for k in range(1995):
        g.step(1)
        if k % 500 == 499:
            g.synchronize()
            banner(f"  step {k + 1}")
    s = g.download()
    f1995 = g.reduce()["phase_frac"][0]
    assert f1995 > f0
    o = oracle.OracleDomain(d.nx, d.ny, d.nz, d.depth, 1)
    o.set_params(oracle.as_orc_params(p))
    o.set_state(s["index"], s["value"], s["phase"], s["conc"])
    g.step(5)
    o.step(5)
    assert_same(g.download(), o.get_state(), d.depth, "256^3 steps 1996-2000")
    g.close()

def test_fec_3d_256_late_steps_mixed_bitwise(tmp_path):
    """The same horizon in the mixed form: the device runs 1995 mixed steps from the seeded state;
    that state (FP32 values and compensation codes) is uploaded into a fresh mixed domain and handed
    to the oracle's mixed form, whose rounding keys both count from the hand-over, and both run
    steps 1996-1998; everything, compensations included, agrees bitwise."""
    banner("fec_3d_256 mixed: steps 1996-1998 vs oracle")
    m = Model(configs.write_inputs(str(tmp_path), "fec_3d_256"))
    p = m.update_parameter()
    d = m.desc()
    f = Domain.from_desc(d)
    f.set_params(p)
    f.construct(m.voronoi_seeds(), m.get_real("reduction"), m.get_int("num_init_iterations"),
                [m.get_real(f"init_conc{i}") for i in range(3)])
    f.nucleate(m.get_int("num_nuclei"), m.get_real("nucleus_radius"),
               (m.get_real("nucleus_conc"), -1.0), m.get_int("nucleation_seed"))
    s0 = f.download()
    f.close()
    d.precision = 1
    g = Domain.from_desc(d)
    g.set_params(p)
    g.upload(s0["index"], s0["value"], s0["phase"], s0["conc"])
    f0 = g.reduce()["phase_frac"][0]
    for k in range(1995):
        g.step(1)
        if k % 500 == 499:
            g.synchronize()
            banner(f"  step {k + 1}")
    s = g.download()
    assert g.reduce()["phase_frac"][0] > f0
    assert np.any(s["conc_comp"] != 0.0)
    g.close()
    h = Domain.from_desc(d)
    h.set_params(p)
    h.upload(s["index"], s["value"], s["phase"], s["conc"], s["conc_comp"])
    o = oracle.OracleDomain(d.nx, d.ny, d.nz, d.depth, 1, oracle.ORC_MIXED)
    o.set_params(oracle.as_orc_params(p))
    o.set_state(s["index"], s["value"], s["phase"], s["conc"])
    o.set_comp(s["conc_comp"])
    h.step(3)
    o.step(3)
    hs, os_ = h.download(), o.get_state()
    assert_same(hs, os_, d.depth, "256^3 mixed steps 1996-1998")
    assert np.array_equal(hs["conc_comp"], os_["conc_comp"])
    h.close()
