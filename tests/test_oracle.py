"""Pins the CPU oracle (oracle/pf_oracle.cpp) with checks that do not restate its formulas:
closed forms, conservation laws, symmetries, brute force on tiny inputs, the C++ standard's
generator value, and finite differences for the dual-number derivatives.

The reference ships no tests, golden outputs or fixtures for this path (its only data files are
the example inputs, used by test_host_model.py), so these checks are the oracle's anchor.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests import helpers

pytestmark = pytest.mark.usefixtures("orc")


@pytest.fixture(scope="module")
def model3d(tmp_path_factory):
    d = tmp_path_factory.mktemp("m3")
    return helpers.small_model(str(d), "fec_3d_256", nx=16, ny=12, nz=10, grains=5, nuclei=3,
                               steps=3)


@pytest.fixture(scope="module")
def params3d(model3d):
    return helpers.model_params(model3d)


def test_mt19937_standard_value():
    # C++ [rand.predef]: the 10000th consecutive output of a default-constructed mt19937
    assert oracle.lib().orc_mt19937_nth(5489, 10000) == 4123659995


def bulk_state(nx, ny, nz, K, grain=200001):
    N = nx * ny * nz
    ids = np.zeros((N, K), np.int32)
    val = np.zeros((N, K))
    ids[:, 0] = grain
    val[:, 0] = 1.0
    phase = np.zeros((N, 3))
    phase[:, 1 if grain >= 200000 else 0] = 1.0
    return ids.reshape(-1), val.reshape(-1), phase.reshape(-1)


@pytest.mark.parametrize("shape,mode", [((32, 16, 8), (1, 2, 1)), ((24, 20, 1), (3, 1, 0))])
def test_diffusion_fourier_mode_decay(params3d, shape, mode):
    """Single-phase ferrite: c <- c + r*sum(c_nb - c) with r = D*dt/dx^2, so a Fourier mode decays
    by exactly g = 1 - 4r*sum_d sin^2(pi m_d/N_d) per sub-step."""
    nx, ny, nz = shape
    K = 4
    p = oracle.as_orc_params(params3d)
    p.dims = 3 if nz > 1 else 2
    od = oracle.OracleDomain(nx, ny, nz, K, 1)
    od.set_params(p)
    ids, val, phase = bulk_state(nx, ny, nz, K)
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    arg = 2 * np.pi * (mode[0] * x / nx + mode[1] * y / ny + mode[2] * z / nz)
    c0 = 0.5 + 0.1 * np.cos(arg)
    c0 = c0.transpose(2, 1, 0).reshape(-1)  # cell = x + nx*(y + ny*z)
    od.set_state(ids, val, phase, c0)
    D = p.diffusivity[0][1]
    dt_d = 6.0e-6
    nsub = 50
    od.diffuse(0, nsub, dt_d)
    r = D * dt_d / (p.dx * p.dx)
    g = 1.0 - 4.0 * r * sum(np.sin(np.pi * m / n) ** 2 for m, n in zip(mode, shape))
    expect = 0.5 + (c0 - 0.5) * g ** nsub
    got = od.get_state()["conc"]
    assert np.max(np.abs(got - expect)) < 1e-13


def test_diffusion_conserves_mass(params3d, rng):
    """The flux between two cells is antisymmetric, so sum(c) is invariant up to rounding."""
    nx, ny, nz, K = 12, 10, 8, 4
    ids, val, phase, conc = helpers.random_state(rng, nx, ny, nz, K)
    od = oracle.OracleDomain(nx, ny, nz, K, 1)
    od.set_params(params3d)
    od.set_state(ids, val, phase, conc)
    m0 = od.reduce()["mass"][0]
    od.diffuse(0, 200, params3d.dt_d[0])
    m1 = od.reduce()["mass"][0]
    assert abs(m1 - m0) / m0 < 1e-13


def test_partition_identities(params3d):
    """Single-phase cells return the cell's concentration for their own phase; a two-phase cell
    splits c so that phi_1 c_1 + phi_2 c_2 = c and c_2 = k(1,2) c_1 (c(1,2) = 0 up to rounding)."""
    K = 4
    od = oracle.OracleDomain(3, 1, 1, K, 1)
    p = oracle.as_orc_params(params3d)
    p.dims = 1
    od.set_params(p)
    ids = np.zeros(3 * K, np.int32)
    val = np.zeros(3 * K)
    ids[0], val[0] = 200001, 1.0
    ids[K], val[K] = 100001, 1.0
    ids[2 * K], ids[2 * K + 1], val[2 * K], val[2 * K + 1] = 100001, 200001, 0.3, 0.7
    phase = np.array([0, 1, 0, 1, 0, 0, 0.3, 0.7, 0], float)
    c = np.array([0.02, 0.8, 0.1])
    od.set_state(ids, val, phase, c)
    pc = od.get_state()["phase_conc"].reshape(3, 3)
    assert pc[0, 1] == 0.02 and pc[1, 0] == 0.8
    k12 = p.part_k[0][0][1]
    assert abs(0.3 * pc[2, 0] + 0.7 * pc[2, 1] - 0.1) < 1e-15
    assert abs(pc[2, 1] - k12 * pc[2, 0]) < 1e-15
    assert abs(k12 - 0.04) < 1e-12


def _state_from(od):
    s = od.get_state()
    return s["index"].copy(), s["value"].copy(), s["phase"].copy(), s["conc"].copy()


@pytest.fixture(scope="module")
def setup_state(model3d, params3d):
    od = helpers.oracle_initial_state(model3d, params3d)
    return _state_from(od), model3d.desc()


def test_pf_sweep_invariants(setup_state, params3d):
    """After a sweep: sum of slots = 1 within rounding, 0 <= phi <= 1, sorted descending,
    empty slots have phi = 0, and single-grain cells are copied bitwise."""
    (ids, val, phase, conc), d = setup_state
    K = d.depth
    od = oracle.OracleDomain(d.nx, d.ny, d.nz, K, 1)
    od.set_params(params3d)
    od.set_state(ids, val, phase, conc)
    od.sweep()
    s = od.get_state()
    v = s["value"].reshape(-1, K)
    i = s["index"].reshape(-1, K)
    assert np.all(v >= 0) and np.all(v <= 1)
    assert np.max(np.abs(v.sum(1) - 1)) < 4e-16 * K
    assert np.all(np.diff(v, axis=1) <= 0)
    assert np.all(v[i == 0] == 0)
    # a cell whose own slots and neighbours' slots all hold one grain is copied unchanged
    i0 = ids.reshape(d.nz, d.ny, d.nx, K)
    single = (i0[..., 1] == 0)
    for ax in (0, 1, 2):
        for sh in (1, -1):
            nb = np.roll(i0, sh, axis=ax)
            single &= (nb[..., 0] == i0[..., 0]) & (nb[..., 1] == 0)
    single = single.reshape(-1)
    assert single.sum() > 0
    assert np.array_equal(v[single], val.reshape(-1, K)[single])
    assert np.array_equal(s["phase"].reshape(-1, 3)[single], phase.reshape(-1, 3)[single])


def _run(d, p, st, n_steps=1):
    od = oracle.OracleDomain(d.nx, d.ny, d.nz, d.depth, 1)
    od.set_params(p)
    od.set_state(*st)
    od.step(n_steps)
    return od.get_state()


def test_relabel_invariance(setup_state, params3d):
    """Grain ids only enter through equality tests: a permutation of the ids within each phase
    gives bitwise the same fields and the permuted id maps."""
    (ids, val, phase, conc), d = setup_state
    p = oracle.as_orc_params(params3d)
    p.n_sub[0] = 20
    a = _run(d, p, (ids, val, phase, conc), 2)
    uniq = np.unique(ids[ids > 0])
    rng = np.random.default_rng(7)
    perm = {}
    for base in (100000, 200000):
        grp = uniq[(uniq > base) & (uniq < base + 100000)]
        for g, h in zip(grp, rng.permutation(grp)):
            perm[int(g)] = int(h)
    ids2 = np.array([perm.get(int(g), 0) for g in ids], np.int32)
    b = _run(d, p, (ids2, val, phase, conc), 2)
    assert np.array_equal(a["value"], b["value"])
    assert np.array_equal(a["conc"], b["conc"])
    assert np.array_equal(np.array([perm.get(int(g), 0) for g in a["index"]]), b["index"])


def test_periodic_shift_invariance(setup_state, params3d):
    """On a periodic domain the step commutes with a cyclic shift of the whole state (bitwise)."""
    (ids, val, phase, conc), d = setup_state
    K = d.depth
    p = oracle.as_orc_params(params3d)
    p.n_sub[0] = 20
    a = _run(d, p, (ids, val, phase, conc), 2)

    def shift(arr, w):
        return np.roll(arr.reshape(d.nz, d.ny, d.nx, w), (3, -2, 5), axis=(0, 1, 2)).reshape(-1)

    b = _run(d, p, (shift(ids, K), shift(val, K), shift(phase, 3), shift(conc, 1)), 2)
    assert np.array_equal(shift(a["index"], K), b["index"])
    assert np.array_equal(shift(a["value"], K), b["value"])
    assert np.array_equal(shift(a["conc"], 1), b["conc"])


def test_transpose_symmetry(tmp_path, params3d):
    """Swapping x and y permutes the stencil faces, which only reorders sums: after one step the
    fields agree to rounding.  (Over several steps exact ties in Sort_xyz's first-max rule are
    broken by the union order, which the swap changes, so only one step is compared.)"""
    m = helpers.small_model(str(tmp_path), "fec_3d_256", nx=12, ny=12, nz=6, grains=4, nuclei=2)
    p = helpers.model_params(m)
    od = helpers.oracle_initial_state(m, p)
    ids, val, phase, conc = _state_from(od)
    d = m.desc()
    K = d.depth
    pp = oracle.as_orc_params(p)
    pp.n_sub[0] = 10
    a = _run(d, pp, (ids, val, phase, conc), 1)

    def tr(arr, w):
        return arr.reshape(d.nz, d.ny, d.nx, w).transpose(0, 2, 1, 3).reshape(-1)

    b = _run(d, pp, (tr(ids, K), tr(val, K), tr(phase, 3), tr(conc, 1)), 1)
    assert np.max(np.abs(tr(a["value"], K) - b["value"])) < 1e-14
    assert np.max(np.abs(tr(a["conc"], 1) - b["conc"])) < 1e-14


def test_tessellation_brute_force(rng):
    """Nearest seed under per-axis periodic distance with the reduction stretch, lowest index on
    ties, against a vectorised numpy evaluation."""
    nx, ny, nz, G = 17, 11, 7, 9
    seeds = np.empty((G, 3))
    seeds[:, 0] = rng.uniform(1, nx, G)
    seeds[:, 1] = rng.uniform(1, ny, G)
    seeds[:, 2] = rng.uniform(1, nz, G)
    seeds[3] = seeds[1]  # an exact tie: the lower index must win
    for red in (0.0, 0.3):
        got = oracle.tessellation(nx, ny, nz, seeds.reshape(-1), red)
        k, j, i = np.meshgrid(np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1),
                              indexing="ij")
        d2 = np.empty((G,) + i.shape)
        for n in range(G):
            dd = []
            for coord, ext, s in ((i, nx, seeds[n, 0]), (j, ny, seeds[n, 1]), (k, nz, seeds[n, 2])):
                a = np.abs(s - coord)
                dd.append(np.where(a < 0.5 * ext, a, ext - a))
            d2[n] = (dd[0] / (1 - red)) ** 2 + (dd[1] * (1 - red)) ** 2 + dd[2] ** 2
        expect = np.argmin(d2, axis=0).reshape(-1) + 1
        assert np.array_equal(got, expect)
        assert not np.any(got == 4)


def test_update_parameter_closed_forms(model3d):
    """Q = 0 gives M = M0 and D = D0 exactly; pfm_dt = dx^2/dims*0.2/max(M*sigma);
    diff_dt = dx^2/max(D)/dims*0.2; 200 diffusion sub-steps for the workloads' physics."""
    mob0 = np.zeros((3, 3))
    mob0[:2, :2] = 1e12
    sig = np.zeros((3, 3))
    sig[:2, :2] = 0.5e-12
    ent = np.array([[1, 2, 0, -0.2e-12], [1, 3, 0, 0], [2, 3, 0, 0]], float)
    para = np.array([[1, 2, 0, 0.45 / -172, 0.45 / 172 * 1185], [2, 1, 0, 0.018 / -172, 0.018 / 172 * 1185],
                     [1, 3, 0, -1e-3, 1.5], [3, 1, 0, -1e-3, 1.2], [2, 3, 0, -1e-3, 1.4],
                     [3, 2, 0, -1e-3, 1.1]])
    for dims, nsub_expect in ((3, 200), (2, 200)):
        out, pdt, ddt = oracle.update_parameter(1013.0, 0.1, dims, mob0, np.zeros(9), sig,
                                                [0.5, 100.0, 0.0], np.zeros(3), ent, para)
        assert np.array_equal(np.ctypeslib.as_array(out.mobility), mob0)
        assert pdt == 0.1 * 0.1 / dims * 0.2 / 0.5
        assert ddt == 0.1 * 0.1 / 100.0 / dims * 0.2
        assert max(int(np.ceil(pdt / ddt)), 1) == nsub_expect
        k = np.ctypeslib.as_array(out.part_k)[0]
        assert abs(k[0, 1] - 0.04) < 1e-14 and abs(k[1, 0] - 25.0) < 1e-12
        assert np.all(np.diag(k) == 1.0)
        assert abs(out.eq_conc[0][1] - 0.45) < 1e-13 and abs(out.eq_conc[1][0] - 0.018) < 1e-14


def test_nucleation_brute_force(model3d, params3d):
    """Nuclei occupy exactly the cells within `radius` (periodic distance) of the successive
    first maxima of the flag; later nuclei overwrite earlier ones; nucleus carbon applied."""
    d = model3d.desc()
    od = oracle.OracleDomain(d.nx, d.ny, d.nz, d.depth, 1)
    od.set_params(params3d)
    g = oracle.tessellation(d.nx, d.ny, d.nz, model3d.voronoi_seeds())
    od.construct(g, 10, [0.8, 0.02, 0.0])
    before = od.get_state()
    od.nucleate(3, 2.5, (0.8,), 99)
    after = od.get_state()
    K = d.depth
    ia = after["index"].reshape(-1, K)
    nuc = ia[:, 0] > 100000
    nuc &= ia[:, 0] < 200000
    assert 3 <= len(np.unique(ia[nuc, 0])) <= 3
    assert np.all(after["conc"][nuc] == 0.8)
    assert np.array_equal(after["conc"][~nuc], before["conc"][~nuc])
    # every nucleus is a periodic ball of radius 2.5 around one cell
    for q in np.unique(ia[nuc, 0]):
        cells = np.nonzero(ia[:, 0] == q)[0]
        assert 1 <= len(cells) <= 81


def test_gamma_nucleus_grows(tmp_path):
    """Driving-force sign: carbon-rich austenite (0.8 > 0.45 wt%) grows into ferrite."""
    m = helpers.small_model(str(tmp_path), "fec_3d_256", nx=16, ny=16, nz=16, grains=3, nuclei=2)
    p = helpers.model_params(m)
    od = helpers.oracle_initial_state(m, p)
    f0 = od.reduce()["phase_frac"][0]
    pp = oracle.as_orc_params(p)
    pp.n_sub[0] = 20
    od.set_params(pp)
    od.step(5)
    assert od.reduce()["phase_frac"][0] > f0


def test_noflux_equals_mirrored_periodic(params3d, rng):
    """A no-flux z domain of depth n evolves like the periodic domain of depth 2n built by
    mirroring it (the ghost of a boundary cell is the cell itself).  Diffusion sub-steps agree
    bitwise; full steps agree to rounding (the mirrored cell meets its z faces in the opposite
    order when it builds its active set, which reorders two sums)."""
    nx, ny, nz, K = 8, 6, 5, 4
    ids, val, phase, conc = helpers.random_state(rng, nx, ny, nz, K)
    p = oracle.as_orc_params(params3d)
    p.n_sub[0] = 15

    def mirror(arr, w):
        x = arr.reshape(nz, ny, nx, w)
        return np.concatenate([x, x[::-1]], axis=0).reshape(-1)

    half = nx * ny * nz
    for full in (False, True):
        a = oracle.OracleDomain(nx, ny, nz, K, 1, oracle.ORC_FP64, 1)
        a.set_params(p)
        a.set_state(ids, val, phase, conc)
        b = oracle.OracleDomain(nx, ny, 2 * nz, K, 1, oracle.ORC_FP64, 0)
        b.set_params(p)
        b.set_state(mirror(ids, K), mirror(val, K), mirror(phase, 3), mirror(conc, 1))
        for o in (a, b):
            if full:
                o.step(2)
            else:
                o.diffuse(0, 40, p.dt_d[0])
        sa, sb = a.get_state(), b.get_state()
        assert np.array_equal(sa["index"], sb["index"][: half * K])
        tol = 1e-14 if full else 0.0
        assert np.max(np.abs(sa["value"] - sb["value"][: half * K])) <= tol
        assert np.max(np.abs(sa["conc"] - sb["conc"][:half])) <= tol


def test_compensation_rounding_is_unbiased_and_exact():
    """The mixed form's 4-bit Kahan compensation (multiples of 2^-3 ulp of the value, saturating
    at -8..7) is rounded stochastically with thresholds from hashed keys.
    - Over all 2^24 thresholds the mean of the stored compensation equals the compensation exactly
      (exhaustive, in integers), also for compensations with bits far below a quantum; thresholds
      with 16 fraction bits would give floor(f 2^16)/2^16 of a quantum for the fraction f instead.
    - Over hashed keys the mean is the compensation within sampling noise (rounding to nearest
      would return 0 for 0.3 quanta every time).
    - A compensation already on the grid is kept by every key, key 0 rounds down, and the range
      saturates."""
    L = oracle.lib()
    q = lambda comp, value, key: L.orc_quantize_comp(comp, value, key)
    keys = [L.orc_mix32(i) for i in range(1 << 15)]
    for value, ms in ((0.02, (0x4CCCCD + 1, -0x2AAAAB, 3)),
                      (1.7, (0x4CCCCD, 5 * (1 << 24) + 0x34568, -(7 << 24) - (7 << 6),
                             6 * (1 << 24) + 0xFFFFF0))):
        e = int(np.frexp(np.float32(value))[1])
        quantum = float(np.ldexp(1.0, e - 27))
        for m in ms:
            comp = float(np.float32(m * 2.0 ** -24 * quantum))
            assert comp == m * 2.0 ** -24 * quantum  # exactly representable
            assert L.orc_quantize_comp_mean(comp, value) == comp, (value, m)
    for value in (0.5, 0.75, 0.02, 1.7, 3.0e-3):
        e = int(np.frexp(np.float32(value))[1])
        quantum = float(np.ldexp(1.0, e - 27))
        for frac in (0.3, -0.3, 0.71, -1.5, 5.125):
            comp = float(np.float32(frac * quantum))
            mean = np.mean([q(comp, value, k) for k in keys]) / quantum
            assert abs(mean - comp / quantum) < 5e-3, (value, frac, mean)
            assert q(comp, value, 0) == np.floor(comp / quantum) * quantum
        for n in (-8, -5, -1, 0, 1, 4, 7):
            comp = n * quantum
            assert all(q(comp, value, k) == comp for k in keys[:512])
        assert q(20 * quantum, value, keys[7]) == 7 * quantum
        assert q(-20 * quantum, value, keys[7]) == -8 * quantum


def test_mixed_form_close_to_fp64(setup_state, params3d):
    """The mixed form (FP32 + Kahan, FP64 decisions) tracks the FP64 form closely."""
    (ids, val, phase, conc), d = setup_state
    p = oracle.as_orc_params(params3d)
    p.n_sub[0] = 50
    out = []
    for form in (oracle.ORC_FP64, oracle.ORC_MIXED):
        od = oracle.OracleDomain(d.nx, d.ny, d.nz, d.depth, 1, form)
        od.set_params(p)
        od.set_state(ids, val, phase, conc)
        od.step(3)
        out.append((od.get_state(), od.reduce()))
    (a, ra), (b, rb) = out
    ceff = b["conc"] - b["conc_comp"]
    assert np.max(np.abs(ceff - a["conc"])) < 1e-5
    assert abs(ra["phase_frac"][0] - rb["phase_frac"][0]) < 1e-6
    assert abs(ra["mass"][0] - rb["mass"][0]) / ra["mass"][0] < 1e-6


def test_dual_derivatives_match_central_differences(tmp_path):
    """Forward-mode derivatives of f_gamma and of c after a few steps against central finite
    differences of the FP64 form.  The step is piecewise smooth (clamp at 1, prune below the
    cutoff, top-K selection); on this tiny grid a relative step of 1e-4 already moves single
    cells across those thresholds, so the differences use 1e-6, inside one smooth piece."""
    m = helpers.small_model(str(tmp_path), "fec_3d_256", nx=10, ny=10, nz=8, grains=3, nuclei=2)
    p = helpers.model_params(m)
    od = helpers.oracle_initial_state(m, p)
    st = _state_from(od)
    d = m.desc()
    pp = oracle.as_orc_params(p)
    pp.n_sub[0] = 10
    rng = np.random.default_rng(3)
    direction = rng.standard_normal(d.nx * d.ny * d.nz) * 1e-2
    dual = oracle.OracleDomain(d.nx, d.ny, d.nz, d.depth, 1, oracle.ORC_DUAL)
    dual.set_params(pp)
    dual.set_state(*st)
    dual.dual_seed(direction)
    dual.step(3)
    f, df = dual.dual_phase_fraction()
    c, dc = dual.dual_conc()
    w = rng.standard_normal(c.size)

    def run(scale_m=1.0, scale_s=1.0, dgam=0.0, eps_c=0.0):
        q = oracle.as_orc_params(pp)
        mob = np.ctypeslib.as_array(q.mobility)
        mob *= scale_m
        sg = np.ctypeslib.as_array(q.sigma)
        sg *= scale_s
        q.diffusivity[0][0] += dgam
        o = oracle.OracleDomain(d.nx, d.ny, d.nz, d.depth, 1)
        o.set_params(q)
        o.set_state(st[0], st[1], st[2], st[3] + eps_c * direction)
        o.step(3)
        s = o.get_state()
        return o.reduce()["phase_frac"][0], float(w @ s["conc"])

    h = 1e-6
    dg = pp.diffusivity[0][0]
    cases = [(dict(scale_m=1 + h), dict(scale_m=1 - h), h, 0),
             (dict(scale_s=1 + h), dict(scale_s=1 - h), h, 1),
             (dict(dgam=h * dg), dict(dgam=-h * dg), h * dg, 2),
             (dict(eps_c=h), dict(eps_c=-h), h, 3)]
    for plus, minus, step, k in cases:
        fp, cp = run(**plus)
        fm, cm = run(**minus)
        fd_f = (fp - fm) / (2 * step)
        fd_c = (cp - cm) / (2 * step)
        ad_c = float(w @ dc[k])
        scale_f = max(abs(fd_f), 1e-12)
        assert abs(df[k] - fd_f) <= 1e-5 * scale_f + 1e-12, (k, df[k], fd_f)
        assert abs(ad_c - fd_c) <= 1e-5 * max(abs(fd_c), 1e-12) + 1e-12, (k, ad_c, fd_c)


# ---- flat interfaces: closed forms of the continuum model pin the pair term ----------------------
# For two grains with phi_a + phi_b = 1 the sweep is explicit Euler for
#   dphi/dt = M [sigma (lap(phi) + pi^2/eta^2 (phi - 1/2)) + pi/eta sqrt(phi (1 - phi)) dG]
# ((S_a phi_b - S_b phi_a)/dx^2 is the discrete Laplacian when phi_a + phi_b = 1), whose
# travelling-wave solution is the sine profile phi = 1/2 + 1/2 sin(pi (x - v t)/eta) of width eta
# moving at v = M dG, for any dG.

def _slab(nx, ny, K, lo, hi, width, id_in, id_out):
    """Grain id_in on x in [lo, hi] with sine-profile interfaces of the given width, id_out
    elsewhere; uniform along y."""
    x = np.arange(nx, dtype=float)

    def prof(d):
        return np.where(d >= width / 2, 1.0,
                        np.where(d <= -width / 2, 0.0, 0.5 + 0.5 * np.sin(np.pi * d / width)))

    a = np.minimum(prof(x - lo), prof(hi - x))
    ids = np.zeros((ny, nx, K), np.int32)
    val = np.zeros((ny, nx, K))
    for i in range(nx):
        ent = sorted([(v, g) for v, g in ((a[i], id_in), (1 - a[i], id_out)) if v > 0],
                     key=lambda t: -t[0])
        for s, (v, g) in enumerate(ent):
            ids[:, i, s] = g
            val[:, i, s] = v
    return a, ids.reshape(-1), val.reshape(-1)


def _grain_profile(st, gid, nx, ny, K):
    ii = st["index"].reshape(ny, nx, K)
    vv = st["value"].reshape(ny, nx, K)
    return np.where(ii == gid, vv, 0.0).sum(-1)[0]


@pytest.mark.parametrize("start_width", [1.5, 0.6])
def test_flat_interface_relaxes_to_sine_profile(params3d, start_width):
    """Two ferrite grains (dG = 0): a profile of the wrong width relaxes onto the width-eta sine
    profile and the grain keeps its volume (the two interfaces are mirror images)."""
    nx, ny, K = 64, 2, 4
    q = oracle.as_orc_params(params3d)
    q.dims = 2
    w = q.eta / q.dx
    a0, ids, val = _slab(nx, ny, K, 16, 48, w * start_width, 200001, 200002)
    ref, _, _ = _slab(nx, ny, K, 16, 48, w, 200001, 200002)
    phase = np.zeros((nx * ny, 3))
    phase[:, 1] = 1.0
    o = oracle.OracleDomain(nx, ny, 1, K, 1)
    o.set_params(q)
    o.set_state(ids, val, phase.reshape(-1), np.full(nx * ny, 0.02))
    for _ in range(600):
        o.sweep()
    a = _grain_profile(o.get_state(), 200001, nx, ny, K)
    assert np.max(np.abs(a0 - ref)) > 0.1  # the start is far from the equilibrium shape
    assert np.max(np.abs(a - ref)) < 0.01
    assert abs(a.sum() - ref.sum()) < 1e-9


@pytest.mark.parametrize("dG", [0.25e-12, 0.5e-12])
def test_flat_interface_velocity(params3d, dG):
    """Austenite/ferrite slab with a constant driving force (partition k = 1, c = 0, so the local
    concentration is the total one; dG(1,2) = +dG, dG(2,1) = -dG): the austenite grows, each front
    moving at v = M dG.  At eta = 6 dx the discrete front runs 5-6% slow."""
    nx, ny, K = 64, 2, 4
    r = oracle.as_orc_params(params3d)
    r.dims = 2
    for i in range(3):
        for j in range(3):
            r.mobility[i][j] = r.sigma[i][j] = r.entropy[i][j] = r.eq_conc[i][j] = 0.0
            r.eq_slope[i][j] = 1.0
            for s in range(2):
                r.part_k[s][i][j], r.part_c[s][i][j] = 1.0, 0.0
    M = 1e12
    for (i, j) in ((0, 1), (1, 0)):
        r.mobility[i][j] = M
        r.sigma[i][j] = 0.5e-12
    r.entropy[0][1], r.eq_slope[0][1] = dG, 1.0
    r.entropy[1][0], r.eq_slope[1][0] = dG, -1.0
    w = r.eta / r.dx
    a0, ids, val = _slab(nx, ny, K, 16, 48, w, 100001, 200001)
    phase = np.zeros((nx * ny, 3))
    phase[:, 0] = np.tile(a0, ny)
    phase[:, 1] = 1.0 - phase[:, 0]
    o = oracle.OracleDomain(nx, ny, 1, K, 1)
    o.set_params(r)
    o.set_state(ids, val, phase.reshape(-1), np.full(nx * ny, 1.0))
    vols = []
    for it in range(600):
        o.sweep()
        if it in (199, 599):
            vols.append(_grain_profile(o.get_state(), 100001, nx, ny, K).sum())
    v = (vols[1] - vols[0]) * r.dx / 2.0 / (400 * r.pfm_dt)
    assert 0.9 < v / (M * dG) < 1.0, v / (M * dG)


# ---- three behaviours of the reference's step that the checks above leave open ------------------

@pytest.mark.parametrize("phi1", [0.3, 0.55])
def test_two_phase_mixture_fourier_decay(params3d, phi1):
    """The diffusion weight is sqrt(phi_p(nb) * phi_p) (solver.cpp:363-364).  On a uniform two-phase
    mixture (phi_1, phi_2) = (a, 1 - a) with partition tables k(i,j) = 1, c(i,j) = 0, every phase
    carries the cell's concentration and the weight of phase p is sqrt(a^2) = a exactly, so the
    sub-step is a one-phase sub-step with D_eff = a D_1 + (1 - a) D_2: a Fourier mode decays by
    g = 1 - 4 r_eff sum sin^2(pi m/N) per sub-step.  A weight without the square root gives
    a^2 D_1 + (1 - a)^2 D_2, and a swapped phase index (1 - a) D_1 + a D_2; both miss by far more
    than the tolerance."""
    nx, ny, nz, K = 24, 12, 8, 4
    mode = (2, 1, 1)
    p = oracle.as_orc_params(params3d)
    for i in range(3):
        for j in range(3):
            for s in range(2):
                p.part_k[s][i][j], p.part_c[s][i][j] = 1.0, 0.0
    D1, D2 = 60.0, 100.0
    p.diffusivity[0][0], p.diffusivity[0][1], p.diffusivity[0][2] = D1, D2, 7.0
    N = nx * ny * nz
    ids = np.zeros((N, K), np.int32)
    val = np.zeros((N, K))
    hi, lo = (100001, phi1), (200001, 1 - phi1)
    a, b = (hi, lo) if phi1 >= 0.5 else (lo, hi)
    ids[:, 0], val[:, 0] = a
    ids[:, 1], val[:, 1] = b
    phase = np.zeros((N, 3))
    phase[:, 0], phase[:, 1] = phi1, 1 - phi1
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    arg = 2 * np.pi * (mode[0] * x / nx + mode[1] * y / ny + mode[2] * z / nz)
    c0 = (0.4 + 0.1 * np.cos(arg)).transpose(2, 1, 0).reshape(-1)
    od = oracle.OracleDomain(nx, ny, nz, K, 1)
    od.set_params(p)
    od.set_state(ids.reshape(-1), val.reshape(-1), phase.reshape(-1), c0)
    dt_d, nsub = 4.0e-6, 40
    od.diffuse(0, nsub, dt_d)
    got = od.get_state()["conc"]
    s2 = sum(np.sin(np.pi * m / n) ** 2 for m, n in zip(mode, (nx, ny, nz)))

    def decayed(deff):
        g = 1.0 - 4.0 * deff * dt_d / (p.dx * p.dx) * s2
        return 0.4 + (c0 - 0.4) * g ** nsub

    err = np.max(np.abs(got - decayed(phi1 * D1 + (1 - phi1) * D2)))
    assert err < 1e-13, err
    for wrong in (phi1 ** 2 * D1 + (1 - phi1) ** 2 * D2, (1 - phi1) * D1 + phi1 * D2):
        assert np.max(np.abs(got - decayed(wrong))) > 1e3 * max(err, 1e-16)


def _uniform_state(N, K, slots):
    ids = np.zeros((N, K), np.int32)
    val = np.zeros((N, K))
    for s, (g, v) in enumerate(slots):
        ids[:, s], val[:, s] = g, v
    phase = np.zeros((N, 3))
    for g, v in slots:
        phase[:, 0 if g < 200000 else 1] += v
    return ids.reshape(-1), val.reshape(-1), phase.reshape(-1)


@pytest.mark.parametrize("case", ["tie_first", "tie_behind_max"])
def test_sort_ties_follow_first_max_and_swap(params3d, case):
    """Sort_xyz (solver.cpp:426-432) is a selection sort that moves the first maximum of D[i..10]
    to position i by a swap (solver.cpp:273), so ties are not kept in union order in general.
    Every cell of a uniform domain holds two ferrite grains A, C with equal phi (their updates are
    the same operations on the same operands, so they tie bitwise) and an austenite grain B.
    - tie_first: B is smallest after the update and comes last in the union; the tie stays in
      union order A, C.
    - tie_behind_max: B leads after the update but comes last in the union (A, C, B); the first
      pass swaps B with A, giving B, C, A.  A stable sort or a last-maximum rule gives B, A, C."""
    nx, ny, nz, K = 4, 4, 4, 4
    N = nx * ny * nz
    A, C, B = 200001, 200002, 100001
    if case == "tie_first":
        slots, conc = [(A, 0.4), (C, 0.4), (B, 0.2)], 0.02
    else:
        slots, conc = [(A, 0.34), (C, 0.34), (B, 0.32)], 0.8
    ids, val, phase = _uniform_state(N, K, slots)
    od = oracle.OracleDomain(nx, ny, nz, K, 1)
    od.set_params(params3d)
    od.set_state(ids, val, phase, np.full(N, conc))
    od.sweep()
    s = od.get_state()
    oi = s["index"].reshape(N, K)
    ov = s["value"].reshape(N, K)
    if case == "tie_first":
        assert np.all(oi[:, :3] == [A, C, B])
        assert np.all(ov[:, 0] == ov[:, 1]) and np.all(ov[:, 2] < ov[:, 0])
    else:
        assert np.all(oi[:, :3] == [B, C, A])
        assert np.all(ov[:, 1] == ov[:, 2]) and np.all(ov[:, 0] > ov[:, 1])


@pytest.mark.parametrize("own", [6, 5])
def test_union_cap_drops_the_eleventh_grain(params3d, own):
    """The active-set union holds at most 10 entries; a neighbour grain met when it is full is
    dropped silently (solver.cpp:197-202).  K = 6: the centre cell owns `own` ferrite grains, its -x
    neighbour four more, and every other cell of the domain holds grain G alone, so G reaches the
    union through five faces, after those ten (own = 6) or as the tenth entry (own = 5).  As the
    tenth entry G takes the largest neighbour sum and enters the centre's output at slot 1; as an
    eleventh it never enters it."""
    nx, ny, nz, K = 5, 5, 5, 6
    N = nx * ny * nz
    G = 200099
    ids = np.zeros((N, K), np.int32)
    val = np.zeros((N, K))
    ids[:, 0], val[:, 0] = G, 1.0
    c = 2 + nx * (2 + ny * 2)
    xm = c - 1
    ids[c], val[c] = 0, 0.0
    for s in range(own):
        ids[c, s], val[c, s] = 200001 + s, 1.0 / own
    ids[xm], val[xm] = 0, 0.0
    for s in range(4):
        ids[xm, s], val[xm, s] = 200011 + s, 0.25
    phase = np.zeros((N, 3))
    phase[:, 1] = 1.0
    od = oracle.OracleDomain(nx, ny, nz, K, 1)
    od.set_params(params3d)
    od.set_state(ids.reshape(-1), val.reshape(-1), phase.reshape(-1), np.full(N, 0.02))
    od.sweep()
    s = od.get_state()
    out = s["index"].reshape(N, K)[c]
    v = s["value"].reshape(N, K)[c]
    if own == 6:
        assert G not in out
        assert set(out[out > 0]) <= set(range(200001, 200007)) | set(range(200011, 200015))
    else:
        assert out[0] == G and v[0] == v.max()
    assert abs(v.sum() - 1.0) < 1e-15


def test_set_comp_hands_over_the_mixed_compensation(params3d):
    """orc_set_comp (the hand-over of a mixed device state's stored compensations): the values come
    back unchanged, the FP64 form ignores them, and the next sub-step takes them as the Kahan update
    does: with no flux anywhere, the effective value c - comp of every cell is kept exactly (a
    compensation on the code grid of up to half an ulp is re-stored without rounding noise)."""
    rng = np.random.default_rng(5)
    shape = (6, 5, 4)
    N = 6 * 5 * 4
    st = helpers.random_state(rng, *shape, 4)
    # a uniform concentration and a single phase: no flux anywhere, so del = 0 in every cell
    idx, val, ph, _ = st
    idx = np.zeros_like(idx)
    idx.reshape(N, 4)[:, 0] = 200001
    val = np.zeros_like(val)
    val.reshape(N, 4)[:, 0] = 1.0
    ph = np.zeros_like(ph)
    ph.reshape(N, 3)[:, 1] = 1.0
    conc = np.full(N, 0.75)
    quantum = 2.0 ** (0 - 27)  # 0.75 = 0.75 * 2^0
    comp = quantum * rng.integers(-4, 5, size=N).astype(np.float64)
    o = oracle.OracleDomain(*shape, 4, 1, oracle.ORC_MIXED)
    o.set_params(oracle.as_orc_params(params3d))
    o.set_state(idx, val, ph, conc)
    o.set_comp(comp)
    s = o.get_state()
    assert np.array_equal(s["conc_comp"], comp)
    assert np.array_equal(s["conc"], conc)
    # one sub-step with zero flux: y = -comp, t = c + y (|y| < half an ulp of 0.75 rounds back to c),
    # comp' = (t - c) - y = comp: the stored state is unchanged and the effective value kept
    o.diffuse(0, 1, oracle.as_orc_params(params3d).dt_d[0])
    s1 = o.get_state()
    assert np.array_equal(s1["conc"] - s1["conc_comp"], conc - comp)
    f = oracle.OracleDomain(*shape, 4, 1, oracle.ORC_FP64)
    f.set_state(idx, val, ph, conc)
    f.set_comp(comp)
    assert not np.any(f.get_state()["conc_comp"])
