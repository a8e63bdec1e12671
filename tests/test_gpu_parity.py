"""GPU parity: the HIP kernels (through the C-ABI) against the CPU oracle on identical inputs.

FP64 form: bitwise identical active-set maps, phi, value_phase and concentrations (the build has no
FMA contraction and gfx950's division and square root round correctly), which is stricter than the
1e-10 relative bar.  Mixed form: against the oracle's mixed form, bitwise as well.
Shapes include partial 64-wide x tiles, partial y tiles and z-chunk edges.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from phasefield_amd import Domain, gpu_cleanup, gpu_diffusion_solver, gpu_init
from tests import helpers

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def base_params(tmp_path_factory):
    m = helpers.small_model(str(tmp_path_factory.mktemp("gp")), "fec_3d_256", nx=16, ny=16,
                            nz=16, grains=4, nuclei=2)
    return helpers.model_params(m)


def with_phase3(p):
    """Give the placeholder phase nonzero interface energy/mobility so P=3 paths do work."""
    q = oracle.as_orc_params(p)
    for (i, j) in ((1, 2), (2, 1), (2, 2)):
        q.sigma[i][j] = 0.3e-12
        q.mobility[i][j] = 0.5e12
    q.diffusivity[0][2] = 5.0
    return q


def assert_same(gs, os_, K, tag, tol=0.0):
    assert np.array_equal(gs["index"], os_["index"]), f"{tag}: active-set maps differ"
    for key in ("value", "phase", "conc"):
        diff = np.max(np.abs(gs[key] - os_[key])) if gs[key].size else 0.0
        assert diff <= tol, f"{tag}: {key} differs by {diff}"


def run_pair(shape, K, P, S, bc, form, state, params, fn_gpu, fn_orc):
    nx, ny, nz = shape
    g = Domain(nx, ny, nz, depth=K, nphase=P, nsolute=S,
               precision="mixed" if form == oracle.ORC_MIXED else "fp64", bc_z=bc)
    pp = oracle.as_orc_params(params)
    pp.dims = 3 - (nz == 1) - (ny == 1)
    from phasefield_amd._capi import PfParams
    gp = PfParams.from_buffer_copy(pp)
    g.set_params(gp)
    g.upload(*state)
    o = oracle.OracleDomain(nx, ny, nz, K, S, form, bc)
    o.set_params(pp)
    o.set_state(*state)
    fn_gpu(g, gp)
    fn_orc(o, pp)
    gs = g.download()
    os_ = o.get_state()
    if form == oracle.ORC_MIXED:
        gs["conc"] = gs["conc"]  # stored value; compensation compared separately
        assert np.array_equal(gs["conc_comp"], os_["conc_comp"])
    g.close()
    return gs, os_


SHAPES = [((32, 32, 1), 0), ((70, 13, 19), 0), ((64, 16, 40), 0), ((40, 24, 18), 1)]


@pytest.mark.parametrize("shape,bc", SHAPES)
@pytest.mark.parametrize("P", [2, 3])
def test_diffusion_substeps_bitwise(shape, bc, P, base_params, rng):
    banner(f"diffusion {shape} bc={bc} P={P}")
    K = 4
    params = with_phase3(base_params) if P == 3 else base_params
    st = helpers.random_state(rng, *shape, K, nphase=P)
    gs, os_ = run_pair(shape, K, P, 1, bc, oracle.ORC_FP64, st, params,
                       lambda g, p: g.diffuse(0, 37, p.dt_d[0]),
                       lambda o, p: o.diffuse(0, 37, p.dt_d[0]))
    assert_same(gs, os_, K, "diffusion")


@pytest.mark.parametrize("shape,bc", SHAPES)
@pytest.mark.parametrize("K,P", [(4, 2), (6, 2), (4, 3), (6, 3)])
def test_phase_field_sweep_bitwise(shape, bc, K, P, base_params, rng):
    banner(f"sweep {shape} bc={bc} K={K} P={P}")
    params = with_phase3(base_params) if P == 3 else base_params
    st = helpers.random_state(rng, *shape, K, nphase=P)
    gs, os_ = run_pair(shape, K, P, 1, bc, oracle.ORC_FP64, st, params,
                       lambda g, p: g.sweep(), lambda o, p: o.sweep())
    assert_same(gs, os_, K, "sweep")


@pytest.mark.parametrize("shape,bc", SHAPES)
@pytest.mark.parametrize("K,P", [(4, 2), (6, 2), (4, 3), (6, 3)])
def test_mixed_steps_random_states_bitwise(shape, bc, K, P, base_params, rng):
    """The mixed form's full steps (sweep + sub-steps) from random active-set states, against the
    oracle's mixed form: every depth and phase count, periodic and no-flux z, 2D, partial tiles;
    up to K grains per cell, so unions beyond four entries take the sweep's general path."""
    banner(f"mixed steps {shape} bc={bc} K={K} P={P}")
    params = with_phase3(base_params) if P == 3 else base_params
    st = helpers.random_state(rng, *shape, K, nphase=P)
    gs, os_ = run_pair(shape, K, P, 1, bc, oracle.ORC_MIXED, st, params,
                       lambda g, p: g.step(2), lambda o, p: o.step(2))
    assert_same(gs, os_, K, "mixed steps")


@pytest.mark.parametrize("form", [oracle.ORC_FP64, oracle.ORC_MIXED])
@pytest.mark.parametrize("shape,bc", [((2, 2, 2), 0), ((2, 2, 2), 1), ((3, 5, 4), 0), ((4, 4, 2), 1),
                                      ((5, 3, 3), 1), ((2, 6, 1), 0), ((8, 2, 3), 0)])
def test_tiny_domains_bitwise(shape, bc, form, base_params, rng):
    """Domains of two to eight cells per axis, where a periodic neighbour wraps onto the cell's
    other neighbour or (no-flux, nz = 2) the two z faces meet the same cell, and every cell is at
    a tile and chunk edge: full steps from random states, both forms, bitwise."""
    banner(f"tiny domain {shape} bc={bc} form={form}")
    st = helpers.random_state(rng, *shape, 4)
    gs, os_ = run_pair(shape, 4, 2, 1, bc, form, st, base_params,
                       lambda g, p: g.step(2), lambda o, p: o.step(2))
    assert_same(gs, os_, 4, "tiny domain")


def _setup_state(tmp_path, **over):
    m = helpers.small_model(str(tmp_path), over.pop("workload", "fec_3d_256"), **over)
    p = helpers.model_params(m)
    od = helpers.oracle_initial_state(m, p)
    s = od.get_state()
    return m, p, (s["index"], s["value"], s["phase"], s["conc"])


@pytest.mark.parametrize("sweep_split", ["1", "0"])
@pytest.mark.parametrize("shape,bc,K", [((70, 20, 18), "periodic", 4), ((48, 40, 1), "periodic", 6),
                                        ((40, 36, 24), "noflux", 4)])
def test_full_steps_bitwise(tmp_path, monkeypatch, shape, bc, K, sweep_split):
    """Several complete Update() steps (sweep + 200 sub-steps) from a seeded microstructure, with
    the sweep's copy cells classified apart (and their unchanged rows not rewritten) or not."""
    banner(f"full steps {shape} {bc} K={K} sweep_split={sweep_split}")
    monkeypatch.setenv("PFAMD_SWEEP_SPLIT", sweep_split)
    nx, ny, nz = shape
    m, p, st = _setup_state(tmp_path, nx=nx, ny=ny, nz=nz, grains=6, nuclei=4, depth=K, bc_z=bc)
    bcn = 1 if bc == "noflux" else 0
    gs, os_ = run_pair(shape, K, 2, 1, bcn, oracle.ORC_FP64, st, p,
                       lambda g, q: g.step(3), lambda o, q: o.step(3))
    assert_same(gs, os_, K, "full step")


@pytest.mark.parametrize("form", [oracle.ORC_FP64, oracle.ORC_MIXED])
@pytest.mark.parametrize("rows", ["0", "1", "5", "-1"])
@pytest.mark.parametrize("shape", [(70, 20, 18), (64, 23, 18)])
def test_sweep_list_orders_bitwise(tmp_path, monkeypatch, shape, rows, form):
    """The sweep's list in storage order (rows 0) and in row tiles of 1, 5 and the size-chosen
    number of rows: 23 rows leave a 3-row last tile, nx = 70 takes the per-position list build
    (rows not starting at multiples of 16 positions) and nx = 64 the 16-flag loads.  The cells'
    updates are independent, so every order gives the oracle's result bitwise."""
    banner(f"sweep list order rows={rows} shape={shape} form={form}")
    monkeypatch.setenv("PFAMD_SWEEP_ROWS", rows)
    nx, ny, nz = shape
    m, p, st = _setup_state(tmp_path, nx=nx, ny=ny, nz=nz, grains=6, nuclei=4)
    gs, os_ = run_pair(shape, 4, 2, 1, 0, form, st, p,
                       lambda g, q: g.step(3), lambda o, q: o.step(3))
    assert_same(gs, os_, 4, f"sweep list rows={rows}")


@pytest.mark.parametrize("shape,bc", [((33, 20, 12), 0), ((35, 18, 1), 0), ((41, 16, 10), 1)])
def test_mixed_listed_pass_odd_nx_bitwise(tmp_path, shape, bc):
    """The mixed listed pass with nx odd, where the two cells of a code byte can lie in different
    rows (no shared stencil): each listed cell of a byte pair gets its own stencil."""
    banner(f"mixed listed pass odd nx shape={shape} bc={bc}")
    nx, ny, nz = shape
    m, p, st = _setup_state(tmp_path, nx=nx, ny=ny, nz=nz, grains=5, nuclei=4,
                            bc_z="noflux" if bc else "periodic")
    gs, os_ = run_pair(shape, 4, 2, 1, bc, oracle.ORC_MIXED, st, p,
                       lambda g, q: g.step(2), lambda o, q: o.step(2))
    assert_same(gs, os_, 4, "mixed listed pass, odd nx")


@pytest.mark.parametrize("form", [oracle.ORC_FP64, oracle.ORC_MIXED])
@pytest.mark.parametrize("shape,bc", [((64, 64, 16), 0), ((128, 32, 8), 0), ((64, 32, 32), 1),
                                      ((256, 16, 8), 0), ((256, 12, 6), 1)])
def test_sweep_classify_xcd_order_bitwise(tmp_path, shape, bc, form):
    """Shapes whose copy-cell classification runs a multiple of 128 workgroups, the case in which
    each XCD takes interleaved groups of 16 workgroups instead of the plain order (N / 256 =
    256, 128, 256 and 128 here; periodic and no-flux z), and rows of 256 cells, whose workgroups
    take their y and z from the workgroup (256 x 12 x 6: 72 workgroups, plain order)."""
    banner(f"sweep classify XCD order shape={shape} bc={bc} form={form}")
    nx, ny, nz = shape
    m, p, st = _setup_state(tmp_path, nx=nx, ny=ny, nz=nz, grains=6, nuclei=4,
                            bc_z="noflux" if bc else "periodic")
    gs, os_ = run_pair(shape, 4, 2, 1, bc, form, st, p,
                       lambda g, q: g.step(3), lambda o, q: o.step(3))
    assert_same(gs, os_, 4, "sweep classify XCD order")


@pytest.mark.parametrize("form", [oracle.ORC_FP64, oracle.ORC_MIXED])
@pytest.mark.parametrize("queue", ["1", "0"])
@pytest.mark.parametrize("shape", [(70, 20, 18), (64, 23, 18)])
def test_sweep_queue_walk_bitwise(tmp_path, monkeypatch, shape, queue, form):
    """The listed sweep's two walks: wavefronts taking entries from per-XCD counters (the default,
    PFAMD_SWEEP_QUEUE=1) and a fixed stride (0).  Every cell is updated exactly once either way."""
    banner(f"sweep walk queue={queue} shape={shape} form={form}")
    monkeypatch.setenv("PFAMD_SWEEP_QUEUE", queue)
    nx, ny, nz = shape
    m, p, st = _setup_state(tmp_path, nx=nx, ny=ny, nz=nz, grains=6, nuclei=4)
    gs, os_ = run_pair(shape, 4, 2, 1, 0, form, st, p,
                       lambda g, q: g.step(3), lambda o, q: o.step(3))
    assert_same(gs, os_, 4, "sweep queue walk")


@pytest.mark.parametrize("form", [oracle.ORC_FP64, oracle.ORC_MIXED])
@pytest.mark.parametrize("kernel,rows,split,vec,march,zchunk", [
    ("bulk", "-1", "1", "1", "1", "0"),
    ("bulk", "8", "1", "1", "1", "5"), ("bulk", "16", "1", "1", "1", "3"),
    ("bulk", "32", "1", "1", "1", "7"),
    ("bulk", "-1", "1", "1", "0", "0"),
    ("bulk", "0", "1", "0", "0", "0"), ("bulk", "4", "1", "0", "0", "0"),
    ("bulk", "8", "1", "0", "0", "0"), ("bulk", "16", "1", "0", "0", "0"),
    ("bulk", "0", "0", "0", "0", "0"), ("bulk", "4", "0", "0", "0", "0"),
    ("bulk", "8", "0", "0", "0", "0"), ("bulk", "16", "0", "0", "0", "0"),
    ("lds", "-1", "0", "0", "0", "0")])
def test_diffusion_kernel_variants_bitwise(tmp_path, monkeypatch, kernel, rows, split, vec, march,
                                           zchunk, form):
    """Every diffusion kernel against the oracle on a seeded microstructure (bulk and interface
    cells), in both forms: the plane-marching bulk kernel (timed choice of tile and z-chunk, and
    8/16-row tiles with 5/3-plane chunks, 32-row tiles with two rows per wavefront and 7-plane
    chunks), the batched 16-byte-run bulk kernel, storage order and
    column-marching with 4/8/16-row tiles (full-formula cells inline or in their own listed
    pass), and the LDS-tiled full formula."""
    banner(f"diffusion variant {kernel} rows={rows} split={split} vec={vec} march={march} "
           f"zchunk={zchunk} form={form}")
    monkeypatch.setenv("PFAMD_DIFF_KERNEL", kernel)
    monkeypatch.setenv("PFAMD_DIFF_BY", rows)
    monkeypatch.setenv("PFAMD_DIFF_SPLIT", split)
    monkeypatch.setenv("PFAMD_DIFF_VEC", vec)
    monkeypatch.setenv("PFAMD_DIFF_MARCH", march)
    monkeypatch.setenv("PFAMD_DIFF_ZCHUNK", zchunk)
    shape = (72, 36, 22)  # nx a multiple of 4 and not of 64: partial vector tiles
    m, p, st = _setup_state(tmp_path, nx=shape[0], ny=shape[1], nz=shape[2], grains=6, nuclei=5)
    gs, os_ = run_pair(shape, 4, 2, 1, 0, form, st, p,
                       lambda g, q: g.step(2), lambda o, q: o.step(2))
    assert_same(gs, os_, 4, f"{kernel}/{rows}/{split}/{vec}/{march}/{zchunk}")


@pytest.mark.parametrize("form", [oracle.ORC_FP64, oracle.ORC_MIXED])
@pytest.mark.parametrize("rows,zchunk", [("8", "5"), ("16", "4"), ("32", "6"), ("-1", "0")])
def test_march_noflux_bitwise(tmp_path, monkeypatch, rows, zchunk, form):
    """The plane-marching bulk kernel with no-flux z faces: chunk end planes clamp to the cell
    itself at the domain's z ends and read the neighbour chunk's plane inside."""
    banner(f"march noflux rows={rows} zchunk={zchunk} form={form}")
    monkeypatch.setenv("PFAMD_DIFF_BY", rows)
    monkeypatch.setenv("PFAMD_DIFF_ZCHUNK", zchunk)
    shape = (40, 36, 24)
    m, p, st = _setup_state(tmp_path, nx=shape[0], ny=shape[1], nz=shape[2], grains=6, nuclei=4,
                            bc_z="noflux")
    gs, os_ = run_pair(shape, 4, 2, 1, 1, form, st, p,
                       lambda g, q: g.step(2), lambda o, q: o.step(2))
    assert_same(gs, os_, 4, f"march noflux {rows}/{zchunk}")


@pytest.mark.parametrize("form", [oracle.ORC_FP64, oracle.ORC_MIXED])
@pytest.mark.parametrize("gamma_cols", [48, 16, 0])
def test_bulk_phase_choice_bitwise(base_params, rng, form, gamma_cols):
    """The plane-marching kernel applies the bulk update of the phase with the most bulk cells to
    every cell and the listed pass rewrites the rest.  A single austenite block of gamma_cols of the
    64 columns and a ferrite block elsewhere: austenite is the bulk phase at 48 columns, ferrite at
    16, and at 0 every cell is ferrite bulk.  Random concentrations, 37 sub-steps, bitwise."""
    banner(f"bulk phase choice gamma_cols={gamma_cols} form={form}")
    nx, ny, nz, K = 64, 24, 16, 4
    N = nx * ny * nz
    x = np.tile(np.arange(nx), ny * nz)
    gamma = x < gamma_cols
    ids = np.zeros((N, K), np.int32)
    val = np.zeros((N, K))
    ids[:, 0] = np.where(gamma, 100001, 200001)
    val[:, 0] = 1.0
    phase = np.zeros((N, 3))
    phase[gamma, 0] = 1.0
    phase[~gamma, 1] = 1.0
    conc = 0.02 + 0.8 * rng.random(N)
    st = (ids.reshape(-1), val.reshape(-1), phase.reshape(-1), conc)
    gs, os_ = run_pair((nx, ny, nz), K, 2, 1, 0, form, st, base_params,
                       lambda g, p: g.diffuse(0, 37, p.dt_d[0]),
                       lambda o, p: o.diffuse(0, 37, p.dt_d[0]))
    assert_same(gs, os_, K, f"bulk phase {gamma_cols}")


def test_mixed_checkpoint_keeps_rounding_keys(tmp_path):
    """The mixed form rounds its compensation with keys counted in sub-steps since the state was
    set: a checkpoint carries the count, an upload restarts it.  2 steps + checkpoint + 3 steps +
    restore + 2 steps equal 4 straight steps, bitwise, compensation codes included."""
    banner("mixed checkpoint keys")
    m, p, st = _setup_state(tmp_path, workload="fecmn_3d_512", nx=40, ny=24, nz=20, grains=5,
                            nuclei=3)
    pp = oracle.as_orc_params(p)
    from phasefield_amd._capi import PfParams
    gp = PfParams.from_buffer_copy(pp)
    a = Domain(40, 24, 20, depth=4, nphase=2, nsolute=2, precision="mixed")
    a.set_params(gp)
    a.upload(*st)
    a.step(2)
    blob = a.checkpoint()
    a.step(3)
    a.restore(blob)
    a.step(2)
    b = Domain(40, 24, 20, depth=4, nphase=2, nsolute=2, precision="mixed")
    b.set_params(gp)
    b.upload(*st)
    b.step(4)
    ra, rb = a.download(), b.download()
    assert_same(ra, rb, 4, "restore")
    assert np.array_equal(ra["conc_comp"], rb["conc_comp"])
    a.upload(*st)
    a.step(4)
    ra = a.download()
    assert_same(ra, rb, 4, "re-upload")
    assert np.array_equal(ra["conc_comp"], rb["conc_comp"])
    a.close()
    b.close()


def test_bulk_reduction_needs_identity_partition(tmp_path):
    """With k(2,2) != 1 the partition of a pure ferrite cell is not its total concentration, so the
    bulk-cell reduction must not be taken for ferrite: results stay bitwise equal to the oracle."""
    banner("bulk reduction guard")
    m, p, st = _setup_state(tmp_path, nx=40, ny=36, nz=20, grains=5, nuclei=4)
    pp = oracle.as_orc_params(p)
    pp.part_k[0][1][1] = 1.1
    gs, os_ = run_pair((40, 36, 20), 4, 2, 1, 0, oracle.ORC_FP64, st, pp,
                       lambda g, q: g.step(2), lambda o, q: o.step(2))
    assert_same(gs, os_, 4, "k(2,2)=1.1")


def test_two_solutes_bitwise(tmp_path):
    """Fe-C-Mn: carbon sub-cycled, manganese one explicit update per step."""
    banner("two solutes")
    m, p, st = _setup_state(tmp_path, workload="fecmn_3d_512", nx=36, ny=28, nz=20, grains=5,
                            nuclei=3, precision="fp64")
    gs, os_ = run_pair((36, 28, 20), 4, 2, 2, 0, oracle.ORC_FP64, st, p,
                       lambda g, q: g.step(2), lambda o, q: o.step(2))
    assert_same(gs, os_, 4, "two solutes")


def test_mixed_form_matches_oracle_mixed(tmp_path):
    banner("mixed form")
    m, p, st = _setup_state(tmp_path, workload="fecmn_3d_512", nx=40, ny=24, nz=20, grains=5,
                            nuclei=3)
    gs, os_ = run_pair((40, 24, 20), 4, 2, 2, 0, oracle.ORC_MIXED, st, p,
                       lambda g, q: g.step(2), lambda o, q: o.step(2))
    assert_same(gs, os_, 4, "mixed")


def test_device_setup_matches_oracle(tmp_path):
    """Voronoi + Iso_growth + partition + nucleation + Mn noise on the device vs the oracle."""
    banner("device setup")
    m = helpers.small_model(str(tmp_path), "fecmn_3d_512", nx=48, ny=40, nz=24, grains=9,
                            nuclei=6, precision="fp64")
    p = helpers.model_params(m)
    od = helpers.oracle_initial_state(m, p)
    d = m.desc()
    g = Domain.from_desc(d)
    g.set_params(p)
    ic = [m.get_real(f"init_conc{i}") for i in range(3)] + [m.get_real("mn_conc")] * 3
    g.construct(m.voronoi_seeds(), m.get_real("reduction"), m.get_int("num_init_iterations"), ic)
    g.solute_noise(1, m.get_real("mn_conc"), m.get_real("mn_noise"), m.get_int("mn_seed"))
    g.nucleate(m.get_int("num_nuclei"), m.get_real("nucleus_radius"),
               (m.get_real("nucleus_conc"), m.get_real("mn_nucleus_conc")),
               m.get_int("nucleation_seed"))
    assert_same(g.download(), od.get_state(), d.depth, "setup")


def test_reduce_matches_oracle(tmp_path):
    banner("reduce")
    m, p, st = _setup_state(tmp_path, nx=50, ny=30, nz=20, grains=5, nuclei=3)
    g = Domain(50, 30, 20, depth=4)
    g.set_params(p)
    g.upload(*st)
    o = oracle.OracleDomain(50, 30, 20, 4, 1)
    o.set_params(p)
    o.set_state(*st)
    rg, ro = g.reduce(), o.reduce()
    # the device sums in a fixed tree, the oracle sequentially: both against the exactly rounded
    # sum of the same values
    import math
    exact = math.fsum(st[3])
    assert abs(rg["mass"][0] - exact) <= 1e-14 * exact
    assert abs(ro["mass"][0] - exact) <= 1e-11 * exact
    for i in range(3):
        assert abs(rg["phase_frac"][i] - ro["phase_frac"][i]) <= 1e-14
    assert rg["max_conc"][0] == ro["max_conc"][0] and rg["min_conc"][0] == ro["min_conc"][0]
    assert rg["n_interface"] == ro["n_interface"]
    assert abs(rg["max_sum_err"] - ro["max_sum_err"]) <= 1e-15


def test_roundtrips_and_checkpoint(tmp_path):
    banner("round trips")
    m, p, st = _setup_state(tmp_path, nx=30, ny=20, nz=10, grains=4, nuclei=2)
    g = Domain(30, 20, 10, depth=4)
    g.set_params(p)
    g.upload(*st)
    s = g.download(phase_conc=True)
    assert np.array_equal(s["index"], st[0]) and np.array_equal(s["value"], st[1])
    assert np.array_equal(s["phase"], st[2]) and np.array_equal(s["conc"], st[3])
    o = oracle.OracleDomain(30, 20, 10, 4, 1)
    o.set_params(p)
    o.set_state(*st)
    assert np.array_equal(s["phase_conc"], o.get_state()["phase_conc"])
    blob = g.checkpoint()
    g.step(2)
    a = g.download()
    g.restore(blob)
    g.step(2)
    b = g.download()
    for k in ("index", "value", "phase", "conc"):
        assert np.array_equal(a[k], b[k])


def test_determinism(tmp_path):
    banner("determinism")
    m, p, st = _setup_state(tmp_path, nx=66, ny=18, nz=20, grains=5, nuclei=3)
    outs = []
    for _ in range(2):
        g = Domain(66, 18, 20, depth=4)
        g.set_params(p)
        g.upload(*st)
        g.step(2)
        outs.append((g.download(), g.reduce()))
        g.close()
    for k in ("index", "value", "phase", "conc"):
        assert np.array_equal(outs[0][0][k], outs[1][0][k])
    assert outs[0][1] == outs[1][1]


def test_compat_shim_matches_oracle(base_params, rng):
    """GPU_Init / GPU_Diffusion_solver / GPU_Cleanup with host pointers (gpu_solver.h:18-23):
    sub-step count and dt/dx^2 computed as gpu_solver.cpp:205-209 (including its 0.2f)."""
    banner("compat shim")
    nx, ny, nz = 40, 30, 12
    K = 4
    ids, val, phase, conc = helpers.random_state(rng, nx, ny, nz, K)
    p = base_params
    D = np.array([p.diffusivity[0][i] for i in range(3)])
    pk = np.array([[p.part_k[0][i][j] for j in range(3)] for i in range(3)])
    pc = np.array([[p.part_c[0][i][j] for j in range(3)] for i in range(3)])
    assert gpu_init(nx, ny, nz) == 1
    c = conc.copy()
    dt, dx = p.pfm_dt, p.dx
    gpu_diffusion_solver(D, pk.T.reshape(-1), pc.T.reshape(-1), phase, c, nx, ny, nz, dt, dx)
    gpu_cleanup()
    del_t = dx * dx / D.max() / 3 * float(np.float32(0.2))
    nloop = int(np.ceil(dt / del_t))
    o = oracle.OracleDomain(nx, ny, nz, K, 1)
    o.set_params(p)
    o.set_state(ids, val, phase, conc)
    o.diffuse_factor(0, nloop, dt / nloop / (dx * dx))
    assert np.array_equal(c, o.get_state()["conc"])


def test_reference_cxx_caller_matches_oracle(base_params, rng, tmp_path):
    """A C++ program that declares GPU_Init / GPU_Diffusion_solver / GPU_Cleanup exactly as the
    reference's gpu_solver.h:18-23 does (C++ linkage, no header of this library), linked with g++
    against libpfamd.so, gives the oracle's concentrations bitwise (2D and 3D shapes)."""
    import subprocess
    banner("reference-shaped C++ caller")
    exe = helpers.build_reference_shim_program(str(tmp_path))
    p = base_params
    D = np.array([p.diffusivity[0][i] for i in range(3)])
    pk = np.array([[p.part_k[0][i][j] for j in range(3)] for i in range(3)])
    pc = np.array([[p.part_c[0][i][j] for j in range(3)] for i in range(3)])
    for (nx, ny, nz) in ((36, 20, 1), (24, 18, 10)):
        K = 4
        ids, val, phase, conc = helpers.random_state(rng, nx, ny, nz, K)
        dt, dx = p.pfm_dt, p.dx
        inp, out = tmp_path / f"in{nz}.bin", tmp_path / f"out{nz}.bin"
        helpers.write_shim_input(str(inp), nx, ny, nz, dt, dx, D, pk.T.reshape(-1),
                                 pc.T.reshape(-1), phase, conc)
        res = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True,
                             timeout=120)
        assert res.returncode == 0, res.stdout + res.stderr
        got = np.fromfile(str(out), np.float64)
        dims = 3 if nz > 1 else 2
        del_t = dx * dx / D.max() / dims * float(np.float32(0.2))
        nloop = int(np.ceil(dt / del_t))
        q = oracle.as_orc_params(p)
        q.dims = 3  # Diff3D sums six faces; the z faces of a one-cell-thick domain carry no flux
        o = oracle.OracleDomain(nx, ny, nz, K, 1)
        o.set_params(q)
        o.set_state(ids, val, phase, conc)
        o.diffuse_factor(0, nloop, dt / nloop / (dx * dx))
        assert np.array_equal(got, o.get_state()["conc"]), (nx, ny, nz)


@pytest.mark.slow
def test_fec_3d_256_steps_bitwise(tmp_path):
    """The 256^3 Fe-C parity workload at full size (128 grains, 200 nuclei, K = 4): the device-built
    seeded state is handed to the oracle, both run 5 complete steps (sweep + 200 sub-steps each),
    and active-set maps, phi, value_phase and concentrations must agree bit for bit."""
    from phasefield_amd import Model, configs
    banner("fec_3d_256: 5 full steps vs oracle")
    m = Model(configs.write_inputs(str(tmp_path), "fec_3d_256"))
    p = m.update_parameter()
    d = m.desc()
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
    g.step(5)
    o.step(5)
    assert_same(g.download(), o.get_state(), d.depth, "256^3 x 5 steps")


def test_sweep_history_across_restore(tmp_path):
    """The sweep skips rewriting copy cells whose destination row already holds them; a restored
    checkpoint (and an upload) must drop that knowledge: 2 + restore + 2 steps equal 4 steps."""
    banner("sweep history across restore")
    m, p, st = _setup_state(tmp_path, nx=64, ny=24, nz=20, grains=5, nuclei=4)
    pp = oracle.as_orc_params(p)
    from phasefield_amd._capi import PfParams
    a = Domain(64, 24, 20, depth=4, nphase=2)
    a.set_params(PfParams.from_buffer_copy(pp))
    a.upload(*st)
    a.step(2)
    blob = a.checkpoint()
    a.step(3)
    a.restore(blob)
    a.step(2)
    b = Domain(64, 24, 20, depth=4, nphase=2)
    b.set_params(PfParams.from_buffer_copy(pp))
    b.upload(*st)
    b.step(4)
    assert_same(a.download(), b.download(), 4, "restore")
    a.upload(*st)  # upload after history: starts over
    a.step(4)
    assert_same(a.download(), b.download(), 4, "re-upload")


def _host_diffusion_listed(phase, shape, P, noflux, params, mixed):
    """Cells the diffusion sub-step lists with the plane-marching kernel, counted on the host from
    the stored value_phase: a cell is bulk of phase p when its value_phase and its six neighbours'
    are exactly e_p and phase p's partition is the identity for every solute; every cell that is not
    bulk of the bulk phase (the phase with the most bulk cells, the first on a tie) is listed."""
    nx, ny, nz = shape
    pp = oracle.as_orc_params(params)
    conv = np.float32 if mixed else np.float64
    mask = [q + 1 for q in range(P)
            if all(conv(pp.part_k[s][q][q]) == 1 and conv(pp.part_c[s][q][q]) == 0
                   for s in range(1))]
    ph = phase.reshape(nz, ny, nx, 3)[..., :P]
    ones = (ph == 1.0).sum(-1)
    zeros = (ph == 0.0).sum(-1)
    unit = np.where((ones == 1) & (zeros == P - 1), np.argmax(ph == 1.0, axis=-1) + 1, 0)
    unit = np.where(np.isin(unit, mask), unit, 0)
    flag = unit.copy()
    for axis in (2, 1, 0):
        for d in (-1, 1):
            if axis == 0 and noflux:
                nb = (np.concatenate([unit[:1], unit[:-1]]) if d < 0
                      else np.concatenate([unit[1:], unit[-1:]]))
            else:
                nb = np.roll(unit, -d, axis)
            flag = np.where(nb == unit, flag, 0)
    counts = [int((flag == q).sum()) for q in (1, 2, 3)]
    bulk = 0
    for q in (1, 2):
        if counts[q] > counts[bulk]:
            bulk = q
    return nx * ny * nz - counts[bulk]


@pytest.mark.parametrize("form", [oracle.ORC_FP64, oracle.ORC_MIXED])
@pytest.mark.parametrize("bc", [0, 1])
def test_diffusion_list_length(tmp_path, form, bc):
    """The diffusion list's cell count (pf_list_lengths) against a host count from the downloaded
    value_phase.  The mixed form's list holds byte pairs and counts its cells per list block, summed
    when asked: asking twice gives the same count."""
    shape = (64, 20, 12)
    banner(f"diffusion list length form={form} bc={bc}")
    nx, ny, nz = shape
    m, p, st = _setup_state(tmp_path, nx=nx, ny=ny, nz=nz, grains=5, nuclei=4,
                            bc_z="noflux" if bc else "periodic")
    g = Domain(nx, ny, nz, depth=4, nphase=2, nsolute=1,
               precision="mixed" if form == oracle.ORC_MIXED else "fp64", bc_z=bc)
    pp = oracle.as_orc_params(p)
    from phasefield_amd._capi import PfParams
    gp = PfParams.from_buffer_copy(pp)
    g.set_params(gp)
    g.upload(*st)
    g.step(1)  # a sweep, then sub-steps: the list of the stored value_phase
    g.diffuse(0, 1, gp.dt_d[0])
    _, listed = g.list_lengths()
    _, again = g.list_lengths()
    gs = g.download()
    g.close()
    want = _host_diffusion_listed(gs["phase"], shape, 2, bool(bc), p, form == oracle.ORC_MIXED)
    assert 0 < want < nx * ny * nz
    assert listed == want, f"listed {listed}, host count {want}"
    assert again == listed


@pytest.mark.parametrize("form", [oracle.ORC_FP64, oracle.ORC_MIXED])
def test_cutoff_change_between_split_sweeps(tmp_path, form):
    """The split sweep's classification reads, from the previous sweep's per-cell byte, how many own
    slots are above the cutoff; a new cutoff between steps must make it count them again."""
    shape = (64, 20, 12)
    banner(f"cutoff change between split sweeps form={form}")
    nx, ny, nz = shape
    m, p, st = _setup_state(tmp_path, nx=nx, ny=ny, nz=nz, grains=5, nuclei=4)

    def gpu(g, q):
        g.step(2)
        q.cutoff = 0.02
        g.set_params(q)
        g.step(2)

    def orc(o, q):
        o.step(2)
        q.cutoff = 0.02
        o.set_params(q)
        o.step(2)

    gs, os_ = run_pair(shape, 4, 2, 1, 0, form, st, p, gpu, orc)
    assert_same(gs, os_, 4, "cutoff change")


@pytest.mark.parametrize("form", [oracle.ORC_FP64, oracle.ORC_MIXED])
def test_bulk_mask_change_between_steps(tmp_path, form):
    """The split sweep keeps the diffusion flag pass's unit-phase bytes current from step to step; a
    partition table that stops being the identity for a phase (k(2,2) = 1.1) changes the bulk mask,
    which must recompute them, and so must the change back."""
    shape = (64, 20, 12)
    banner(f"bulk mask change between steps form={form}")
    nx, ny, nz = shape
    m, p, st = _setup_state(tmp_path, nx=nx, ny=ny, nz=nz, grains=5, nuclei=4)

    def run(d, q, setp):
        d.step(2)
        q.part_k[0][1][1] = 1.1
        setp(q)
        d.step(2)
        q.part_k[0][1][1] = 1.0
        setp(q)
        d.step(2)

    gs, os_ = run_pair(shape, 4, 2, 1, 0, form, st, p, lambda g, q: run(g, q, g.set_params),
                       lambda o, q: run(o, q, o.set_params))
    assert_same(gs, os_, 4, "bulk mask change")
