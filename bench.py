"""Benchmark: cell-updates/s of the multi-phase-field + solute-diffusion step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload fecmn_3d_512] [--no-cpu]

One step = phase_field::Update(): one phase-field sweep + 200 carbon diffusion sub-steps (+ one
manganese update for two-solute workloads), with the per-step parameter update on the host.
Default workload: 512^3 Fe-C-Mn, 512 Voronoi ferrite grains + 1000 austenite nuclei, K = 4.
`value` is the mixed form (FP32 storage/arithmetic, Kahan-compensated concentration, FP64
pruning/renormalisation); the FP64 form runs the same steps from the same seeded state and is
reported beside it.  The state is resident in HBM before the timed region; the region is bracketed
by stream synchronisation.

Roofline of the dominant kernel, the plane-marching diffusion kernel (about 80 % of a step): the
bytes it must move per launch (every cell's stored concentration read once and written once: FP64
8 + 8 B, mixed 4 + 1 B each way) over its mean launch time from HIP event pairs around a sample of
its launches inside the timed region, on the stream it runs on.  `substep` adds the listed pass
(the cells next to a phase boundary, full formula) with the list length read from the device, and
the gaps between the two launches.  `model_8d` states the same run in the compulsory-traffic model
of the scope document (every cell reads value_phase; 24 / 32 B per cell per sub-step).
After the timed steps both forms continue to step 1000 from the same state and the mixed form's
error against FP64 is reported along the way.  cpu_baseline times the oracle's OpenMP path
(oracle/, a restatement of the reference's CPU loop) on the host cores: a bounded sample of this
workload, and the reference's 2D 256^2 case in full.
"""
from __future__ import annotations

import argparse
import csv
import glob
import hashlib
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Gcell-updates/s FP32-mix vs FP64 at stated error; %HBM roofline; adjoint/fwd"
# the sources of the diffusion kernels whose HBM traffic profiles/ records (a changed source makes
# the recorded traffic stale)
TRAFFIC_SOURCES = ("phasefield_amd/csrc/pf_diffuse.hip", "phasefield_amd/csrc/pf_diffuse_common.h",
                   "phasefield_amd/csrc/pf_device.h")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def setup_fp64(m, p):
    from phasefield_amd import Domain, PF_FP64

    d = m.desc()
    d.precision = PF_FP64
    g = Domain.from_desc(d)
    g.set_params(p)
    ic = [m.get_real(f"init_conc{i}") for i in range(3)]
    if d.nsolute > 1:
        ic += [m.get_real("mn_conc")] * 3
    g.construct(m.voronoi_seeds(), m.get_real("reduction"), m.get_int("num_init_iterations"), ic)
    if d.nsolute > 1 and m.get_real("mn_noise") > 0:
        g.solute_noise(1, m.get_real("mn_conc"), m.get_real("mn_noise"), m.get_int("mn_seed"))
    g.nucleate(m.get_int("num_nuclei"), m.get_real("nucleus_radius"),
               (m.get_real("nucleus_conc"), m.get_real("mn_nucleus_conc")),
               m.get_int("nucleation_seed"))
    g.synchronize()
    return g


# ---- byte models ----------------------------------------------------------------------------------
def march_bytes_per_cell(mixed: bool) -> int:
    """Bytes the plane-marching kernel must move per cell and launch: the stored concentration read
    once and written once (FP64 8 B each way; mixed FP32 value + 4-bit compensation code, 4.5 B)."""
    return 9 if mixed else 8 + 8


def listed_bytes_per_cell(mixed: bool, nphase: int) -> float:
    """Bytes of one listed cell's full-formula update with no reuse between cells: its own and its six
    neighbours' stored concentration and value_phase read, its concentration written."""
    b_c, b_phi = (4.5, 4) if mixed else (8, 8)
    return 7 * (b_c + nphase * b_phi) + b_c


def model_8d_bytes_per_cell_step(mixed: bool, nphase: int, nsolute: int, n_sub: int) -> int:
    """The scope document's compulsory model for one whole step: a sweep (K = 4 slots of int16 id
    + phi, read and written, value_phase read and written, each solute's concentration read) and,
    per sub-step, value_phase + concentration read and the concentration written (mixed: FP32
    value + FP32 Kahan term), manganese once per step."""
    b_phi, b_c, K = (4, 8, 4) if mixed else (8, 8, 4)
    sweep = 2 * K * (2 + b_phi) + 2 * nphase * b_phi + nsolute * b_c
    sub = nphase * b_phi + 2 * b_c
    return sweep + n_sub * sub + (nsolute - 1) * sub


def kernel_source_hash() -> str:
    h = hashlib.sha256()
    for rel in TRAFFIC_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def measured_traffic(workload: str) -> dict:
    """Memory bytes per launch measured with rocprofv3: reads as the L2's read requests to the
    fabric by size (TCC_EA0_RDREQ_32B/64B/128B), writes as WRITE_SIZE, both checked on a plain
    copy of known size, from the `TRAFFIC {...}` line of a summary in profiles/ for this
    workload.  The line carries the hash of the diffusion kernels' sources it was
    measured with; `stale` is set when the sources have changed since."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return {}
    for name in sorted(os.listdir(d)):
        if not name.endswith(".md"):
            continue
        with open(os.path.join(d, name)) as f:
            for line in f:
                if line.startswith("TRAFFIC "):
                    rec = json.loads(line[len("TRAFFIC "):])
                    if rec.get("workload") == workload:
                        rec["file"] = f"profiles/{name}"
                        rec["stale"] = rec.get("sources") != kernel_source_hash()
                        return rec
    return {}


def march_kernel_name(mixed: bool, nphase: int, noflux: bool, rows: int) -> str:
    form = "FormMixed" if mixed else "FormF64"
    return (f"diffuse_bulk_march_kernel<{form}, {nphase}, 6, {'true' if noflux else 'false'}, "
            f"{rows}, {4 if mixed else 2}, 1>")


def timed_run(dom, model, steps: int, warmup: int):
    """warmup untimed steps, then exactly `steps` timed steps with HIP-event kernel timing."""
    for _ in range(warmup):
        dom.set_params(model.update_parameter())
        dom.step(1)
    dom.synchronize()
    dom.timing_reset(True)
    dom.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dom.set_params(model.update_parameter())
        dom.step(1)
    dom.synchronize()
    wall = time.perf_counter() - t0
    t = {"sub": dom.timing(0), "sweep": dom.timing(1), "march": dom.timing(2)}
    dom.timing_reset(False)
    return wall, t


# ---- host identification and CPU baselines -------------------------------------------------------
def host_info() -> dict:
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return {"cpu_model": model, "cpus": os.cpu_count(), "main_thread_affinity_cpus": affinity,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
            "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
            "OMP_PLACES": os.environ.get("OMP_PLACES")}


def _oracle_setup(m, p):
    import oracle

    d = m.desc()
    od = oracle.OracleDomain(d.nx, d.ny, d.nz, d.depth, d.nsolute)
    od.set_params(p)
    grains = oracle.tessellation(d.nx, d.ny, d.nz, m.voronoi_seeds(), 0.0)
    ic = [m.get_real(f"init_conc{i}") for i in range(3)]
    if d.nsolute > 1:
        ic += [m.get_real("mn_conc")] * 3
    od.construct(grains, m.get_int("num_init_iterations"), ic)
    od.nucleate(m.get_int("num_nuclei"), m.get_real("nucleus_radius"), (0.8, -1.0), 1)
    return od


def cpu_baseline(workload: str, tmp: str, steps: int = 12):
    """Oracle (OpenMP, reference loop schedules) on a 128^3 sample of the workload with the grain
    and nucleus densities kept: setup, then `steps` full steps timed (sweep + all sub-steps)."""
    import oracle
    from phasefield_amd import Model, configs

    over = dict(nx=128, ny=128, nz=128, grains=8, nuclei=16, precision="fp64")
    m = Model(configs.write_inputs(os.path.join(tmp, "cpu"), workload, **over))
    p = m.update_parameter()
    d = m.desc()
    od = _oracle_setup(m, p)
    p2 = m.update_parameter()
    od.set_params(p2)
    t0 = time.perf_counter()
    od.step(steps)
    dt = time.perf_counter() - t0
    n = d.nx * d.ny * d.nz
    return {"value": n * steps / dt / 1e9, "unit": "Gcell-updates/s",
            "cores": oracle.num_threads(), "kind": "port",
            "sample": f"{steps} full steps (1 sweep + {p2.n_sub[0]} carbon sub-steps"
                      f"{' + 1 Mn update' if d.nsolute > 1 else ''} each) of a 128^3 cut of the "
                      f"workload (8 grains, 16 nuclei), FP64, {dt:.1f} s",
            "host": host_info()}


def c0_baselines(tmp: str) -> dict:
    """The reference's CPU-runnable case (2D 256^2 Fe-C, 12 grains + 20 nuclei, K = 6, FP64) in
    full: 1000 steps on the oracle's OpenMP path, only the step loop timed (main.cpp:29-41), and
    the same 1000 steps on the GPU in both forms."""
    import oracle
    from phasefield_amd import Domain, Model, configs

    m = Model(configs.write_inputs(os.path.join(tmp, "c0"), "fec_2d_256"))
    p = m.update_parameter()
    d = m.desc()
    steps = 1000
    od = _oracle_setup(m, p)
    s0 = od.get_state()
    p2 = m.update_parameter()
    od.set_params(p2)
    log(f"cpu: fec_2d_256, {steps} steps on {oracle.num_threads()} threads")
    t0 = time.perf_counter()
    od.step(steps)
    cpu_s = time.perf_counter() - t0
    n = d.nx * d.ny * d.nz
    out = {"cpu": {"value": n * steps / cpu_s / 1e9, "unit": "Gcell-updates/s",
                   "cores": oracle.num_threads(), "kind": "port", "seconds": round(cpu_s, 2),
                   "sample": f"fec_2d_256 in full: {steps} steps of 256x256 Fe-C (12 grains, 20 "
                             f"nuclei, K = 6, FP64, {p2.n_sub[0]} sub-steps per step), step loop "
                             f"only", "host": host_info()}}
    for prec in ("fp64", "mixed"):
        dd = m.desc()
        dd.precision = 0 if prec == "fp64" else 1
        g = Domain.from_desc(dd)
        g.set_params(p2)
        g.upload(s0["index"], s0["value"], s0["phase"], s0["conc"])
        g.step(5)
        g.synchronize()
        t0 = time.perf_counter()
        g.step(steps)
        g.synchronize()
        gs = time.perf_counter() - t0
        g.close()
        out[f"gpu_{prec}"] = {"value": n * steps / gs / 1e9, "unit": "Gcell-updates/s",
                              "seconds": round(gs, 3)}
    return out


# ---- adjoint --------------------------------------------------------------------------------------
def adjoint_section(tmp: str, steps: int | None, fd: list):
    """Discrete adjoint on the 128^3 Fe-C adjoint workload: targets from a forward run with
    mobility x1.1 and D_gamma x0.9, loss = sum over 20 sample times of (f_gamma - target)^2
    + 1e-3 ||c_final - c_target||^2, checkpoints every 20 steps.  Reports the reverse sweep's cost
    against a plain forward run of the same steps, and central finite differences of the four
    gradients at the given relative steps, with the number of cells whose final active-set maps
    differ between the +h and -h runs."""
    import numpy as np

    from phasefield_amd import Model, configs, params_from_dict, params_to_dict
    from phasefield_amd.adjoint import AdjointProblem

    wl = "fec_3d_128_adjoint"
    nsteps = steps or configs.WORKLOADS[wl]["steps"]
    ck = max(1, nsteps // 20)
    m = Model(configs.write_inputs(os.path.join(tmp, "adj"), wl))
    p0 = m.update_parameter()
    log(f"adjoint: setup {wl}, {nsteps} steps, checkpoints every {ck}")
    g = setup_fp64(m, p0)
    s0 = g.download()
    st = (s0["index"], s0["value"], s0["phase"], s0["conc"].copy())
    N = g.N
    K = m.desc().depth
    samples = [ck * (q + 1) for q in range(nsteps // ck)][:20]
    base = params_to_dict(p0)

    def forward_loss(theta, c0, targets=None, c_target=None, wc=1e-3):
        prob.domain.set_params(params_from_dict(prob.params_at(theta)))
        g.upload(st[0], st[1], st[2], c0)
        fr, loss = [], 0.0
        for k in range(1, nsteps + 1):
            g.step(1)
            if k in samples:
                fr.append(g.reduce()["phase_frac"][0])
        s = g.download()
        if targets is not None:
            loss = sum((a - b) ** 2 for a, b in zip(fr, targets)) + wc * float(
                np.sum((s["conc"] - c_target) ** 2))
        return loss, fr, s

    prob = AdjointProblem(g, base, st[0], st[1], st[2], nsteps, samples, [0.0] * len(samples), 1e-3,
                          None, ck)
    th0 = prob.theta0().numpy()
    # targets: mobility x1.1, D_gamma x0.9
    _, f_t, s_t = forward_loss(th0 * np.array([1.1, 1.0, 0.9]), st[3])
    c_t = s_t["conc"]
    prob.target_fraction = f_t
    prob.target_conc = c_t
    # plain forward of the same steps (no sampling), timed
    g.set_params(params_from_dict(prob.params_at(th0)))
    g.upload(*st)
    g.synchronize()
    t0 = time.perf_counter()
    g.step(nsteps)
    g.synchronize()
    fwd_s = time.perf_counter() - t0
    r = prob.run(th0, st[3])
    out = {"workload": f"{wl}: 128^3 Fe-C, 32 grains + 40 nuclei, {nsteps} steps, "
                       f"{p0.n_sub[0]} sub-steps/step, checkpoints every {ck}, 20 samples",
           "gradient": "exact derivative of the computed discrete map with the step's discrete "
                       "decisions (active-set union, clamps, prunes, top-K cut, sort order) held "
                       "fixed at the forward run's; dt and n_sub frozen",
           "loss": r["loss"], "forward_ms_per_step": round(fwd_s * 1e3 / nsteps, 4),
           "adjoint_forward_ms_per_step": round(r["forward_ms"] / nsteps, 4),
           "adjoint_reverse_ms_per_step": round(r["reverse_ms"] / nsteps, 4),
           "reverse_over_forward": round(r["reverse_ms"] / (fwd_s * 1e3), 3),
           "forward_plus_reverse_over_forward": round((r["forward_ms"] + r["reverse_ms"]) /
                                                      (fwd_s * 1e3), 3),
           "grad": {"mobility_scale": r["grad_theta"][0], "sigma_scale": r["grad_theta"][1],
                    "D_gamma": r["grad_theta"][2],
                    "c0_norm": float(np.linalg.norm(r["d_conc0"]))}}
    log(f"adjoint: forward {out['forward_ms_per_step']} ms/step, reverse "
        f"{out['adjoint_reverse_ms_per_step']} ms/step ({out['reverse_over_forward']}x)")
    if fd:
        rng = np.random.default_rng(2024)
        grads = list(r["grad_theta"])
        direction = st[3] * rng.standard_normal(N)
        grads.append(float(r["d_conc0"] @ direction))
        out["fd_check"] = {"params": ["mobility_scale", "sigma_scale", "D_gamma",
                                      "c0 . (c0 * N(0,1))"], "adjoint": grads, "by_rel_step": {}}
        for h in fd:
            rows = []
            for k in range(4):
                if k < 3:
                    e = np.zeros(3)
                    e[k] = h * th0[k]
                    lp, _, sp = forward_loss(th0 + e, st[3], f_t, c_t)
                    lm, _, sm = forward_loss(th0 - e, st[3], f_t, c_t)
                    fdv = (lp - lm) / (2 * e[k])
                else:
                    lp, _, sp = forward_loss(th0, st[3] + h * direction, f_t, c_t)
                    lm, _, sm = forward_loss(th0, st[3] - h * direction, f_t, c_t)
                    fdv = (lp - lm) / (2 * h)
                differ = int(np.any(sp["index"].reshape(N, K) != sm["index"].reshape(N, K),
                                    axis=1).sum())
                rows.append({"fd": fdv, "rel_err": abs(grads[k] - fdv) / max(abs(fdv), 1e-300),
                             "cells_with_different_maps": differ})
            out["fd_check"]["by_rel_step"][f"{h:g}"] = rows
            log(f"adjoint: FD h={h:g} rel errors {[round(x['rel_err'], 6) for x in rows]} "
                f"map differences {[x['cells_with_different_maps'] for x in rows]}")
    g.close()
    return out


# ---- accuracy -------------------------------------------------------------------------------------
def accuracy_rows(m_mixed, m_fp64, mixed, g64, st0, st0_mixed, nsolute, start: int, steps: int,
                  every: int):
    """Step both forms from step `start` to step `steps` (each with its own per-step parameter
    update) and report the mixed form's error against the FP64 form every `every` steps:
    austenite fraction, max c per solute, solute-mass drift of each form."""
    rows = []
    t0 = time.perf_counter()

    def row(k):
        smx, s64 = mixed.reduce(), g64.reduce()
        fg64, fgmx = s64["phase_frac"][0], smx["phase_frac"][0]
        r = {"step": k, "austenite_fraction": fgmx, "austenite_fraction_fp64": fg64,
             "austenite_fraction_rel": abs(fgmx - fg64) / max(abs(fg64), 1e-300),
             "max_c_rel": [abs(a - b) / abs(b) for a, b in
                           zip(smx["max_conc"][:nsolute], s64["max_conc"][:nsolute])],
             "mass_drift_rel_mixed": [(a - b) / b for a, b in
                                      zip(smx["mass"][:nsolute], st0_mixed["mass"][:nsolute])],
             "mass_drift_rel_fp64": [(a - b) / b for a, b in
                                     zip(s64["mass"][:nsolute], st0["mass"][:nsolute])]}
        log(f"accuracy step {k}: f_gamma rel {r['austenite_fraction_rel']:.3e} max c rel "
            f"{[f'{x:.2e}' for x in r['max_c_rel']]} mass drift mixed "
            f"{[f'{x:.2e}' for x in r['mass_drift_rel_mixed']]} ({time.perf_counter() - t0:.0f} s)")
        return r

    rows.append(row(start))
    for k in range(start + 1, steps + 1):
        for dom, model in ((mixed, m_mixed), (g64, m_fp64)):
            dom.set_params(model.update_parameter())
            dom.step(1)
        if k % every == 0 or k == steps:
            rows.append(row(k))
    return rows


# ---- rocprofv3 counter files ----------------------------------------------------------------------
def pmc_summary(directory: str) -> dict:
    """Mean of each counter per kernel name over the dispatches in the rocprofv3 counter-collection
    CSV files under `directory` (the per-dispatch values the profiles/ summaries quote)."""
    acc = {}
    for path in glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as f:
            for rec in csv.DictReader(f):
                key = (rec.get("Kernel_Name", ""), rec.get("Counter_Name", ""))
                v = acc.setdefault(key, [0.0, 0])
                v[0] += float(rec.get("Counter_Value", 0) or 0)
                v[1] += 1
    out = {}
    for (kern, ctr), (tot, n) in sorted(acc.items()):
        out.setdefault(kern, {})[ctr] = {"mean": tot / n, "dispatches": n}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="fecmn_3d_512")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="after the timed steps, continue both forms to this step and report the "
                         "mixed form's error against FP64 along the way (0: off)")
    ap.add_argument("--accuracy-every", type=int, default=100)
    ap.add_argument("--adjoint-only", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=0)
    ap.add_argument("--adjoint-fd", default="1e-4,1e-8", metavar="H1,H2,..",
                    help="central finite differences of the adjoint gradients at these relative "
                         "steps (8 forward runs each); empty for none")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (the known-size "
                         "kernel that checks the rocprofv3 byte counters)")
    ap.add_argument("--pmc-summary", metavar="DIR",
                    help="only print the per-kernel mean of each counter in the rocprofv3 "
                         "counter-collection files under DIR")
    args = ap.parse_args()
    if args.pmc_summary:
        print(json.dumps(pmc_summary(args.pmc_summary), indent=1))
        return 0
    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes": n * 8}))
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "phasefield_amd runs one domain on one GPU; --gpus must be 1 "
                                   "(multi-GPU halo exchange is not built)"}))
        return 2
    # the oracle's OpenMP runtime reads these when it starts (cpu_baseline)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from phasefield_amd import Domain, Model, configs

    tmp = tempfile.mkdtemp(prefix="pfbench_")
    fd_steps = [float(x) for x in args.adjoint_fd.split(",") if x.strip()]
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_section(tmp, args.adjoint_steps or None,
                                                     fd_steps)}), flush=True)
        return 0
    wl = configs.WORKLOADS[args.workload]
    m_mixed = Model(configs.write_inputs(os.path.join(tmp, "mixed"), args.workload,
                                         precision="mixed"))
    m_fp64 = Model(configs.write_inputs(os.path.join(tmp, "fp64"), args.workload,
                                        precision="fp64"))
    d = m_mixed.desc()
    N = d.nx * d.ny * d.nz
    p0 = m_mixed.update_parameter()  # Construct's parameter update (time += pfm_dt)
    m_fp64.update_parameter()
    log(f"setup {args.workload} {d.nx}x{d.ny}x{d.nz} on the device")
    t0 = time.perf_counter()
    g64 = setup_fp64(m_mixed, p0)
    log(f"setup {time.perf_counter() - t0:.1f} s")
    mixed = Domain.from_desc(d)
    mixed.set_params(p0)
    mixed.copy_from(g64)
    st0 = g64.reduce()
    st0_mixed = mixed.reduce()  # after rounding the FP64 state to FP32

    res = {}
    traffic = measured_traffic(args.workload)
    noflux = d.bc_z == 1
    for name, dom, model in (("mixed", mixed, m_mixed), ("fp64", g64, m_fp64)):
        if name == "fp64" and args.no_fp64:
            continue
        is_mixed = name == "mixed"
        form = "FormMixed" if is_mixed else "FormF64"
        log(f"{name}: {args.warmup} warmup + {args.steps} timed steps")
        wall, t = timed_run(dom, model, args.steps, args.warmup)
        sweep_listed, diff_listed = dom.list_lengths()
        rows, zc = dom.march_shape()
        st = dom.reduce()
        sub_us = t["sub"][0] / t["sub"][1] * 1e3
        march_us = t["march"][0] / max(t["march"][1], 1) * 1e3
        sweep_us = t["sweep"][0] / t["sweep"][1] * 1e3
        bpc = march_bytes_per_cell(is_mixed)
        achieved = N * bpc / (march_us * 1e-6) / 1e9
        kname = march_kernel_name(is_mixed, d.nphase, noflux, rows)
        # the summary records each profiled tile shape's bytes by kernel name; the tuner's pick
        # on this device selects one (none when it picked a shape the summary did not profile)
        measured = (traffic.get("march", {}).get(kname)
                    if traffic and not traffic.get("stale") else None)
        lbpc = listed_bytes_per_cell(is_mixed, d.nphase)
        sub_bytes = N * bpc + diff_listed * lbpc
        res[name] = {
            "wall": wall, "ms_per_step": wall / args.steps * 1e3, "value": N * args.steps / wall / 1e9,
            "stats": st,
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "traffic": measured,
                         "kernel": kname, "tile_rows": rows, "z_chunk": zc,
                         "bytes_per_cell": bpc, "cells_per_launch": N,
                         "launch_us": round(march_us, 2), "launches_timed": t["march"][1],
                         "traffic_source": (f"{traffic.get('file')} (kernel sources "
                                            f"{traffic.get('sources')}"
                                            f"{', stale: not used' if traffic.get('stale') else ''})"
                                            if traffic else None),
                         "traffic_frac": (round(measured / (march_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
                                          if measured else None)},
            "substep": {"us": round(sub_us, 2), "launches_timed": t["sub"][1],
                        "listed_cells": diff_listed,
                        "listed_fraction": round(diff_listed / N, 6),
                        "bytes": sub_bytes, "listed_bytes_per_cell": lbpc,
                        "achieved_gbs": round(sub_bytes / (sub_us * 1e-6) / 1e9, 1),
                        "frac": round(sub_bytes / (sub_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)},
            "sweep": {"us": round(sweep_us, 1), "listed_cells": sweep_listed,
                      "listed_fraction": round(sweep_listed / N, 6)},
        }
        log(f"{name}: {res[name]['ms_per_step']:.2f} ms/step, march {march_us:.1f} us "
            f"({achieved:.0f} GB/s), sub-step {sub_us:.1f} us ({res[name]['substep']['frac']:.3f} of "
            f"peak, {diff_listed} listed), sweep {sweep_us:.0f} us ({sweep_listed} listed)")

    mx = res["mixed"]
    n_sub = p0.n_sub[0]
    out = {
        "metric": METRIC, "value": round(mx["value"], 5), "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(mx["ms_per_step"], 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded Voronoi + nucleation)",
        "config": {"workload": f"{args.workload}: {d.nx}^3 Fe-C{'-Mn' if d.nsolute > 1 else ''}, "
                               f"{wl['grains']} Voronoi ferrite grains + {wl['nuclei']} austenite "
                               f"nuclei (r=4), {'no-flux z' if noflux else 'periodic'}, "
                               f"K={d.depth}, P={d.nphase}, "
                               f"{n_sub} carbon sub-steps"
                               f"{' + 1 Mn update' if d.nsolute > 1 else ''} per step",
                   "form": "mixed (FP32 + Kahan, FP64 prune/renormalise)",
                   "nx": d.nx, "ny": d.ny, "nz": d.nz},
        "roofline": mx["roofline"],
        "substep": mx["substep"],
        "sweep": mx["sweep"],
    }
    for name in res:
        b8 = model_8d_bytes_per_cell_step(name == "mixed", d.nphase, d.nsolute, n_sub)
        gbs = N * b8 / (res[name]["ms_per_step"] * 1e-3) / 1e9
        res[name]["model_8d"] = {"bytes_per_cell_step": b8, "achieved_gbs": round(gbs, 1),
                                 "ratio_to_peak": round(gbs / HBM_PEAK_GBS, 4)}
    out["model_8d"] = mx["model_8d"]
    if "fp64" in res:
        f64 = res["fp64"]
        out["fp64"] = {"value": round(f64["value"], 5), "ms_per_step": round(f64["ms_per_step"], 3),
                       "dtype": "f64", "roofline": f64["roofline"], "substep": f64["substep"],
                       "sweep": f64["sweep"], "model_8d": f64["model_8d"]}
        out["mixed_over_fp64"] = round(mx["value"] / f64["value"], 4)
        out["march_fp64_over_mixed"] = round(f64["roofline"]["launch_us"] / mx["roofline"]["launch_us"], 4)
        done = args.warmup + args.steps
        if args.accuracy_steps > done:
            log(f"accuracy: both forms to step {args.accuracy_steps}")
            rows = accuracy_rows(m_mixed, m_fp64, mixed, g64, st0, st0_mixed, d.nsolute, done,
                                 args.accuracy_steps, args.accuracy_every)
        else:
            rows = accuracy_rows(m_mixed, m_fp64, mixed, g64, st0, st0_mixed, d.nsolute, done,
                                 done, 1)
        last = rows[-1]
        out["error_vs_fp64"] = {
            "after_steps": last["step"],
            "austenite_fraction_rel": last["austenite_fraction_rel"],
            "max_austenite_fraction_rel": max(r["austenite_fraction_rel"] for r in rows),
            "max_c_rel": last["max_c_rel"],
            "mass_drift_rel_mixed": last["mass_drift_rel_mixed"],
            "mass_fp32_rounding_of_initial_state_rel": [
                abs(a - b) / b for a, b in zip(st0_mixed["mass"][: d.nsolute],
                                               st0["mass"][: d.nsolute])],
            "mass_drift_rel_fp64": last["mass_drift_rel_fp64"],
            "trajectory": [{"step": r["step"], "f_gamma_rel": r["austenite_fraction_rel"],
                            "max_c_rel": r["max_c_rel"],
                            "mass_drift_mixed": r["mass_drift_rel_mixed"],
                            "mass_drift_fp64": r["mass_drift_rel_fp64"]} for r in rows],
        }
    mixed.close()
    g64.close()
    if not args.no_adjoint:
        out["adjoint"] = adjoint_section(tmp, args.adjoint_steps or None, fd_steps)
    if not args.no_cpu:
        log("cpu baseline (oracle, OpenMP)")
        out["cpu_baseline"] = cpu_baseline(args.workload, tmp)
        out["configs0"] = c0_baselines(tmp)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
