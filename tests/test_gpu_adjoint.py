"""Discrete adjoint on the GPU (pf_adjoint_run) against the oracle's forward-mode dual numbers
(exact derivatives of the same FP64 computation, independent code) and against central finite
differences of the GPU forward; checkpoint spacing must not change the result; the torch wrapper
must hand out the same gradients.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from phasefield_amd import Domain
from phasefield_amd._capi import PfParams
from phasefield_amd.adjoint import AdjointProblem, adjoint_loss
from tests import helpers

pytestmark = pytest.mark.gpu

N_SUB = 20
STEPS = 3


def banner(msg):
    print(f"[gpu] {msg}", flush=True)


def _case(tmp_path, shape, bc, K, P):
    nx, ny, nz = shape
    m = helpers.small_model(str(tmp_path), "fec_3d_256", nx=nx, ny=ny, nz=nz, grains=4, nuclei=3,
                            depth=K, bc_z=bc)
    p = helpers.model_params(m)
    od = helpers.oracle_initial_state(m, p)
    s = od.get_state()
    pp = oracle.as_orc_params(p)
    pp.n_sub[0] = N_SUB
    pp.dims = 3 - (nz == 1) - (ny == 1)
    if P == 3:
        for (i, j) in ((1, 2), (2, 1), (2, 2)):
            pp.sigma[i][j] = 0.3e-12
            pp.mobility[i][j] = 0.5e12
        pp.diffusivity[0][2] = 5.0
    st = (s["index"].copy(), s["value"].copy(), s["phase"].copy(), s["conc"].copy())
    return pp, st


def _loss_spec(st, rng, f0):
    samples = list(range(1, STEPS + 1))
    targets = [f0 - 0.01 * (q + 1) for q in range(STEPS)]
    target_c = st[3] * (1.0 + 0.05 * rng.standard_normal(st[3].size))
    return samples, targets, 3.0, target_c


def _gpu_domain(shape, bc, K, P, pp):
    g = Domain(*shape, depth=K, nphase=P, nsolute=1, precision="fp64", bc_z=bc)
    g.set_params(PfParams.from_buffer_copy(pp))
    return g


CASES = [((20, 16, 12), "periodic", 4, 2, "1"), ((32, 28, 1), "periodic", 6, 2, "1"),
         ((18, 14, 10), "noflux", 4, 3, "1"), ((21, 16, 10), "periodic", 4, 2, "1"),
         ((20, 16, 12), "periodic", 4, 2, "0")]


@pytest.mark.parametrize("shape,bc,K,P,split", CASES)
def test_adjoint_matches_dual_numbers(tmp_path, monkeypatch, shape, bc, K, P, split):
    """Reverse sub-steps split (bulk cells by the forward bulk kernel on lambda: 16-byte runs, or the
    scalar column kernel when nx is odd; listed cells apart) and unsplit."""
    banner(f"adjoint vs dual {shape} {bc} K={K} P={P} split={split}")
    monkeypatch.setenv("PFAMD_ADJ_SPLIT", split)
    rng = np.random.default_rng(7)
    pp, st = _case(tmp_path, shape, bc, K, P)
    bcn = 1 if bc == "noflux" else 0
    direction = rng.standard_normal(st[3].size) * 1e-2
    dual = oracle.OracleDomain(*shape, K, 1, oracle.ORC_DUAL, bcn)
    dual.set_params(pp)
    dual.set_state(*st)
    f0 = float(st[2].reshape(-1, 3)[:, 0].mean())
    samples, targets, wc, target_c = _loss_spec(st, rng, f0)
    dual.dual_seed(direction)
    expect = np.zeros(4)
    fr = []
    for q in range(STEPS):
        dual.step(1)
        f, df = dual.dual_phase_fraction()
        fr.append(f)
        expect += 2.0 * (f - targets[q]) * np.asarray(df)
    c, dc = dual.dual_conc()
    dc = np.asarray(dc).reshape(4, -1)
    expect += 2.0 * wc * dc @ (c - target_c)
    loss_ref = sum((a - b) ** 2 for a, b in zip(fr, targets)) + wc * float(np.sum((c - target_c) ** 2))

    g = _gpu_domain(shape, bcn, K, P, pp)
    g.upload(*st)
    r = g.adjoint(STEPS, samples, targets, wc, target_c, checkpoint_every=2)
    got = np.array([r["d_mobility"], r["d_sigma"], r["d_diff_gamma"],
                    float(r["d_conc0"] @ direction)])
    assert np.allclose(r["fraction"], fr, rtol=0, atol=1e-13)
    assert abs(r["loss"] - loss_ref) <= 1e-12 * abs(loss_ref)
    banner(f"adjoint {got.tolist()} dual {expect.tolist()}")
    assert np.all(np.abs(expect) > 0), "degenerate case: a gradient is exactly zero"
    for k, name in enumerate(("mobility", "sigma", "D_gamma", "c0 . dir")):
        err = abs(got[k] - expect[k])
        assert err <= 1e-8 * max(abs(expect[k]), 1e-30) + 1e-14, \
            f"{name}: adjoint {got[k]!r} vs dual {expect[k]!r}"


def test_adjoint_matches_central_differences(tmp_path):
    """dL/dtheta from the reverse sweep against (L(theta+h) - L(theta-h)) / 2h of the GPU forward.
    h = 1e-6 relative keeps the perturbed runs inside one smooth piece of the step."""
    shape, bc, K, P = (24, 20, 8), "periodic", 4, 2
    banner(f"adjoint vs FD {shape}")
    rng = np.random.default_rng(11)
    pp, st = _case(tmp_path, shape, bc, K, P)
    g = _gpu_domain(shape, 0, K, P, pp)
    f0 = float(st[2].reshape(-1, 3)[:, 0].mean())
    samples, targets, wc, target_c = _loss_spec(st, rng, f0)
    prob = AdjointProblem(g, PfParams.from_buffer_copy(pp), st[0], st[1], st[2], STEPS, samples,
                          targets, wc, target_c)
    th = prob.theta0().numpy()
    r = prob.run(th, st[3])
    grad = r["grad_theta"].copy()
    gc = r["d_conc0"].copy()

    def L(theta, conc):
        return prob.run(theta, conc)["loss"]

    h = 1e-6
    for k in range(3):
        e = np.zeros(3)
        e[k] = h * th[k]
        fd = (L(th + e, st[3]) - L(th - e, st[3])) / (2 * e[k])
        assert abs(fd - grad[k]) <= 1e-4 * abs(grad[k]) + 1e-10, f"theta[{k}]: {grad[k]} vs {fd}"
    d = rng.standard_normal(st[3].size) * 1e-2
    fd = (L(th, st[3] + h * d) - L(th, st[3] - h * d)) / (2 * h)
    assert abs(fd - gc @ d) <= 1e-4 * abs(fd) + 1e-10


def test_checkpoint_spacing_and_final_state(tmp_path, monkeypatch):
    """Gradients do not depend on the checkpoint spacing or on whether a segment's sub-step
    concentrations are stored or recomputed per step (the recomputation is deterministic), and the
    domain is left in the plain forward run's final state."""
    shape, bc, K, P = (20, 18, 10), "periodic", 4, 2
    banner("adjoint checkpoint spacing")
    rng = np.random.default_rng(5)
    pp, st = _case(tmp_path, shape, bc, K, P)
    f0 = float(st[2].reshape(-1, 3)[:, 0].mean())
    samples, targets, wc, target_c = _loss_spec(st, rng, f0)
    g = _gpu_domain(shape, 0, K, P, pp)
    res = []
    for ck, mode in ((1, "segment"), (2, "segment"), (STEPS, "segment"), (2, "step")):
        monkeypatch.setenv("PFAMD_ADJ_RECOMPUTE", mode)
        g.upload(*st)
        r = g.adjoint(STEPS, samples, targets, wc, target_c, checkpoint_every=ck)
        res.append(r)
        fin = g.download()
    ref = _gpu_domain(shape, 0, K, P, pp)
    ref.upload(*st)
    ref.step(STEPS)
    plain = ref.download()
    for key in ("index", "value", "phase", "conc"):
        assert np.array_equal(fin[key], plain[key]), key
    for r in res[1:]:
        for key in ("loss", "d_mobility", "d_sigma", "d_diff_gamma"):
            assert r[key] == res[0][key], key
        assert np.array_equal(r["d_conc0"], res[0]["d_conc0"])


def test_torch_autograd_wrapper(tmp_path):
    shape, K, P = (16, 16, 8), 4, 2
    banner("adjoint torch wrapper")
    rng = np.random.default_rng(9)
    pp, st = _case(tmp_path, shape, "periodic", K, P)
    f0 = float(st[2].reshape(-1, 3)[:, 0].mean())
    samples, targets, wc, target_c = _loss_spec(st, rng, f0)
    g = _gpu_domain(shape, 0, K, P, pp)
    prob = AdjointProblem(g, PfParams.from_buffer_copy(pp), st[0], st[1], st[2], STEPS, samples,
                          targets, wc, target_c)
    theta = (prob.theta0() * torch.tensor([1.5, 0.8, 1.0], dtype=torch.float64)).requires_grad_()
    conc = torch.from_numpy(st[3].copy()).to("cuda").requires_grad_()
    loss = adjoint_loss(theta, conc, prob)
    (2.0 * loss).backward()
    r = prob.run(theta.detach().numpy(), st[3])
    assert loss.item() == r["loss"]
    assert np.array_equal(theta.grad.numpy(), 2.0 * r["grad_theta"])
    assert np.array_equal(conc.grad.cpu().numpy(), 2.0 * r["d_conc0"])
    # theta scales the base parameters: d/d(theta_0) at 1.5 by central differences
    h = 1e-6
    th = theta.detach().numpy()
    e = np.array([h * th[0], 0, 0])
    fd = (prob.run(th + e, st[3])["loss"] - prob.run(th - e, st[3])["loss"]) / (2 * e[0])
    assert abs(fd - r["grad_theta"][0]) <= 1e-4 * abs(fd) + 1e-10


def test_adjoint_rejects_bad_spec():
    g = Domain(8, 8, 4, depth=4, nphase=2, precision="fp64")
    with pytest.raises(Exception):
        g.adjoint(2)  # no parameters / state
    gm = Domain(8, 8, 4, depth=4, nphase=2, precision="mixed")
    with pytest.raises(Exception):
        gm.adjoint(1)
