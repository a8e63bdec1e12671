"""Gradient harness over the discrete adjoint of the FP64 step (pf_adjoint_run).

`AdjointProblem` fixes a domain, its starting microstructure (active-set map, slot values,
value_phase) and a loss; `AdjointLoss` is a torch.autograd.Function whose forward runs the GPU
forward + reverse sweep once and whose backward hands out the gradients that sweep produced:

    theta = (mobility scale, interface-energy scale, carbon diffusivity of phase 1)
    loss  = sum_k (f_gamma(t_k) - target_k)^2 + weight_conc * ||c(t_end) - c_target||^2

The adjoint differentiates the arithmetic of every step with its discrete structure (active-set
union, clamp/prune decisions, the sort permutation, the top-K cut) held fixed, as a finite
difference small enough to stay inside one smooth piece sees it.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Domain, params_from_dict, params_to_dict

__all__ = ["AdjointProblem", "AdjointLoss", "adjoint_loss"]


class AdjointProblem:
    def __init__(self, domain: Domain, params, index, value, phase, n_steps: int,
                 sample_steps=(), target_fraction=(), weight_conc: float = 0.0,
                 target_conc=None, checkpoint_every: int = 1):
        if domain.mixed:
            raise ValueError("the adjoint runs on FP64 domains")
        self.domain = domain
        self.base = params_to_dict(params) if not isinstance(params, dict) else dict(params)
        self.index = np.ascontiguousarray(index, np.int32).reshape(-1)
        self.value = np.ascontiguousarray(value, np.float64).reshape(-1)
        self.phase = np.ascontiguousarray(phase, np.float64).reshape(-1)
        self.n_steps = int(n_steps)
        self.sample_steps = list(sample_steps)
        self.target_fraction = list(target_fraction)
        self.weight_conc = float(weight_conc)
        self.target_conc = None if target_conc is None else np.asarray(target_conc, np.float64)
        self.checkpoint_every = int(checkpoint_every)
        self.last: dict | None = None

    def theta0(self) -> torch.Tensor:
        return torch.tensor([1.0, 1.0, float(self.base["diffusivity"][0][0])], dtype=torch.float64)

    def params_at(self, theta) -> dict:
        s_m, s_s, d_g = (float(x) for x in theta)
        d = {k: (np.array(v, copy=True) if isinstance(v, np.ndarray) else v)
             for k, v in self.base.items()}
        d["mobility"] = d["mobility"] * s_m
        d["sigma"] = d["sigma"] * s_s
        d["diffusivity"][0][0] = d_g
        return d

    def run(self, theta, conc0) -> dict:
        """Set parameters and state, run forward + reverse; gradients refer to theta."""
        theta = [float(x) for x in theta]
        self.domain.set_params(params_from_dict(self.params_at(theta)))
        self.domain.upload(self.index, self.value, self.phase,
                           np.ascontiguousarray(conc0, np.float64).reshape(-1))
        r = self.domain.adjoint(self.n_steps, self.sample_steps, self.target_fraction,
                                self.weight_conc, self.target_conc, self.checkpoint_every)
        # the library's scale derivatives are taken at the parameters it was given; theta scales
        # the base parameters, so d/d(theta) = d/d(scale at 1) / theta
        r["grad_theta"] = np.array([r["d_mobility"] / theta[0] if theta[0] != 0 else 0.0,
                                    r["d_sigma"] / theta[1] if theta[1] != 0 else 0.0,
                                    r["d_diff_gamma"]])
        self.last = r
        return r


class AdjointLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta: torch.Tensor, conc0: torch.Tensor, problem: AdjointProblem):
        r = problem.run(theta.detach().cpu().numpy(), conc0.detach().cpu().numpy())
        ctx.save_for_backward(torch.from_numpy(r["grad_theta"]).to(theta.device, theta.dtype),
                              torch.from_numpy(r["d_conc0"]).to(conc0.device, conc0.dtype)
                              .reshape(conc0.shape))
        return torch.tensor(r["loss"], dtype=torch.float64, device=theta.device)

    @staticmethod
    def backward(ctx, grad_out):
        g_theta, g_conc = ctx.saved_tensors
        return grad_out * g_theta, grad_out * g_conc, None


def adjoint_loss(theta: torch.Tensor, conc0: torch.Tensor, problem: AdjointProblem) -> torch.Tensor:
    return AdjointLoss.apply(theta, conc0, problem)
