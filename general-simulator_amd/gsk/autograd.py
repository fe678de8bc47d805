"""torch.autograd binding of the library's adjoint (gsk_adjoint_forward / gsk_adjoint_backward).

``rollout(sim, eta0, params, nsteps, K)`` returns eta(T) as a torch tensor whose gradient
flows back to ``eta0`` and to ``params = [s_1..s_V, L, kappa]``: the eigenstrain magnitudes
s_p (eigenstrain p is s_p times the base eigenstrain the rollout was built with), the
mobility L and the gradient coefficient kappa. The forward steps the FP32 fused path and
keeps a checkpoint every K steps; the backward is the hand-written reverse sweep on the GPU.
The reference has no reverse mode; this differentiates the martensite step it defines
(src/dynamics_mart.cu:270-303 elasticity, the semi-implicit update of the step).
"""
from __future__ import annotations

import numpy as np
import torch

from . import Sim


class MartRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eta0, params, sim: Sim, C, eps_base, nsteps: int, K: int, step0: int):
        p = params.detach().cpu().double().numpy()
        V = sim.nvar
        s = p[:V]
        sim.build_btensor(C, eps_base * s[:, None, None])
        sim.set_param("L", float(p[V]))
        sim.set_param("kappa", float(p[V + 1]))
        sim.set_field("eta", eta0.detach().cpu().float().numpy())
        sim.set_step(step0)  # noise stream position of the first step
        sim.adjoint_forward(nsteps, K)
        ctx.sim, ctx.s = sim, s
        ctx.dev, ctx.dtype = eta0.device, eta0.dtype
        return torch.from_numpy(sim.get_field("eta")).to(device=eta0.device, dtype=eta0.dtype)

    @staticmethod
    def backward(ctx, g):
        ge, gp = ctx.sim.adjoint_backward(g.detach().cpu().float().numpy())
        V = ctx.sim.nvar
        gp = gp.copy()
        gp[:V] = gp[:V] / ctx.s  # the library returns d/ds relative to the built tensor (s = 1)
        g_eta = torch.from_numpy(ge).to(device=ctx.dev, dtype=ctx.dtype)
        g_par = torch.from_numpy(gp).to(device=ctx.dev, dtype=torch.float64)
        return g_eta, g_par, None, None, None, None, None, None


def rollout(sim: Sim, eta0: torch.Tensor, params: torch.Tensor, C: np.ndarray, eps_base: np.ndarray, nsteps: int,
            K: int = 20, step0: int = 0) -> torch.Tensor:
    """eta(T) after nsteps of the FP32 fused martensite step from step index step0,
    differentiable in eta0 and params."""
    return MartRollout.apply(eta0, params, sim, C, np.asarray(eps_base, dtype=np.float64), nsteps, K, step0)
