"""torch.autograd.Function around one implicit time step of the SSA solver.

forward(beta, H_prev, B, smb) -> (u, v, H): sets beta, B, smb and H_prev on
the finest level, restarts the state from the saved warm start (u0, v0,
H_prev), runs FASCD V-cycles (to a relative residual tolerance or a fixed
count) on the GPU.

backward(dL/du, dL/dv, dL/dH): the forward saves only its inputs; backward
re-solves the step from them to recover the state it linearises about
(nothing from inside the nonlinear solve is stored across steps), then runs
adjoint V-cycles (glide/solver.py:160-220) until ||f_adj - J^T lambda|| is
below adj_tol * ||f_adj|| and returns
    dL/dbeta   = lambda . dR/dbeta              (glide/cuda/grad.cu:1-74, exact transpose)
    dL/dH_prev = -lambda_H / dt  on cells outside the adjoint's active set
                 (f_H = H_prev/dt + smb, glide/physics.py:173; masked cells'
                 rows are H - gamma and do not depend on H_prev)
    dL/dsmb    = -lambda_H       on the same cells
    dL/dB      = lambda . dR/dB  (through the viscosity, d eta / d B = eta / B)
Only the gradients of inputs that require them are formed.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from .physics import IcePhysics


@dataclass
class SolveConfig:
    dt: float = 10.0
    tol: float | None = 1e-12      # relative residual; None = fixed n_vcycles
    n_vcycles: int = 3
    max_vcycles: int = 60
    adj_tol: float = 1e-12
    max_adj_vcycles: int = 200


class SSAStep(torch.autograd.Function):
    @staticmethod
    def forward(ctx, beta, H_prev, B, smb, physics: IcePhysics, cfg: SolveConfig, u0, v0):
        g = physics.grid
        _solve(physics, cfg, beta, H_prev, B, smb, u0, v0)
        ctx.physics, ctx.cfg = physics, cfg
        ctx.save_for_backward(beta, H_prev, B, smb, u0, v0)
        return g.u.clone(), g.v.clone(), g.H.clone()

    @staticmethod
    def backward(ctx, gu, gv, gH):
        physics, cfg = ctx.physics, ctx.cfg
        beta, H_prev, B, smb, u0, v0 = ctx.saved_tensors
        # recompute the linearisation point from the saved inputs
        _solve(physics, cfg, beta, H_prev, B, smb, u0, v0)
        g = physics.grid
        z = torch.zeros
        gu = gu if gu is not None else z(g.u.shape, dtype=torch.float64, device=g.U.device)
        gv = gv if gv is not None else z(g.v.shape, dtype=torch.float64, device=g.U.device)
        gH = gH if gH is not None else z(g.H.shape, dtype=torch.float64, device=g.U.device)
        physics.adjoint_solve(gu, gv, gH, tol=cfg.adj_tol, max_vcycles=cfg.max_adj_vcycles)
        need = ctx.needs_input_grad
        grad_beta = physics.grad_beta(reference_edges=False) if need[0] else None
        grad_Hprev = physics.grad_H_prev() if need[1] else None
        grad_B = physics.grad_B() if need[2] else None
        grad_smb = physics.grad_smb() if need[3] else None
        return grad_beta, grad_Hprev, grad_B, grad_smb, None, None, None, None


def _solve(physics: IcePhysics, cfg: SolveConfig, beta, H_prev, B, smb, u0, v0):
    g = physics.grid
    g.u.copy_(u0)
    g.v.copy_(v0)
    g.H.copy_(H_prev)
    g.H_prev.copy_(H_prev)
    g.gamma.fill_(physics.thklim)
    physics.set_parameters(beta=beta, B=B, smb=smb)
    if cfg.tol is None:
        physics.forward(cfg.dt, n_vcycles=cfg.n_vcycles, update_geometry=False)
    else:
        physics.forward(cfg.dt, tol=cfg.tol, max_vcycles=cfg.max_vcycles, update_geometry=False)


def ssa_step(physics: IcePhysics, beta: torch.Tensor, H_prev: torch.Tensor, cfg: SolveConfig | None = None,
             u0: torch.Tensor | None = None, v0: torch.Tensor | None = None, B: torch.Tensor | None = None,
             smb: torch.Tensor | None = None):
    """Differentiable implicit step: returns (u, v, H) with gradients to beta,
    H_prev, B and smb (B and smb default to the hierarchy's current fields,
    taken as constants).  u0, v0 = warm start (default zeros)."""
    cfg = cfg or SolveConfig()
    g = physics.grid
    if u0 is None:
        u0 = torch.zeros_like(g.u)
    if v0 is None:
        v0 = torch.zeros_like(g.v)
    if B is None:
        B = g.B.detach().clone()
    if smb is None:
        smb = g.smb.detach().clone()
    return SSAStep.apply(beta, H_prev, B, smb, physics, cfg, u0, v0)
