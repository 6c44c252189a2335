"""Adjoint inversion of the basal friction beta (the objective of
examples/greenland/greenland_inverse.py:176-262, driven by scipy's L-BFGS-B on
the host with every forward and adjoint solve on the GPU).

Two objective forms:
  reference : |misfit| (abs_loss), Tikhonov on beta added to the log-beta
              gradient (greenland_inverse.py:194-215), 5 forward V-cycles
              from a reset state, one adjoint V-cycle, the reference's
              edge handling of grad beta;
  l2        : 1/2 ||u - u_obs||^2 + 1/2 ||v - v_obs||^2 + alpha * Tikhonov(log
              beta), forward solved to a relative-residual tolerance and the
              adjoint to a tolerance, exact-transpose gradient.
Both optimise x = log beta with bounds (-6, 5) (greenland_inverse.py:239-247).
"""
from __future__ import annotations

import numpy as np
import torch
from scipy.optimize import minimize

from .physics import IcePhysics, abs_grad, abs_loss, tikhonov_regularization


class BetaInversion:
    def __init__(self, physics: IcePhysics, u_obs, v_obs, H_prev, dt=10.0, alpha=1e-3, mode="l2",
                 fwd_tol=1e-12, adj_tol=1e-12, max_vcycles=60, n_fwd_vcycles=5, n_adj_vcycles=1):
        self.ph = physics
        dev = physics.device
        self.u_obs = torch.as_tensor(np.asarray(u_obs, dtype=np.float64), device=dev)
        self.v_obs = torch.as_tensor(np.asarray(v_obs, dtype=np.float64), device=dev)
        self.H_prev = torch.as_tensor(np.asarray(H_prev, dtype=np.float64), device=dev)
        self.dt, self.alpha, self.mode = float(dt), float(alpha), mode
        self.fwd_tol, self.adj_tol, self.max_vcycles = fwd_tol, adj_tol, max_vcycles
        self.n_fwd_vcycles, self.n_adj_vcycles = n_fwd_vcycles, n_adj_vcycles
        self.history: list[dict] = []

    # ---------------------------------------------------------------- solves
    def _forward(self, beta: torch.Tensor):
        g = self.ph.grid
        g.u.zero_()
        g.v.zero_()
        g.H.copy_(self.H_prev)
        g.H_prev.copy_(self.H_prev)
        self.ph.set_parameters(beta=beta)
        if self.mode == "reference":
            self.ph.forward(self.dt, n_vcycles=self.n_fwd_vcycles, update_geometry=False)
        else:
            self.ph.forward(self.dt, tol=self.fwd_tol, max_vcycles=self.max_vcycles, update_geometry=False)

    def misfit(self, log_beta: np.ndarray) -> float:
        beta = torch.exp(torch.as_tensor(log_beta.reshape(self.ph.ny, self.ph.nx), device=self.ph.device))
        self._forward(beta)
        g = self.ph.grid
        return float(0.5 * ((g.u - self.u_obs) ** 2).sum() + 0.5 * ((g.v - self.v_obs) ** 2).sum())

    def objective(self, log_beta_flat: np.ndarray):
        ph = self.ph
        g = ph.grid
        log_beta = torch.as_tensor(log_beta_flat.reshape(ph.ny, ph.nx), dtype=torch.float64, device=ph.device)
        beta = torch.exp(log_beta)
        self._forward(beta)
        zH = torch.zeros_like(g.H)
        if self.mode == "reference":
            J = float(abs_loss(g.u, g.v, self.u_obs, self.v_obs))
            dJu, dJv = abs_grad(g.u, g.v, self.u_obs, self.v_obs)
            grad_beta = ph.adjoint(dJu, dJv, zH, n_vcycles=self.n_adj_vcycles, reference_edges=True)
            reg, reg_grad = tikhonov_regularization(beta)
        else:
            du, dv = g.u - self.u_obs, g.v - self.v_obs
            J = float(0.5 * (du ** 2).sum() + 0.5 * (dv ** 2).sum())
            ph.adjoint_solve(du, dv, zH, tol=self.adj_tol)
            grad_beta = ph.grad_beta(reference_edges=False)
            reg, reg_grad = tikhonov_regularization(log_beta)
        total = J + self.alpha * reg
        grad = beta * grad_beta + self.alpha * reg_grad
        self.history.append({"misfit": J, "reg": self.alpha * reg, "total": total})
        return total, grad.reshape(-1).cpu().numpy()

    def run(self, log_beta0: np.ndarray, maxiter=50, bounds=(-6.0, 5.0)):
        x0 = np.asarray(log_beta0, dtype=np.float64).reshape(-1)
        res = minimize(self.objective, x0, jac=True, method="L-BFGS-B",
                       bounds=[bounds] * x0.size, options={"maxiter": maxiter, "maxcor": 15})
        return res
