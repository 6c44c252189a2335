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

MultiResolutionInversion runs the reference objective coarse to fine over the
levels of one hierarchy (greenland_inverse.py:123-262): observations are
restricted to every level by the face restrictions, each level is optimised
with L-BFGS-B (factr 1e11, m 15) on the sub-hierarchy below it, and its beta
is prolonged bilinearly to the next finer level as that level's start.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import torch
from scipy.optimize import fmin_l_bfgs_b, minimize

from . import _lib
from .physics import IcePhysics, _as_f64, abs_grad, abs_loss, tikhonov_regularization


class BetaInversion:
    def __init__(self, physics: IcePhysics, u_obs, v_obs, H_prev, dt=10.0, alpha=1e-3, mode="l2",
                 fwd_tol=1e-12, adj_tol=1e-12, max_vcycles=60, n_fwd_vcycles=5, n_adj_vcycles=1):
        self.ph = physics
        dev = physics.device
        self.u_obs = _as_f64(u_obs, dev).clone()
        self.v_obs = _as_f64(v_obs, dev).clone()
        self.H_prev = _as_f64(H_prev, dev).clone()
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


def restrict_observations(physics: IcePhysics, u_obs, v_obs) -> list:
    """(u_obs, v_obs) on every level of ``physics``, finest first: face
    restriction u_c[J, I] = (u[2J, 2I] + u[2J+1, 2I]) / 2 and its v analogue
    on the GPU (glide_restrict_u / glide_restrict_v; greenland_inverse.py:
    128-135 with restrict_vfacet / restrict_hfacet, kernels.py:90-128)."""
    dev, st = physics.device, physics._stream()
    u = torch.as_tensor(np.asarray(u_obs, dtype=np.float64), device=dev).contiguous()
    v = torch.as_tensor(np.asarray(v_obs, dtype=np.float64), device=dev).contiguous()
    g0 = physics.grid
    if tuple(u.shape) != (g0.ny, g0.nx + 1) or tuple(v.shape) != (g0.ny + 1, g0.nx):
        raise ValueError(f"observations must be u ({g0.ny}, {g0.nx + 1}) and v ({g0.ny + 1}, {g0.nx})")
    out = [(u, v)]
    for g in physics.grids[1:]:
        uc = torch.empty((g.ny, g.nx + 1), dtype=torch.float64, device=dev)
        vc = torch.empty((g.ny + 1, g.nx), dtype=torch.float64, device=dev)
        _lib.call("glide_restrict_u", ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(uc.data_ptr()), g.ny, g.nx, st)
        _lib.call("glide_restrict_v", ctypes.c_void_p(v.data_ptr()), ctypes.c_void_p(vc.data_ptr()), g.ny, g.nx, st)
        out.append((uc, vc))
        u, v = uc, vc
    return out


class MultiResolutionInversion:
    """Coarse-to-fine inversion of beta (examples/greenland/greenland_inverse.py:123-262).

    For level k = n_levels - 1, ..., 0 (``levels``): the objective of
    BetaInversion (``mode``, default the reference's) runs on
    ``physics.level_view(k)`` against the observations restricted to level k;
    the level's momentum forcing is zero, as on a level that was never a
    coarse grid of a V-cycle; L-BFGS-B (fmin_l_bfgs_b with factr 1e11,
    m = 15, bounds (-6, 5)) starts from the level's current log beta; the
    optimum is written back as beta_k and prolonged to level k - 1 by the
    bilinear cell prolongation (prolongate_cell_centered with smooth = True,
    i.e. prolongate_H_smooth; greenland_inverse.py:257-260).

    ``dt`` defaults to 1: the reference example never sets the time step, so
    its hierarchy keeps the Grid default (glide/physics.py:70-71).
    ``callback(k, iteration, view)`` is called after every L-BFGS iteration
    (the example writes VTI snapshots there).
    """

    def __init__(self, physics: IcePhysics, u_obs, v_obs, reg_weight=1e-4, dt=1.0, mode="reference",
                 n_fwd_vcycles=5, n_adj_vcycles=1, factr=1e11, m=15, bounds=(-6.0, 5.0), callback=None,
                 **objective_kw):
        self.ph = physics
        self.obs = restrict_observations(physics, u_obs, v_obs)
        self.reg_weight, self.dt, self.mode = float(reg_weight), float(dt), mode
        self.n_fwd_vcycles, self.n_adj_vcycles = int(n_fwd_vcycles), int(n_adj_vcycles)
        self.factr, self.m, self.bounds = float(factr), int(m), tuple(bounds)
        self.callback = callback
        self.objective_kw = objective_kw
        self.results: list[dict] = []

    def level_objective(self, k: int) -> BetaInversion:
        """The BetaInversion that the optimiser of level k calls."""
        view = self.ph.level_view(k)
        g = view.grid
        fu, fv, _ = g.fields(g.f)
        fu.zero_()
        fv.zero_()
        uo, vo = self.obs[k]
        inv = BetaInversion(view, uo, vo, g.H_prev, dt=self.dt,
                            alpha=self.reg_weight, mode=self.mode, n_fwd_vcycles=self.n_fwd_vcycles,
                            n_adj_vcycles=self.n_adj_vcycles, **self.objective_kw)
        return inv

    def run(self, maxiter=None, levels=None) -> list[dict]:
        """Optimise the given levels (default all, coarsest first).
        ``maxiter``: L-BFGS iterations per level, an int or one per level
        index (default: scipy's, i.e. until factr stops it)."""
        n = self.ph.n_levels
        order = sorted(range(n) if levels is None else levels, reverse=True)
        for k in order:
            view = self.ph.level_view(k)
            g = view.grid
            inv = self.level_objective(k)
            x0 = torch.log(g.beta).reshape(-1).cpu().numpy().astype(np.float64)
            it = [0]

            def cb(_x, k=k, view=view, it=it):
                it[0] += 1
                if self.callback is not None:
                    self.callback(k, it[0], view)

            kw = {}
            mi = maxiter[k] if isinstance(maxiter, (list, tuple, dict)) else maxiter
            if mi is not None:
                kw["maxiter"] = int(mi)
            torch.cuda.synchronize(self.ph.device)
            t0 = time.perf_counter()
            x, f, info = fmin_l_bfgs_b(inv.objective, x0, bounds=[self.bounds] * x0.size, factr=self.factr,
                                       m=self.m, callback=cb, **kw)
            torch.cuda.synchronize(self.ph.device)
            secs = time.perf_counter() - t0
            beta = torch.exp(torch.as_tensor(x.reshape(g.ny, g.nx), dtype=torch.float64, device=self.ph.device))
            g.beta.copy_(beta)
            g.mask.zero_()
            if k > 0:
                parent = self.ph.grids[k - 1]
                _lib.call("glide_prolong_H", ctypes.c_void_p(g.beta.data_ptr()), ctypes.c_void_p(parent.beta.data_ptr()),
                          parent.ny, parent.nx, self.ph._stream())
            self.results.append({"level": k, "shape": (g.ny, g.nx), "nit": info["nit"], "nfev": info["funcalls"],
                                 "loss0": inv.history[0]["total"], "loss": float(f),
                                 "misfit0": inv.history[0]["misfit"], "misfit": inv.history[-1]["misfit"],
                                 "seconds": secs, "warnflag": info["warnflag"]})
        return self.results
