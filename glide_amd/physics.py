"""Public API: IcePhysics (counterpart of glide/physics.py:19-259) plus the
misfit / regularisation helpers of the inversion (glide/physics.py:262-411).

Every solver step runs on the GPU through the C ABI of libglide_hip.so; this
module only orders the calls, moves inputs into the level buffers and exposes
views of the state.
"""
from __future__ import annotations

import ctypes
import json
import math
import time

import numpy as np
import torch

from . import _lib
from .grid import Grid, descriptors, stream_handle

RHO_ICE = 917.0
G = 9.81

# V-cycle settings for solving a step to a residual tolerance (the reference
# only runs a fixed number of V-cycles): the restricted residual's
# boundary-face rows weighted 0.5, without which FASCD stalls wherever the
# solution carries normal stress on the domain edge, and 5 coarsest-level
# sweeps instead of 200.  On seeded 512^2 ice sheets at dt = 1 a the V-cycle
# count to 1e-8 is set by the finest levels' smoothing: 2 to 200 coarsest
# sweeps, and 4 to 6 levels, give the same counts (one more on 2 of 20
# seeds), while each coarsest sweep costs as much as one on a 128^2 level
# (the per-cell Newton chain, not the level size, bounds it).  The adjoint
# V-cycle likewise takes 10 coarsest sweeps instead of 100: on the 512^2
# inversion slab the adjoint still reaches 1e-10 in the same 4 V-cycles and
# the gradient agrees with the reference schedule's to 1e-13, at 4.6 instead
# of 7.8 ms per adjoint V-cycle.  The adjoint keeps the reference's residual
# restriction (its linear V-cycle converges with it).
TOLERANCE_SCHEDULE = dict(boundary_weight=0.5, coarse=5, adj_coarse=10)


def _as_f64(x, device) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=torch.float64)
    a = np.asarray(x, dtype=np.float64)
    if not a.flags.writeable:   # read-only input (a view of a constant, a memory map): copy, never alias it
        a = a.copy()
    return torch.as_tensor(a, device=device)


class IcePhysics:
    """Shallow-shelf ice-flow model on a MAC grid, FP64 on one MI355X.

    Same constructor, set_geometry / set_parameters / forward / adjoint
    semantics as the reference (glide/physics.py:49-242).  forward() returns
    views (u, v, H) of the finest level's state; u and v warm-start the next
    call.
    """

    def __init__(self, ny, nx, dx, n_levels=5, n=3.0, eps_reg=1e-5, thklim=0.1, water_drag=1e-3,
                 calving_rate=1.0, gl_sigmoid_c=0.1, gl_derivatives=False, device=None):
        if ny % (1 << (n_levels - 1)) or nx % (1 << (n_levels - 1)):
            raise ValueError(f"grid {ny}x{nx} is not divisible by 2^(n_levels-1) = {1 << (n_levels - 1)}")
        _lib.load()
        self.ny, self.nx, self.dx, self.n_levels = int(ny), int(nx), float(dx), int(n_levels)
        self.n, self.eps_reg, self.thklim = float(n), float(eps_reg), float(thklim)
        self.water_drag, self.calving_rate = float(water_drag), float(calving_rate)
        self.gl_sigmoid_c, self.gl_derivatives = float(gl_sigmoid_c), bool(gl_derivatives)
        self.device = torch.device(device if device is not None else "cuda")
        self.grid = Grid(ny, nx, dx, 1.0, self.device, n=n, eps_reg=eps_reg, water_drag=water_drag,
                         calving_rate=calving_rate, gl_sigmoid_c=gl_sigmoid_c, gl_derivatives=gl_derivatives)
        self.grids = [self.grid]
        for _ in range(n_levels - 1):
            self.grids.append(self.grids[-1].spawn_child())
        self.schedule = _lib.schedule()   # the reference's counts and restriction, idle-pass check on
        self._work = None                 # GlideVcycleWork of the idle-pass check (created on first use)
        self._desc = None
        self.last_history: list[float] = []
        self.steps = 0                # forward steps taken (numbers the log records)
        self._nonfinite = torch.zeros(1, dtype=torch.int64, device=self.device)

    # ------------------------------------------------------------ helpers
    def _levels(self):
        if self._desc is None:
            self._desc = descriptors(self.grids)
        for k, g in enumerate(self.grids):
            self._desc[k].g.dt = g.dt
        return self._desc

    def _stream(self):
        return ctypes.c_void_p(stream_handle(self.device))

    def set_schedule(self, pre=10, post=10, coarse=200, local_passes=10, local_sweeps=10, n_newton=10,
                     adj_sweeps=10, adj_coarse=100, boundary_weight=1.0, local_skip=True):
        """V-cycle settings.  The defaults are the reference's sweep counts
        (glide/solver.py:71-220) and its residual restriction
        (boundary_weight=1.0).  boundary_weight=0.5 halves the restricted
        residual's boundary-face rows, whose one-sided stress term scales
        like 1/dx where interior rows scale like 1/dx^2: with 1.0 the coarse
        correction overshoots there and FASCD stalls on states with normal
        stress at the domain edge; 0.5 converges (at the cost of departing
        from the reference's iterates).  local_skip stops a level's local
        passes at the first idle pass (same results up to the sign of
        zeros; the V-cycle then waits on the GPU between passes)."""
        self.schedule = _lib.schedule(pre, post, coarse, local_passes, local_sweeps, n_newton, adj_sweeps,
                                      adj_coarse, boundary_weight, local_skip)

    # ------------------------------------------------------------ inputs
    def _check_finite(self, name: str, t: torch.Tensor):
        """Raise ValueError when t holds NaN or +-Inf (counted on the GPU)."""
        n = ctypes.c_ulonglong(0)
        t = t.contiguous()
        _lib.call("glide_count_nonfinite", t.data_ptr(), t.numel(), ctypes.c_void_p(self._nonfinite.data_ptr()),
                  ctypes.byref(n), self._stream())
        if n.value:
            raise ValueError(f"{name} has {n.value} non-finite value(s)")

    def set_geometry(self, bed, thickness):
        """glide/physics.py:82-99; non-finite bed or thickness is rejected
        (ValueError) before the hierarchy is touched."""
        g = self.grid
        bed_t = _as_f64(bed, self.device).reshape(g.ny, g.nx)
        H_t = _as_f64(thickness, self.device).reshape(g.ny, g.nx)
        self._check_finite("bed", bed_t)
        self._check_finite("thickness", H_t)
        g.bed.copy_(bed_t)
        g.H.copy_(H_t)
        g.H_prev.copy_(g.H)
        g.gamma.fill_(self.thklim)
        _lib.call("glide_propagate_geometry", self._levels(), self.n_levels, self.thklim, self._stream())

    def set_parameters(self, B=None, beta=None, smb=None):
        """glide/physics.py:101-129; non-finite values are rejected
        (ValueError) before any of them is stored."""
        g = self.grid
        Bt = _as_f64(B, self.device) if B is not None else None
        bt = _as_f64(beta, self.device).reshape(g.ny, g.nx) if beta is not None else None
        st = _as_f64(smb, self.device) if smb is not None else None
        for name, t in (("B", Bt), ("beta", bt), ("smb", st)):
            if t is not None:
                self._check_finite(name, t.reshape(-1))
        if Bt is not None:
            if Bt.ndim == 0:
                g.B.fill_(float(Bt))
            else:
                g.B.copy_(Bt.reshape(g.ny, g.nx))
        if bt is not None:
            g.beta.copy_(bt)
        if st is not None:
            if st.ndim == 0:
                g.smb.fill_(float(st))
            else:
                g.smb.copy_(st.reshape(g.ny, g.nx))
        _lib.call("glide_restrict_parameters", self._levels(), self.n_levels, self._stream())

    # ------------------------------------------------------------ solve
    def residual_norm(self) -> float:
        out = ctypes.c_double(0.0)
        _lib.call("glide_residual_norm", ctypes.byref(self._levels()[0]), ctypes.byref(out), self._stream())
        return out.value

    def vcycle(self):
        work = None
        if self.schedule.local_skip:
            if self._work is None:
                self._work = _lib.VcycleWork()
            work = self._work.handle
        _lib.call("glide_fascd_vcycle", self._levels(), self.n_levels, self.thklim, ctypes.byref(self.schedule), work,
                  self._stream())

    def local_pass_counts(self, reset: bool = False):
        """(local passes run, local passes skipped) by this hierarchy's
        V-cycles with the idle-pass check, since the last reset."""
        if self._work is None:
            return 0, 0
        return self._work.counts(reset)

    def forward(self, dt, n_vcycles=3, verbose=False, update_geometry=True, tol=None, max_vcycles=None, log=None):
        """One implicit time step (glide/physics.py:141-209).

        Runs ``n_vcycles`` FASCD V-cycles; with ``tol`` it instead runs until
        the relative residual ||(r_u, r_v, FB(r_H, H - gamma))|| / r0 drops
        to ``tol`` (at most ``max_vcycles``).  The relative residual after
        each cycle is kept in ``last_history`` when it was computed; a
        non-finite one raises FloatingPointError.  ``log`` (a path or a
        callable) receives one JSON record per step: step, dt, V-cycles,
        the residual history and the wall time (the step is synchronised
        for it).
        """
        t0 = time.perf_counter() if log is not None else 0.0
        for g in self.grids:
            g.dt = float(dt)
        levels = self._levels()
        _lib.call("glide_forward_setup", levels, self.n_levels, float(dt), self._stream())
        track = verbose or tol is not None
        self.last_history = []
        r0 = self.residual_norm() if track else 0.0
        if track and not math.isfinite(r0):
            raise FloatingPointError(f"initial residual is {r0}")
        if verbose:
            print(f"  Initial: |r| = {r0:.2e}")
        n_max = n_vcycles if tol is None else (max_vcycles or max(n_vcycles, 50))
        done = 0
        for it in range(n_max):
            self.vcycle()
            done = it + 1
            if track:
                rel = self.residual_norm() / r0 if r0 > 0 else 0.0
                self.last_history.append(rel)
                if not math.isfinite(rel):
                    raise FloatingPointError(f"residual became {rel} in V-cycle {it}")
                if verbose:
                    print(f"  V-cycle {it}: |r|/|r0| = {rel:.2e}")
                if tol is not None and rel <= tol:
                    break
        if update_geometry:
            self.grid.H_prev.copy_(self.grid.H)
        self.steps += 1
        if log is not None:
            torch.cuda.synchronize(self.device)
            rec = {"step": self.steps, "dt": float(dt), "vcycles": done, "r0": r0 if track else None,
                   "rel_residual": list(self.last_history), "ms": (time.perf_counter() - t0) * 1e3,
                   "grid": [self.ny, self.nx], "n_levels": self.n_levels}
            if callable(log):
                log(rec)
            else:
                with open(log, "a") as fh:
                    fh.write(json.dumps(rec) + "\n")
        return self.grid.u, self.grid.v, self.grid.H

    # ------------------------------------------------------------ adjoint
    def adjoint_vcycle(self):
        _lib.call("glide_adjoint_vcycle", self._levels(), self.n_levels, ctypes.byref(self.schedule), self._stream())

    def grad_beta(self, reference_edges: bool = True) -> torch.Tensor:
        """lambda . dR/dbeta at the finest level (glide/grid.py:336-355).
        reference_edges reproduces the reference's skipped edge-cell terms
        (glide/cuda/grad.cu:52-53,70-71); False gives the exact transpose."""
        g = self.grid
        lev = self._levels()
        grad = torch.empty((g.ny, g.nx), dtype=torch.float64, device=self.device)
        _lib.call("glide_compute_mask", ctypes.byref(lev[0]), self._stream())
        geom, prm = g.geom(), g.cparams()
        _lib.call("glide_grad_beta", ctypes.byref(geom), ctypes.byref(prm), grad.data_ptr(), g.U.data_ptr(),
                  g.Lambda.data_ptr(), g.bed.data_ptr(), g.beta.data_ptr(), int(bool(reference_edges)),
                  self._stream())
        g.mask.zero_()
        return grad

    def adjoint(self, dL_du, dL_dv, dL_dH, n_vcycles=2, reference_edges=True):
        """Gradient of a loss w.r.t. beta by the adjoint V-cycles
        (glide/physics.py:211-242): f_adj = -dL/dU, Lambda = 0, n adjoint
        V-cycles about the current (converged) state, then lambda . dR/dbeta."""
        g = self.grid
        fu, fv, fH = g.fields(g.f_adj)
        fu.copy_(-_as_f64(dL_du, self.device).reshape(fu.shape))
        fv.copy_(-_as_f64(dL_dv, self.device).reshape(fv.shape))
        fH.copy_(-_as_f64(dL_dH, self.device).reshape(fH.shape))
        g.Lambda.zero_()
        g.Lambda_out.zero_()
        for _ in range(n_vcycles):
            self.adjoint_vcycle()
        return self.grad_beta(reference_edges)

    def adjoint_solve(self, dL_du, dL_dv, dL_dH, tol=1e-12, max_vcycles=200):
        """Adjoint V-cycles until ||f_adj - J^T Lambda|| <= tol * ||f_adj||
        (or max_vcycles); the relative residual history is kept in
        last_adj_history."""
        g = self.grid
        fu, fv, fH = g.fields(g.f_adj)
        fu.copy_(-_as_f64(dL_du, self.device).reshape(fu.shape))
        fv.copy_(-_as_f64(dL_dv, self.device).reshape(fv.shape))
        fH.copy_(-_as_f64(dL_dH, self.device).reshape(fH.shape))
        g.Lambda.zero_()
        g.Lambda_out.zero_()
        r0 = float(torch.linalg.norm(g.f_adj))
        self.last_adj_history = []
        if r0 == 0.0:
            return
        for _ in range(max_vcycles):
            self.adjoint_vcycle()
            rel = self.adjoint_residual_norm() / r0
            self.last_adj_history.append(rel)
            if rel <= tol:
                break

    def grad_H_prev(self) -> torch.Tensor:
        """dL/dH_prev = -Lambda_H / dt outside the adjoint's active set
        (f_H = H_prev / dt + smb, glide/physics.py:173)."""
        g = self.grid
        lev = self._levels()
        _lib.call("glide_compute_mask", ctypes.byref(lev[0]), self._stream())
        _, _, lH = g.fields(g.Lambda)
        out = -lH / g.dt * (1.0 - g.mask)
        g.mask.zero_()
        return out

    def grad_smb(self) -> torch.Tensor:
        """dL/dsmb = -Lambda_H outside the adjoint's active set (f_H =
        H_prev / dt + smb, glide/physics.py:173; masked rows H - gamma do not
        depend on smb).  No reference counterpart."""
        return self.grad_H_prev() * self.grid.dt

    def grad_B(self) -> torch.Tensor:
        """dL/dB = Lambda . dR/dB at the finest level (B enters through the
        viscosity only: d eta / d B = eta / B).  No reference counterpart."""
        g = self.grid
        grad = torch.empty((g.ny, g.nx), dtype=torch.float64, device=self.device)
        geom, prm = g.geom(), g.cparams()
        _lib.call("glide_grad_B", ctypes.byref(geom), ctypes.byref(prm), grad.data_ptr(), g.U.data_ptr(),
                  g.Lambda.data_ptr(), g.B.data_ptr(), self._stream())
        return grad

    def adjoint_residual_norm(self) -> float:
        """||f_adj - J^T Lambda|| at the finest level with the adjoint's mask."""
        g = self.grid
        lev = self._levels()
        _lib.call("glide_compute_mask", ctypes.byref(lev[0]), self._stream())
        _lib.call("glide_adjoint_residual", ctypes.byref(lev[0]), self._stream())
        g.mask.zero_()
        return float(torch.linalg.norm(g.r_adj))

    # ------------------------------------------------------------ checkpoint
    _STATE = ("u", "v", "H", "H_prev", "bed", "B", "beta", "smb")

    def state_dict(self) -> dict:
        """Everything a later process needs to continue the time stepping
        bit for bit: the finest level's state (u, v, H), H_prev, the inputs
        (bed, B, beta, smb), the step count and the constructor arguments.
        Tensors are copies on the host."""
        g = self.grid
        sd = {k: getattr(g, k).detach().to("cpu", copy=True) for k in self._STATE}
        sd["config"] = {"ny": self.ny, "nx": self.nx, "dx": self.dx, "n_levels": self.n_levels, "n": self.n,
                        "eps_reg": self.eps_reg, "thklim": self.thklim, "water_drag": self.water_drag,
                        "calving_rate": self.calving_rate, "gl_sigmoid_c": self.gl_sigmoid_c,
                        "gl_derivatives": self.gl_derivatives, "steps": self.steps}
        return sd

    def load_state_dict(self, sd: dict):
        """Restore a state_dict into this hierarchy (same grid and physics
        parameters): the inputs are set and restricted as set_geometry and
        set_parameters do, then H and the velocities are restored."""
        c = sd["config"]
        mine = self.state_dict()["config"]
        for k in ("ny", "nx", "dx", "n_levels", "n", "eps_reg", "thklim", "water_drag", "calving_rate",
                  "gl_sigmoid_c", "gl_derivatives"):
            if c[k] != mine[k]:
                raise ValueError(f"checkpoint {k} = {c[k]} does not match this IcePhysics ({mine[k]})")
        g = self.grid
        self.set_geometry(sd["bed"], sd["H_prev"])
        self.set_parameters(B=sd["B"], beta=sd["beta"], smb=sd["smb"])
        for k in ("u", "v", "H"):
            t = _as_f64(sd[k], self.device)
            self._check_finite(k, t)
            getattr(g, k).copy_(t.reshape(getattr(g, k).shape))
        self.steps = int(c["steps"])

    def save(self, path):
        """torch.save of state_dict()."""
        torch.save(self.state_dict(), path)

    @classmethod
    def load(cls, path, device=None) -> "IcePhysics":
        """A new IcePhysics continued from a file written by save()
        (loaded with weights_only=True: plain tensors and numbers only)."""
        sd = torch.load(path, map_location="cpu", weights_only=True)
        c = sd["config"]
        ph = cls(c["ny"], c["nx"], c["dx"], n_levels=c["n_levels"], n=c["n"], eps_reg=c["eps_reg"],
                 thklim=c["thklim"], water_drag=c["water_drag"], calving_rate=c["calving_rate"],
                 gl_sigmoid_c=c["gl_sigmoid_c"], gl_derivatives=c["gl_derivatives"], device=device)
        ph.load_state_dict(sd)
        return ph

    # ------------------------------------------------------------ levels
    def level_view(self, k: int) -> "IcePhysics":
        """An IcePhysics over levels k.. of this hierarchy, sharing their
        buffers (nothing is allocated): level k acts as the finest level of
        forward, adjoint and set_parameters.  The multi-resolution inversion
        solves on these views (examples/greenland/greenland_inverse.py:137-
        142, where each level's Grid is used as the top of its own V-cycle)."""
        if not 0 <= k < self.n_levels:
            raise ValueError(f"level {k} outside 0..{self.n_levels - 1}")
        if k == 0:
            return self
        view = object.__new__(IcePhysics)
        view.__dict__.update(self.__dict__)
        g = self.grids[k]
        view.grid, view.grids = g, self.grids[k:]
        view.ny, view.nx, view.dx, view.n_levels = g.ny, g.nx, g.dx, self.n_levels - k
        view._desc = None
        view.last_history = []
        return view

    # ------------------------------------------------------------ misc API
    def reset_solution(self):
        """glide/physics.py:244-248"""
        self.grid.u.zero_()
        self.grid.v.zero_()
        self.grid.H.copy_(self.grid.H_prev)

    def get_surface(self):
        """glide/physics.py:250-253"""
        base = torch.maximum(self.grid.bed, -RHO_ICE / 1000.0 * self.grid.H)
        return self.grid.H + base

    def get_velocities_cell_centered(self):
        """glide/physics.py:255-259"""
        u_c = 0.5 * (self.grid.u[:, 1:] + self.grid.u[:, :-1])
        v_c = 0.5 * (self.grid.v[1:] + self.grid.v[:-1])
        return u_c, v_c


# ---------------------------------------------------------------- losses
# Misfit and regularisation helpers of the inversion (glide/physics.py:262-411).
# Unobserved entries (NaN in the observations) give a zero misfit; as in the
# reference, they still count in the means, and both gradient components are
# divided by the number of u entries.
def _misfit(model: torch.Tensor, obs: torch.Tensor) -> torch.Tensor:
    d = model - obs
    return torch.where(torch.isnan(d), torch.zeros_like(d), d)


def huber_loss(u, v, u_obs, v_obs, eps=10.0):
    """Pseudo-Huber misfit: mean sqrt(du^2 + eps) + mean sqrt(dv^2 + eps)
    (glide/physics.py:262-287)."""
    return sum(torch.sqrt(_misfit(m, o) ** 2 + eps).mean() for m, o in ((u, u_obs), (v, v_obs)))


def huber_grad(u, v, u_obs, v_obs, eps=10.0):
    """d huber_loss / d(u, v) with the reference's normaliser u.size for
    both components (glide/physics.py:290-318)."""
    n = u.numel()
    return tuple(d / torch.sqrt(d ** 2 + eps) / n for d in (_misfit(u, u_obs), _misfit(v, v_obs)))


def abs_loss(u, v, u_obs, v_obs):
    """Mean absolute misfit of u plus that of v, in FP64 (glide/physics.py:320-349)."""
    return sum(_misfit(m.double(), o.double()).abs().mean() for m, o in ((u, u_obs), (v, v_obs)))


def abs_grad(u, v, u_obs, v_obs, eps=10.0):
    """sign(misfit) / u.size for u and v (glide/physics.py:352-379); eps is
    accepted for the reference's signature and unused."""
    n = u.numel()
    return tuple(torch.sign(d) / n for d in (_misfit(u, u_obs), _misfit(v, v_obs)))


def tikhonov_regularization(field: torch.Tensor):
    """Gradient-smoothness penalty 0.5 (|Dx f|^2 + |Dy f|^2) with Dx, Dy the
    forward differences between neighbouring cells along rows and columns,
    and its gradient Dx^T Dx f + Dy^T Dy f (glide/physics.py:382-411):
    each difference pushes its two cells apart, so the ends of a row or
    column get one term and interior cells two."""
    gx = field[:, 1:] - field[:, :-1]
    gy = field[1:, :] - field[:-1, :]
    loss = 0.5 * (float((gx * gx).sum()) + float((gy * gy).sum()))
    grad = torch.zeros_like(field)
    grad[:, 1:] += gx
    grad[:, :-1] -= gx
    grad[1:, :] += gy
    grad[:-1, :] -= gy
    return loss, grad
