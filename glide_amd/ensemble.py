"""Batched ensembles of independent ice-sheet solves on one GPU.

Each member is an :class:`IcePhysics` of its own size, level count, geometry
and basal friction.  One ``Ensemble.forward`` runs the FASCD time step of all
members together: every operation of the V-cycle is a single persistent
launch over all members' tiles (libglide_hip.so, ``glide_batch_*``), so the
launch count per V-cycle does not grow with the member count and the
latency-bound coarse levels of many members share the GPU.  A member's
trajectory equals the one ``IcePhysics.forward`` gives it alone, bit for bit.

The reference runs one domain per process and has no batched path; the
per-member semantics follow glide/physics.py:141-209 (forward) and
glide/solver.py:71-157 (V-cycle).
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .physics import IcePhysics

BK_VANKA, BK_VANKA_LOCAL, BK_RESIDUAL, BK_RESIDUAL_CHILD = 0, 1, 2, 3


@dataclass
class EnsembleStats:
    """Outcome of one ensemble time step."""
    vcycles: list = field(default_factory=list)      # V-cycles run per member
    latency_s: list = field(default_factory=list)    # wall time until each member finished
    history: list = field(default_factory=list)      # relative residual after each V-cycle (tol mode)
    converged: list = field(default_factory=list)
    wall_s: float = 0.0
    r_start: object = None                            # per-member norm at the start (tol mode)
    converged_at_start: bool = False


class Ensemble:
    """A batch plan over ``members`` (list of IcePhysics on one device)."""

    def __init__(self, members: list[IcePhysics]):
        if not members:
            raise ValueError("an ensemble needs at least one member")
        dev = members[0].device
        if any(m.device != dev for m in members):
            raise ValueError("all members must live on one device")
        self.members = list(members)
        self.device = dev
        self.n = len(members)
        descs = [m._levels() for m in members]
        self._descs = descs          # kept alive: the plan copies them, but keep for clarity
        ptrs = (ctypes.POINTER(_lib.GlideLevel) * self.n)(
            *[ctypes.cast(d, ctypes.POINTER(_lib.GlideLevel)) for d in descs])
        nlev = (ctypes.c_int * self.n)(*[m.n_levels for m in members])
        h = ctypes.c_void_p()
        _lib.call("glide_batch_create", ptrs, nlev, self.n, ctypes.byref(h))
        self._h = h
        n_ph = ctypes.c_int()
        ent = (ctypes.c_int * 32)()
        cells = (ctypes.c_int * 32)()
        _lib.call("glide_batch_info", self._h, ctypes.byref(n_ph), ent, cells)
        self.n_phases = n_ph.value
        self.entries_per_phase = list(ent[: self.n_phases])
        self.cells_per_phase = list(cells[: self.n_phases])
        self.tiles: dict = {}
        self.schedule = _lib.schedule()   # the reference's counts and restriction, idle-pass check on
        self.thklim = members[0].thklim

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.call("glide_batch_destroy", self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ----------------------------------------------------------- helpers
    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def set_schedule(self, pre=10, post=10, coarse=200, local_passes=10, local_sweeps=10, n_newton=10,
                     adj_sweeps=10, adj_coarse=100, boundary_weight=1.0, local_skip=True):
        """V-cycle settings (the same fields as IcePhysics.set_schedule)."""
        self.schedule = _lib.schedule(pre, post, coarse, local_passes, local_sweeps, n_newton, adj_sweeps,
                                      adj_coarse, boundary_weight, local_skip)

    def set_enabled(self, on):
        arr = (ctypes.c_ubyte * self.n)(*[1 if x else 0 for x in on])
        _lib.call("glide_batch_set_enabled", self._h, arr, self._stream())

    def set_tiles(self, kernel: int, phase: int, ty: int, wg_per_cu: int = 0):
        _lib.call("glide_batch_set_tiles", self._h, int(kernel), int(phase), int(ty), int(wg_per_cu))
        self.tiles[(int(kernel), int(phase))] = (int(ty), int(wg_per_cu))

    def time_kernel(self, kernel: int, phase: int, ty: int, wg_per_cu: int = 0, reps: int = 5) -> float:
        ms = ctypes.c_double()
        _lib.call("glide_batch_time_kernel", self._h, int(kernel), int(phase), int(ty), int(wg_per_cu), int(reps),
                  ctypes.byref(ms), self._stream())
        return ms.value

    def time_restriction(self, phase: int, reps: int = 20) -> tuple[float, float]:
        """(ms per launch, compulsory bytes) of the batched residual
        restriction of one phase (every member with a child level)."""
        ms, nb = ctypes.c_double(), ctypes.c_double()
        _lib.call("glide_batch_time_restriction", self._h, int(phase), int(reps), ctypes.byref(ms), ctypes.byref(nb),
                  self._stream())
        return ms.value, nb.value

    # ----------------------------------------------------------- solve
    def forward_setup(self, dt: float):
        for m in self.members:
            for g in m.grids:
                g.dt = float(dt)
        _lib.call("glide_batch_forward_setup", self._h, float(dt), self._stream())

    def vcycle(self):
        _lib.call("glide_batch_vcycle", self._h, float(self.thklim), ctypes.byref(self.schedule), self._stream())

    def residual_norms(self) -> np.ndarray:
        out = (ctypes.c_double * self.n)()
        _lib.call("glide_batch_residual_norms", self._h, out, self._stream())
        return np.array(out[:], dtype=np.float64)

    def residual(self, phase: int | None = None):
        """r = F(U) - f on every entry of ``phase`` (default: every finest level)."""
        _lib.call("glide_batch_residual", self._h, self.n_phases if phase is None else int(phase), self._stream())

    def forward(self, dt: float, n_vcycles: int = 3, tol: float | None = None, max_vcycles: int | None = None,
                update_geometry: bool = True, r0=None, setup: bool = True) -> EnsembleStats:
        """One implicit time step of every member (glide/physics.py:141-209).

        Without ``tol`` every member runs ``n_vcycles`` V-cycles.  With ``tol``
        members iterate until their residual norm reaches ``tol`` times ``r0``
        (per member; default: the norm at the start of the step) in at most
        ``max_vcycles``; a converged member is disabled, so the remaining
        launches skip its tiles.  ``setup=False`` keeps a forcing already set
        by forward_setup (e.g. before seeding the velocities)."""
        st = EnsembleStats()
        if setup:
            self.forward_setup(dt)
        on = [True] * self.n
        self.set_enabled(on)
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        if tol is None:
            for _ in range(n_vcycles):
                self.vcycle()
            torch.cuda.synchronize(self.device)
            wall = time.perf_counter() - t0
            st.vcycles = [n_vcycles] * self.n
            st.latency_s = [wall] * self.n
            st.converged = [None] * self.n
            st.history = [[] for _ in range(self.n)]
        else:
            n_max = max_vcycles or max(n_vcycles, 50)
            r_start = self.residual_norms()
            r0 = r_start if r0 is None else np.asarray(r0, dtype=np.float64)
            st.r_start = r_start
            for m in range(self.n):
                if r_start[m] <= tol * r0[m]:
                    on[m] = False
                    st.converged_at_start = True
            if not all(on):
                self.set_enabled(on)
            st.vcycles = [0] * self.n
            st.latency_s = [None if on[m] else 0.0 for m in range(self.n)]
            st.converged = [not on[m] for m in range(self.n)]
            st.history = [[] for _ in range(self.n)]
            for it in range(n_max if any(on) else 0):
                self.vcycle()
                r = self.residual_norms()
                t = time.perf_counter() - t0
                for m in range(self.n):
                    if not on[m]:
                        continue
                    rel = r[m] / r0[m] if r0[m] > 0 else 0.0
                    st.history[m].append(float(rel))
                    st.vcycles[m] = it + 1
                    if rel <= tol:
                        on[m] = False
                        st.converged[m] = True
                        st.latency_s[m] = t
                if not any(on):
                    break
                self.set_enabled(on)
            t = time.perf_counter() - t0
            st.latency_s = [x if x is not None else t for x in st.latency_s]
            self.set_enabled([True] * self.n)
            wall = t
        st.wall_s = wall
        if update_geometry:
            for m in self.members:
                m.grid.H_prev.copy_(m.grid.H)
        return st


@dataclass
class StreamStats:
    """Outcome of a continuously batched run (stream_solve)."""
    jobs: list = field(default_factory=list)   # per job: slot, vcycles, converged, used_net, fallback, times
    wall_s: float = 0.0
    batched_vcycles: int = 0                    # batched V-cycle launches over the whole run
    occupancy: float = 0.0                      # mean fraction of slots solving a job per V-cycle


def stream_solve(ens: Ensemble, load, n_jobs: int, dt: float, tol: float = 1e-8, max_vcycles: int = 100,
                 net=None, on_done=None, refill_min: int = 1) -> StreamStats:
    """Continuous batching: the ensemble's members are slots, and job k (one
    implicit step of one domain, written into a slot by ``load(k, physics)``
    through set_geometry / set_parameters) runs in the first free slot.  Each
    job starts from u = v = 0 and H = H_prev, optionally seeded by ``net``
    (the learned warm start with the fallback rule: the prediction is kept
    only if its residual norm is at most the cold start's), and iterates in
    the batched V-cycles until ||r|| <= tol ||r0_cold|| or max_vcycles; then
    its slot takes the next job at once instead of idling until the slowest
    member of a fixed batch finishes.  Every slot must have the size the
    jobs need.  A job's states are those of solving it alone: the batched
    V-cycle is the single-domain one, member by member, bit for bit.
    ``on_done(k, physics)`` is called when job k ends, before its slot is
    reused.  ``refill_min``: freed slots wait until at least this many are
    free (or nothing else is running), so that new jobs start together; a
    started job's first V-cycles run every local pass of the batch.  Every slot must have the grid size and dx of the jobs it takes
    (the plan fixes both when it is built)."""
    from .warmstart import IN_CH, N_OUT, apply_prediction, features, net_scales

    S = ens.n
    for m in ens.members:
        for g in m.grids:
            g.dt = float(dt)
    ens.forward_setup(dt)
    slot_job = [None] * S
    vc = [0] * S
    r0 = np.zeros(S)
    st = StreamStats()
    jobs = {}
    next_job = 0
    busy_sum = 0
    t0 = time.perf_counter()

    def refill(free):
        nonlocal next_job
        new = []
        for m in free:
            if next_job >= n_jobs:
                break
            k = next_job
            next_job += 1
            ph = ens.members[m]
            load(k, ph)
            g = ph.grid
            g.u.zero_()
            g.v.zero_()
            g.H.copy_(g.H_prev)
            _lib.call("glide_forward_setup", ph._levels(), ph.n_levels, float(dt), ph._stream())
            slot_job[m], vc[m] = k, 0
            jobs[k] = {"job": k, "slot": m, "t_start": time.perf_counter() - t0, "used_net": False,
                       "fallback": False}
            new.append(m)
        if not new:
            return
        ens.set_enabled([slot_job[m] is not None for m in range(S)])
        r = ens.residual_norms()
        for m in new:
            r0[m] = r[m]
        if net is not None:
            mem = [ens.members[m] for m in new]
            g0 = mem[0].grid
            X = torch.empty((len(new), g0.ny * g0.nx * IN_CH), dtype=torch.int16, device=g0.device)
            for q, ph in enumerate(mem):
                features(ph, out=X[q])
            pred = net.forward(X, len(new), g0.ny, g0.nx)
            n_px = g0.ny * g0.nx * N_OUT
            for q, ph in enumerate(mem):
                apply_prediction(ph, pred[q * n_px:(q + 1) * n_px], *net_scales(net))
            rw = ens.residual_norms()
            for m in new:
                k = slot_job[m]
                if rw[m] <= r0[m]:
                    jobs[k]["used_net"] = True
                else:
                    g = ens.members[m].grid
                    g.u.zero_()
                    g.v.zero_()
                    g.H.copy_(g.H_prev)
                    jobs[k]["fallback"] = True

    refill(list(range(S)))
    pending = []
    while any(j is not None for j in slot_job):
        ens.vcycle()
        st.batched_vcycles += 1
        busy_sum += sum(j is not None for j in slot_job)
        r = ens.residual_norms()
        free = []
        for m in range(S):
            k = slot_job[m]
            if k is None:
                continue
            vc[m] += 1
            rel = r[m] / r0[m] if r0[m] > 0 else 0.0
            if rel <= tol or vc[m] >= max_vcycles:
                jobs[k].update(vcycles=vc[m], converged=bool(rel <= tol), final_rel_residual=float(rel),
                               t_end=time.perf_counter() - t0)
                if on_done is not None:
                    on_done(k, ens.members[m])
                slot_job[m] = None
                free.append(m)
        pending += free
        if pending and (len(pending) >= refill_min or all(j is None for j in slot_job)):
            refill(pending)
            pending = []
        if free:
            ens.set_enabled([j is not None for j in slot_job])
    torch.cuda.synchronize(ens.device)
    st.wall_s = time.perf_counter() - t0
    st.jobs = [jobs[k] for k in sorted(jobs)]
    st.occupancy = busy_sum / (S * st.batched_vcycles) if st.batched_vcycles else 0.0
    ens.set_enabled([True] * S)
    return st
