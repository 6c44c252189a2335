"""Tile-shape autotuner of the batched ensemble kernels.

For every phase of an :class:`Ensemble` plan the tuner times each tunable
kernel over its candidate tile heights (64-wide tiles: one wave64 per tile
row) on the plan's own members, in place, with HIP events, and keeps the
fastest.  Timings are state-preserving (the Vanka timing writes only the
per-cell update scratch, the residual timings only residual buffers), so the
tuner may run on a live ensemble between time steps.  Tile shape never
changes results: every shape computes the same per-point arithmetic.

Choices are cached per device and plan signature (member sizes and level
counts) in a JSON file: ``$GLIDE_AUTOTUNE_CACHE`` or
``~/.cache/glide_amd/autotune.json``.
"""
from __future__ import annotations

import json
import os

import torch

from .ensemble import BK_RESIDUAL, BK_RESIDUAL_CHILD, BK_VANKA, BK_VANKA_LOCAL, Ensemble

CANDIDATES = {
    BK_VANKA: (1, 2, 4, 8),
    BK_RESIDUAL: (2, 4, 8, 16),
    BK_RESIDUAL_CHILD: (2, 4, 8, 16),
}
NAMES = {BK_VANKA: "vanka", BK_VANKA_LOCAL: "vanka_local", BK_RESIDUAL: "residual",
         BK_RESIDUAL_CHILD: "residual_child"}


def cache_path() -> str:
    return os.environ.get("GLIDE_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "glide_amd", "autotune.json"))


def signature(ens: Ensemble) -> str:
    dev = torch.cuda.get_device_name(ens.device)
    members = sorted((m.ny, m.nx, m.n_levels) for m in ens.members)
    return f"{dev}|" + ";".join(f"{a}x{b}L{c}" for a, b, c in members)


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, data):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass   # a read-only home only loses the cache


def tune(ens: Ensemble, reps: int = 5, use_cache: bool = True, path: str | None = None) -> dict:
    """Choose tile heights for every (kernel, phase) of ``ens``; returns
    {"choice": {name: {phase: ty}}, "times_ms": {name: {phase: {ty: ms}}}}."""
    path = path or cache_path()
    key = signature(ens)
    db = _load(path) if use_cache else {}
    if use_cache and key in db:
        rec = db[key]
        _apply(ens, rec["choice"])
        rec["cached"] = True
        return rec
    phases = list(range(ens.n_phases + 1))   # the last one is every member's finest level
    choice: dict = {NAMES[k]: {} for k in (BK_VANKA, BK_VANKA_LOCAL, BK_RESIDUAL, BK_RESIDUAL_CHILD)}
    times: dict = {NAMES[k]: {} for k in CANDIDATES}
    for p in phases:
        for k, cands in CANDIDATES.items():
            if k == BK_VANKA and p == ens.n_phases:
                continue                      # no sweeps run on the finest-level pseudo-phase
            if k == BK_RESIDUAL_CHILD and p >= ens.n_phases - 1:
                continue                      # the coarsest phase has no child
            t = {ty: ens.time_kernel(k, p, ty, 0, reps) for ty in cands}
            best = min(t, key=t.get)
            times[NAMES[k]][str(p)] = {str(ty): ms for ty, ms in t.items()}
            choice[NAMES[k]][str(p)] = best
            if k == BK_VANKA:
                choice[NAMES[BK_VANKA_LOCAL]][str(p)] = best
    _apply(ens, choice)
    rec = {"choice": choice, "times_ms": times, "cached": False}
    if use_cache:
        db[key] = {"choice": choice, "times_ms": times}
        _store(path, db)
    return rec


def _apply(ens: Ensemble, choice: dict):
    ids = {v: k for k, v in NAMES.items()}
    for name, per_phase in choice.items():
        for p, ty in per_phase.items():
            ens.set_tiles(ids[name], int(p), int(ty), 0)
