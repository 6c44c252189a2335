"""One multigrid level of the MAC-grid hierarchy (counterpart of glide/grid.py:14-355).

All fields of a level live in ONE contiguous FP64 allocation in HBM (torch is
only the allocator and the carrier): the state-sized vectors [u | v | H]
(U, f, r, w, z and the adjoint's Lambda, Lambda_out, f_adj, r_adj), the cell
fields (H_prev, bed, B, beta, smb, gamma, chi, mask) and the Vanka scratch
(5 values per cell).  Field views mirror the reference's names.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib

VECS = ["U", "f", "r", "w", "z", "Lambda", "Lambda_out", "f_adj", "r_adj"]
CELLS = ["H_prev", "bed", "B", "beta", "smb", "gamma", "chi", "mask"]
RB = 14   # local-smoother block size of the reference launch tiling (glide/grid.py:182-188)


def sizes(ny: int, nx: int):
    nu, nv, nc = ny * (nx + 1), (ny + 1) * nx, ny * nx
    return nu, nv, nc, nu + nv + nc


class Grid:
    """A level: geometry, physical parameters and device buffers."""

    def __init__(self, ny: int, nx: int, dx: float, dt: float = 1.0, device=None, n: float = 3.0,
                 eps_reg: float = 1e-5, water_drag: float = 1e-3, calving_rate: float = 1.0,
                 gl_sigmoid_c: float = 0.1, gl_derivatives: bool = False):
        if ny < 2 or nx < 2:
            raise ValueError("grid must be at least 2 x 2")
        self.ny, self.nx, self.dx, self.dt = int(ny), int(nx), float(dx), float(dt)
        self.device = torch.device(device if device is not None else "cuda")
        self.params = dict(n=float(n), eps_reg=float(eps_reg), water_drag=float(water_drag),
                           calving_rate=float(calving_rate), gl_sigmoid_c=float(gl_sigmoid_c),
                           gl_derivatives=bool(gl_derivatives))
        self.nu, self.nv, self.nh, self.n_total = sizes(self.ny, self.nx)
        total = len(VECS) * self.n_total + len(CELLS) * self.nh + 5 * self.nh
        self._buf = torch.zeros(total, dtype=torch.float64, device=self.device)
        off = 0
        for name in VECS:
            setattr(self, name, self._buf[off:off + self.n_total])
            off += self.n_total
        for name in CELLS:
            setattr(self, name, self._buf[off:off + self.nh].view(self.ny, self.nx))
            off += self.nh
        self.scratch = self._buf[off:off + 5 * self.nh]
        self.nby, self.nbx = self.ny // RB + 1, self.nx // RB + 1
        self.active = torch.zeros(self.nby * self.nbx, dtype=torch.uint8, device=self.device)
        self.u, self.v, self.H = self.fields(self.U)
        self.child = None
        self.parent = None

    # views [u | v | H] of a state-sized vector (glide/grid.py:160-165)
    def fields(self, x: torch.Tensor):
        return (x[:self.nu].view(self.ny, self.nx + 1), x[self.nu:self.nu + self.nv].view(self.ny + 1, self.nx),
                x[self.nu + self.nv:].view(self.ny, self.nx))

    def spawn_child(self) -> "Grid":
        """2x coarser child; like the reference (glide/grid.py:167-180) it keeps n,
        eps_reg, water_drag, calving_rate but uses the default sigmoid settings."""
        if self.ny % 2 or self.nx % 2:
            raise ValueError("odd grid cannot be coarsened")
        p = self.params
        child = Grid(self.ny // 2, self.nx // 2, self.dx * 2, self.dt, self.device, n=p["n"], eps_reg=p["eps_reg"],
                     water_drag=p["water_drag"], calving_rate=p["calving_rate"])
        child.parent = self
        self.child = child
        return child

    # ---- C-ABI descriptors
    def geom(self) -> _lib.GlideGeom:
        return _lib.GlideGeom(self.ny, self.nx, self.dx, self.dt)

    def cparams(self) -> _lib.GlideParams:
        p = self.params
        return _lib.GlideParams(p["n"], p["eps_reg"], p["water_drag"], p["calving_rate"], p["gl_sigmoid_c"],
                                int(p["gl_derivatives"]))

    def descriptor(self) -> _lib.GlideLevel:
        d = _lib.GlideLevel()
        d.g = self.geom()
        d.p = self.cparams()
        for name in VECS + CELLS:
            setattr(d, name, getattr(self, name).data_ptr())
        d.scratch = self.scratch.data_ptr()
        d.active = self.active.data_ptr()
        return d


def stream_handle(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def descriptors(grids) -> ctypes.Array:
    arr = (_lib.GlideLevel * len(grids))()
    for k, g in enumerate(grids):
        arr[k] = g.descriptor()
    return arr
