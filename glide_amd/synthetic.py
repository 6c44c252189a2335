"""Seeded synthetic ice-cap geometries and parameters.

The reference's drivers load BedMachine / MAR data from Google Drive
(glide/data.py:24-79); offline, the workloads use these generators instead.
Quantities follow the reference's head-normalised convention: stresses are
divided by rho_i * g (glide/physics.py:15-16, examples/greenland/
greenland_forward.py:91), so B and beta carry a 1/(917 * 9.81) factor while
H, bed and smb stay in metres (per year).

Workloads (sizes as used by the tests and the benchmark):
  dome_128   : 128 x 128, dx = 2000 m, 4 levels, dome radius 100 km
  cap_512    : 512 x 512, dx = 1500 m, 5 levels, radius 300 km plus a 50 m
               floating margin ring
  any size   : radius scales as 100/256 of the smaller domain side.
"""
from __future__ import annotations

import numpy as np

RHO_I = 917.0
G = 9.81
RHO_G = RHO_I * G                      # 8995.77
GLEN_N = 3.0
A_RATE = 1e-16                         # Pa^-3 a^-1
B_HEAD = A_RATE ** (-1.0 / GLEN_N) / RHO_G   # 23.9494...


def gaussian_field(ny: int, nx: int, dx: float, corr: float, rng: np.random.Generator) -> np.ndarray:
    """Zero-mean, unit-variance Gaussian random field: white noise convolved
    (via FFT) with a Gaussian kernel of standard deviation ``corr`` metres."""
    noise = rng.standard_normal((ny, nx))
    ky = 2.0 * np.pi * np.fft.fftfreq(ny, d=dx)
    kx = 2.0 * np.pi * np.fft.fftfreq(nx, d=dx)
    k2 = ky[:, None] ** 2 + kx[None, :] ** 2
    f = np.real(np.fft.ifft2(np.fft.fft2(noise) * np.exp(-0.5 * k2 * corr ** 2)))
    f -= f.mean()
    s = f.std()
    return f / s if s > 0 else f


def ice_cap(ny: int, nx: int, dx: float, seed: int, radius: float | None = None,
            floating_margin: bool = False, thklim: float = 0.1, smb: float = 0.3) -> dict:
    """Dome H = 2000 sqrt(max(0, 1 - (r/R)^2)) m on bed -200 + 0.002 x m plus a
    GRF (sigma 100 m, 20 km), log10 beta ~ 3 + 0.5 * (unit GRF at 10 km) Pa a/m,
    uniform B from A = 1e-16 Pa^-3 a^-1 and uniform smb.  With
    ``floating_margin`` a ring of seeded width (15 km +- 5 km) just outside
    the dome carries 50 m of ice over bed lowered to at most -100 m, so it
    floats (0.917 * 50 < 100)."""
    rng = np.random.default_rng(seed)
    Ly, Lx = ny * dx, nx * dx
    if radius is None:
        radius = (100.0 / 256.0) * min(Lx, Ly)
    y = (np.arange(ny) + 0.5) * dx
    x = (np.arange(nx) + 0.5) * dx
    X, Y = np.meshgrid(x, y)
    r = np.hypot(X - 0.5 * Lx, Y - 0.5 * Ly)
    H = 2000.0 * np.sqrt(np.maximum(0.0, 1.0 - (r / radius) ** 2))
    bed = -200.0 + 0.002 * X + 100.0 * gaussian_field(ny, nx, dx, 20e3, rng)
    logb = 3.0 + 0.5 * gaussian_field(ny, nx, dx, 10e3, rng)
    if floating_margin:
        theta = np.arctan2(Y - 0.5 * Ly, X - 0.5 * Lx)
        phases = rng.uniform(0, 2 * np.pi, 3)
        width = 15e3 + 5e3 * (np.sin(2 * theta + phases[0]) * 0.5 + np.sin(3 * theta + phases[1]) * 0.3
                               + np.sin(5 * theta + phases[2]) * 0.2)
        ring = (r >= radius * np.sqrt(1.0 - (50.0 / 2000.0) ** 2)) & (r < radius + width)
        H = np.where(ring, 50.0, H)
        bed = np.where(ring, np.minimum(bed, -100.0), bed)
    H = np.maximum(H, thklim)
    beta = 10.0 ** logb / RHO_G
    B = np.full((ny, nx), B_HEAD)
    return {"bed": np.ascontiguousarray(bed), "H": np.ascontiguousarray(H),
            "beta": np.ascontiguousarray(beta), "B": B, "smb": np.full((ny, nx), float(smb)),
            "dx": float(dx), "radius": float(radius)}


WORKLOADS = {
    # name: (ny, nx, dx, n_levels, seed, floating_margin, dt, n_vcycles, n_steps)
    "dome_128": (128, 128, 2000.0, 4, 0, False, 10.0, 3, 5),
    "cap_512": (512, 512, 1500.0, 5, 1, True, 10.0, 3, 20),
}


def workload(name: str) -> dict:
    ny, nx, dx, nlev, seed, margin, dt, nvc, nsteps = WORKLOADS[name]
    d = ice_cap(ny, nx, dx, seed, floating_margin=margin)
    d.update(ny=ny, nx=nx, n_levels=nlev, seed=seed, dt=dt, n_vcycles=nvc, n_steps=nsteps, name=name)
    return d


def max_levels(ny: int, nx: int, coarsest: int = 16) -> int:
    """Largest level count with both sides divisible by 2^(L-1) and the
    coarsest side >= ``coarsest``."""
    L = 1
    while (ny % (1 << L) == 0 and nx % (1 << L) == 0 and
           min(ny, nx) >> L >= coarsest):
        L += 1
    return L


def ice_slab(ny: int, nx: int, dx: float, seed: int, smb: float = 0.3) -> dict:
    """Ice-covered, grounded slab (no margin): H = 800 + 200 sin cos m on a
    bed 300 m + 0.004 x + a 20 m GRF, log10 beta ~ 3 + 0.3 * (unit GRF at
    8 km).  On these slabs FASCD converges to about 1e-13 relative residual
    with the reference's restriction too, which the finite-difference
    gradient checks need."""
    rng = np.random.default_rng(seed)
    y = np.arange(ny)[:, None]
    x = np.arange(nx)[None, :]
    H = 800.0 + 200.0 * np.sin(y * (2 * np.pi / max(ny, 8)) * 1.3) * np.cos(x * (2 * np.pi / max(nx, 8)) * 0.9)
    bed = 300.0 + 0.004 * (x + 0.5) * dx + 20.0 * gaussian_field(ny, nx, dx, 4 * dx, rng)
    logb = 3.0 + 0.3 * gaussian_field(ny, nx, dx, 8e3, rng)
    return {"bed": np.ascontiguousarray(np.broadcast_to(bed, (ny, nx))), "H": np.ascontiguousarray(np.broadcast_to(H, (ny, nx))),
            "beta": 10.0 ** logb / RHO_G, "B": np.full((ny, nx), B_HEAD), "smb": np.full((ny, nx), float(smb)),
            "dx": float(dx)}


# Batched ensemble: 16 members each at 256^2, 384 x 512, 512^2 and 768^2,
# seeds 1000-1063 (in that order), dx = 1500 m, the dome / bed / beta
# generators of the 128^2 workload (radius 100/256 of the smaller side), and
# the largest level count that keeps the coarsest side >= 16.
ENSEMBLE_SIZES = [(256, 256)] * 16 + [(384, 512)] * 16 + [(512, 512)] * 16 + [(768, 768)] * 16
ENSEMBLE_DX = 1500.0
ENSEMBLE_SEED0 = 1000


def ensemble_members(sizes=None, seed0: int = ENSEMBLE_SEED0, dx: float = ENSEMBLE_DX) -> list[dict]:
    """Inputs of each ensemble member: ice_cap(ny, nx, dx, seed) plus its
    level count (max_levels) and seed."""
    sizes = ENSEMBLE_SIZES if sizes is None else sizes
    out = []
    for k, (ny, nx) in enumerate(sizes):
        d = ice_cap(ny, nx, dx, seed0 + k)
        d.update(ny=ny, nx=nx, seed=seed0 + k, n_levels=max_levels(ny, nx))
        out.append(d)
    return out


def ice_sheet(ny: int, nx: int, dx: float, seed: int, smb: float = 0.3) -> dict:
    """Seeded ice-covered, grounded sheet (the warm-start network's training
    and test geometries): H = 700 + 250 * (unit GRF at 30 km), clipped to
    [250, 1500] m, on a bed 300 m + 0.003 x + 80 * (unit GRF at 20 km)
    (grounded everywhere: bed > 0 > -0.917 H), log10 beta ~ 3 + 0.4 * (unit
    GRF at 8 km), uniform B and smb.  No ice margin and no active thickness
    constraint; with the restricted residual's boundary rows weighted 0.5
    every measured sheet's forward step reaches 1e-8 (with the reference's
    weight 1 about a quarter stall above it, at the domain edge)."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx)[None, :] + 0.5) * dx
    H = np.clip(700.0 + 250.0 * gaussian_field(ny, nx, dx, 30e3, rng), 250.0, 1500.0)
    bed = np.maximum(300.0 + 0.003 * x + 80.0 * gaussian_field(ny, nx, dx, 20e3, rng), 10.0)
    logb = 3.0 + 0.4 * gaussian_field(ny, nx, dx, 8e3, rng)
    return {"bed": np.ascontiguousarray(bed), "H": np.ascontiguousarray(H), "beta": 10.0 ** logb / RHO_G,
            "B": np.full((ny, nx), B_HEAD), "smb": np.full((ny, nx), float(smb)), "dx": float(dx)}
