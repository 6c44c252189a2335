"""Shared test inputs: small random MAC-grid states (seeded)."""
import numpy as np

from oracle import oracle as O


def random_state(ny=8, nx=10, dx=2000.0, seed=0, thin_frac=0.2, float_frac=0.3, vel=300.0):
    """A random but physically plausible state on an ny x nx level: thickness
    from 0.1 m to ~1500 m (some cells at/near the 0.1 m limit), a bed that
    leaves part of the domain floating, random velocities."""
    rng = np.random.default_rng(seed)
    H = rng.uniform(50.0, 1500.0, (ny, nx))
    thin = rng.random((ny, nx)) < thin_frac
    H[thin] = rng.uniform(0.1, 2.0, thin.sum())
    bed = rng.uniform(-300.0, 800.0, (ny, nx))
    fl = rng.random((ny, nx)) < float_frac
    bed[fl] = rng.uniform(-1500.0, -400.0, fl.sum())
    u = rng.normal(0.0, vel, (ny, nx + 1))
    v = rng.normal(0.0, vel, (ny + 1, nx))
    beta = 10.0 ** rng.normal(3.0, 0.5, (ny, nx)) / 8995.77
    B = np.full((ny, nx), 23.9494) * rng.uniform(0.8, 1.2, (ny, nx))
    f = O.join(rng.normal(0, 1.0, (ny, nx + 1)), rng.normal(0, 1.0, (ny + 1, nx)),
               rng.normal(0, 0.1, (ny, nx)))
    gamma = np.full((ny, nx), 0.1)
    mask = np.zeros((ny, nx))
    return dict(ny=ny, nx=nx, dx=dx, dt=10.0, U=O.join(u, v, H), f=f, bed=bed, B=B, beta=beta,
                mask=mask, gamma=gamma, H=H)


def res(s, prm, U=None, f=None, beta=None, mask=None):
    return O.residual(s["ny"], s["nx"], s["dx"], s["dt"], prm, s["U"] if U is None else U,
                      s["f"] if f is None else f, s["bed"], s["B"], s["beta"] if beta is None else beta,
                      s["mask"] if mask is None else mask, s["gamma"])
