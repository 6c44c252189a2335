"""Input preparation on the GPU (glide_amd.inputs, libglide_hip.so
glide_interp_grid / glide_bedmachine_thickness).

- interpolate_to_grid against scipy's RegularGridInterpolator, the routine
  the reference calls (glide/data.py:386-431): ascending and descending
  axes, targets off the data's extent (fill 0), on nodes and on the last
  node, NaN targets and NaN data.  The FP64 results are bitwise equal (the
  kernel forms the interval search, the normalised distances and the four
  corner terms in the interpolator's order), and so is the float32 result
  the reference returns.
- bedmachine_thickness against its definition (glide/data.py:257-258) and
  the hydrostatic invariant it exists for.
- A BedMachine-like mini pipeline (descending y, crop, thickness,
  observations interpolated) feeding a forward step.
"""
import time

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from scipy.interpolate import RegularGridInterpolator  # noqa: E402

from glide_amd.inputs import bedmachine_thickness, interpolate_to_grid, prepare_grid  # noqa: E402
from glide_amd.physics import IcePhysics  # noqa: E402

DEV = torch.device("cuda")


def scipy_interp(data, x_data, y_data, x_t, y_t):
    X, Y = np.meshgrid(x_t, y_t)
    f = RegularGridInterpolator((y_data, x_data), data, bounds_error=False, fill_value=0.0)
    return f((Y, X))


def compare(got64, want64):
    both_nan = np.isnan(got64) & np.isnan(want64)
    assert np.array_equal(np.isnan(got64), np.isnan(want64))
    g, w = got64[~both_nan], want64[~both_nan]
    ulps = np.abs(g - w) / np.maximum(np.spacing(np.abs(w)), np.finfo(float).tiny)
    return float(np.mean(g == w)), float(ulps.max()) if ulps.size else 0.0


@pytest.mark.parametrize("y_desc", [False, True])
def test_interpolation_matches_scipy(y_desc):
    rng = np.random.default_rng(5 + int(y_desc))
    ny, nx = 57, 83
    x = np.cumsum(rng.uniform(0.5, 2.0, nx)) * 100.0            # uneven, ascending
    y = np.cumsum(rng.uniform(0.5, 2.0, ny)) * 100.0
    if y_desc:
        y = y[::-1].copy()
    data = rng.standard_normal((ny, nx)) * 50.0
    lo_x, hi_x, lo_y, hi_y = x.min(), x.max(), y.min(), y.max()
    xt = np.concatenate([rng.uniform(lo_x - 300, hi_x + 300, 150), x[::7], [hi_x, lo_x, np.nan]])
    yt = np.concatenate([rng.uniform(lo_y - 300, hi_y + 300, 120), y[::5], [hi_y, lo_y, np.nan]])
    got = interpolate_to_grid(data, x, y, xt, yt, dtype=torch.float64).cpu().numpy()
    want = scipy_interp(data, x, y, xt, yt)
    eq, ulps = compare(got, want)
    print(f"y descending={y_desc}: {got.size} targets, {100 * eq:.2f}% bitwise equal, max {ulps:.1f} ulp", flush=True)
    assert eq == 1.0
    g32 = interpolate_to_grid(data, x, y, xt, yt).cpu().numpy()
    w32 = want.astype(np.float32)
    assert g32.dtype == np.float32
    assert np.array_equal(g32, w32, equal_nan=True)
    # outside the extent: the fill value, exactly
    out = (yt[:, None] < lo_y) | (yt[:, None] > hi_y) | (xt[None, :] < lo_x) | (xt[None, :] > hi_x)
    assert np.all(got[out & ~np.isnan(got)] == 0.0)


def test_nan_data_propagates_like_scipy():
    rng = np.random.default_rng(9)
    x, y = np.linspace(0, 10, 11), np.linspace(0, 5, 6)
    data = rng.standard_normal((6, 11))
    data[2, 3] = np.nan
    xt, yt = rng.uniform(0.01, 9.99, 40), rng.uniform(0.01, 4.99, 30)      # no node hits
    got = interpolate_to_grid(data, x, y, xt, yt, dtype=torch.float64).cpu().numpy()
    want = scipy_interp(data, x, y, xt, yt)
    eq, ulps = compare(got, want)
    assert np.isnan(got).any() and eq == 1.0


def test_interpolation_rejects_bad_axes():
    data = np.zeros((3, 4))
    with pytest.raises(ValueError):
        interpolate_to_grid(data, [0, 1, 1, 2], [0, 1, 2], [0.5], [0.5])
    with pytest.raises(ValueError):
        interpolate_to_grid(data, [0, 1, 2], [0, 1, 2], [0.5], [0.5])      # shape mismatch


def test_thickness_definition_and_hydrostatic_floating_ice():
    rng = np.random.default_rng(3)
    n = 100_000
    bed = rng.uniform(-1500, 800, n)
    surface = rng.uniform(-5, 2500, n)
    surface[:10] = [np.nan, 0, 1, 2, 3, 4, 5, 6, 7, 8]
    bed[1] = np.nan
    H = bedmachine_thickness(bed, surface, thklim=0.1).cpu().numpy()
    c = 1.0 - 1.0 / (1.0 - 0.917)
    base = np.maximum(bed, surface * c)
    assert np.array_equal(np.isnan(H), np.isnan(base))
    ok = ~np.isnan(H)
    assert np.array_equal(H[ok], (surface - base + 0.1)[ok])
    grounded = ok & (bed >= surface * c)
    floating = ok & ~grounded
    assert np.array_equal(H[grounded], (surface - bed + 0.1)[grounded])
    # floating: the ice below sea level is 0.917 of the column
    Hf = H[floating] - 0.1
    assert np.allclose(surface[floating], (1 - 0.917) * Hf, rtol=1e-12, atol=1e-9)


def test_bedmachine_like_pipeline_runs_a_step():
    rng = np.random.default_rng(4)
    ny, nx, dx = 150, 141, 2000.0
    x = np.arange(nx) * dx
    y = 2e6 - np.arange(ny) * dx                              # BedMachine stores y descending
    Y, X = np.meshgrid(y, x, indexing="ij")
    r = np.hypot(X - x.mean(), Y - y.mean())
    bed = 300.0 - r / 400.0 + 20.0 * rng.standard_normal((ny, nx))
    # ice dome; ice-free land at its bed, the sea at 0 (surface >= bed on land, as in BedMachine)
    surface = np.maximum(2000.0 * (1 - (r / 1.2e5) ** 2), np.maximum(bed, 0.0))
    H = bedmachine_thickness(bed, surface, thklim=0.1)
    assert float(H.min()) >= 0.1 - 1e-12
    geo = prepare_grid({"x": x, "y": y, "bed": bed, "thickness": H, "surface": surface, "dx": dx}, n_levels=4)
    assert (geo["ny"], geo["nx"]) == (144, 128)
    vel = rng.standard_normal((len(y[::2]), len(x[::2]))) * 10.0
    u_cell = interpolate_to_grid(vel, x[::2], y[::2], geo["x"], geo["y"], dtype=torch.float64)
    assert tuple(u_cell.shape) == (144, 128) and torch.isfinite(u_cell).all()
    ph = IcePhysics(geo["ny"], geo["nx"], dx, n_levels=4)
    ph.set_geometry(geo["bed"], geo["thickness"])
    ph.set_parameters(B=1e-16 ** (-1 / 3) / (917 * 9.81), beta=np.full((144, 128), 0.05), smb=0.0)
    u, v, Hn = ph.forward(1.0, n_vcycles=2)
    assert torch.isfinite(u).all() and torch.isfinite(Hn).all()


def test_interpolation_timing_2048_to_1024():
    rng = np.random.default_rng(6)
    n_d, n_t = 2048, 1024
    x = np.arange(n_d) * 250.0
    y = x[::-1].copy()
    data = torch.as_tensor(rng.standard_normal((n_d, n_d)), device=DEV)
    xt = torch.as_tensor(np.linspace(100.0, x[-1] - 100.0, n_t), device=DEV)
    yt = torch.as_tensor(np.linspace(100.0, y[0] - 100.0, n_t), device=DEV)
    xd, yd = torch.as_tensor(x, device=DEV), torch.as_tensor(y, device=DEV)
    interpolate_to_grid(data, xd, yd, xt, yt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        interpolate_to_grid(data, xd, yd, xt, yt, dtype=torch.float64)
    torch.cuda.synchronize()
    t_gpu = (time.perf_counter() - t0) / 10
    dn = data.cpu().numpy()
    t0 = time.perf_counter()
    scipy_interp(dn, x, y, xt.cpu().numpy(), yt.cpu().numpy())
    t_cpu = time.perf_counter() - t0
    print(f"interpolate 2048^2 -> 1024^2: GPU {t_gpu * 1e3:.2f} ms per call (incl. host checks), "
          f"scipy {t_cpu * 1e3:.1f} ms", flush=True)
