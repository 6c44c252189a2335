"""Input preparation on the CPU: the multigrid crop (glide/data.py:433-473)
against its closed form, and the coordinate checks of the interpolation."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from glide_amd.inputs import _direction, prepare_grid  # noqa: E402


def geometry(ny, nx):
    rng = np.random.default_rng(0)
    return {"x": np.arange(nx) * 150.0 - 7e5, "y": 3e6 - np.arange(ny) * 150.0,
            "bed": rng.normal(size=(ny, nx)), "thickness": rng.normal(size=(ny, nx)),
            "surface": rng.normal(size=(ny, nx)), "dx": 150.0}


@pytest.mark.parametrize("ny,nx,levels", [(45, 70, 3), (64, 64, 5), (33, 100, 5), (200, 129, 4)])
def test_prepare_grid_centres_the_largest_multiple(ny, nx, levels):
    g = geometry(ny, nx)
    out = prepare_grid(g, n_levels=levels)
    f = 2 ** levels
    nyt, nxt = ny - ny % f, nx - nx % f
    y0, x0 = (ny % f) // 2, (nx % f) // 2          # the remainder split, the extra row/column at the end
    assert (out["ny"], out["nx"]) == (nyt, nxt)
    for k in ("bed", "thickness", "surface"):
        assert np.array_equal(out[k], g[k][y0:y0 + nyt, x0:x0 + nxt]), k
    assert np.array_equal(out["x"], g["x"][x0:x0 + nxt]) and np.array_equal(out["y"], g["y"][y0:y0 + nyt])
    assert out["dx"] == 150.0


def test_prepare_grid_works_on_tensors_and_rejects_tiny_grids():
    g = {k: (torch.as_tensor(v) if isinstance(v, np.ndarray) else v) for k, v in geometry(40, 50).items()}
    out = prepare_grid(g, n_levels=3)
    assert tuple(out["bed"].shape) == (40, 48) and torch.equal(out["bed"], g["bed"][:, 1:49])
    with pytest.raises(ValueError):
        prepare_grid(geometry(20, 40), n_levels=5)


def test_coordinate_direction_checks():
    assert _direction(np.arange(5.0), "x") == 0
    assert _direction(np.arange(5.0)[::-1], "y") == 1
    for bad in (np.array([0.0, 1.0, 1.0]), np.array([0.0, 2.0, 1.0]), np.array([1.0]), np.zeros((2, 2))):
        with pytest.raises(ValueError):
            _direction(bad, "x")
