"""Learned warm start: the network input, the face mapping, the fallback rule,
and what the trained network (glide_amd/data/warmstart_unet.npz) gains over
the cold start on unseen seeds."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from glide_amd import synthetic as S  # noqa: E402
from glide_amd.train_warmstart import TEST_SEED0, member  # noqa: E402
from glide_amd.warmstart import IN_CH, VEL_SCALE, UNet, WarmStart, features, targets  # noqa: E402
from oracle import unet_oracle as U  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "glide_amd", "data", "warmstart_unet.npz")


def test_features_match_numpy():
    ph = member(TEST_SEED0, 32, dx=1000.0, n_levels=2)
    x = features(ph)
    got = U.from_bits(x.cpu().numpy().view(np.uint16)).reshape(32, 32, IN_CH)
    d = S.ice_sheet(32, 32, 1000.0, TEST_SEED0)
    H, bed = d["H"], d["bed"]
    s = np.maximum(bed, -0.917 * H) + H
    sx = np.zeros_like(s)
    sy = np.zeros_like(s)
    sx[:, 1:-1] = (s[:, 2:] - s[:, :-2]) / 2000.0
    sx[:, 0] = (s[:, 1] - s[:, 0]) / 1000.0
    sx[:, -1] = (s[:, -1] - s[:, -2]) / 1000.0
    sy[1:-1] = (s[2:] - s[:-2]) / 2000.0
    sy[0] = (s[1] - s[0]) / 1000.0
    sy[-1] = (s[-1] - s[-2]) / 1000.0
    ref = np.stack([H / 1000.0, bed / 1000.0, 100.0 * np.sqrt(sx ** 2 + sy ** 2), np.log(d["beta"]) + 2.0,
                    d["B"] / 24.0 - 1.0], -1).astype(np.float32)
    ref = U.bf16(ref)
    assert np.all(got[..., 5:] == 0)
    diff = np.abs(got[..., :5] - ref)
    assert np.all(diff <= np.abs(ref) * 2.0 ** -7 * 1.01 + 1e-30)
    assert np.mean(diff == 0) > 0.99


def test_targets_and_faces_round_trip_constant_field():
    ph = member(TEST_SEED0 + 1, 16, dx=1000.0, n_levels=1)
    g = ph.grid
    g.u.fill_(120.0)
    g.v.fill_(-35.0)
    t = targets(ph)
    assert torch.allclose(t.view(-1, 2)[:, 0], torch.tensor(120.0 * VEL_SCALE, device=t.device))
    g.u.zero_()
    g.v.zero_()
    from glide_amd import _lib
    import ctypes
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.call("glide_unet_to_faces", ctypes.c_void_p(t.data_ptr()), g.ny, g.nx, VEL_SCALE, g.u.data_ptr(),
              g.v.data_ptr(), st)
    assert torch.allclose(g.u, torch.full_like(g.u, 120.0)) and torch.allclose(g.v, torch.full_like(g.v, -35.0))


class _Seeded(WarmStart):
    """A warm start whose 'prediction' is a given (u, v)."""

    def __init__(self, u, v):
        super().__init__(None)
        self.pu, self.pv = u, v

    def predict_into(self, physics):
        physics.grid.u.copy_(self.pu)
        physics.grid.v.copy_(self.pv)


def test_exact_seed_needs_no_vcycles_and_garbage_falls_back():
    ph = member(TEST_SEED0 + 2, 64, dx=1000.0, n_levels=3)
    cold = WarmStart(None).solve(ph, 10.0, tol=1e-8, max_vcycles=200, use_net=False)
    assert cold["converged"] and cold["vcycles"] > 5
    u_sol, v_sol = ph.grid.u.clone(), ph.grid.v.clone()
    ph.grid.u.zero_()
    ph.grid.v.zero_()
    exact = _Seeded(u_sol, v_sol).solve(ph, 10.0, tol=1e-8, max_vcycles=200, use_net=True)
    assert exact["used_net"] and not exact["fallback"] and exact["vcycles"] <= 2
    ph.grid.u.zero_()
    ph.grid.v.zero_()
    bad = _Seeded(u_sol * 0 + 1e4, v_sol * 0 - 1e4).solve(ph, 10.0, tol=1e-8, max_vcycles=200, use_net=True)
    assert bad["fallback"] and not bad["used_net"]
    assert bad["vcycles"] == cold["vcycles"]          # the cold start, unchanged


def test_trained_network_saves_vcycles():
    assert os.path.exists(WEIGHTS), "trained weights are part of the product"
    net = UNet.load(WEIGHTS)
    ws = WarmStart(net)
    saved = []
    for k in range(3):
        ph = member(TEST_SEED0 + 10 + k, 256, n_levels=5)
        cold = ws.solve(ph, 10.0, tol=1e-8, max_vcycles=300, use_net=False)
        ph.grid.u.zero_()
        ph.grid.v.zero_()
        warm = ws.solve(ph, 10.0, tol=1e-8, max_vcycles=300, use_net=True)
        print(f"seed {TEST_SEED0 + 10 + k}: cold {cold['vcycles']} warm {warm['vcycles']} "
              f"r0 ratio {warm['r0_warm'] / cold['r0_cold']:.3e}", flush=True)
        assert cold["converged"] and warm["converged"] and not warm["fallback"]
        saved.append(cold["vcycles"] - warm["vcycles"])
    assert min(saved) >= 1
