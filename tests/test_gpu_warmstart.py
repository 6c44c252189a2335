"""Learned warm start: the network input, the face mapping, the fallback rule,
and what the trained network (glide_amd/data/warmstart_unet.npz) gains over
the cold start on unseen seeds."""
import os
import time

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from glide_amd import synthetic as S  # noqa: E402
from glide_amd.physics import TOLERANCE_SCHEDULE  # noqa: E402
from glide_amd.train_warmstart import TEST_SEED0, member  # noqa: E402
from glide_amd.warmstart import (DH_SCALE, IN_CH, N_OUT, VEL_SCALE, UNet, WarmStart,  # noqa: E402
                                 apply_prediction, features, targets)
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
    ref = np.stack([H / 1000.0, np.tanh((bed + 0.917 * H) / 200.0), 100.0 * np.sqrt(sx ** 2 + sy ** 2),
                    np.log(d["beta"]) + 2.0, d["B"] / 24.0 - 1.0, 100.0 * sx, 100.0 * sy], -1).astype(np.float32)
    ref = U.bf16(ref)
    assert np.all(got[..., 7:] == 0)
    diff = np.abs(got[..., :7] - ref)
    assert np.all(diff <= np.abs(ref) * 2.0 ** -7 * 1.01 + 1e-30)
    assert np.mean(diff == 0) > 0.99


def test_targets_and_prediction_round_trip_constant_field():
    ph = member(TEST_SEED0 + 1, 16, dx=1000.0, n_levels=1)
    g = ph.grid
    g.u.fill_(120.0)
    g.v.fill_(-35.0)
    g.H.copy_(g.H_prev + 7.5)
    t = targets(ph)
    tv = t.view(-1, N_OUT)
    assert torch.allclose(tv[:, 0], torch.tensor(120.0 * VEL_SCALE, device=t.device))
    assert torch.allclose(tv[:, 2], torch.tensor(7.5 * DH_SCALE, device=t.device))
    g.u.zero_()
    g.v.zero_()
    g.H.copy_(g.H_prev)
    apply_prediction(ph, t)
    assert torch.allclose(g.u, torch.full_like(g.u, 120.0)) and torch.allclose(g.v, torch.full_like(g.v, -35.0))
    assert torch.allclose(g.H, g.H_prev + 7.5)
    # the thickness seed never goes below gamma
    tv[:, 2] = -1e9
    apply_prediction(ph, t)
    assert torch.equal(g.H, torch.maximum(g.H, g.gamma)) and float((g.H - g.gamma).abs().max()) == 0.0


class _Seeded(WarmStart):
    """A warm start whose 'prediction' is a given state (u, v, H)."""

    def __init__(self, u, v, H):
        super().__init__(None)
        self.pu, self.pv, self.pH = u, v, H

    def predict_into(self, physics):
        physics.grid.u.copy_(self.pu)
        physics.grid.v.copy_(self.pv)
        physics.grid.H.copy_(self.pH)


def test_exact_seed_needs_no_vcycles_and_garbage_falls_back():
    ph = member(TEST_SEED0 + 2, 64, dx=1000.0, n_levels=3)
    cold = WarmStart(None).solve(ph, 10.0, tol=1e-8, max_vcycles=200, use_net=False)
    assert cold["converged"] and cold["vcycles"] > 5
    u_sol, v_sol, H_sol = ph.grid.u.clone(), ph.grid.v.clone(), ph.grid.H.clone()
    ph.grid.u.zero_()
    ph.grid.v.zero_()
    exact = _Seeded(u_sol, v_sol, H_sol).solve(ph, 10.0, tol=1e-8, max_vcycles=200, use_net=True)
    assert exact["used_net"] and not exact["fallback"] and exact["vcycles"] <= 1
    ph.grid.u.zero_()
    ph.grid.v.zero_()
    bad = _Seeded(u_sol * 0 + 1e4, v_sol * 0 - 1e4, H_sol).solve(ph, 10.0, tol=1e-8, max_vcycles=200, use_net=True)
    assert bad["fallback"] and not bad["used_net"]
    assert bad["vcycles"] == cold["vcycles"]          # the cold start, unchanged


def test_trained_network_saves_vcycles_and_time():
    """The shipped network on unseen 512^2 ice sheets (dx 1500 m, 5 levels,
    dt = 1 a, the tolerance schedule: the headline's workload, seeds past the
    benchmark's), no seed skipped: every solve reaches 1e-8 from both starts,
    the prediction cuts the initial residual with no fallback, and over the
    seeds the warm start needs fewer V-cycles and less time (inference
    included)."""
    assert os.path.exists(WEIGHTS), "trained weights are part of the product"
    net = UNet.load(WEIGHTS)
    ws = WarmStart(net)
    vc_cold = vc_warm = 0
    t_cold = t_warm = 0.0
    for k in range(6):
        seed = TEST_SEED0 + 40 + k
        ph = member(seed, 512, dx=1500.0, n_levels=5)
        ph.set_schedule(**TOLERANCE_SCHEDULE)
        cold = ws.solve(ph, 1.0, tol=1e-8, max_vcycles=100, use_net=False)     # warm-up of this seed's buffers
        for rep in range(2):
            ph.grid.u.zero_()
            ph.grid.v.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = ws.solve(ph, 1.0, tol=1e-8, max_vcycles=100, use_net=bool(rep))
            torch.cuda.synchronize()
            dt_ms = (time.perf_counter() - t0) * 1e3
            if rep == 0:
                cold, tc = r, dt_ms
            else:
                warm, tw = r, dt_ms
        ratio = warm["r0_warm"] / cold["r0_cold"]
        print(f"seed {seed}: cold {cold['vcycles']} V-cycles {tc:.1f} ms, warm {warm['vcycles']} V-cycles "
              f"{tw:.1f} ms, r0 ratio {ratio:.3e}", flush=True)
        assert cold["converged"] and warm["converged"] and not warm["fallback"] and ratio < 0.5
        vc_cold += cold["vcycles"]
        vc_warm += warm["vcycles"]
        t_cold += tc
        t_warm += tw
    print(f"total: cold {vc_cold} V-cycles {t_cold:.1f} ms, warm {vc_warm} V-cycles {t_warm:.1f} ms", flush=True)
    assert vc_warm < vc_cold and t_warm < t_cold
