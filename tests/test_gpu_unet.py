"""Warm-start network kernels (libglide_hip.so glide_unet_*) against the CPU
oracle (oracle/unet_oracle.py): the MFMA fragment layout, each layer, and the
whole network's loss and parameter gradients.

Tolerances: the oracle sums in float64 where the GPU accumulates in FP32, so a
BF16-rounded output can differ by one BF16 unit where the FP32 sum lands next
to a rounding boundary; FP32 outputs computed from identical BF16 inputs
agree to ~1e-6; whole-network quantities, where one-unit activation
differences propagate, agree to ~1e-2.
"""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from glide_amd import _lib  # noqa: E402
from glide_amd import warmstart as WS  # noqa: E402
from oracle import unet_oracle as U  # noqa: E402

DEV = torch.device("cuda")
_KEEP = []


def P(t):
    _KEEP.append(t)
    return ctypes.c_void_p(t.data_ptr())


def S():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def bits_t(x):
    """float32 array of BF16 values -> device int16 tensor of bits."""
    return torch.from_numpy(U.to_bits(x).view(np.int16).copy()).to(DEV)


def from_bits_t(t, shape):
    return U.from_bits(t.cpu().numpy().view(np.uint16)).reshape(shape)


def f32_t(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(DEV)


@pytest.fixture(autouse=True)
def _sync():
    yield
    torch.cuda.synchronize()
    _KEEP.clear()


def bf16_rand(rng, shape, scale=1.0):
    return U.bf16((rng.standard_normal(shape) * scale).astype(np.float32))


def test_mfma_fragment_layout_gemm():
    rng = np.random.default_rng(0)
    M, N, K = 48, 32, 96
    A, B = bf16_rand(rng, (M, K)), bf16_rand(rng, (N, K))
    C = torch.empty(M * N, dtype=torch.float32, device=DEV)
    _lib.call("glide_unet_gemm_bf16", P(bits_t(A)), P(bits_t(B)), P(C), M, N, K, S())
    ref = A.astype(np.float64) @ B.astype(np.float64).T
    got = C.cpu().numpy().reshape(M, N)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))


@pytest.mark.parametrize("cb,cout", [(0, 64), (32, 96), (64, 32)])
def test_conv_forward_matches_oracle(cb, cout):
    rng = np.random.default_rng(cout + cb)
    N, H, W, CA = 2, 8, 12, 32
    xa, xb = bf16_rand(rng, (N, H, W, CA)), bf16_rand(rng, (N, H, W, max(cb, 32)))[..., :cb]
    Wm = (rng.standard_normal((9, cout, CA + cb)) * 0.1).astype(np.float32)
    b = (rng.standard_normal(cout) * 0.1).astype(np.float32)
    wb = torch.empty(Wm.size, dtype=torch.int16, device=DEV)
    _lib.call("glide_unet_prep_weights", P(f32_t(Wm)), cout, CA + cb, P(wb), None, S())
    out = torch.empty(N * H * W * cout, dtype=torch.int16, device=DEV)
    xbt = bits_t(xb) if cb else None
    _lib.call("glide_unet_conv3x3", P(bits_t(xa)), CA, P(xbt) if cb else None, cb, N, H, W, P(wb), cout,
              P(f32_t(b)), 1, P(out), S())
    got = from_bits_t(out, (N, H, W, cout))
    x = np.concatenate([xa, xb], axis=-1) if cb else xa
    ref = U.conv3x3(x, U.bf16(Wm), b, relu=True)
    diff = np.abs(got - ref)
    ulp = np.maximum(np.abs(ref), 1e-30) * 2.0 ** -7
    assert np.all(diff <= ulp * 1.01 + 1e-30)
    assert np.mean(diff == 0) > 0.98


def test_conv_gradients_match_oracle():
    rng = np.random.default_rng(5)
    N, H, W, CA, CB, Co = 2, 8, 16, 32, 64, 32
    xa, xb = bf16_rand(rng, (N, H, W, CA)), bf16_rand(rng, (N, H, W, CB))
    dz = bf16_rand(rng, (N, H, W, Co))
    Wm = (rng.standard_normal((9, Co, CA + CB)) * 0.1).astype(np.float32)
    wb = torch.empty(Wm.size, dtype=torch.int16, device=DEV)
    wf = torch.empty(Wm.size, dtype=torch.int16, device=DEV)
    _lib.call("glide_unet_prep_weights", P(f32_t(Wm)), Co, CA + CB, P(wb), P(wf), S())
    ga = torch.full((N * H * W * CA,), 1.0, dtype=torch.float32, device=DEV)   # accumulate into ones
    gb = torch.empty(N * H * W * CB, dtype=torch.float32, device=DEV)
    _lib.call("glide_unet_conv3x3_dgrad", P(bits_t(dz)), Co, N, H, W, P(wf), CA, P(ga), 1, CB, P(gb), 0, S())
    ref = U.conv3x3_dgrad(dz, U.bf16(Wm))
    got = np.concatenate([ga.cpu().numpy().reshape(N, H, W, CA) - 1.0, gb.cpu().numpy().reshape(N, H, W, CB)], -1)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))
    nz = _lib.fn("glide_unet_wgrad_chunks")(N, H, W)
    part = torch.empty(nz * 9 * Co * (CA + CB), dtype=torch.float32, device=DEV)
    dw = torch.empty(9 * Co * (CA + CB), dtype=torch.float32, device=DEV)
    _lib.call("glide_unet_conv3x3_wgrad", P(bits_t(dz)), Co, P(bits_t(xa)), CA, P(bits_t(xb)), CB, N, H, W, P(part),
              P(dw), S())
    ref = U.conv3x3_wgrad(dz, np.concatenate([xa, xb], -1))
    got = dw.cpu().numpy().reshape(9, Co, CA + CB)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))


@pytest.mark.parametrize("C", [32, 4])
def test_pool_and_upsample_bitwise(C):
    """C = 32 takes the 8-channel vector kernels, C = 4 the per-element ones."""
    rng = np.random.default_rng(5 + C)
    N, H, W = 2, 8, 6
    x = bf16_rand(rng, (N, H, W, C))
    y = torch.empty(N * H * W * C // 4, dtype=torch.int16, device=DEV)
    _lib.call("glide_unet_pool2", P(bits_t(x)), N, H, W, C, P(y), S())
    assert np.array_equal(from_bits_t(y, (N, H // 2, W // 2, C)), U.pool2(x))
    up = torch.empty(N * H * W * C * 4, dtype=torch.int16, device=DEV)
    _lib.call("glide_unet_up2", P(bits_t(x)), N, H, W, C, P(up), S())
    assert np.array_equal(from_bits_t(up, (N, 2 * H, 2 * W, C)), U.up2(x))


def test_elementwise_layers():
    rng = np.random.default_rng(6)
    N, H, W, C = 2, 8, 6, 32
    x = bf16_rand(rng, (N, H, W, C))
    dy = rng.standard_normal((N, H // 2, W // 2, C)).astype(np.float32)
    dx = torch.full((N * H * W * C,), 2.0, dtype=torch.float32, device=DEV)
    _lib.call("glide_unet_pool2_bwd", P(f32_t(dy)), N, H, W, C, P(dx), 1, S())
    assert np.allclose(dx.cpu().numpy().reshape(N, H, W, C) - 2.0, U.pool2_bwd(dy), rtol=0, atol=1e-6)
    dy2 = rng.standard_normal((N, 2 * H, 2 * W, C)).astype(np.float32)
    dx2 = torch.empty(N * H * W * C, dtype=torch.float32, device=DEV)
    _lib.call("glide_unet_up2_bwd", P(f32_t(dy2)), N, H, W, C, P(dx2), S())
    assert np.allclose(dx2.cpu().numpy().reshape(N, H, W, C), U.up2_bwd(dy2), rtol=1e-6, atol=1e-6)
    g = rng.standard_normal(x.shape).astype(np.float32)
    dz = torch.empty(x.size, dtype=torch.int16, device=DEV)
    _lib.call("glide_unet_relu_bwd", P(f32_t(g)), P(bits_t(x)), x.size, P(dz), S())
    assert np.array_equal(from_bits_t(dz, x.shape), U.bf16(np.where(x > 0, g, 0.0)))


def test_network_loss_and_gradients_match_oracle():
    rng = np.random.default_rng(7)
    net = WS.UNet(DEV, seed=3)
    for k in net.params:
        if k.endswith(".b"):
            net.params[k].copy_(torch.as_tensor(rng.standard_normal(net.params[k].shape) * 0.05))
    net.sync_weights()
    params = net.state_dict()
    N, H, W = 2, 16, 24
    x0 = bf16_rand(rng, (N, H, W, WS.IN_CH))
    tgt = rng.standard_normal((N, H, W, WS.N_OUT)).astype(np.float32)
    xt, tt = bits_t(x0), f32_t(tgt)
    _KEEP.extend([xt, tt])
    pred = net.forward(xt, N, H, W, train=True)
    loss = float(net.loss_and_grad(tt, N, H, W))
    net.backward(N, H, W)
    torch.cuda.synchronize()
    rp, acts = U.forward(params, x0)
    rl, dpred = U.mse(rp, tgt)
    rg = U.backward(params, acts, dpred)
    gp = pred.cpu().numpy().reshape(N, H, W, WS.N_OUT)
    assert np.linalg.norm(gp - rp) <= 1e-2 * np.linalg.norm(rp)
    assert abs(loss - rl) <= 1e-2 * rl
    worst = {}
    for k, gt in net.grads.items():
        got = gt.cpu().numpy().reshape(rg[k].shape)
        worst[k] = np.linalg.norm(got - rg[k]) / max(np.linalg.norm(rg[k]), 1e-30)
    print("relative gradient errors", {k: f"{v:.1e}" for k, v in worst.items()}, flush=True)
    assert max(worst.values()) <= 3e-2, worst


def test_training_reduces_loss():
    # regression of a smooth function of the inputs: the loss falls by 5x
    rng = np.random.default_rng(8)
    N, H, W = 4, 32, 32
    x = np.zeros((N, H, W, WS.IN_CH), dtype=np.float32)
    yy, xx = np.meshgrid(np.linspace(-1, 1, H), np.linspace(-1, 1, W), indexing="ij")
    for n in range(N):
        a, b = rng.uniform(-1, 1, 2)
        x[n, :, :, 0] = a * yy + b * xx
        x[n, :, :, 1] = np.sin(3 * xx + n)
    x = U.bf16(x)
    tgt = np.stack([x[..., 0] * 2 + x[..., 1], x[..., 0] - x[..., 1] ** 2, x[..., 1] * 0.5], -1).astype(np.float32)
    net = WS.UNet(DEV, seed=1)
    xt, tt = bits_t(x), f32_t(tgt)
    losses = []
    for it in range(150):
        net.forward(xt, N, H, W, train=True)
        losses.append(float(net.loss_and_grad(tt, N, H, W)))
        net.backward(N, H, W)
        net.adam_step(2e-3)
    print("loss", losses[0], "->", losses[-1], flush=True)
    assert losses[-1] < losses[0] / 5


@pytest.mark.parametrize("C", [32, 12])
def test_head_matches_oracle(C):
    """The 1x1 head (FP32 sums over BF16 inputs) against the oracle's float64
    sums; C = 32 takes the 8-channel vector kernel, C = 12 the per-element
    one (both sum the channels in ascending order)."""
    rng = np.random.default_rng(40 + C)
    Pn, NO = 1000, 3
    x = bf16_rand(rng, (Pn, C))
    Wh = rng.standard_normal((NO, C)).astype(np.float32) * 0.2
    bh = rng.standard_normal(NO).astype(np.float32)
    out = torch.empty(Pn * NO, dtype=torch.float32, device=DEV)
    _lib.call("glide_unet_head", P(bits_t(x)), Pn, C, NO, P(f32_t(Wh)), P(f32_t(bh)), P(out), S())
    ref = U.head(x, Wh, bh)
    got = out.cpu().numpy().reshape(Pn, NO)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))
