"""Pins of the warm-start network's CPU oracle (oracle/unet_oracle.py):
BF16 rounding on hand-picked ties, the convolution against brute-force loops,
adjoint (dot-product) identities of every backward operator, and finite
differences of the whole network's loss with rounding disabled."""
import numpy as np

from oracle import unet_oracle as U


def test_bf16_round_to_nearest_even():
    one = 1.0
    ulp = 2.0 ** -7                 # BF16 spacing just above 1
    x = np.array([one + ulp / 2,    # tie -> even (1.0)
                  one + 3 * ulp / 2,  # tie -> even (1 + 2 ulp)
                  one + ulp / 2 + 2.0 ** -20,  # above the tie -> up
                  -(one + ulp / 2),  # symmetric
                  np.finfo(np.float32).max,  # above the largest BF16: rounds to inf
                  0.0, -0.0], dtype=np.float32)
    y = U.bf16(x)
    assert y[0] == 1.0 and y[1] == np.float32(one + 2 * ulp) and y[2] == np.float32(one + ulp)
    assert y[3] == -1.0 and np.isinf(y[4])
    assert y[5] == 0.0 and np.signbit(y[6])
    assert np.array_equal(U.from_bits(U.to_bits(y)), y)
    # random values: the result is one of the two neighbouring BF16 values, and the nearer one
    r = np.random.default_rng(0).standard_normal(10000).astype(np.float32)
    b = U.bf16(r)
    lo = U.from_bits(U.to_bits(r))                      # truncation toward zero
    step = np.abs(lo) * 2.0 ** -7
    assert np.all(np.abs(b - r) <= 0.5 * step * (1 + 1e-6) + 1e-30)


def _brute_conv(x, W, b):
    N, H, Wd, Ci = x.shape
    Co = W.shape[1]
    out = np.zeros((N, H, Wd, Co))
    for n in range(N):
        for y in range(H):
            for xx in range(Wd):
                for co in range(Co):
                    s = b[co]
                    for t in range(9):
                        yy, xs = y + t // 3 - 1, xx + t % 3 - 1
                        if 0 <= yy < H and 0 <= xs < Wd:
                            s += np.dot(W[t, co], x[n, yy, xs])
                    out[n, y, xx, co] = s
    return out


def test_conv_matches_brute_force():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 3, 4, 2))
    W = rng.standard_normal((9, 3, 2))
    b = rng.standard_normal(3)
    got = U.conv3x3(x, W, b, relu=False, rounding=False)
    assert np.allclose(got, _brute_conv(x, W, b), rtol=1e-13, atol=1e-13)


def test_backward_operators_are_adjoints():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 4, 6, 3))
    W = rng.standard_normal((9, 5, 3))
    dz = rng.standard_normal((2, 4, 6, 5))
    y = U.conv3x3(x, W, np.zeros(5), relu=False, rounding=False)
    # <conv_W(x), dz> = <x, dgrad(dz)> = <W, wgrad(dz, x)>
    a = np.sum(y * dz)
    assert abs(a - np.sum(x * U.conv3x3_dgrad(dz, W))) <= 1e-12 * abs(a)
    assert abs(a - np.sum(W * U.conv3x3_wgrad(dz, x))) <= 1e-12 * abs(a)
    # pooling and upsampling
    p = rng.standard_normal((2, 4, 6, 3))
    q = rng.standard_normal((2, 2, 3, 3))
    assert abs(np.sum(U._pool64(p) * q) - np.sum(p * U.pool2_bwd(q))) < 1e-12
    assert abs(np.sum(U.up2(q) * p) - np.sum(q * U.up2_bwd(p))) < 1e-12


def test_network_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    params = {k: v.astype(np.float64) for k, v in U.init_params(4).items()}
    for k in params:
        if k.endswith(".b"):
            params[k] = rng.standard_normal(params[k].shape) * 0.1
    x0 = rng.standard_normal((1, 8, 8, U.IN_CH))
    tgt = rng.standard_normal((1, 8, 8, U.N_OUT))

    def loss(p):
        pred, _ = U.forward(p, x0, rounding=False)
        return U.mse(pred, tgt)[0]

    pred, acts = U.forward(params, x0, rounding=False)
    _, dpred = U.mse(pred, tgt)
    grads = U.backward(params, acts, dpred, rounding=False)
    assert set(grads) == set(params)
    for name in ["e1a.W", "e3.W", "b.W", "d3a.W", "d2a.b", "d1.W", "head.W", "head.b"]:
        d = rng.standard_normal(params[name].shape)
        h = 1e-6
        pp = dict(params)
        pp[name] = params[name] + h * d
        lp = loss(pp)
        pp[name] = params[name] - h * d
        lm = loss(pp)
        fd = (lp - lm) / (2 * h)
        ad = float(np.sum(grads[name] * d))
        assert abs(fd - ad) <= 1e-6 * max(abs(fd), 1e-12), (name, fd, ad)
