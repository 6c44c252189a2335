"""CPU oracle of the warm-start network (test infrastructure only).

A numpy restatement of the network that glide_amd/warmstart.py trains on the
GPU through libglide_hip.so (glide_amd/csrc/unet.hip): the same layers, the
same BF16 round-to-nearest-even points (every stored activation and every
ReLU-masked gradient), sums in float64 instead of the GPU's FP32 MFMA
accumulation.  The reference has no learned component, so there is no
reference file to follow; the layer algebra is the textbook one:

  conv3x3  y[p][co] = act(b[co] + sum_{t,ci} W[t][co][ci] x[p+off_t][ci])
  dgrad    dx[q][ci] = sum_{t,co} dz[q-off_t][co] W[t][co][ci]
  wgrad    dW[t][co][ci] = sum_p dz[p][co] x[p+off_t][ci]
  pool2    2x2 mean;  up2  nearest x2;  head  1x1 linear, 3 outputs (FP32)

With ``rounding=False`` every rounding point is the identity, which the
finite-difference tests use to pin the backward formulas.  Pins: tests/
test_unet_oracle.py (brute-force loops for conv/dgrad/wgrad on tiny inputs,
finite differences of the whole network's loss, BF16 rounding vs the
IEEE definition on hand-picked ties).
"""
from __future__ import annotations

import numpy as np

# ------------------------------------------------------------------ BF16
def bf16(x: np.ndarray) -> np.ndarray:
    """float32 -> bf16 (round to nearest even) -> float32, elementwise."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def to_bits(x: np.ndarray) -> np.ndarray:
    """bf16 values (float32 holding bf16) -> uint16 bit patterns."""
    return (np.ascontiguousarray(x, dtype=np.float32).view(np.uint32) >> 16).astype(np.uint16)


def from_bits(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)


# ------------------------------------------------------------------ layers
OFFS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def shift(x: np.ndarray, dy: int, dx: int) -> np.ndarray:
    """out[n, y, x] = x[n, y+dy, x+dx] (0 outside); x is NHWC."""
    N, H, W, C = x.shape
    out = np.zeros_like(x)
    ys, yd = (slice(dy, H), slice(0, H - dy)) if dy >= 0 else (slice(0, H + dy), slice(-dy, H))
    xs, xd = (slice(dx, W), slice(0, W - dx)) if dx >= 0 else (slice(0, W + dx), slice(-dx, W))
    out[:, yd, xd, :] = x[:, ys, xs, :]
    return out


def conv3x3(x, W, b, relu=True, rounding=True):
    """x NHWC (float), W [9][Co][Ci], b [Co]."""
    x64 = x.astype(np.float64)
    acc = np.zeros(x.shape[:3] + (W.shape[1],), dtype=np.float64)
    for t, (dy, dx) in enumerate(OFFS):
        acc += shift(x64, dy, dx) @ W[t].astype(np.float64).T
    acc += b.astype(np.float64)
    if relu:
        acc = np.maximum(acc, 0.0)
    return bf16(acc) if rounding else acc


def conv3x3_dgrad(dz, W):
    """dz NHWC (Co channels) -> dx NHWC (Ci channels), float64."""
    dz64 = dz.astype(np.float64)
    out = np.zeros(dz.shape[:3] + (W.shape[2],), dtype=np.float64)
    for t, (dy, dx) in enumerate(OFFS):
        out += shift(dz64, -dy, -dx) @ W[t].astype(np.float64)
    return out


def conv3x3_wgrad(dz, x):
    """dW[t][co][ci] = sum_p dz[p][co] x[p+off_t][ci], float64."""
    Co, Ci = dz.shape[-1], x.shape[-1]
    dzf = dz.astype(np.float64).reshape(-1, Co)
    out = np.zeros((9, Co, Ci), dtype=np.float64)
    for t, (dy, dx) in enumerate(OFFS):
        out[t] = dzf.T @ shift(x.astype(np.float64), dy, dx).reshape(-1, Ci)
    return out


def pool2(x):
    """2x2 mean in float32 ((a + b) + c) + d, times 0.25, rounded to BF16."""
    x = x.astype(np.float32)
    s = ((x[:, 0::2, 0::2] + x[:, 0::2, 1::2]) + x[:, 1::2, 0::2]) + x[:, 1::2, 1::2]
    return bf16(np.float32(0.25) * s)


def pool2_bwd(dy):
    return 0.25 * np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2)


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up2_bwd(dy):
    return ((dy[:, 0::2, 0::2] + dy[:, 0::2, 1::2]) + dy[:, 1::2, 0::2]) + dy[:, 1::2, 1::2]


def head(x, Wh, bh):
    return x.astype(np.float64) @ Wh.astype(np.float64).T + bh.astype(np.float64)


# ------------------------------------------------------------------ network
# name, inputs (concatenated along channels), Cout; every conv is followed by ReLU
LAYERS = [
    ("e1a", ["x0"], 32), ("e1", ["e1a"], 32),
    ("e2a", ["p1"], 64), ("e2", ["e2a"], 64),
    ("e3a", ["p2"], 96), ("e3", ["e3a"], 96),
    ("b", ["p3"], 96),
    ("d3a", ["u3", "e3"], 64), ("d3", ["d3a"], 64),
    ("d2a", ["u2", "e2"], 32), ("d2", ["d2a"], 32),
    ("d1a", ["u1", "e1"], 32), ("d1", ["d1a"], 32),
]
POOLS = {"p1": "e1", "p2": "e2", "p3": "e3"}      # pooled tensor <- source
UPS = {"u3": "b", "u2": "d3", "u1": "d2"}         # upsampled tensor <- source
IN_CH = 32
HEAD_IN = "d1"
N_OUT = 3          # u, v, H - H_prev


def channels():
    ch = {"x0": IN_CH}
    for name, ins, co in LAYERS:
        for src in (POOLS, UPS):
            for k, v in src.items():
                if k in ins and k not in ch:
                    ch[k] = ch[v]
        ch[name] = co
    return ch


def layer_shapes():
    ch = {"x0": IN_CH}
    shapes = {}
    for name, ins, co in LAYERS:
        for i in ins:
            if i not in ch:
                ch[i] = ch[POOLS.get(i) or UPS.get(i)]
        ci = sum(ch[i] for i in ins)
        shapes[name] = (co, ci)
        ch[name] = co
    return shapes


def init_params(seed: int = 0) -> dict:
    """He-normal 3x3 weights (fan-in 9 Cin), zero biases, small head."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, (co, ci) in layer_shapes().items():
        p[name + ".W"] = (rng.standard_normal((9, co, ci)) * np.sqrt(2.0 / (9 * ci))).astype(np.float32)
        p[name + ".b"] = np.zeros(co, dtype=np.float32)
    p["head.W"] = (rng.standard_normal((N_OUT, 32)) * np.sqrt(1.0 / 32)).astype(np.float32)
    p["head.b"] = np.zeros(N_OUT, dtype=np.float32)
    return p


def forward(params: dict, x0: np.ndarray, rounding: bool = True):
    """x0: N x H x W x 32 (float32 holding bf16 values).  Returns (pred, acts)."""
    rnd = bf16 if rounding else (lambda a: a)
    acts = {"x0": x0}
    for name, ins, co in LAYERS:
        for i in ins:
            if i in POOLS and i not in acts:
                acts[i] = pool2(acts[POOLS[i]]) if rounding else _pool64(acts[POOLS[i]])
            if i in UPS and i not in acts:
                acts[i] = up2(acts[UPS[i]])
        x = np.concatenate([acts[i] for i in ins], axis=-1)
        W = rnd(params[name + ".W"]) if rounding else params[name + ".W"].astype(np.float64)
        acts[name] = conv3x3(x, W, params[name + ".b"], relu=True, rounding=rounding)
    pred = head(acts[HEAD_IN], params["head.W"], params["head.b"])
    return pred, acts


def _pool64(x):
    x = x.astype(np.float64)
    return 0.25 * (((x[:, 0::2, 0::2] + x[:, 0::2, 1::2]) + x[:, 1::2, 0::2]) + x[:, 1::2, 1::2])


def mse(pred, target, mask=None):
    """loss = sum(m (pred - target)^2) / count, count = number of values."""
    m = 1.0 if mask is None else mask[..., None]
    d = (pred - target) * m
    n = d.size
    return float(np.sum(d * d) / n), 2.0 * d / n


def backward(params: dict, acts: dict, dpred: np.ndarray, rounding: bool = True) -> dict:
    """Gradients of all parameters given d(loss)/d(pred)."""
    rnd = bf16 if rounding else (lambda a: a)
    g = {}
    grads = {}
    d1 = acts[HEAD_IN]
    grads["head.W"] = np.einsum("nhwo,nhwc->oc", dpred, d1.astype(np.float64))
    grads["head.b"] = dpred.reshape(-1, dpred.shape[-1]).sum(axis=0)
    g[HEAD_IN] = dpred @ params["head.W"].astype(np.float64)
    shapes = layer_shapes()
    ch = channels()
    for name, ins, co in reversed(LAYERS):
        y = acts[name]
        dz = rnd(np.where(y > 0, g[name], 0.0))
        x = np.concatenate([acts[i] for i in ins], axis=-1)
        grads[name + ".W"] = conv3x3_wgrad(dz, x)
        grads[name + ".b"] = dz.reshape(-1, co).sum(axis=0)
        if ins == ["x0"]:
            continue
        W = rnd(params[name + ".W"]) if rounding else params[name + ".W"].astype(np.float64)
        dx = conv3x3_dgrad(dz, W)
        c0 = 0
        for i in ins:
            part = dx[..., c0:c0 + ch[i]]
            c0 += ch[i]
            g[i] = g[i] + part if i in g else part
        # propagate through pooling / upsampling as soon as a pooled / upsampled tensor's gradient is complete
        for i in ins:
            if i in UPS:
                src = UPS[i]
                g[src] = g.get(src, 0.0) + up2_bwd(g[i])
            if i in POOLS:
                src = POOLS[i]
                g[src] = g.get(src, 0.0) + pool2_bwd(g[i])
    assert set(shapes) <= {k[:-2] for k in grads}
    return grads
