"""Learned warm start of the FASCD forward solve.

A 4-level convolutional U-Net (32/64/96/96 channels, 516,931 parameters)
maps the cell fields of a domain (H, the bed as flotation margin, |grad s|,
log beta, B and the signed surface slopes) to the cell-centred (u, v) and
the thickness change H - H_prev of a forward step; the prediction, mapped to
the MAC faces and to H = max(H_prev + dH, gamma), replaces the cold initial
state of the step.  Plain FASCD remains the solver and the arbiter: the step iterates V-cycles until the reference's residual norm
(glide/physics.py:176-203) drops below ``tol`` times the COLD start's initial
norm, and if the predicted state's initial residual exceeds the cold one the
prediction is discarded (fallback to the cold start).

Every network operation runs in libglide_hip.so (glide_amd/csrc/unet.hip):
BF16 weights and activations, FP32 accumulation on the matrix cores, FP32
master weights and Adam.  torch only allocates the buffers.  The reference has
no learned component; the solve it seeds is IcePhysics.forward
(glide/physics.py:141-209).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib

# name, inputs (concatenated along channels), output channels; every conv is followed by ReLU
LAYERS = [
    ("e1a", ["x0"], 32), ("e1", ["e1a"], 32),
    ("e2a", ["p1"], 64), ("e2", ["e2a"], 64),
    ("e3a", ["p2"], 96), ("e3", ["e3a"], 96),
    ("b", ["p3"], 96),
    ("d3a", ["u3", "e3"], 64), ("d3", ["d3a"], 64),
    ("d2a", ["u2", "e2"], 32), ("d2", ["d2a"], 32),
    ("d1a", ["u1", "e1"], 32), ("d1", ["d1a"], 32),
]
POOLS = {"p1": "e1", "p2": "e2", "p3": "e3"}
UPS = {"u3": "b", "u2": "d3", "u1": "d2"}
IN_CH = 32
HEAD_IN = "d1"
VEL_SCALE = 0.01          # network units: m/a x 0.01
DH_SCALE = 0.05           # network units of the step's thickness change: m x 0.05
N_OUT = 3                 # outputs per cell: u, v, H - H_prev
DOWNSAMPLE = 8            # domain sides must be multiples of this


def _ch_level():
    ch, lv = {"x0": IN_CH}, {"x0": 0}
    for name, ins, co in LAYERS:
        for i in ins:
            if i in POOLS and i not in ch:
                ch[i], lv[i] = ch[POOLS[i]], lv[POOLS[i]] + 1
            if i in UPS and i not in ch:
                ch[i], lv[i] = ch[UPS[i]], lv[UPS[i]] - 1
        ch[name], lv[name] = co, lv[ins[0]]
    return ch, lv


CH, LEVEL = _ch_level()
SHAPES = {name: (co, sum(CH[i] for i in ins)) for name, ins, co in LAYERS}


def n_params() -> int:
    return sum(9 * co * ci + co for co, ci in SHAPES.values()) + N_OUT * CH[HEAD_IN] + N_OUT


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class UNet:
    """The warm-start network: FP32 master parameters, BF16 working copies,
    Adam state and per-shape activation / gradient buffers on one device."""

    def __init__(self, device=None, seed: int = 0):
        _lib.load()
        self.device = torch.device(device if device is not None else "cuda")
        rng = np.random.default_rng(seed)
        self.params: dict[str, torch.Tensor] = {}
        for name, (co, ci) in SHAPES.items():
            w = rng.standard_normal((9, co, ci)) * math.sqrt(2.0 / (9 * ci))
            self.params[name + ".W"] = torch.tensor(w, dtype=torch.float32, device=self.device)
            self.params[name + ".b"] = torch.zeros(co, dtype=torch.float32, device=self.device)
        c = CH[HEAD_IN]
        self.params["head.W"] = torch.tensor(rng.standard_normal((N_OUT, c)) * math.sqrt(1.0 / c),
                                             dtype=torch.float32, device=self.device)
        self.params["head.b"] = torch.zeros(N_OUT, dtype=torch.float32, device=self.device)
        self.grads = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.m = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.step_count = 0
        self.wb = {n: torch.empty(9 * co * ci, dtype=torch.int16, device=self.device) for n, (co, ci) in SHAPES.items()}
        self.wf = {n: torch.empty(9 * co * ci, dtype=torch.int16, device=self.device) for n, (co, ci) in SHAPES.items()}
        self._bufs: dict = {}
        self.sync_weights()

    # ------------------------------------------------------------ helpers
    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def sync_weights(self):
        """BF16 (and flipped) working copies of the FP32 master weights."""
        for name, (co, ci) in SHAPES.items():
            _lib.call("glide_unet_prep_weights", _p(self.params[name + ".W"]), co, ci, _p(self.wb[name]),
                      _p(self.wf[name]), self._stream())

    def _buffers(self, N, H, W, train: bool):
        key = (N, H, W, train)
        if key in self._bufs:
            return self._bufs[key]
        if H % DOWNSAMPLE or W % DOWNSAMPLE:
            raise ValueError(f"network input {H}x{W} must have sides divisible by {DOWNSAMPLE}")
        dims = {lv: (H >> lv, W >> lv) for lv in range(4)}
        acts = {}
        for name in CH:
            if name == "x0":
                continue
            h, w = dims[LEVEL[name]]
            acts[name] = torch.empty(N * h * w * CH[name], dtype=torch.int16, device=self.device)
        b = {"acts": acts, "dims": dims,
             "pred": torch.empty(N * H * W * N_OUT, dtype=torch.float32, device=self.device)}
        if train:
            g = {}
            for name in CH:
                if name == "x0":
                    continue
                h, w = dims[LEVEL[name]]
                g[name] = torch.empty(N * h * w * CH[name], dtype=torch.float32, device=self.device)
            b["g"] = g
            b["dz"] = torch.empty(max(a.numel() for a in acts.values()), dtype=torch.int16, device=self.device)
            part = 0
            for name, ins, co in LAYERS:
                h, w = dims[LEVEL[name]]
                nz = _lib.fn("glide_unet_wgrad_chunks")(N, h, w)
                part = max(part, nz * 9 * co * SHAPES[name][1], _lib.fn("glide_unet_bias_chunks")(N * h * w) * co)
            nk = N_OUT * CH[HEAD_IN] + N_OUT
            part = max(part, _lib.fn("glide_unet_bias_chunks")(N * H * W) * nk,
                       _lib.fn("glide_unet_mse_blocks")(N * H * W * N_OUT))
            b["partial"] = torch.empty(part, dtype=torch.float32, device=self.device)
            b["dpred"] = torch.empty(N * H * W * N_OUT, dtype=torch.float32, device=self.device)
            b["head"] = torch.empty(nk, dtype=torch.float32, device=self.device)
        self._bufs[key] = b
        return b

    def release(self):
        self._bufs.clear()

    # ------------------------------------------------------------ forward / backward
    def forward(self, x0: torch.Tensor, N: int, H: int, W: int, train: bool = False) -> torch.Tensor:
        """x0: N x H x W x 32 BF16 bits (int16 tensor).  Returns N x H x W x 3 FP32."""
        b = self._buffers(N, H, W, train)
        acts, dims, st = b["acts"], b["dims"], self._stream()
        acts["x0"] = x0
        for name, ins, co in LAYERS:
            for i in ins:
                if i in POOLS:
                    src = POOLS[i]
                    h, w = dims[LEVEL[src]]
                    _lib.call("glide_unet_pool2", _p(acts[src]), N, h, w, CH[src], _p(acts[i]), st)
                if i in UPS:
                    src = UPS[i]
                    h, w = dims[LEVEL[src]]
                    _lib.call("glide_unet_up2", _p(acts[src]), N, h, w, CH[src], _p(acts[i]), st)
            h, w = dims[LEVEL[name]]
            xa, xb = acts[ins[0]], (acts[ins[1]] if len(ins) > 1 else None)
            _lib.call("glide_unet_conv3x3", _p(xa), CH[ins[0]], _p(xb), CH[ins[1]] if xb is not None else 0, N, h, w,
                      _p(self.wb[name]), co, _p(self.params[name + ".b"]), 1, _p(acts[name]), st)
        _lib.call("glide_unet_head", _p(acts[HEAD_IN]), N * H * W, CH[HEAD_IN], N_OUT, _p(self.params["head.W"]),
                  _p(self.params["head.b"]), _p(b["pred"]), st)
        return b["pred"]

    def loss_and_grad(self, target: torch.Tensor, N: int, H: int, W: int) -> torch.Tensor:
        """MSE of the last forward's prediction; d(loss)/d(pred) into the
        training buffers.  Returns the loss as a 0-d device tensor."""
        b = self._bufs[(N, H, W, True)]
        n = N * H * W * N_OUT
        nb = _lib.fn("glide_unet_mse_blocks")(n)
        _lib.call("glide_unet_mse", _p(b["pred"]), _p(target), None, n, ctypes.c_float(1.0 / n), _p(b["dpred"]),
                  _p(b["partial"]), self._stream())
        return b["partial"][:nb].sum() / n

    def backward(self, N: int, H: int, W: int, dpred: torch.Tensor | None = None):
        """Gradients of every parameter (self.grads) from d(loss)/d(pred)."""
        b = self._bufs[(N, H, W, True)]
        acts, g, dims, st = b["acts"], b["g"], b["dims"], self._stream()
        dpred = b["dpred"] if dpred is None else dpred
        P0 = N * H * W
        C = CH[HEAD_IN]
        _lib.call("glide_unet_head_bwd", _p(dpred), _p(acts[HEAD_IN]), P0, C, N_OUT, _p(self.params["head.W"]),
                  _p(g[HEAD_IN]), _p(b["partial"]), _p(b["head"]), st)
        self.grads["head.W"].copy_(b["head"][:N_OUT * C].view(N_OUT, C))
        self.grads["head.b"].copy_(b["head"][N_OUT * C:])
        dz = b["dz"]
        for name, ins, co in reversed(LAYERS):
            h, w = dims[LEVEL[name]]
            n = N * h * w * co
            _lib.call("glide_unet_relu_bwd", _p(g[name]), _p(acts[name]), n, _p(dz), st)
            xa, xb = acts[ins[0]], (acts[ins[1]] if len(ins) > 1 else None)
            ca, cb = CH[ins[0]], (CH[ins[1]] if xb is not None else 0)
            _lib.call("glide_unet_conv3x3_wgrad", _p(dz), co, _p(xa), ca, _p(xb), cb, N, h, w, _p(b["partial"]),
                      _p(self.grads[name + ".W"]), st)
            _lib.call("glide_unet_bias_grad", _p(dz), N * h * w, co, _p(b["partial"]), _p(self.grads[name + ".b"]), st)
            if ins == ["x0"]:
                continue
            gb = g[ins[1]] if xb is not None else None
            _lib.call("glide_unet_conv3x3_dgrad", _p(dz), co, N, h, w, _p(self.wf[name]), ca, _p(g[ins[0]]), 0, cb,
                      _p(gb), 0, st)
            for i in ins:
                if i in UPS:
                    src = UPS[i]
                    hs, ws = dims[LEVEL[src]]
                    _lib.call("glide_unet_up2_bwd", _p(g[i]), N, hs, ws, CH[src], _p(g[src]), st)
                if i in POOLS:
                    src = POOLS[i]
                    hs, ws = dims[LEVEL[src]]
                    _lib.call("glide_unet_pool2_bwd", _p(g[i]), N, hs, ws, CH[src], _p(g[src]), 1, st)

    def adam_step(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8, weight_decay: float = 0.0):
        self.step_count += 1
        st = self._stream()
        for k, w in self.params.items():
            _lib.call("glide_unet_adam", _p(w), _p(self.grads[k]), _p(self.m[k]), _p(self.v[k]), w.numel(),
                      ctypes.c_float(lr), ctypes.c_float(b1), ctypes.c_float(b2), ctypes.c_float(eps),
                      self.step_count, ctypes.c_float(weight_decay), st)
        self.sync_weights()

    # ------------------------------------------------------------ persistence
    def state_dict(self) -> dict:
        return {k: v.detach().cpu().numpy().copy() for k, v in self.params.items()}

    def load_state_dict(self, sd: dict):
        for k, v in self.params.items():
            a = np.asarray(sd[k], dtype=np.float32)
            if a.shape != tuple(v.shape):
                raise ValueError(f"{k}: shape {a.shape} != {tuple(v.shape)}")
            v.copy_(torch.from_numpy(a))
        self.sync_weights()

    def save(self, path: str, **meta):
        np.savez(path, **self.state_dict(), **{"meta." + k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path: str, device=None) -> "UNet":
        with np.load(path, allow_pickle=False) as z:
            net = cls(device)
            net.load_state_dict({k: z[k] for k in z.files if not k.startswith("meta.")})
            net.meta = {k[5:]: z[k] for k in z.files if k.startswith("meta.")}
        return net


# ---------------------------------------------------------------- solver side
def features(physics, out: torch.Tensor | None = None) -> torch.Tensor:
    """Network input of the finest level of ``physics`` (1 x ny x nx x 32 BF16)."""
    g = physics.grid
    x = out if out is not None else torch.empty(g.ny * g.nx * IN_CH, dtype=torch.int16, device=g.device)
    st = ctypes.c_void_p(torch.cuda.current_stream(g.device).cuda_stream)
    _lib.call("glide_unet_features", g.H_prev.data_ptr(), g.bed.data_ptr(), g.beta.data_ptr(), g.B.data_ptr(), g.ny,
              g.nx, g.dx, _p(x), st)
    return x


def targets(physics, out: torch.Tensor | None = None, scale: float = VEL_SCALE,
            hscale: float = DH_SCALE) -> torch.Tensor:
    """Training targets of a solved step: cell-centred (u, v) x scale and
    (H - H_prev) x hscale, 3 FP32 values per cell."""
    g = physics.grid
    t = out if out is not None else torch.empty(g.ny * g.nx * N_OUT, dtype=torch.float32, device=g.device)
    st = ctypes.c_void_p(torch.cuda.current_stream(g.device).cuda_stream)
    _lib.call("glide_unet_targets", g.u.data_ptr(), g.v.data_ptr(), g.H.data_ptr(), g.H_prev.data_ptr(), g.ny, g.nx,
              float(scale), float(hscale), _p(t), st)
    return t


def apply_prediction(physics, pred: torch.Tensor, scale: float = VEL_SCALE, hscale: float = DH_SCALE):
    """Write a prediction (ny x nx x 3) into the finest level: u, v faces and
    H = max(H_prev + dH, gamma)."""
    g = physics.grid
    st = ctypes.c_void_p(torch.cuda.current_stream(g.device).cuda_stream)
    _lib.call("glide_unet_apply_prediction", _p(pred), g.ny, g.nx, float(scale), float(hscale), g.u.data_ptr(),
              g.v.data_ptr(), g.H.data_ptr(), g.H_prev.data_ptr(), g.gamma.data_ptr(), st)


def net_scales(net: UNet | None) -> tuple[float, float]:
    """(velocity, thickness-change) scales of a network's output units: the
    values stored with its weights, else VEL_SCALE and DH_SCALE."""
    meta = getattr(net, "meta", None) or {}
    vs = float(meta["vel_scale"]) if "vel_scale" in meta else VEL_SCALE
    hs = float(meta["dh_scale"]) if "dh_scale" in meta else DH_SCALE
    return vs, hs


class WarmStart:
    """Seeds IcePhysics forward steps with the network's (u, v) and, with
    ``seed_thickness``, its H = max(H_prev + dH, gamma) (else H = H_prev).
    The output scales default to those stored with the network's weights."""

    def __init__(self, net: UNet, scale: float | None = None, hscale: float | None = None,
                 seed_thickness: bool = True):
        vs, hs = net_scales(net)
        self.net = net
        self.scale = vs if scale is None else float(scale)
        self.hscale = hs if hscale is None else float(hscale)
        self.seed_thickness = bool(seed_thickness)

    def predict_into(self, physics):
        g = physics.grid
        x = features(physics)
        pred = self.net.forward(x, 1, g.ny, g.nx)
        apply_prediction(physics, pred, self.scale, self.hscale)
        if not self.seed_thickness:
            g.H.copy_(g.H_prev)

    def solve(self, physics, dt: float, tol: float = 1e-8, max_vcycles: int = 200, use_net: bool = True) -> dict:
        """One forward step (H starts at H_prev) to ||r|| <= tol * r0_cold.
        The cold start is the state's current (u, v) with H = H_prev."""
        g = physics.grid
        for lv in physics.grids:
            lv.dt = float(dt)
        u0, v0 = g.u.clone(), g.v.clone()
        g.H.copy_(g.H_prev)
        _lib.call("glide_forward_setup", physics._levels(), physics.n_levels, float(dt), physics._stream())
        r_cold = physics.residual_norm()
        out = {"r0_cold": r_cold, "r0_warm": None, "fallback": False, "used_net": False}
        if use_net:
            self.predict_into(physics)
            r_warm = physics.residual_norm()
            out["r0_warm"] = r_warm
            if not (r_warm <= r_cold):
                g.u.copy_(u0)
                g.v.copy_(v0)
                g.H.copy_(g.H_prev)
                out["fallback"] = True
            else:
                out["used_net"] = True
        target = tol * r_cold
        hist = []
        k = 0
        r = physics.residual_norm()
        while r > target and k < max_vcycles:
            physics.vcycle()
            k += 1
            r = physics.residual_norm()
            hist.append(r / r_cold)
        out.update(vcycles=k, converged=r <= target, history=hist)
        return out
