"""Train the warm-start network on solves the solver itself generates.

    python -m glide_amd.train_warmstart --n-train 2000 --out glide_amd/data/warmstart_unet.npz

1. Data: ``n_train`` seeded ice sheets (glide_amd.synthetic.ice_sheet, seeds
   10000...), ``size`` x ``size`` cells, dx alternating 1000 / 1500 m, one
   forward step of ``dt`` (1 a) each, solved by batched FASCD (Ensemble, ``batch`` members per plan) from a cold
   start (u = v = 0) to ``tol`` relative residual; the network input is the
   member's cell fields (glide_unet_features) and the target its converged,
   cell-centred (u, v) and thickness change H - H_prev; solves that do not reach ``tol`` in ``max_vcycles``
   are left out.  Everything stays in HBM.
2. Training: Adam on the GPU (BF16 MFMA forward and backward kernels, FP32
   master weights), cosine learning-rate decay, 5% of the samples held out.
3. Evaluation: on test seeds (20000...) the number of V-cycles a forward step
   needs to reach ``tol`` times the cold start's initial residual, cold vs
   warm-started, and the prediction's error against the cold solve.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

import numpy as np
import torch

from . import synthetic as S
from .ensemble import Ensemble
from .physics import TOLERANCE_SCHEDULE, IcePhysics
from .warmstart import IN_CH, N_OUT, VEL_SCALE, UNet, WarmStart, features, targets

TRAIN_SEED0 = 10000
TEST_SEED0 = 20000


def dx_of(seed: int) -> float:
    return 1000.0 if seed % 2 == 0 else 1500.0


def member(seed: int, size: int, dx: float | None = None, n_levels: int | None = None, device=None) -> IcePhysics:
    dx = dx_of(seed) if dx is None else dx
    d = S.ice_sheet(size, size, dx, seed)
    ph = IcePhysics(size, size, dx, n_levels=n_levels or S.max_levels(size, size), device=device)
    ph.set_geometry(d["bed"], d["H"])
    ph.set_parameters(B=d["B"], beta=d["beta"], smb=d["smb"])
    return ph


def generate(n: int, size: int, seed0: int = TRAIN_SEED0, dt: float = 1.0, tol: float = 1e-8, batch: int = 64,
             max_vcycles: int = 80, device=None, log=print, drop_unconverged: bool = True,
             hscale: float = 1.0, sched: dict | None = None):
    """Solve n seeded sheets (batched FASCD with the tolerance schedule);
    returns (X int16 [m, size*size*32], Y fp32 [m, size*size*3], stats),
    keeping only the m solves that reached ``tol`` unless ``drop_unconverged``
    is False.  Targets: cell-centred (u, v) x VEL_SCALE, (H - H_prev) x hscale."""
    sched = TOLERANCE_SCHEDULE if sched is None else sched
    dev = torch.device(device if device is not None else "cuda")
    X = torch.empty((n, size * size * IN_CH), dtype=torch.int16, device=dev)
    Y = torch.empty((n, size * size * N_OUT), dtype=torch.float32, device=dev)
    vcycles, conv = [], []
    t0 = time.perf_counter()
    for c0 in range(0, n, batch):
        seeds = list(range(seed0 + c0, seed0 + min(n, c0 + batch)))
        mem = [member(s, size, device=dev) for s in seeds]
        ens = Ensemble(mem)
        ens.set_schedule(**sched)
        st = ens.forward(dt, tol=tol, max_vcycles=max_vcycles, update_geometry=False)
        for k, ph in enumerate(mem):
            features(ph, out=X[c0 + k])
            targets(ph, out=Y[c0 + k], hscale=hscale)
        vcycles += st.vcycles
        conv += st.converged
        ens.close()
        del mem, ens
        log(f"  solved {min(n, c0 + batch)}/{n} ({time.perf_counter() - t0:.1f} s; "
            f"V-cycles {min(st.vcycles)}-{max(st.vcycles)})")
    torch.cuda.synchronize(dev)
    stats = {"solve_s": time.perf_counter() - t0, "vcycles_mean": float(np.mean(vcycles)),
             "vcycles_max": int(np.max(vcycles)), "converged": int(np.sum(conv)), "n": n}
    if drop_unconverged and not all(conv):
        keep = torch.as_tensor(np.flatnonzero(np.asarray(conv, dtype=bool)), device=dev)
        X, Y = X.index_select(0, keep), Y.index_select(0, keep)
        stats["kept"] = int(keep.numel())
    return X, Y, stats


def train(net: UNet, X: torch.Tensor, Y: torch.Tensor, size: int, epochs: int = 12, batch: int = 16,
          lr: float = 2e-3, val_frac: float = 0.05, seed: int = 0, log=print) -> dict:
    n = X.shape[0]
    n_val = max(1, int(round(val_frac * n)))
    n_tr = n - n_val
    steps_per_epoch = max(1, n_tr // batch)
    total = epochs * steps_per_epoch
    dev = X.device
    xb = torch.empty((batch, X.shape[1]), dtype=X.dtype, device=dev)
    yb = torch.empty((batch, Y.shape[1]), dtype=Y.dtype, device=dev)
    gen = torch.Generator().manual_seed(seed)
    hist = []
    t0 = time.perf_counter()
    step = 0
    for ep in range(epochs):
        perm = torch.randperm(n_tr, generator=gen).to(dev)
        acc = torch.zeros((), dtype=torch.float32, device=dev)
        for s in range(steps_per_epoch):
            idx = perm[s * batch:(s + 1) * batch]
            torch.index_select(X, 0, idx, out=xb)
            torch.index_select(Y, 0, idx, out=yb)
            cur = lr * 0.5 * (1.0 + math.cos(math.pi * step / total))
            net.forward(xb, batch, size, size, train=True)
            acc += net.loss_and_grad(yb, batch, size, size)
            net.backward(batch, size, size)
            net.adam_step(cur)
            step += 1
        tr = float(acc) / steps_per_epoch
        va = evaluate_loss(net, X[n_tr:], Y[n_tr:], size, batch)
        hist.append({"epoch": ep, "train_loss": tr, "val_loss": va, "t": time.perf_counter() - t0})
        log(f"  epoch {ep}: train {tr:.4e} val {va:.4e} ({time.perf_counter() - t0:.1f} s)")
    net.release()
    return {"history": hist, "steps": step, "train_s": time.perf_counter() - t0, "n_train": n_tr, "n_val": n_val}


def evaluate_loss(net: UNet, X, Y, size, batch):
    tot, cnt = 0.0, 0
    for s in range(0, X.shape[0], batch):
        xb, yb = X[s:s + batch], Y[s:s + batch]
        b = xb.shape[0]
        net.forward(xb.contiguous(), b, size, size, train=True)
        tot += float(net.loss_and_grad(yb.contiguous(), b, size, size)) * b
        cnt += b
    return tot / max(cnt, 1)


def evaluate(net: UNet, seeds, size: int, dx: float | None = None, n_levels: int | None = None, dt: float = 1.0,
             tol: float = 1e-8, max_vcycles: int = 200, device=None, sched: dict | None = None) -> list[dict]:
    """Cold vs warm-started forward step on each seed (tolerance schedule):
    V-cycles and wall time to ||r|| <= tol * r0_cold, and the prediction's
    error vs the cold solve."""
    sched = TOLERANCE_SCHEDULE if sched is None else sched
    ws = WarmStart(net)
    out = []
    for seed in seeds:
        ph = member(seed, size, dx=dx, n_levels=n_levels, device=device)
        ph.set_schedule(**sched)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cold = ws.solve(ph, dt, tol=tol, max_vcycles=max_vcycles, use_net=False)
        torch.cuda.synchronize()
        t_cold = time.perf_counter() - t0
        u_ref, v_ref, H_ref = ph.grid.u.clone(), ph.grid.v.clone(), ph.grid.H.clone()
        ph.grid.u.zero_()
        ph.grid.v.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        warm = ws.solve(ph, dt, tol=tol, max_vcycles=max_vcycles, use_net=True)
        torch.cuda.synchronize()
        t_warm = time.perf_counter() - t0
        u_w, v_w = ph.grid.u.clone(), ph.grid.v.clone()
        # prediction error (before any V-cycle) vs the converged cold solve
        ph.grid.u.zero_()
        ph.grid.v.zero_()
        ph.grid.H.copy_(ph.grid.H_prev)
        ws.predict_into(ph)
        num = torch.sqrt(((ph.grid.u - u_ref) ** 2).sum() + ((ph.grid.v - v_ref) ** 2).sum())
        den = torch.sqrt((u_ref ** 2).sum() + (v_ref ** 2).sum())
        dH_ref = H_ref - ph.grid.H_prev
        dH_err = float(torch.linalg.norm(ph.grid.H - H_ref) / torch.linalg.norm(dH_ref))
        diff = torch.sqrt(((u_w - u_ref) ** 2).sum() + ((v_w - v_ref) ** 2).sum())
        out.append({"seed": seed, "size": size, "dx": ph.dx, "levels": ph.n_levels,
                    "vcycles_cold": cold["vcycles"], "vcycles_warm": warm["vcycles"],
                    "converged_cold": bool(cold["converged"]), "converged_warm": bool(warm["converged"]),
                    "ms_cold": t_cold * 1e3, "ms_warm": t_warm * 1e3,
                    "r0_warm_over_cold": (warm["r0_warm"] / cold["r0_cold"]) if warm["r0_warm"] is not None else None,
                    "fallback": warm["fallback"],
                    "prediction_rel_error": float(num / den), "thickness_change_rel_error": dH_err,
                    "solution_rel_diff_warm_vs_cold": float(diff / den)})
        del ph
        torch.cuda.empty_cache()
    return out


def summarize(rows: list[dict]) -> dict:
    c = np.array([r["vcycles_cold"] for r in rows], dtype=float)
    w = np.array([r["vcycles_warm"] for r in rows], dtype=float)
    return {"n": len(rows), "vcycles_cold_mean": float(c.mean()), "vcycles_warm_mean": float(w.mean()),
            "vcycles_saved_mean": float((c - w).mean()),
            "ms_cold_mean": float(np.mean([r["ms_cold"] for r in rows])),
            "ms_warm_mean": float(np.mean([r["ms_warm"] for r in rows])),
            "prediction_rel_error_mean": float(np.mean([r["prediction_rel_error"] for r in rows])),
            "fallbacks": int(sum(r["fallback"] for r in rows)),
            "all_converged": bool(all(r["converged_cold"] and r["converged_warm"] for r in rows))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-train", type=int, default=2000)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--dt", type=float, default=1.0, help="time step of the solved forward steps (a)")
    ap.add_argument("--dh-scale", type=float, nargs="+", default=[0.05],
                    help="network units of the thickness change per m; several values train one network each "
                         "on the same solves (written as <out>.dh<scale>.npz)")
    ap.add_argument("--n-test", type=int, default=4)
    ap.add_argument("--test-size", type=int, nargs="*", default=[256, 512, 1024])
    ap.add_argument("--out", default="warmstart_unet.npz")
    ap.add_argument("--report", default=None)
    args = ap.parse_args()

    def log(msg):
        print(msg, flush=True)

    log(f"== generating {args.n_train} solves at {args.size}^2")
    X, Y, gen = generate(args.n_train, args.size, tol=args.tol, dt=args.dt, log=log, hscale=1.0)
    log(f"   {gen}")
    report = {"generate": gen}
    for hs in args.dh_scale:
        tag = "" if len(args.dh_scale) == 1 else f".dh{hs:g}"
        Yh = Y.clone()
        Yh.view(Yh.shape[0], -1, N_OUT)[:, :, 2] *= hs     # targets were made with 1 unit per m
        net = UNet(seed=0)
        log(f"== training, thickness-change scale {hs:g} ({args.epochs} epochs, batch {args.batch})")
        tr = train(net, X, Yh, args.size, epochs=args.epochs, batch=args.batch, lr=args.lr, log=log)
        del Yh
        torch.cuda.empty_cache()
        out = args.out[:-4] + tag + ".npz" if args.out.endswith(".npz") else args.out + tag
        net.save(out, vel_scale=VEL_SCALE, dh_scale=hs, dt=args.dt, train_size=args.size, n_train=args.n_train,
                 epochs=args.epochs, final_val_loss=tr["history"][-1]["val_loss"])
        net.meta = {"vel_scale": VEL_SCALE, "dh_scale": hs}
        rep_h = {"train": tr, "weights": os.path.basename(out)}
        for size in args.test_size:
            levels = 6 if size == 1024 else (5 if size == 512 else None)
            dx = 1000.0 if size == 1024 else (1500.0 if size == 512 else None)
            log(f"== evaluating at {size}^2")
            rows = evaluate(net, range(TEST_SEED0, TEST_SEED0 + args.n_test), size, dx=dx, n_levels=levels,
                            tol=args.tol, dt=args.dt)
            for r in rows:
                log(f"   {r}")
            rep_h[f"eval_{size}"] = {"rows": rows, "summary": summarize(rows)}
            log(f"   summary {rep_h[f'eval_{size}']['summary']}")
        report[f"dh_scale_{hs:g}"] = rep_h
        net.release()
    del X, Y
    torch.cuda.empty_cache()
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)


if __name__ == "__main__":
    main()
