"""The ensemble and warm-start workloads at their stated sizes.

- The 64-member batched ensemble (16 each at 256^2, 384 x 512, 512^2 and
  768^2, seeds 1000-1063, dx 1500 m, the reference's forward step,
  glide/physics.py:141-209): one V-cycle of the whole plan in batched
  launches, every member bitwise equal to the same V-cycle solved alone
  (the solo solves run one after another to bound memory).
- The learned warm start at 1024^2 (dx 1000 m, 6 levels, dt = 1 a, the
  tolerance schedule) on four unseen seeds: every solve reaches 1e-8 from
  both starts (stalls are counted, not skipped) and the fallback rule holds.
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from glide_amd import synthetic as S  # noqa: E402
from glide_amd.ensemble import Ensemble  # noqa: E402
from glide_amd.physics import TOLERANCE_SCHEDULE, IcePhysics  # noqa: E402
from glide_amd.train_warmstart import TEST_SEED0, member  # noqa: E402
from glide_amd.warmstart import UNet, WarmStart  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "glide_amd", "data", "warmstart_unet.npz")


def _make(d):
    ph = IcePhysics(d["ny"], d["nx"], d["dx"], n_levels=d["n_levels"])
    ph.set_geometry(d["bed"], d["H"])
    ph.set_parameters(B=d["B"], beta=d["beta"], smb=d["smb"])
    return ph


def test_64_member_plan_one_vcycle_bitwise_equal_to_solo():
    members = S.ensemble_members()
    assert len(members) == 64
    batch = [_make(d) for d in members]
    ens = Ensemble(batch)
    assert ens.n_phases == 6
    ens.forward(10.0, n_vcycles=1, update_geometry=False)
    torch.cuda.synchronize()
    states = [ph.grid.U.cpu().numpy() for ph in batch]
    ens.close()
    del batch, ens
    torch.cuda.empty_cache()
    for k, d in enumerate(members):
        solo = _make(d)
        solo.forward(10.0, n_vcycles=1, update_geometry=False)
        torch.cuda.synchronize()
        s = solo.grid.U.cpu().numpy()
        assert np.array_equal(s, states[k]), f"member {k} ({d['ny']}x{d['nx']}, seed {d['seed']})"
        assert np.all(np.isfinite(s))
        del solo
    print("64 members: batched V-cycle bitwise equal to the solo V-cycles", flush=True)


def test_warm_start_1024_converges_and_keeps_the_fallback_rule():
    net = UNet.load(WEIGHTS)
    ws = WarmStart(net)
    stalled, rows = 0, []
    for k in range(4):
        seed = TEST_SEED0 + 50 + k
        ph = member(seed, 1024, dx=1000.0, n_levels=6)
        ph.set_schedule(**TOLERANCE_SCHEDULE)
        cold = ws.solve(ph, 1.0, tol=1e-8, max_vcycles=100, use_net=False)
        u_cold, v_cold = ph.grid.u.clone(), ph.grid.v.clone()
        ph.grid.u.zero_()
        ph.grid.v.zero_()
        warm = ws.solve(ph, 1.0, tol=1e-8, max_vcycles=100, use_net=True)
        stalled += int(not cold["converged"]) + int(not warm["converged"])
        # the fallback rule: the network's state is kept iff its initial
        # residual is at most the cold start's
        assert warm["used_net"] == (warm["r0_warm"] <= warm["r0_cold"]) and warm["fallback"] == (not warm["used_net"])
        if warm["fallback"]:
            assert warm["vcycles"] == cold["vcycles"]
        den = float(torch.sqrt((u_cold ** 2).sum() + (v_cold ** 2).sum()))
        diff = float(torch.sqrt(((ph.grid.u - u_cold) ** 2).sum() + ((ph.grid.v - v_cold) ** 2).sum())) / den
        rows.append((seed, cold["vcycles"], warm["vcycles"], warm["r0_warm"] / warm["r0_cold"], diff))
        print(f"seed {seed}: cold {cold['vcycles']} V-cycles, warm {warm['vcycles']} (r0 ratio "
              f"{rows[-1][3]:.3f}, fallback {warm['fallback']}), warm vs cold solution {diff:.2e}", flush=True)
        del ph
        torch.cuda.empty_cache()
    assert stalled == 0, rows
    # both solves stop at 1e-8 of the same initial residual: the solutions agree to about that
    assert all(r[4] < 1e-6 for r in rows), rows


def _lin(name, g, out, U, D):
    import ctypes

    from glide_amd import _lib
    P = lambda t: ctypes.c_void_p(t.data_ptr())   # noqa: E731
    _lib.call(name, ctypes.byref(g.geom()), ctypes.byref(g.cparams()), P(out), P(U), P(D), P(g.bed), P(g.B),
              P(g.beta), P(g.mask), P(g.gamma), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


@pytest.mark.parametrize("n,levels,dx", [(512, 5, 1500.0), (1024, 6, 1000.0)])
def test_full_size_linearisation_adjoint_and_norm(n, levels, dx):
    """At the benchmark sizes, after one V-cycle of a test sheet:
    <lam, J d> = <J^T lam, d> to 1e-12 between the staged JVP and VJP
    (gl_derivatives off: exact transposes); the per-point forms give the
    same bits as the 64 x 4 staged tiles with their XCD-chunked order; and
    glide_residual_norm equals the norm of the GPU's own residual computed
    with torch (sqrt(|r_u|^2 + |r_v|^2 + |FB(r_H, H - gamma)|^2)) to 1e-13."""
    from glide_amd import _lib

    ph = member(TEST_SEED0 + 70, n, dx=dx, n_levels=levels)
    ph.set_schedule(**TOLERANCE_SCHEDULE)
    ph.forward(1.0, n_vcycles=1, update_geometry=False)
    g = ph.grid
    gen = torch.Generator(device="cuda").manual_seed(n)
    d = torch.randn(g.U.numel(), dtype=torch.float64, device="cuda", generator=gen)
    lam = torch.randn(g.U.numel(), dtype=torch.float64, device="cuda", generator=gen)
    jv, vt = torch.empty_like(d), torch.empty_like(d)
    _lin("glide_jvp", g, jv, g.U, d)
    _lin("glide_vjp", g, vt, g.U, lam)
    a, b = float((lam * jv).sum()), float((vt * d).sum())
    scale = float((lam.abs() * jv.abs()).sum())
    print(f"{n}^2: <lam, Jd> = {a:.15e}, <J^T lam, d> = {b:.15e}, rel {abs(a - b) / scale:.2e}", flush=True)
    assert abs(a - b) <= 1e-12 * scale
    jv0, vt0 = torch.empty_like(d), torch.empty_like(d)
    try:
        _lib.call("glide_set_tile_height", 1, 0)
        _lin("glide_jvp", g, jv0, g.U, d)
        _lin("glide_vjp", g, vt0, g.U, lam)
        torch.cuda.synchronize()
    finally:
        _lib.call("glide_set_tile_height", 1, 4)
    assert torch.equal(jv, jv0) and torch.equal(vt, vt0)
    got = ph.residual_norm()      # also leaves r = F(U) - f in g.r
    nu, nv = g.nu, g.nv
    rH = g.r[nu + nv:]
    x = g.H.reshape(-1) - g.gamma.reshape(-1)
    fb = rH + x - torch.sqrt(rH * rH + x * x)
    ref = float(torch.sqrt((g.r[:nu] ** 2).sum() + (g.r[nu:nu + nv] ** 2).sum() + (fb ** 2).sum()))
    print(f"{n}^2: residual norm {got:.15e}, torch {ref:.15e}", flush=True)
    assert abs(got - ref) <= 1e-13 * ref
    del ph
    torch.cuda.empty_cache()
