"""The hand-written adjoint behind torch.autograd.Function against central
finite differences of the converged forward solve, and the L-BFGS inversion.

The check runs on an ice-covered grounded slab (glide_amd.synthetic.ice_slab)
where FASCD converges to ~1e-13 relative residual: finite differences of a
solver that stops short of convergence would measure the solver, not the
derivative.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from glide_amd import synthetic as S  # noqa: E402
from glide_amd.autograd import SolveConfig, ssa_step  # noqa: E402
from glide_amd.inversion import BetaInversion  # noqa: E402
from glide_amd.physics import IcePhysics  # noqa: E402

DEV = torch.device("cuda")


def _setup(n=32, levels=3, seed=2, gl=True):
    w = S.ice_slab(n, n, 2000.0, seed=seed)
    ph = IcePhysics(n, n, w["dx"], n_levels=levels, gl_derivatives=gl)
    ph.set_geometry(w["bed"], w["H"])
    ph.set_parameters(B=w["B"], beta=w["beta"], smb=w["smb"])
    # the checks need converged solves, not the reference's sweep counts:
    # 20 coarsest sweeps (forward and adjoint) converge as fast, for less time
    ph.set_schedule(coarse=20, adj_coarse=20)
    return w, ph


def test_step_converges_on_slab():
    w, ph = _setup()
    ph.forward(10.0, tol=1e-13, max_vcycles=80, update_geometry=False)
    print("history", ph.last_history[:3], "...", ph.last_history[-1], len(ph.last_history), flush=True)
    assert ph.last_history[-1] <= 1e-13


def test_beta_and_Hprev_gradients_match_finite_differences():
    # gl_derivatives off: on this grounded slab every cell's grounding sigmoid
    # is clamped flat, which is exactly what the linearisation without the
    # sigmoid derivative describes
    w, ph = _setup(gl=False)
    cfg = SolveConfig(dt=10.0, tol=1e-14, max_vcycles=120, adj_tol=1e-14, max_adj_vcycles=400)
    rng = np.random.default_rng(7)
    ny = nx = 32
    u_obs = torch.as_tensor(rng.normal(0, 30.0, (ny, nx + 1)), device=DEV)
    v_obs = torch.as_tensor(rng.normal(0, 30.0, (ny + 1, nx)), device=DEV)
    wH = torch.as_tensor(rng.normal(0, 1.0, (ny, nx)), device=DEV)

    def loss_of(beta, Hp):
        u, v, H = ssa_step(ph, beta, Hp, cfg)
        return 0.5 * ((u - u_obs) ** 2).sum() + 0.5 * ((v - v_obs) ** 2).sum() + (wH * H).sum()

    beta = torch.as_tensor(w["beta"], device=DEV).clone().requires_grad_(True)
    Hp = torch.as_tensor(w["H"], device=DEV).clone().requires_grad_(True)
    L = loss_of(beta, Hp)
    L.backward()
    gb, gH = beta.grad.detach().clone(), Hp.grad.detach().clone()
    assert ph.last_adj_history[-1] <= 1e-13
    worst = 0.0
    with torch.no_grad():
        for k in range(20):
            e_b = torch.as_tensor(rng.normal(size=(ny, nx)), device=DEV) * beta.detach()
            e_h = torch.as_tensor(rng.normal(size=(ny, nx)), device=DEV) * (10.0 if k % 2 else 0.0)
            # fourth-order central difference: truncation O(h^4)
            h = 1e-4   # small enough not to cross upwinding kinks (|u| at faces with u near 0)
            L1p = float(loss_of(beta.detach() + h * e_b, Hp.detach() + h * e_h))
            L1m = float(loss_of(beta.detach() - h * e_b, Hp.detach() - h * e_h))
            L2p = float(loss_of(beta.detach() + 2 * h * e_b, Hp.detach() + 2 * h * e_h))
            L2m = float(loss_of(beta.detach() - 2 * h * e_b, Hp.detach() - 2 * h * e_h))
            fd = (8.0 * (L1p - L1m) - (L2p - L2m)) / (12.0 * h)
            ad = float((gb * e_b).sum() + (gH * e_h).sum())
            rel = abs(fd - ad) / abs(fd)
            worst = max(worst, rel)
            print(f"direction {k}: adjoint {ad:.10e} fd {fd:.10e} rel {rel:.2e}", flush=True)
    assert worst <= 1e-6, worst


def test_B_and_smb_gradients_match_finite_differences():
    """dL/dB (through the viscosity) and dL/dsmb (= -lambda_H) of the same
    loss against fourth-order central differences of the converged solve,
    20 random directions."""
    w, ph = _setup(gl=False)
    cfg = SolveConfig(dt=10.0, tol=1e-14, max_vcycles=120, adj_tol=1e-14, max_adj_vcycles=400)
    rng = np.random.default_rng(11)
    ny = nx = 32
    u_obs = torch.as_tensor(rng.normal(0, 30.0, (ny, nx + 1)), device=DEV)
    v_obs = torch.as_tensor(rng.normal(0, 30.0, (ny + 1, nx)), device=DEV)
    wH = torch.as_tensor(rng.normal(0, 1.0, (ny, nx)), device=DEV)
    beta = torch.as_tensor(w["beta"], device=DEV)
    Hp = torch.as_tensor(w["H"], device=DEV)

    def loss_of(B, smb):
        u, v, H = ssa_step(ph, beta, Hp, cfg, B=B, smb=smb)
        return 0.5 * ((u - u_obs) ** 2).sum() + 0.5 * ((v - v_obs) ** 2).sum() + (wH * H).sum()

    B = torch.as_tensor(w["B"], device=DEV).clone().requires_grad_(True)
    smb = torch.as_tensor(w["smb"], device=DEV).clone().requires_grad_(True)
    L = loss_of(B, smb)
    L.backward()
    gB, gs = B.grad.detach().clone(), smb.grad.detach().clone()
    assert ph.last_adj_history[-1] <= 1e-13
    worst = 0.0
    with torch.no_grad():
        for k in range(20):
            e_B = torch.as_tensor(rng.normal(size=(ny, nx)), device=DEV) * B.detach() * 0.1
            e_s = torch.as_tensor(rng.normal(size=(ny, nx)), device=DEV) * (1.0 if k % 2 else 0.0)
            h = 1e-4
            Ls = [float(loss_of(B.detach() + c * h * e_B, smb.detach() + c * h * e_s)) for c in (1, -1, 2, -2)]
            fd = (8.0 * (Ls[0] - Ls[1]) - (Ls[2] - Ls[3])) / (12.0 * h)
            ad = float((gB * e_B).sum() + (gs * e_s).sum())
            rel = abs(fd - ad) / abs(fd)
            worst = max(worst, rel)
            print(f"direction {k}: adjoint {ad:.10e} fd {fd:.10e} rel {rel:.2e}", flush=True)
    assert worst <= 1e-6, worst


def test_inversion_reduces_misfit():
    n = 32
    w, ph = _setup(n=n, gl=False)
    H_prev = w["H"]
    beta_true = torch.as_tensor(w["beta"], device=DEV)
    ph.set_parameters(beta=beta_true)
    ph.forward(10.0, tol=1e-12, max_vcycles=80, update_geometry=False)
    rng = np.random.default_rng(3)
    u_obs = ph.grid.u.cpu().numpy() * (1 + 0.02 * rng.standard_normal((n, n + 1)))
    v_obs = ph.grid.v.cpu().numpy() * (1 + 0.02 * rng.standard_normal((n + 1, n)))
    inv = BetaInversion(ph, u_obs, v_obs, H_prev, dt=10.0, alpha=1e-3, mode="l2", fwd_tol=1e-10, adj_tol=1e-10)
    x0 = np.full(n * n, np.log(float(beta_true.mean())))
    res = inv.run(x0, maxiter=15)
    m0, m1 = inv.history[0]["misfit"], inv.history[-1]["misfit"]
    print("misfit", m0, "->", m1, "iterations", res.nit, flush=True)
    assert m1 < 0.05 * m0


def test_tolerance_schedule_gives_the_reference_sweeps_gradient():
    """The inversion's converged gradient does not depend on the coarsest
    levels' sweep counts: forward to 1e-12 and adjoint to 1e-12 on a 128^2
    slab with the tolerance schedule (5 forward, 10 adjoint coarsest sweeps,
    boundary weight 0.5) and with the reference's counts (200, 100, weight 1)
    give the same misfit gradient to 1e-9, with the adjoint taking the same
    number of V-cycles."""
    from glide_amd.physics import TOLERANCE_SCHEDULE

    n = 128
    w = S.ice_slab(n, n, 1500.0, seed=2)
    truth = S.ice_slab(n, n, 1500.0, seed=3)["beta"]
    ph = IcePhysics(n, n, w["dx"], n_levels=4, gl_derivatives=False)
    ph.set_geometry(w["bed"], w["H"])
    ph.set_parameters(B=w["B"], beta=truth, smb=w["smb"])
    ph.forward(10.0, tol=1e-12, max_vcycles=80, update_geometry=False)
    rng = np.random.default_rng(3)
    u_obs = ph.grid.u.cpu().numpy() * (1 + 0.02 * rng.standard_normal((n, n + 1)))
    v_obs = ph.grid.v.cpu().numpy() * (1 + 0.02 * rng.standard_normal((n + 1, n)))
    x0 = np.full(n * n, np.log(float(np.mean(truth))))
    out = {}
    for name, sched in (("reference", {}), ("tolerance", TOLERANCE_SCHEDULE)):
        ph.set_schedule(**sched)
        inv = BetaInversion(ph, u_obs, v_obs, w["H"], dt=10.0, alpha=1e-3, mode="l2", fwd_tol=1e-12, adj_tol=1e-12)
        J, g = inv.objective(x0)
        out[name] = (J, g, len(ph.last_history), len(ph.last_adj_history))
        print(name, "forward", out[name][2], "adjoint", out[name][3], "V-cycles", flush=True)
    (Jr, gr, _, ar), (Jt, gt, _, at) = out["reference"], out["tolerance"]
    assert abs(Jt - Jr) <= 1e-10 * abs(Jr)
    assert np.linalg.norm(gt - gr) <= 1e-9 * np.linalg.norm(gr)
    assert at == ar


def test_two_chained_steps_gradient_matches_finite_differences():
    """Two implicit steps chained through autograd: step 2 starts from step
    1's thickness, beta acts in both, and the loss sees only step 2's state.
    The gradient flows back through SSAStep twice (dL/dH_prev of step 2 into
    dL/dH of step 1) and must match 4th-order central differences of the
    twice-solved loss in 6 random beta directions."""
    w, ph = _setup(gl=False)
    cfg = SolveConfig(dt=10.0, tol=1e-14, max_vcycles=120, adj_tol=1e-14, max_adj_vcycles=400)
    rng = np.random.default_rng(11)
    ny = nx = 32
    u_obs = torch.as_tensor(rng.normal(0, 30.0, (ny, nx + 1)), device=DEV)
    v_obs = torch.as_tensor(rng.normal(0, 30.0, (ny + 1, nx)), device=DEV)
    wH = torch.as_tensor(rng.normal(0, 1.0, (ny, nx)), device=DEV)
    H0 = torch.as_tensor(w["H"], device=DEV)

    def loss_of(beta):
        _, _, H1 = ssa_step(ph, beta, H0, cfg)
        u2, v2, H2 = ssa_step(ph, beta, H1, cfg)
        return 0.5 * ((u2 - u_obs) ** 2).sum() + 0.5 * ((v2 - v_obs) ** 2).sum() + (wH * H2).sum()

    beta = torch.as_tensor(w["beta"], device=DEV).clone().requires_grad_(True)
    loss_of(beta).backward()
    gb = beta.grad.detach().clone()
    worst = 0.0
    with torch.no_grad():
        for k in range(6):
            e = torch.as_tensor(rng.normal(size=(ny, nx)), device=DEV) * beta.detach()
            h = 1e-4
            L1p, L1m = float(loss_of(beta.detach() + h * e)), float(loss_of(beta.detach() - h * e))
            L2p, L2m = float(loss_of(beta.detach() + 2 * h * e)), float(loss_of(beta.detach() - 2 * h * e))
            fd = (8.0 * (L1p - L1m) - (L2p - L2m)) / (12.0 * h)
            ad = float((gb * e).sum())
            rel = abs(fd - ad) / abs(fd)
            worst = max(worst, rel)
            print(f"direction {k}: adjoint {ad:.10e} fd {fd:.10e} rel {rel:.2e}", flush=True)
    assert worst <= 1e-6, worst
