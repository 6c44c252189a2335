"""Adjoint inversion of beta at 512 x 512, dx = 1500 m (FP64).

Observations come from a forward solve with a seeded truth beta plus 2%
Gaussian noise; the misfit is L2 plus 1e-3 * Tikhonov on log beta; the
gradient is the hand-written adjoint behind torch.autograd.Function.  The
geometry is the ice-covered grounded slab (glide_amd.synthetic.ice_slab): on
the margin workloads plain FASCD stalls before 1e-8, and a gradient check
needs converged solves.
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
from glide_amd.physics import IcePhysics, tikhonov_regularization  # noqa: E402

DEV = torch.device("cuda")
N = 512


@pytest.fixture(scope="module")
def problem():
    w = S.ice_slab(N, N, 1500.0, seed=2)
    truth = S.ice_slab(N, N, 1500.0, seed=3)["beta"]
    # gl_derivatives off: the slab is grounded everywhere, so the grounding
    # sigmoid is clamped flat and the linearisation without its derivative is exact
    ph = IcePhysics(N, N, w["dx"], n_levels=5, gl_derivatives=False)
    ph.set_geometry(w["bed"], w["H"])
    ph.set_parameters(B=w["B"], beta=truth, smb=w["smb"])
    ph.forward(10.0, tol=1e-12, max_vcycles=60, update_geometry=False)
    rng = np.random.default_rng(3)
    u_obs = ph.grid.u.cpu().numpy() * (1 + 0.02 * rng.standard_normal((N, N + 1)))
    v_obs = ph.grid.v.cpu().numpy() * (1 + 0.02 * rng.standard_normal((N + 1, N)))
    return w, ph, u_obs, v_obs, truth


@pytest.mark.timeout(900)
def test_gradient_fd_20_directions_512(problem):
    w, ph, u_obs, v_obs, truth = problem
    cfg = SolveConfig(dt=10.0, tol=1e-13, max_vcycles=80, adj_tol=1e-13, max_adj_vcycles=200)
    uo = torch.as_tensor(u_obs, device=DEV)
    vo = torch.as_tensor(v_obs, device=DEV)
    Hp = torch.as_tensor(w["H"], device=DEV)
    alpha = 1e-3

    def J(log_beta):
        u, v, _ = ssa_step(ph, torch.exp(log_beta), Hp, cfg)
        reg, _ = tikhonov_regularization(log_beta.detach())
        return 0.5 * ((u - uo) ** 2).sum() + 0.5 * ((v - vo) ** 2).sum() + alpha * reg

    lb = torch.log(torch.as_tensor(w["beta"], device=DEV)).clone().requires_grad_(True)
    L = J(lb)
    L.backward()
    _, reg_grad = tikhonov_regularization(lb.detach())
    g = lb.grad.detach() + alpha * reg_grad
    rng = np.random.default_rng(4)
    worst = 0.0
    with torch.no_grad():
        for k in range(20):
            e = torch.as_tensor(rng.normal(size=(N, N)), device=DEV)
            h = 1e-4
            f = lambda s: float(J(lb.detach() + s * e))  # noqa: E731
            fd = (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)
            ad = float((g * e).sum())
            rel = abs(fd - ad) / abs(fd)
            worst = max(worst, rel)
            print(f"direction {k}: adjoint {ad:.10e} fd {fd:.10e} rel {rel:.2e}", flush=True)
    assert worst <= 1e-6, worst


@pytest.mark.timeout(900)
def test_lbfgs_50_iterations_512(problem):
    w, ph, u_obs, v_obs, truth = problem
    inv = BetaInversion(ph, u_obs, v_obs, w["H"], dt=10.0, alpha=1e-3, mode="l2", fwd_tol=1e-10, adj_tol=1e-10)
    x0 = np.full(N * N, np.log(float(np.mean(truth))))
    res = inv.run(x0, maxiter=50)
    m0, m1 = inv.history[0]["misfit"], inv.history[-1]["misfit"]
    err0 = np.linalg.norm(x0 - np.log(truth).ravel())
    err1 = np.linalg.norm(res.x - np.log(truth).ravel())
    print(f"misfit {m0:.4e} -> {m1:.4e} in {res.nit} iterations ({len(inv.history)} evaluations); "
          f"|log beta - truth| {err0:.3e} -> {err1:.3e}", flush=True)
    assert m1 < 0.01 * m0
    assert err1 < err0
