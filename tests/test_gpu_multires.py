"""Multi-resolution inversion of beta (glide_amd.inversion.MultiResolutionInversion,
the coarse-to-fine scheme of examples/greenland/greenland_inverse.py:123-262).

- The observation hierarchy equals the oracle's face restrictions, bitwise.
- One evaluation of the reference objective on a coarse level k (5 FASCD
  V-cycles from u = v = 0, H = H_prev; mean |misfit|; one adjoint V-cycle;
  the reference's edge handling of grad beta; Tikhonov on beta) equals the
  oracle solving level k as a hierarchy of its own, built from the finest
  inputs restricted by the oracle: loss to 1e-12, gradient to 1e-10 (the
  VJP inside the adjoint V-cycle agrees with the oracle to ~1e-12, not
  bitwise).
- A coarse-to-fine twin experiment on a 128^2 grounded slab lowers the
  finest-level misfit.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from glide_amd import synthetic as S  # noqa: E402
from glide_amd.inversion import BetaInversion, MultiResolutionInversion, restrict_observations  # noqa: E402
from glide_amd.physics import IcePhysics, abs_grad, abs_loss, tikhonov_regularization  # noqa: E402
from oracle import oracle as O  # noqa: E402

DX = 1500.0


def make(n, levels, seed=2, truth_seed=3):
    """A grounded slab and observations from a converged solve with another
    seed's beta (the finest level's forcing path leaves every coarse level's
    f holding FAS data, which the level objective must not see)."""
    w = S.ice_slab(n, n, DX, seed=seed)
    truth = S.ice_slab(n, n, DX, seed=truth_seed)["beta"]
    ph = IcePhysics(n, n, DX, n_levels=levels, gl_derivatives=False)
    ph.set_geometry(w["bed"], w["H"])
    ph.set_parameters(B=w["B"], beta=truth, smb=w["smb"])
    ph.forward(1.0, tol=1e-12, max_vcycles=60, update_geometry=False)
    u_obs, v_obs = ph.grid.u.cpu().numpy().copy(), ph.grid.v.cpu().numpy().copy()
    ph.set_geometry(w["bed"], w["H"])
    ph.set_parameters(B=w["B"], beta=w["beta"], smb=w["smb"])
    return w, ph, u_obs, v_obs, truth


def test_observation_hierarchy_matches_oracle():
    n, levels = 64, 4
    w = S.ice_slab(n, n, DX, seed=1)
    ph = IcePhysics(n, n, DX, n_levels=levels)
    ph.set_geometry(w["bed"], w["H"])
    rng = np.random.default_rng(1)
    uo, vo = rng.normal(size=(n, n + 1)) * 100.0, rng.normal(size=(n + 1, n)) * 100.0
    obs = restrict_observations(ph, uo, vo)
    assert len(obs) == levels
    ru, rv = uo, vo
    for k in range(1, levels):
        ru, rv = O.restrict_u(ru), O.restrict_v(rv)
        assert np.array_equal(obs[k][0].cpu().numpy(), ru), k
        assert np.array_equal(obs[k][1].cpu().numpy(), rv), k
    with pytest.raises(ValueError):
        restrict_observations(ph, uo[:, :-1], vo)


@pytest.mark.parametrize("k", [1, 2])
def test_level_objective_matches_oracle(k):
    n, levels, reg = 64, 3, 1e-4
    w, ph, u_obs, v_obs, _ = make(n, levels)
    mr = MultiResolutionInversion(ph, u_obs, v_obs, reg_weight=reg, dt=1.0)
    inv = mr.level_objective(k)
    g = ph.grids[k]
    nk = n >> k
    # the level's inputs are the oracle's restrictions of the finest inputs
    fields = {"bed": w["bed"], "H_prev": w["H"], "B": np.broadcast_to(w["B"], (n, n)), "smb": w["smb"],
              "beta": w["beta"]}
    fields = {name: np.asarray(a, dtype=np.float64) for name, a in fields.items()}
    for _ in range(k):
        fields = {name: O.restrict_cell(a) for name, a in fields.items()}
    for name in ("bed", "H_prev", "B", "smb", "beta"):
        assert np.array_equal(getattr(g, name).cpu().numpy(), fields[name]), name
    rng = np.random.default_rng(10 + k)
    log_beta = np.log(fields["beta"]) + 0.2 * rng.standard_normal((nk, nk))
    f, grad = inv.objective(log_beta.ravel())

    oh = O.Hierarchy(nk, nk, DX * (1 << k), n_levels=levels - k, thklim=ph.thklim, gl_derivatives=False)
    oh.set_geometry(fields["bed"], fields["H_prev"])
    beta = np.exp(log_beta)
    oh.set_parameters(B=fields["B"], beta=beta, smb=fields["smb"])
    u, v, _ = oh.forward(1.0, n_vcycles=5, update_geometry=False)
    uo, vo = u_obs, v_obs
    for _ in range(k):
        uo, vo = O.restrict_u(uo), O.restrict_v(vo)
    uo, vo = torch.as_tensor(uo), torch.as_tensor(vo)
    ut, vt = torch.as_tensor(u), torch.as_tensor(v)
    J = float(abs_loss(ut, vt, uo, vo))
    dJu, dJv = abs_grad(ut, vt, uo, vo)
    gb = oh.adjoint(dJu.numpy(), dJv.numpy(), np.zeros((nk, nk)), n_vcycles=1, quirk=True)
    tik, tik_grad = tikhonov_regularization(torch.as_tensor(beta))
    total = J + reg * float(tik)
    grad_ref = (beta * gb + reg * tik_grad.numpy()).ravel()
    print(f"level {k}: loss {f:.12e} oracle {total:.12e}; |grad| {np.linalg.norm(grad_ref):.4e} "
          f"rel diff {np.linalg.norm(grad - grad_ref) / np.linalg.norm(grad_ref):.2e}", flush=True)
    assert abs(f - total) <= 1e-12 * abs(total)
    assert np.linalg.norm(grad - grad_ref) <= 1e-10 * np.linalg.norm(grad_ref)


@pytest.mark.timeout(600)
def test_coarse_to_fine_twin_experiment():
    n, levels, reg = 128, 4, 1e-4
    w, ph, u_obs, v_obs, truth = make(n, levels)
    beta0 = np.full((n, n), float(np.exp(np.mean(np.log(w["beta"])))))
    ph.set_parameters(beta=beta0)
    # finest-level misfit of the starting point, by the same objective
    start = BetaInversion(ph, u_obs, v_obs, ph.grid.H_prev, dt=1.0, alpha=reg, mode="reference")
    start.objective(np.log(beta0).ravel())
    m_start = start.history[0]["misfit"]
    order = []
    mr = MultiResolutionInversion(ph, u_obs, v_obs, reg_weight=reg, dt=1.0,
                                  callback=lambda k, it, view: order.append(k))
    res = mr.run(maxiter={3: 20, 2: 20, 1: 15, 0: 10})
    for r in res:
        print(f"level {r['level']} {r['shape']}: {r['nit']} iterations, {r['nfev']} evaluations, "
              f"misfit {r['misfit0']:.4e} -> {r['misfit']:.4e}, {r['seconds']:.2f} s "
              f"({1e3 * r['seconds'] / max(r['nfev'], 1):.1f} ms per evaluation)", flush=True)
    assert [r["level"] for r in res] == [3, 2, 1, 0]
    assert order == sorted(order, reverse=True)      # callbacks come coarse to fine
    assert res[-1]["shape"] == (n, n)
    m_end = res[-1]["misfit"]
    err0 = np.linalg.norm(np.log(beta0) - np.log(truth))
    err1 = np.linalg.norm(np.log(ph.grid.beta.cpu().numpy()) - np.log(truth))
    print(f"finest misfit {m_start:.4e} -> {m_end:.4e}; |log beta - truth| {err0:.3e} -> {err1:.3e}", flush=True)
    assert m_end < 0.5 * m_start
