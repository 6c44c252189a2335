"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no tests, fixtures or golden vectors, so the oracle is
pinned by: an independent vectorised statement of the residual, central
finite differences of the residual (JVP, dense Jacobian, Vanka blocks,
d/dbeta), the dot-product test between JVP and VJP, the symmetry of the
velocity block, LU against numpy.linalg.solve, fixed points of the smoothers,
mass-flux telescoping, closed forms of the transfer operators, and the
accuracy of the deterministic exp/log/pow against numpy.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import numpy_residual as NR
from tests.helpers import random_state, res


def test_detmath_exp_log_pow_accuracy():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-40, 40, 4000), rng.uniform(-700, 700, 500), [0.0, 1e-300, -1e-300]])
    e = np.array([O.exp(x) for x in xs])
    ref = np.exp(xs)
    ulp = np.abs(e - ref) / np.spacing(np.abs(ref))
    assert ulp.max() <= 2.0, ulp.max()
    ls = np.concatenate([10.0 ** rng.uniform(-300, 300, 4000), rng.uniform(0.5, 2.0, 2000), [1.0, 2.0, 5e-324]])
    lg = np.array([O.log(x) for x in ls])
    ref = np.log(ls)
    ok = ref != 0
    ulp = np.abs(lg[ok] - ref[ok]) / np.spacing(np.abs(ref[ok]))
    assert ulp.max() <= 2.0, ulp.max()
    assert O.log(1.0) == 0.0
    xs = 10.0 ** rng.uniform(-12, 4, 3000)
    for p in (-1.0 / 3.0, -4.0 / 3.0, 0.5, -0.2):
        pw = np.array([O.pow(x, p) for x in xs])
        ref = xs ** p
        err = np.abs(pw - ref) / np.spacing(ref)
        bound = np.abs(p * np.log(xs)) * 2 + 4
        assert np.all(err <= bound), (p, err.max())


def test_residual_matches_independent_numpy():
    for seed in range(4):
        s = random_state(ny=9, nx=12, seed=seed)
        s["mask"][2, 3] = 1.0
        s["mask"][5, 7] = 1.0
        for gd in (False,):
            prm = O.params(gl_derivatives=gd)
            r = res(s, prm)
            ru, rv, rH = O.split(r, s["ny"], s["nx"])
            u, v, H = O.split(s["U"], s["ny"], s["nx"])
            fu, fv, fH = O.split(s["f"], s["ny"], s["nx"])
            nu_, nv_, nH_ = NR.residual(u, v, H, fu, fv, fH, s["bed"], s["B"], s["beta"], s["mask"],
                                        s["gamma"], s["dx"], s["dt"])
            for a, b in ((ru, nu_), (rv, nv_), (rH, nH_)):
                scale = np.abs(b).max()
                assert np.abs(a - b).max() <= 1e-11 * scale, np.abs(a - b).max() / scale


def _fd_jvp(s, prm, d, eps):
    return (res(s, prm, U=s["U"] + eps * d) - res(s, prm, U=s["U"] - eps * d)) / (2 * eps)


@pytest.mark.parametrize("gl_derivatives", [False, True])
def test_jvp_matches_central_fd(gl_derivatives):
    """With gl_derivatives the JVP is the exact derivative; without it the
    reference drops d(sigmoid)/dH (stress.cu:265-273, flux.cu:203-208), which
    is exact only where the grounding-line sigmoid is saturated (|0.1 z| > 20),
    so that case is checked on a bed far above flotation."""
    prm = O.params(gl_derivatives=gl_derivatives)
    for seed in range(3):
        s = random_state(ny=7, nx=9, seed=10 + seed, float_frac=0.0 if not gl_derivatives else 0.3)
        if not gl_derivatives:
            s["bed"] = np.abs(s["bed"]) + 300.0
        s["mask"][3, 3] = 1.0
        rng = np.random.default_rng(seed)
        d = rng.normal(size=s["U"].shape)
        jv = O.jvp(s["ny"], s["nx"], s["dx"], s["dt"], prm, s["U"], d, s["bed"], s["B"], s["beta"], s["mask"], s["gamma"])
        fd = _fd_jvp(s, prm, d, 1e-5)
        assert np.linalg.norm(jv - fd) <= 1e-6 * np.linalg.norm(fd), np.linalg.norm(jv - fd) / np.linalg.norm(fd)


def _dense_jacobian(s, prm, eps=1e-6):
    n = s["U"].size
    Jm = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        h = eps * max(1.0, abs(s["U"][k]))
        Jm[:, k] = (res(s, prm, U=s["U"] + h * e) - res(s, prm, U=s["U"] - h * e)) / (2 * h)
    return Jm


def _jvp_vjp_dense(s, prm):
    ny, nx = s["ny"], s["nx"]
    n = s["U"].size
    Jj = np.zeros((n, n))
    Jt = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        Jj[:, k] = O.jvp(ny, nx, s["dx"], s["dt"], prm, s["U"], e, s["bed"], s["B"], s["beta"], s["mask"], s["gamma"])
        Jt[k, :] = O.vjp(ny, nx, s["dx"], s["dt"], prm, s["U"], e, s["bed"], s["B"], s["beta"], s["mask"], s["gamma"])
    return Jj, Jt


def test_dense_jacobian_jvp_vjp_and_symmetry():
    s = random_state(ny=4, nx=5, seed=3)
    s["mask"][1, 2] = 1.0
    ny, nx = s["ny"], s["nx"]
    nu, nv = ny * (nx + 1), (ny + 1) * nx
    # exact derivative (gl_derivatives): JVP columns = FD columns
    prm = O.params(gl_derivatives=True)
    Jfd = _dense_jacobian(s, prm)
    Jj, Jt = _jvp_vjp_dense(s, prm)
    scale = np.abs(Jfd).max()
    assert np.abs(Jj - Jfd).max() <= 1e-6 * scale
    # reference quirk: with gl_derivatives the VJP drops d(tau_b)/dH of the
    # clamped ghost cell on boundary faces (its LDS slot lies outside the
    # domain and is discarded at the flush, residuals.cu:997-1001); every
    # other entry is the exact transpose
    diff = np.abs(Jt - Jj) > 1e-11 * scale
    rows, cols = np.nonzero(diff)
    for r_, c_ in zip(rows, cols):
        assert c_ >= nu + nv, "only H columns may differ"
        if r_ < nu:
            i, j = divmod(r_, nx + 1)
            assert j in (0, nx)
        else:
            i, j = divmod(r_ - nu, nx)
            assert i in (0, ny)
    # without gl_derivatives vjp(e_k) = J^T e_k is exactly row k of J
    prm = O.params()
    Jj, Jt = _jvp_vjp_dense(s, prm)
    assert np.abs(Jt - Jj).max() <= 1e-11 * np.abs(Jj).max()
    Juv = Jj[:nu + nv, :nu + nv]
    assert np.abs(Juv - Juv.T).max() <= 1e-11 * np.abs(Juv).max()


def test_vjp_dot_product():
    prm = O.params()
    for seed in range(3):
        s = random_state(ny=11, nx=8, seed=20 + seed)
        s["mask"][4, 2] = 1.0
        rng = np.random.default_rng(seed)
        d = rng.normal(size=s["U"].shape)
        lam = rng.normal(size=s["U"].shape)
        ny, nx = s["ny"], s["nx"]
        jv = O.jvp(ny, nx, s["dx"], s["dt"], prm, s["U"], d, s["bed"], s["B"], s["beta"], s["mask"], s["gamma"])
        vj = O.vjp(ny, nx, s["dx"], s["dt"], prm, s["U"], lam, s["bed"], s["B"], s["beta"], s["mask"], s["gamma"])
        a, b = lam @ jv, vj @ d
        assert abs(a - b) <= 1e-12 * (np.abs(lam) @ np.abs(jv)), (a, b)


def test_lu5_matches_numpy_solve():
    rng = np.random.default_rng(1)
    for _ in range(50):
        A = rng.normal(size=(5, 5)) + 5 * np.eye(5)
        b = rng.normal(size=5)
        x = O.lu5(A, b)
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-12, atol=1e-13)


def _local_dofs(ny, nx, i, j):
    nu = ny * (nx + 1)
    nv = (ny + 1) * nx
    return [i * (nx + 1) + j, i * (nx + 1) + j + 1, nu + i * nx + j, nu + (i + 1) * nx + j, nu + nv + i * nx + j]


def test_vanka_block_is_frozen_eta_jacobian():
    """The forward block's r equals the global residual at the cell's 5 dofs,
    and J (minus the diag(-1,-1,-1,-1,+1) shift) equals the FD Jacobian of
    that local residual with eta frozen (vanka.cu:56-411)."""
    prm = O.params(gl_derivatives=True)
    s = random_state(ny=8, nx=8, seed=5, thin_frac=0.0)
    ny, nx = s["ny"], s["nx"]
    eta0 = O.eta(ny, nx, s["dx"], prm, s["U"], s["B"])
    rglob = res(s, prm)
    for (i, j) in [(0, 0), (3, 4), (7, 2), (5, 7)]:
        dofs = _local_dofs(ny, nx, i, j)
        J, r = O.vanka_block(ny, nx, s["dx"], s["dt"], prm, s["U"], s["f"], s["bed"], s["B"], s["beta"],
                             s["gamma"], i, j)
        if abs(r[4]) > 0 and s["H"][i, j] - s["dt"] * r[4] <= 0.1:
            continue   # constrained row, checked separately
        assert np.allclose(r, rglob[dofs], rtol=1e-13, atol=1e-12 * np.abs(rglob).max())
        Jfd = np.zeros((5, 5))
        for c, k in enumerate(dofs):
            h = 1e-6 * max(1.0, abs(s["U"][k]))
            Up, Um = s["U"].copy(), s["U"].copy()
            Up[k] += h
            Um[k] -= h
            _, rp = O.vanka_block(ny, nx, s["dx"], s["dt"], prm, Up, s["f"], s["bed"], s["B"], s["beta"],
                                  s["gamma"], i, j, eta_field=eta0)
            _, rm = O.vanka_block(ny, nx, s["dx"], s["dt"], prm, Um, s["f"], s["bed"], s["B"], s["beta"],
                                  s["gamma"], i, j, eta_field=eta0)
            Jfd[:, c] = (rp - rm) / (2 * h)
        shift = np.diag([-1.0, -1.0, -1.0, -1.0, 1.0])
        # The reference assembles only the diagonal shear terms of each vertex
        # (vanka.cu:211-311): the shear cross-couplings u-row/v-dof and
        # v-row/u-dof are left out of the block; each omitted entry is
        # +-etaH_vertex / dx^2 (zero on boundary vertices).
        eH = eta0 * s["H"]
        dx = s["dx"]

        def ehv(a, b):
            if a <= 0 or a >= ny or b <= 0 or b >= nx:
                return 0.0
            return 0.25 * (eH[a - 1, b - 1] + eH[a - 1, b] + eH[a, b - 1] + eH[a, b]) / dx ** 2
        omitted = {(0, 2): ehv(i, j), (0, 3): -ehv(i + 1, j), (1, 2): -ehv(i, j + 1), (1, 3): ehv(i + 1, j + 1),
                   (2, 0): ehv(i, j), (2, 1): -ehv(i, j + 1), (3, 0): -ehv(i + 1, j), (3, 1): ehv(i + 1, j + 1)}
        expect = J - shift
        for (a, b), val in omitted.items():
            expect[a, b] += val
        assert np.abs(expect - Jfd).max() <= 1e-6 * np.abs(Jfd).max()


def test_vanka_fixed_point_and_constraint():
    """With f = F(U) the residual vanishes, so a sweep leaves U unchanged; a
    sweep never takes H below gamma."""
    prm = O.params()
    s = random_state(ny=12, nx=10, seed=7, thin_frac=0.0)
    ny, nx = s["ny"], s["nx"]
    f0 = res(s, prm, f=np.zeros_like(s["f"]))
    U1 = O.vanka_sweep(ny, nx, s["dx"], s["dt"], prm, s["U"], f0, s["bed"], s["B"], s["beta"], s["gamma"])
    assert np.abs(U1 - s["U"]).max() <= 1e-9 * np.abs(s["U"]).max()
    U2 = O.vanka_sweep(ny, nx, s["dx"], s["dt"], prm, s["U"], s["f"], s["bed"], s["B"], s["beta"], s["gamma"])
    _, _, H2 = O.split(U2, ny, nx)
    assert H2.min() >= 0.1 - 1e-12


def test_vanka_sweep_face_averaging():
    """Face update = 1/4 of the sum of adjacent cells' local changes, H = 1/2
    (grid.py:305-310 with omega = 0.5 and the 0.5 factor in vanka.cu:433-437):
    smoothing only one cell changes exactly its four faces and its H."""
    prm = O.params()
    s = random_state(ny=6, nx=6, seed=9, thin_frac=0.0)
    ny, nx = s["ny"], s["nx"]
    act = np.zeros((ny, nx), dtype=np.uint8)
    act[2, 3] = 1
    U1 = O.vanka_sweep(ny, nx, s["dx"], s["dt"], prm, s["U"], s["f"], s["bed"], s["B"], s["beta"], s["gamma"],
                       active=act)
    changed = np.nonzero(U1 != s["U"])[0]
    assert set(changed) <= set(_local_dofs(ny, nx, 2, 3))
    assert len(changed) >= 4


def test_mass_flux_telescoping():
    """Boundary fluxes vanish, so sum(r_H) = sum(H/dt - f_H + calving)."""
    prm = O.params()
    s = random_state(ny=10, nx=13, seed=11)
    ny, nx = s["ny"], s["nx"]
    r = res(s, prm)
    _, _, rH = O.split(r, ny, nx)
    _, _, fH = O.split(s["f"], ny, nx)
    H = s["H"]
    g = NR.sigmoid(s["bed"] + 0.917 * H, 0.1)
    expect = (H / s["dt"] - fH + (1 - g) * H).sum()
    assert abs(rH.sum() - expect) <= 1e-12 * np.abs(H / s["dt"]).sum()


def test_transfers_closed_forms():
    rng = np.random.default_rng(2)
    nyc, nxc = 4, 6
    # restriction of constants and of the hand-computed 4x4 -> 2x2 example
    assert np.all(O.restrict_cell(np.full((8, 12), 3.5)) == 3.5)
    h = np.arange(16.0).reshape(4, 4)
    assert np.array_equal(O.restrict_cell(h), np.array([[2.5, 4.5], [10.5, 12.5]]))
    assert np.array_equal(O.restrict_max(h), np.array([[5.0, 7.0], [13.0, 15.0]]))
    uf = rng.normal(size=(2 * nyc, 2 * nxc + 1))
    uc = O.restrict_u(uf)
    assert np.allclose(uc, 0.5 * (uf[0::2, 0::2] + uf[1::2, 0::2]))
    vf = rng.normal(size=(2 * nyc + 1, 2 * nxc))
    vc = O.restrict_v(vf)
    assert np.allclose(vc, 0.5 * (vf[0::2, 0::2] + vf[0::2, 1::2]))
    hc = rng.normal(size=(nyc, nxc))
    assert np.array_equal(O.prolong_inject(hc), np.kron(hc, np.ones((2, 2))))
    assert np.array_equal(O.prolong_inject(hc, -1.0), -np.kron(hc, np.ones((2, 2))))
    # bilinear prolongation is exact on bilinear functions inside the clamp region
    a, b, c = 0.3, -1.7, 2.2
    J, I = np.meshgrid(np.arange(nyc), np.arange(nxc + 1), indexing="ij")
    ucl = a + b * (2 * J + 1) + c * (2 * I)          # values at coarse u positions (y=2J+1, x=2I) in fine units
    uf = O.prolong_u(ucl)
    j, i = np.meshgrid(np.arange(2 * nyc), np.arange(2 * nxc + 1), indexing="ij")
    exact = a + b * (j + 0.5) + c * i
    inner = (j >= 1) & (j <= 2 * nyc - 2)
    assert np.allclose(uf[inner], exact[inner], rtol=1e-14, atol=1e-13)
    J, I = np.meshgrid(np.arange(nyc + 1), np.arange(nxc), indexing="ij")
    vcl = a + b * (2 * J) + c * (2 * I + 1)
    vfp = O.prolong_v(vcl)
    j, i = np.meshgrid(np.arange(2 * nyc + 1), np.arange(2 * nxc), indexing="ij")
    exact = a + b * j + c * (i + 0.5)
    inner = (i >= 1) & (i <= 2 * nxc - 2)
    assert np.allclose(vfp[inner], exact[inner], rtol=1e-14, atol=1e-13)
    J, I = np.meshgrid(np.arange(nyc), np.arange(nxc), indexing="ij")
    hcl = a + b * (2 * J + 1) + c * (2 * I + 1)
    hf = O.prolong_H(hcl)
    j, i = np.meshgrid(np.arange(2 * nyc), np.arange(2 * nxc), indexing="ij")
    exact = a + b * (j + 0.5) + c * (i + 0.5)
    inner = (i >= 1) & (i <= 2 * nxc - 2) & (j >= 1) & (j <= 2 * nyc - 2)
    assert np.allclose(hf[inner], exact[inner], rtol=1e-14, atol=1e-13)
    # constants are reproduced everywhere (clamped boundary included)
    assert np.allclose(O.prolong_u(np.full((nyc, nxc + 1), 2.0)), 2.0)
    assert np.allclose(O.prolong_v(np.full((nyc + 1, nxc), 2.0)), 2.0)
    assert np.allclose(O.prolong_H(np.full((nyc, nxc), 2.0)), 2.0)


def test_grad_beta_is_lambda_dR_dbeta():
    """grad_beta (exact transpose) = lambda . dR/dbeta by central FD in beta;
    the reference-compatible variant differs only on the domain's edge cells."""
    prm = O.params()
    s = random_state(ny=7, nx=9, seed=13)
    ny, nx = s["ny"], s["nx"]
    rng = np.random.default_rng(4)
    lam = rng.normal(size=s["U"].shape)
    g_exact = O.grad_beta(ny, nx, s["dx"], s["dt"], prm, s["U"], lam, s["bed"], s["beta"], quirk=False)
    g_ref = O.grad_beta(ny, nx, s["dx"], s["dt"], prm, s["U"], lam, s["bed"], s["beta"], quirk=True)
    for _ in range(3):
        e = rng.normal(size=(ny, nx))
        h = 1e-6 * np.abs(s["beta"]).max()
        fd = (lam @ res(s, prm, beta=s["beta"] + h * e) - lam @ res(s, prm, beta=s["beta"] - h * e)) / (2 * h)
        assert abs((g_exact * e).sum() - fd) <= 1e-6 * np.abs(g_exact * e).sum()
    diff = g_exact != g_ref
    # they differ only in the first/last row and column
    assert not diff[1:-1, 1:-1].any()


def test_error_mask_and_local_activation():
    em = np.zeros((40, 30))
    em[20, 5] = 1.0
    act = O.local_activation(em)
    # block (1, 0) (rows 14..27, cols 0..13) owns (20, 5); its footprint does not reach other blocks' cells
    assert act[14:28, 0:14].all()
    assert act.sum() == 14 * 14
    em2 = np.zeros((40, 30))
    em2[13, 14] = 1.0        # on the halo ring of blocks (0,0),(0,1),(1,0),(1,1)
    act2 = O.local_activation(em2)
    assert act2[:28, :28].all() and act2.sum() == 28 * 28


def _fd4(fun, U, d, e):
    return (8 * (fun(U + e * d) - fun(U - e * d)) - (fun(U + 2 * e * d) - fun(U - 2 * e * d))) / (12 * e)


def test_jvp_fourth_order_fd_on_saturated_state():
    """On a grounded state (sigmoid clamped flat everywhere) the JVP without
    grounding-line derivatives is the exact derivative: 4th-order central FD
    agrees to ~1e-10.  With gl_derivatives the reference's sigmoid derivative
    ignores the +-20 clamp (common.cu:67-76), so saturated cells carry a spurious
    c*H*0.917*s(1-s) in the calving Jacobian (flux.cu:205-208): the H columns
    are then off at the 1e-7 level, which this test also pins."""
    n = 24
    from glide_amd import synthetic as S
    w = S.ice_slab(n, n, 2000.0, seed=2)
    rng = np.random.default_rng(0)
    U = O.join(rng.normal(0, 50, (n, n + 1)), rng.normal(0, 50, (n + 1, n)), w["H"])
    f = np.zeros_like(U)
    mask, gamma = np.zeros((n, n)), np.full((n, n), 0.1)
    d = rng.normal(size=U.size)
    d[-n * n:] *= 10.0
    errs = {}
    for gl in (False, True):
        prm = O.params(gl_derivatives=gl)

        def R(X):
            return O.residual(n, n, w["dx"], 10.0, prm, X, f, w["bed"], w["B"], w["beta"], mask, gamma)
        jv = O.jvp(n, n, w["dx"], 10.0, prm, U, d, w["bed"], w["B"], w["beta"], mask, gamma)
        fd = _fd4(R, U, d, 1e-3)
        errs[gl] = np.linalg.norm(jv - fd) / np.linalg.norm(fd)
    assert errs[False] <= 1e-9, errs
    assert 1e-9 < errs[True] <= 1e-5, errs
