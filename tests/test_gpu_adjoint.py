"""GPU parity of the linearisation and adjoint path against the CPU oracle.

JVP, the adjoint Vanka pass and grad_beta are bit-identical to the oracle;
the VJP is a deterministic gather (the oracle scatters in the reference's
per-row order), so it agrees to round-off and is also checked by the
dot-product test <lam, J d> = <J^T lam, d> on the GPU itself.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import random_state

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from glide_amd import _lib  # noqa: E402
from glide_amd import synthetic as S  # noqa: E402
from glide_amd.physics import IcePhysics  # noqa: E402

DEV = torch.device("cuda")
_KEEP = []


def T(a):
    t = torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=DEV)
    _KEEP.append(t)
    return t


def P(t):
    if t is None:
        return None
    _KEEP.append(t)
    return ctypes.c_void_p(t.data_ptr())


@pytest.fixture(autouse=True)
def _release_kept():
    yield
    torch.cuda.synchronize()
    _KEEP.clear()


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def cparams(prm):
    return _lib.GlideParams(prm.n, prm.eps_reg, prm.water_drag, prm.calving_rate, prm.gl_sigmoid_c, prm.gl_derivatives)


def _args(s):
    return [P(T(s[k])) for k in ("bed", "B", "beta", "mask", "gamma")]


@pytest.mark.parametrize("gl", [False, True])
@pytest.mark.parametrize("shape", [(7, 9), (33, 70)])
def test_jvp_bitwise(gl, shape):
    ny, nx = shape
    s = random_state(ny=ny, nx=nx, seed=5 * ny + nx)
    s["mask"][ny // 2, nx // 2] = 1.0
    prm = O.params(gl_derivatives=gl)
    d = np.random.default_rng(ny).normal(size=s["U"].shape)
    ref = O.jvp(ny, nx, s["dx"], s["dt"], prm, s["U"], d, s["bed"], s["B"], s["beta"], s["mask"], s["gamma"])
    out = torch.zeros(ref.size, dtype=torch.float64, device=DEV)
    g = _lib.GlideGeom(ny, nx, s["dx"], s["dt"])
    _lib.call("glide_jvp", ctypes.byref(g), ctypes.byref(cparams(prm)), P(out), P(T(s["U"])), P(T(d)), *_args(s), stream())
    o = out.cpu().numpy()
    assert np.array_equal(o, ref), np.abs(o - ref).max()


@pytest.mark.parametrize("gl", [False, True])
@pytest.mark.parametrize("shape", [(7, 9), (33, 70)])
def test_vjp_matches_oracle_and_dot_product(gl, shape):
    ny, nx = shape
    s = random_state(ny=ny, nx=nx, seed=7 * ny + nx)
    s["mask"][1, 2] = 1.0
    prm = O.params(gl_derivatives=gl)
    rng = np.random.default_rng(nx)
    lam, d = rng.normal(size=s["U"].shape), rng.normal(size=s["U"].shape)
    ref = O.vjp(ny, nx, s["dx"], s["dt"], prm, s["U"], lam, s["bed"], s["B"], s["beta"], s["mask"], s["gamma"])
    g = _lib.GlideGeom(ny, nx, s["dx"], s["dt"])
    out = torch.zeros(ref.size, dtype=torch.float64, device=DEV)
    jv = torch.zeros(ref.size, dtype=torch.float64, device=DEV)
    _lib.call("glide_vjp", ctypes.byref(g), ctypes.byref(cparams(prm)), P(out), P(T(s["U"])), P(T(lam)), *_args(s),
              stream())
    _lib.call("glide_jvp", ctypes.byref(g), ctypes.byref(cparams(prm)), P(jv), P(T(s["U"])), P(T(d)), *_args(s), stream())
    o, j = out.cpu().numpy(), jv.cpu().numpy()
    # gather vs scatter: same terms, different summation order
    scale = np.abs(ref).max()
    assert np.abs(o - ref).max() <= 1e-12 * scale, np.abs(o - ref).max() / scale
    if not gl:   # with gl_derivatives the reference VJP drops ghost-cell terms (see the oracle pins)
        a, b = lam @ j, o @ d
        assert abs(a - b) <= 1e-12 * (np.abs(lam) @ np.abs(j)), (a, b)


@pytest.fixture
def lin_tiles():
    """Set the JVP / VJP tile height for one test; restores the default (4)."""
    yield lambda ty: _lib.call("glide_set_tile_height", 1, ty)
    _lib.call("glide_set_tile_height", 1, 4)


@pytest.mark.parametrize("gl", [False, True])
@pytest.mark.parametrize("shape", [(7, 9), (33, 70), (64, 129), (130, 64)])
def test_lin_tile_heights_same_bits(lin_tiles, gl, shape):
    """The LDS-staged JVP and VJP (tile heights 4 and 8) against the per-point
    form (0): the JVP equals the oracle bit for bit at every height, the VJP
    (plain and as the adjoint residual base - J^T lambda) the per-point VJP."""
    ny, nx = shape
    s = random_state(ny=ny, nx=nx, seed=3 * ny + nx)
    s["mask"][ny // 3, nx // 2] = 1.0
    prm = O.params(gl_derivatives=gl)
    rng = np.random.default_rng(ny * nx)
    d, lam, base = (rng.normal(size=s["U"].shape) for _ in range(3))
    ref = O.jvp(ny, nx, s["dx"], s["dt"], prm, s["U"], d, s["bed"], s["B"], s["beta"], s["mask"], s["gamma"])
    g = _lib.GlideGeom(ny, nx, s["dx"], s["dt"])
    lev = (_lib.GlideLevel * 1)()
    lev[0].g = g
    lev[0].p = cparams(prm)
    Ut, lt, bt = T(s["U"]), T(lam), T(base)
    cells = {k: T(s[k]) for k in ("bed", "B", "beta", "mask", "gamma")}
    for k in ("bed", "B", "beta", "mask", "gamma"):
        setattr(lev[0], k, cells[k].data_ptr())
    lev[0].U, lev[0].Lambda, lev[0].f_adj = Ut.data_ptr(), lt.data_ptr(), bt.data_ptr()
    outs = {}
    for ty in (0, 4, 6, 8):
        lin_tiles(ty)
        jo = torch.zeros(ref.size, dtype=torch.float64, device=DEV)
        vo = torch.zeros(ref.size, dtype=torch.float64, device=DEV)
        ra = torch.zeros(ref.size, dtype=torch.float64, device=DEV)
        lev[0].r_adj = ra.data_ptr()
        _lib.call("glide_jvp", ctypes.byref(g), ctypes.byref(cparams(prm)), P(jo), P(Ut), P(T(d)), *_args(s), stream())
        _lib.call("glide_vjp", ctypes.byref(g), ctypes.byref(cparams(prm)), P(vo), P(Ut), P(lt), *_args(s), stream())
        _lib.call("glide_adjoint_residual", ctypes.byref(lev[0]), stream())
        torch.cuda.synchronize()
        outs[ty] = (jo.cpu().numpy(), vo.cpu().numpy(), ra.cpu().numpy())
        assert np.array_equal(outs[ty][0], ref), (ty, np.abs(outs[ty][0] - ref).max())
    for ty in (4, 6, 8):
        assert np.array_equal(outs[ty][1], outs[0][1]), ty
        assert np.array_equal(outs[ty][2], outs[0][2]), ty
    assert np.array_equal(outs[0][2], base - outs[0][1])


@pytest.mark.parametrize("color", [0, 1])
def test_adjoint_pass_bitwise(color):
    ny, nx = 20, 37
    s = random_state(ny=ny, nx=nx, seed=11 + color)
    s["mask"][3:6, 4:9] = 1.0
    prm = O.params()
    rng = np.random.default_rng(color)
    lam, r_adj = rng.normal(size=s["U"].shape), rng.normal(size=s["U"].shape)
    ref = O.vanka_adjoint_pass(ny, nx, s["dx"], s["dt"], prm, lam, s["mask"], r_adj, s["U"], s["bed"], s["B"], s["beta"],
                               s["gamma"], color, 1.0)
    g = _lib.GlideGeom(ny, nx, s["dx"], s["dt"])
    L = T(lam)
    _lib.call("glide_vanka_adjoint_pass", ctypes.byref(g), ctypes.byref(cparams(prm)), P(L), P(L), P(T(s["mask"])),
              P(T(r_adj)), P(T(s["U"])), P(T(s["bed"])), P(T(s["B"])), P(T(s["beta"])), P(T(s["gamma"])), color, 1.0,
              stream())
    o = L.cpu().numpy()
    assert np.array_equal(o, ref), np.abs(o - ref).max()


@pytest.mark.parametrize("edges", [True, False])
def test_grad_beta_bitwise(edges):
    ny, nx = 17, 23
    s = random_state(ny=ny, nx=nx, seed=31)
    prm = O.params()
    lam = np.random.default_rng(2).normal(size=s["U"].shape)
    ref = O.grad_beta(ny, nx, s["dx"], s["dt"], prm, s["U"], lam, s["bed"], s["beta"], quirk=edges)
    g = _lib.GlideGeom(ny, nx, s["dx"], s["dt"])
    out = torch.zeros((ny, nx), dtype=torch.float64, device=DEV)
    _lib.call("glide_grad_beta", ctypes.byref(g), ctypes.byref(cparams(prm)), P(out), P(T(s["U"])), P(T(lam)),
              P(T(s["bed"])), P(T(s["beta"])), int(edges), stream())
    o = out.cpu().numpy()
    assert np.array_equal(o, ref), np.abs(o - ref).max()


@pytest.mark.parametrize("shape", [(17, 23), (64, 40)])
def test_grad_B_bitwise(shape):
    ny, nx = shape
    s = random_state(ny=ny, nx=nx, seed=37)
    prm = O.params()
    lam = np.random.default_rng(3).normal(size=s["U"].shape)
    ref = O.grad_B(ny, nx, s["dx"], prm, s["U"], lam, s["B"])
    g = _lib.GlideGeom(ny, nx, s["dx"], s["dt"])
    out = torch.zeros((ny, nx), dtype=torch.float64, device=DEV)
    _lib.call("glide_grad_B", ctypes.byref(g), ctypes.byref(cparams(prm)), P(out), P(T(s["U"])), P(T(lam)),
              P(T(s["B"])), stream())
    o = out.cpu().numpy()
    assert np.array_equal(o, ref), np.abs(o - ref).max()


def test_adjoint_vcycle_and_gradient_match_oracle():
    """Forward state from two steps, then two adjoint V-cycles and grad_beta:
    lambda and the gradient agree with the oracle to round-off propagation."""
    w = S.ice_cap(64, 64, 2000.0, seed=4, radius=45e3)
    kw = dict(n_levels=3)
    h = O.Hierarchy(64, 64, w["dx"], **kw)
    ph = IcePhysics(64, 64, w["dx"], **kw)
    for obj in (h, ph):
        obj.set_geometry(w["bed"], w["H"])
        obj.set_parameters(B=w["B"], beta=w["beta"], smb=w["smb"])
        obj.forward(10.0, n_vcycles=2)
    rng = np.random.default_rng(3)
    dLu, dLv, dLH = rng.normal(size=(64, 65)), rng.normal(size=(65, 64)), np.zeros((64, 64))
    g_ref = h.adjoint(dLu, dLv, dLH, n_vcycles=2)
    lam_ref = h.field("Lambda").copy()
    g_gpu = ph.adjoint(dLu, dLv, dLH, n_vcycles=2).cpu().numpy()
    lam = ph.grid.Lambda.cpu().numpy()
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)  # noqa: E731
    print("lambda rel", rel(lam, lam_ref), "grad rel", rel(g_gpu, g_ref), flush=True)
    assert rel(lam, lam_ref) <= 1e-10
    assert rel(g_gpu, g_ref) <= 1e-10
