"""GPU parity of the forward path: HIP kernels through the C ABI against the
CPU oracle on the same seeded inputs.

Kernel-level results must be bit-identical (same FP64 operation order,
no contraction, same deterministic exp/log on both sides); multi-level
V-cycles and time steps are compared by relative L2 on u, v and H.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import random_state

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover - collected only where a GPU is present
    pytest.skip("no GPU", allow_module_level=True)

from glide_amd import _lib  # noqa: E402
from glide_amd import synthetic as S  # noqa: E402
from glide_amd.physics import IcePhysics  # noqa: E402

DEV = torch.device("cuda")


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=DEV)


_KEEP = []


def P(t):
    """Device pointer of t; t is kept alive for the rest of the test session so
    that torch's caching allocator cannot hand its block to a later tensor
    before the enqueued kernel has read it."""
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
    return _lib.GlideParams(prm.n, prm.eps_reg, prm.water_drag, prm.calving_rate, prm.gl_sigmoid_c,
                            prm.gl_derivatives)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.mark.parametrize("shape", [(8, 10), (37, 45), (64, 129), (12, 64), (33, 128)])
@pytest.mark.parametrize("gl", [False, True])
def test_residual_bitwise(shape, gl):
    ny, nx = shape
    s = random_state(ny=ny, nx=nx, seed=ny + nx)
    s["mask"][ny // 2, nx // 3] = 1.0
    prm = O.params(gl_derivatives=gl)
    ref = O.residual(ny, nx, s["dx"], s["dt"], prm, s["U"], s["f"], s["bed"], s["B"], s["beta"], s["mask"], s["gamma"])
    g = _lib.GlideGeom(ny, nx, s["dx"], s["dt"])
    r = torch.zeros(ref.size, dtype=torch.float64, device=DEV)
    args = [T(s[k]) for k in ("U", "f", "bed", "B", "beta", "mask", "gamma")]
    U, f, bed, B, beta, mask, gamma = args
    _lib.call("glide_residual", ctypes.byref(g), ctypes.byref(cparams(prm)), P(r), P(U), P(f), None, P(bed), P(B),
              P(beta), P(mask), P(gamma), stream())
    out = r.cpu().numpy()
    assert np.array_equal(out, ref), np.abs(out - ref).max()
    # F(U) - sub in one pass == (F(U) - 0) - sub
    sub = np.random.default_rng(1).normal(size=ref.size)
    refF = O.residual(ny, nx, s["dx"], s["dt"], prm, s["U"], np.zeros_like(s["f"]), s["bed"], s["B"], s["beta"],
                      s["mask"], s["gamma"])
    _lib.call("glide_residual", ctypes.byref(g), ctypes.byref(cparams(prm)), P(r), P(U), None, P(T(sub)), P(bed),
              P(B), P(beta), P(mask), P(gamma), stream())
    assert np.array_equal(r.cpu().numpy(), refF - sub)


@pytest.mark.parametrize("ty", [4, 8, 16])
@pytest.mark.parametrize("shape", [(8, 10), (37, 45), (64, 129), (33, 128), (300, 260)])
def test_residual_tile_heights_bitwise(ty, shape):
    """The residual on 64 x 4, 64 x 8 and 64 x 16 tiles equals the oracle bit for bit."""
    ny, nx = shape
    s = random_state(ny=ny, nx=nx, seed=2 * ny + nx)
    s["mask"][ny // 2, nx // 3] = 1.0
    prm = O.params(gl_derivatives=True)
    ref = O.residual(ny, nx, s["dx"], s["dt"], prm, s["U"], s["f"], s["bed"], s["B"], s["beta"], s["mask"], s["gamma"])
    g = _lib.GlideGeom(ny, nx, s["dx"], s["dt"])
    r = torch.zeros(ref.size, dtype=torch.float64, device=DEV)
    U, f, bed, B, beta, mask, gamma = (T(s[k]) for k in ("U", "f", "bed", "B", "beta", "mask", "gamma"))
    _lib.call("glide_set_tile_height", 0, ty)
    try:
        _lib.call("glide_residual", ctypes.byref(g), ctypes.byref(cparams(prm)), P(r), P(U), P(f), None, P(bed), P(B),
                  P(beta), P(mask), P(gamma), stream())
        torch.cuda.synchronize()
    finally:
        _lib.call("glide_set_tile_height", 0, 0)
    out = r.cpu().numpy()
    assert np.array_equal(out, ref), np.abs(out - ref).max()


@pytest.fixture
def small_levels():
    """Set the small-level Vanka variant for one test; restores the library
    default (levels up to 128 x 128 cells, one lane per cell with the apply
    pass folded in) afterwards."""
    yield lambda cells, variant: _lib.call("glide_set_small_levels", cells, variant)
    _lib.call("glide_set_small_levels", 128 * 128, 2)


# Every shape sweeps with each kernel: the row split and the full-register
# per-cell tile (both below the small-level limit), and the two-waves-per-SIMD
# tile (limit 0).  131 x 150 has interior tiles, which take the specialised
# assembly of the per-cell tiles.
@pytest.mark.parametrize("variant", ["rows", "small", "large"])
@pytest.mark.parametrize("shape", [(9, 12), (40, 70), (131, 150)])
def test_vanka_sweep_bitwise(shape, variant, small_levels):
    ny, nx = shape
    small_levels(*{"rows": (1 << 20, 1), "small": (1 << 20, 0), "large": (0, 0)}[variant])
    s = random_state(ny=ny, nx=nx, seed=3 * ny + nx, vel=100.0)
    prm = O.params()
    ref = O.vanka_sweep(ny, nx, s["dx"], s["dt"], prm, s["U"], s["f"], s["bed"], s["B"], s["beta"], s["gamma"])
    g = _lib.GlideGeom(ny, nx, s["dx"], s["dt"])
    U = T(s["U"])
    scratch = torch.zeros(5 * ny * nx, dtype=torch.float64, device=DEV)
    _lib.call("glide_vanka_sweep", ctypes.byref(g), ctypes.byref(cparams(prm)), P(U), P(T(s["f"])), P(T(s["bed"])),
              P(T(s["B"])), P(T(s["beta"])), P(T(s["gamma"])), None, 1, 10, P(scratch), stream())
    out = U.cpu().numpy()
    assert np.array_equal(out, ref), np.abs(out - ref).max()


# The folded-apply small-level kernel alternates the state between U and the
# scratch, so it runs for two or more sweeps: odd counts start with one
# in-place sweep.  Shapes: owned column blocks of 62 that end exactly at the
# last column (124) or one short of it, odd row counts, interior tiles.
@pytest.mark.parametrize("n_sweeps", [2, 3])
@pytest.mark.parametrize("shape", [(9, 12), (7, 124), (40, 123), (131, 150)])
def test_vanka_sweeps_folded_apply_bitwise(shape, n_sweeps, small_levels):
    ny, nx = shape
    small_levels(1 << 20, 2)
    s = random_state(ny=ny, nx=nx, seed=5 * ny + nx, vel=100.0)
    prm = O.params()
    ref = s["U"]
    for _ in range(n_sweeps):
        ref = O.vanka_sweep(ny, nx, s["dx"], s["dt"], prm, ref, s["f"], s["bed"], s["B"], s["beta"], s["gamma"])
    g = _lib.GlideGeom(ny, nx, s["dx"], s["dt"])
    U = T(s["U"])
    scratch = torch.zeros(5 * ny * nx, dtype=torch.float64, device=DEV)
    _lib.call("glide_vanka_sweep", ctypes.byref(g), ctypes.byref(cparams(prm)), P(U), P(T(s["f"])), P(T(s["bed"])),
              P(T(s["B"])), P(T(s["beta"])), P(T(s["gamma"])), None, n_sweeps, 10, P(scratch), stream())
    out = U.cpu().numpy()
    assert np.array_equal(out, ref), np.abs(out - ref).max()


@pytest.mark.parametrize("variant", ["rows", "small", "large", "folded"])
def test_local_activation_and_local_sweep_bitwise(variant, small_levels):
    small_levels(*{"rows": (1 << 20, 1), "small": (1 << 20, 0), "large": (0, 0), "folded": (1 << 20, 2)}[variant])
    ny, nx = 48, 60
    s = random_state(ny=ny, nx=nx, seed=77, vel=50.0)
    prm = O.params()
    r = O.residual(ny, nx, s["dx"], s["dt"], prm, s["U"], s["f"], s["bed"], s["B"], s["beta"], s["mask"], s["gamma"])
    # shrink the residual below the 0.01 threshold, then re-arm three cells so
    # that only the blocks whose footprint holds them are smoothed
    r = r * 1e-9
    nu, nv = ny * (nx + 1), (ny + 1) * nx
    for (ci, cj) in ((3, 5), (20, 33), (41, 59)):
        r[nu + nv + ci * nx + cj] = 1.0
    em = O.error_mask(ny, nx, r, s["U"], s["gamma"])
    act_cells = O.local_activation(em)
    g = _lib.GlideGeom(ny, nx, s["dx"], s["dt"])
    nby, nbx = ny // 14 + 1, nx // 14 + 1
    blk = torch.zeros(nby * nbx, dtype=torch.uint8, device=DEV)
    _lib.call("glide_local_activation", ctypes.byref(g), P(T(r)), P(T(s["U"])), P(T(s["gamma"])), P(blk), stream())
    b = blk.cpu().numpy().reshape(nby, nbx)
    cells = np.array([[b[i // 14, j // 14] for j in range(nx)] for i in range(ny)])
    assert np.array_equal(cells.astype(np.uint8), act_cells)
    assert 0 < cells.sum() < ny * nx
    # the folded-apply kernel runs from the second sweep on
    n_sweeps = 2 if variant == "folded" else 1
    ref = s["U"]
    for _ in range(n_sweeps):
        ref = O.vanka_sweep(ny, nx, s["dx"], s["dt"], prm, ref, s["f"], s["bed"], s["B"], s["beta"], s["gamma"],
                            active=act_cells)
    U = T(s["U"])
    scratch = torch.zeros(5 * ny * nx, dtype=torch.float64, device=DEV)
    _lib.call("glide_vanka_sweep", ctypes.byref(g), ctypes.byref(cparams(prm)), P(U), P(T(s["f"])), P(T(s["bed"])),
              P(T(s["B"])), P(T(s["beta"])), P(T(s["gamma"])), P(blk), n_sweeps, 10, P(scratch), stream())
    assert np.array_equal(U.cpu().numpy(), ref)


def test_transfers_bitwise():
    rng = np.random.default_rng(9)
    nyc, nxc = 13, 22
    uf, vf, hf = rng.normal(size=(2 * nyc, 2 * nxc + 1)), rng.normal(size=(2 * nyc + 1, 2 * nxc)), rng.normal(size=(2 * nyc, 2 * nxc))
    for name, src, ref, shape in (("glide_restrict_u", uf, O.restrict_u(uf), (nyc, nxc + 1)),
                                  ("glide_restrict_v", vf, O.restrict_v(vf), (nyc + 1, nxc)),
                                  ("glide_restrict_cell", hf, O.restrict_cell(hf), (nyc, nxc)),
                                  ("glide_restrict_max", hf, O.restrict_max(hf), (nyc, nxc))):
        out = torch.zeros(shape, dtype=torch.float64, device=DEV)
        _lib.call(name, P(T(src)), P(out), nyc, nxc, stream())
        assert np.array_equal(out.cpu().numpy(), ref), name
    uc, vc, hc = rng.normal(size=(nyc, nxc + 1)), rng.normal(size=(nyc + 1, nxc)), rng.normal(size=(nyc, nxc))
    nyf, nxf = 2 * nyc, 2 * nxc
    for name, src, ref, shape in (("glide_prolong_u", uc, O.prolong_u(uc), (nyf, nxf + 1)),
                                  ("glide_prolong_v", vc, O.prolong_v(vc), (nyf + 1, nxf)),
                                  ("glide_prolong_H", hc, O.prolong_H(hc), (nyf, nxf))):
        out = torch.zeros(shape, dtype=torch.float64, device=DEV)
        _lib.call(name, P(T(src)), P(out), nyf, nxf, stream())
        assert np.array_equal(out.cpu().numpy(), ref), name
    out = torch.zeros((nyf, nxf), dtype=torch.float64, device=DEV)
    _lib.call("glide_prolong_inject", P(T(hc)), P(out), nyf, nxf, -1.0, stream())
    assert np.array_equal(out.cpu().numpy(), O.prolong_inject(hc, -1.0))


def _pair(w, n_levels, **kw):
    ny, nx = w["bed"].shape
    h = O.Hierarchy(ny, nx, w["dx"], n_levels=n_levels, **kw)
    h.set_geometry(w["bed"], w["H"])
    h.set_parameters(B=w["B"], beta=w["beta"], smb=w["smb"])
    ph = IcePhysics(ny, nx, w["dx"], n_levels=n_levels, **kw)
    ph.set_geometry(w["bed"], w["H"])
    ph.set_parameters(B=w["B"], beta=w["beta"], smb=w["smb"])
    return h, ph


def test_one_vcycle_bitwise_small():
    w = S.ice_cap(64, 64, 2000.0, seed=3, radius=45e3)
    h, ph = _pair(w, 3)
    h.set_schedule(pre=2, post=2, coarse=5, local_passes=2, local_sweeps=2)
    ph.set_schedule(pre=2, post=2, coarse=5, local_passes=2, local_sweeps=2)
    (u0, v0, H0) = h.forward(10.0, n_vcycles=1)
    u, v, H = ph.forward(10.0, n_vcycles=1)
    torch.cuda.synchronize()
    for a, b in ((u, u0), (v, v0), (H, H0)):
        assert np.array_equal(a.cpu().numpy(), b), np.abs(a.cpu().numpy() - b).max()


def test_forward_dome_128_five_steps():
    """Workload dome_128 (128 x 128, dx 2 km, 4 levels, 3 V-cycles, dt 10 a, 5
    steps): relative L2 on u, v, H <= 1e-10 against the oracle."""
    w = S.workload("dome_128")
    h, ph = _pair(w, w["n_levels"])
    for step in range(w["n_steps"]):
        u0, v0, H0 = h.forward(w["dt"], n_vcycles=w["n_vcycles"])
        u, v, H = ph.forward(w["dt"], n_vcycles=w["n_vcycles"])
        errs = [rel(a.cpu().numpy(), b) for a, b in ((u, u0), (v, v0), (H, H0))]
        print(f"step {step}: rel L2 u {errs[0]:.2e} v {errs[1]:.2e} H {errs[2]:.2e}", flush=True)
        assert max(errs) <= 1e-10, errs


def test_boundary_weight_vcycles_bitwise():
    """Halved boundary rows of the restricted residual: GPU V-cycles equal
    the oracle's bit for bit."""
    w = S.ice_sheet(64, 64, 1500.0, seed=20000)
    h, ph = _pair(w, 3)
    sched = dict(pre=4, post=4, coarse=20, local_passes=2, local_sweeps=2, boundary_weight=0.5)
    h.set_schedule(**sched)
    ph.set_schedule(**sched)
    u0, v0, H0 = h.forward(1.0, n_vcycles=2)
    u, v, H = ph.forward(1.0, n_vcycles=2)
    torch.cuda.synchronize()
    for a, b in ((u, u0), (v, v0), (H, H0)):
        assert np.array_equal(a.cpu().numpy(), b), np.abs(a.cpu().numpy() - b).max()


def test_boundary_weight_converges_where_the_reference_stalls():
    """128^2 ice sheet, 4 levels, dt = 1 a, full schedule: to 1e-8 with
    boundary_weight 0.5 (each V-cycle's state equal to the oracle's), while
    the reference restriction stays above 1e-6."""
    w = S.ice_sheet(128, 128, 1500.0, seed=20000)
    h, ph = _pair(w, 4)
    h.set_schedule(boundary_weight=0.5)
    ph.set_schedule(boundary_weight=0.5)
    (u0, v0, H0), hist0 = h.forward(1.0, n_vcycles=20, tol=1e-8, history=True)
    u, v, H = ph.forward(1.0, tol=1e-8, max_vcycles=20)
    torch.cuda.synchronize()
    print("weighted:", len(ph.last_history), ph.last_history[-1], flush=True)
    assert len(ph.last_history) == len(hist0) <= 14 and ph.last_history[-1] <= 1e-8
    np.testing.assert_allclose(ph.last_history, hist0, rtol=1e-9)
    for a, b in ((u, u0), (v, v0), (H, H0)):
        assert np.array_equal(a.cpu().numpy(), b)
    _, ref = _pair(w, 4)
    ref.forward(1.0, tol=1e-8, max_vcycles=20)
    assert len(ref.last_history) == 20 and min(ref.last_history) > 1e-6


@pytest.mark.parametrize("shape", [(8, 10), (64, 64), (130, 70)])
def test_residual_norm_matches_oracle(shape):
    """glide_residual_norm against the oracle's norm (same residual, FP64
    sums in another order) on random states with cells at the thickness
    limit, so that the FB branch matters (glide/physics.py:176-203)."""
    ny, nx = shape
    s = random_state(ny, nx, seed=ny + nx)
    H = s["H"].copy()
    H.ravel()[::7] = 0.1
    s["U"][-ny * nx:] = H.ravel()
    ph = IcePhysics(ny, nx, s["dx"], n_levels=1)
    ph.set_geometry(s["bed"], H)
    ph.set_parameters(B=s["B"], beta=s["beta"], smb=0.0)
    for g in ph.grids:
        g.dt = s["dt"]
    ph.grid.U.copy_(T(s["U"]))
    ph.grid.f.copy_(T(s["f"]))
    h = O.Hierarchy(ny, nx, s["dx"], n_levels=1)
    h.set_geometry(s["bed"], H)
    h.set_parameters(B=s["B"], beta=s["beta"], smb=0.0)
    h.forward(s["dt"], n_vcycles=0, update_geometry=False)
    h.field("U")[:] = s["U"]
    h.field("f")[:] = s["f"]
    ref = h.residual_norm()
    got = ph.residual_norm()
    print(shape, got, ref, flush=True)
    assert abs(got - ref) <= 1e-13 * ref
    r = ph.grid.r.cpu().numpy()
    r0 = O.residual(ny, nx, s["dx"], s["dt"], O.params(), s["U"], s["f"], s["bed"], s["B"], s["beta"],
                    np.zeros((ny, nx)), np.full((ny, nx), 0.1))
    assert np.array_equal(r, r0)
