"""glide_vanka_sweep_noalloc (Vanka scratch in a static device array)
against glide_vanka_sweep (caller-provided scratch): bitwise equal states on
two calls in a row, the second on a smaller level, full and local sweeps;
sizes beyond the static scratch are rejected before any device work."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import random_state

torch = pytest.importorskip("torch")

from glide_amd import _lib  # noqa: E402

DEV = torch.device("cuda")


def test_rejects_levels_beyond_the_static_scratch():
    g = _lib.GlideGeom(1024, 1025, 1000.0, 1.0)
    p = _lib.GlideParams(3.0, 1e-5, 1e-3, 1.0, 0.1, 0)
    with pytest.raises(_lib.GlideError) as e:
        _lib.call("glide_vanka_sweep_noalloc", ctypes.byref(g), ctypes.byref(p), 8, 8, 8, 8, 8, 8, None, 1, 10, None)
    assert e.value.code == _lib.GLIDE_E_SIZE and "static scratch" in str(e.value)


def _sweep(name, s, active, n_sweeps):
    t = {k: torch.as_tensor(np.ascontiguousarray(s[k], dtype=np.float64), device=DEV)
         for k in ("U", "f", "bed", "B", "beta", "gamma")}
    g = _lib.GlideGeom(s["ny"], s["nx"], s["dx"], s["dt"])
    prm = O.params()
    p = _lib.GlideParams(prm.n, prm.eps_reg, prm.water_drag, prm.calving_rate, prm.gl_sigmoid_c, prm.gl_derivatives)
    act = None
    if active is not None:
        act_t = torch.as_tensor(active, dtype=torch.uint8, device=DEV)
        act = ctypes.c_void_p(act_t.data_ptr())
    ptrs = [ctypes.c_void_p(t[k].data_ptr()) for k in ("U", "f", "bed", "B", "beta", "gamma")]
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if name == "glide_vanka_sweep":
        scratch = torch.zeros(5 * s["ny"] * s["nx"], dtype=torch.float64, device=DEV)
        _lib.call(name, ctypes.byref(g), ctypes.byref(p), *ptrs, act, n_sweeps, 10, ctypes.c_void_p(scratch.data_ptr()),
                  st)
    else:
        _lib.call(name, ctypes.byref(g), ctypes.byref(p), *ptrs, act, n_sweeps, 10, st)
    torch.cuda.synchronize()
    return t["U"].cpu().numpy()


@pytest.mark.gpu
@pytest.mark.skipif(not torch.cuda.is_available(), reason="no GPU")
def test_noalloc_sweep_equals_sweep_on_two_calls():
    for k, (ny, nx) in enumerate([(96, 130), (40, 56)]):       # second call on a smaller level
        s = random_state(ny, nx, seed=70 + k)
        nby, nbx = ny // 14 + 1, nx // 14 + 1
        active = (np.arange(nby * nbx) % 3 != 1).astype(np.uint8)
        for act, n in ((None, 2), (active, 3)):
            print(f"call {k}: {ny}x{nx}, {'local' if act is not None else 'full'} sweeps x{n}", flush=True)
            a = _sweep("glide_vanka_sweep", s, act, n)
            b = _sweep("glide_vanka_sweep_noalloc", s, act, n)
            assert np.array_equal(a, b)
            assert not np.array_equal(a, s["U"])
