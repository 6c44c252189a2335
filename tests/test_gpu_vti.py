"""Snapshot files written through the GPU packer (glide_amd.io, libglide_hip.so
glide_vti_pack) are byte-identical to the oracle's restatement of the
reference writer (oracle/vti_oracle.py, glide/io.py:20-202), including the
FP64 -> float32 rounding corner cases, a series with its manifest written
from a solve as the reference examples do, and the error paths."""
import ctypes
import time

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

from glide_amd import _lib  # noqa: E402
from glide_amd import io as gio  # noqa: E402
from glide_amd import synthetic as S  # noqa: E402
from glide_amd.physics import IcePhysics  # noqa: E402
from oracle import vti_oracle as V  # noqa: E402

DEV = torch.device("cuda")

CORNERS = [1.0 + 2.0 ** -24,          # tie: rounds to even (1.0)
           1.0 + 3 * 2.0 ** -24,      # tie: rounds to even (1 + 2^-22)
           1e-40, -3e-42,             # float32 subnormals
           1e-310,                    # float64 subnormal -> 0
           3.4028235e38, 3.4028236e38, 1e39, -1e39,   # largest float32, overflow to +-inf
           np.nan, 0.1, -0.0]


def fields(rng, ny, nx, scale=100.0):
    a = rng.standard_normal((ny, nx)) * scale
    flat = a.reshape(-1)
    flat[: len(CORNERS)] = CORNERS
    return a


def on_dev(data):
    return {k: ([torch.as_tensor(c, device=DEV) for c in v] if isinstance(v, list) else torch.as_tensor(v, device=DEV))
            for k, v in data.items()}


@pytest.mark.parametrize("flip,time_value", [(True, None), (True, 4.5), (False, 0.0)])
def test_bytes_match_oracle(tmp_path, flip, time_value):
    rng = np.random.default_rng(11)
    ny, nx = 37, 53
    data = {"vel": [fields(rng, ny, nx), fields(rng, ny, nx)], "thk": fields(rng, ny, nx),
            "w": [fields(rng, ny, nx) for _ in range(3)], "srf": fields(rng, ny, nx, 1e4)}
    path = gio.write_vti(tmp_path / "s.vti", on_dev(data), 1500.0, origin=(-7.0, 3.0), time_value=time_value,
                         flip_y=flip)
    got = path.read_bytes()
    want = V.vti_bytes(data, 1500.0, origin=(-7.0, 3.0), time_value=time_value, flip_y=flip)
    assert len(got) == len(want)
    assert got == want


def test_numpy_inputs_and_strided_views(tmp_path):
    rng = np.random.default_rng(12)
    u = torch.as_tensor(rng.standard_normal((16, 25)), device=DEV)
    v = torch.as_tensor(rng.standard_normal((17, 24)), device=DEV)
    u_c = u[:, 1:]                       # a strided view, as glide/physics.py:255-259 slices faces
    v_c = v[1:]
    H = rng.uniform(0, 100, (16, 24))    # host array
    gio.write_vti(tmp_path / "a.vti", {"H": H, "vel": [u_c, v_c]}, 2.0)
    want = V.vti_bytes({"H": H, "vel": [u_c.cpu().numpy(), v_c.cpu().numpy()]}, 2.0)
    assert (tmp_path / "a.vti").read_bytes() == want
    with pytest.raises(ValueError):
        gio.write_vti(tmp_path / "b.vti", {"H": H, "u": u}, 2.0)     # shapes differ


def test_series_from_a_solve(tmp_path):
    w = S.workload("dome_128")
    ph = IcePhysics(w["ny"], w["nx"], w["dx"], n_levels=w["n_levels"])
    ph.set_geometry(w["bed"], w["H"])
    ph.set_parameters(B=w["B"], beta=w["beta"], smb=w["smb"])
    writer = gio.VTIWriter(tmp_path, base="dome", dx=w["dx"])
    gio.write_vti(tmp_path / "bed.vti", {"bed": ph.grid.bed}, w["dx"])
    expect = {"bed.vti": V.vti_bytes({"bed": ph.grid.bed.cpu().numpy()}, w["dx"])}
    t = 0.0
    for step in range(2):
        _, _, H = ph.forward(w["dt"], n_vcycles=1)
        t += w["dt"]
        u_c, v_c = ph.get_velocities_cell_centered()
        srf = ph.get_surface()
        writer.write_step(step, t, {"thk": H, "srf": srf, "vel": [u_c, v_c]})
        writer.write_pvd()
        expect[f"dome_{step:04d}.vti"] = V.vti_bytes(
            {"thk": H.cpu().numpy(), "srf": srf.cpu().numpy(), "vel": [u_c.cpu().numpy(), v_c.cpu().numpy()]},
            w["dx"], time_value=t)
    for name, blob in expect.items():
        assert (tmp_path / name).read_bytes() == blob, name
    assert (tmp_path / "dome.pvd").read_text() == V.pvd_text([(w["dt"], "dome_0000.vti"),
                                                               (2 * w["dt"], "dome_0001.vti")])


def test_pack_rejects_bad_layouts():
    x = torch.zeros(4, 4, dtype=torch.float64, device=DEV)
    out = torch.empty(1024, dtype=torch.uint8, device=DEV)
    arr = (_lib.GlideVtiArray * 1)()
    arr[0].comp[0] = x.data_ptr()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for n_comp, n_out, code in ((4, 1024, -2), (1, 10, -2), (2, 1024, -1)):
        arr[0].n_comp = n_comp
        with pytest.raises(_lib.GlideError) as e:
            _lib.call("glide_vti_pack", arr, 1, 4, 4, 1, ctypes.c_void_p(out.data_ptr()), n_out, st)
        assert e.value.code == code, (n_comp, n_out)
    arr[0].n_comp = 1
    assert _lib.fn("glide_vti_block_bytes")(arr, 1, 4, 4) == 4 + 64
    assert _lib.fn("glide_vti_block_bytes")(arr, 0, 4, 4) == 0


def test_snapshot_1024_timing(tmp_path):
    """thk, srf and vel = [u_c, v_c] at 1024^2: the GPU-packed write against
    the NumPy writer on the same fields (already on the host), both to the
    same file system; the pack kernel alone timed with HIP events."""
    rng = np.random.default_rng(13)
    n = 1024
    host = {"thk": rng.uniform(0, 3000, (n, n)), "srf": rng.uniform(0, 3000, (n, n)),
            "vel": [rng.standard_normal((n, n)), rng.standard_normal((n, n))]}
    dev = on_dev(host)
    gio.write_vti(tmp_path / "warm.vti", dev, 1000.0)
    reps = 5
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(reps):
        gio.write_vti(tmp_path / f"g{k}.vti", dev, 1000.0, time_value=k)
    t_gpu = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for k in range(2):
        (tmp_path / f"c{k}.vti").write_bytes(V.vti_bytes(host, 1000.0, time_value=k))
    t_cpu = (time.perf_counter() - t0) / 2
    assert (tmp_path / "g1.vti").read_bytes() == (tmp_path / "c1.vti").read_bytes()
    comps = [dev["thk"], dev["srf"], dev["vel"][0], dev["vel"][1]]
    arr = (_lib.GlideVtiArray * 3)()
    for a, cs in enumerate([[comps[0]], [comps[1]], comps[2:]]):
        arr[a].n_comp = len(cs)
        for q, c in enumerate(cs):
            arr[a].comp[q] = c.data_ptr()
    nbytes = _lib.fn("glide_vti_block_bytes")(arr, 3, n, n)
    out = torch.empty(nbytes, dtype=torch.uint8, device=DEV)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(20):
        _lib.call("glide_vti_pack", arr, 3, n, n, 1, ctypes.c_void_p(out.data_ptr()), nbytes,
                  ctypes.c_void_p(st.cuda_stream))
    e1.record(st)
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) / 20 * 1e3
    moved = 4 * n * n * 8 + nbytes
    print(f"1024^2 snapshot (4 components, {nbytes / 1e6:.1f} MB file block): GPU-packed write "
          f"{t_gpu * 1e3:.1f} ms, NumPy writer {t_cpu * 1e3:.1f} ms; pack kernel {us:.1f} us = "
          f"{moved / us / 1e3:.0f} GB/s of {moved / 1e6:.1f} MB", flush=True)
