"""The C-ABI library loads and exports every entry point include/glide_hip.h
declares, and the ctypes table binds exactly those (no GPU work here)."""
import ctypes
import os
import re

import pytest

from glide_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


HEADERS = ("glide_hip.h", "glide_unet.h", "glide_io.h")


def declared():
    names = set()
    for h in HEADERS:
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(glide_[A-Za-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    assert sorted(_lib.SIGNATURES) == declared()


def test_errors_are_reported_without_a_gpu():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    # argument validation runs before any device work
    with pytest.raises(_lib.GlideError) as e:
        _lib.call("glide_residual", None, None, None, None, None, None, None, None, None, None, None, None)
    assert e.value.code == _lib.GLIDE_E_NULL
    g = _lib.GlideGeom(0, 4, 1.0, 1.0)
    p = _lib.GlideParams(3.0, 1e-5, 1e-3, 1.0, 0.1, 0)
    with pytest.raises(_lib.GlideError) as e:
        _lib.call("glide_residual", ctypes.byref(g), ctypes.byref(p), 1, 1, None, None, 1, 1, 1, 1, 1, None)
    assert e.value.code == _lib.GLIDE_E_SIZE
    g = _lib.GlideGeom(8, 8, -1.0, 1.0)
    with pytest.raises(_lib.GlideError) as e:
        _lib.call("glide_residual", ctypes.byref(g), ctypes.byref(p), 1, 1, None, None, 1, 1, 1, 1, 1, None)
    assert e.value.code == _lib.GLIDE_E_ARG
    n = ctypes.c_ulonglong()
    with pytest.raises(_lib.GlideError) as e:
        _lib.call("glide_count_nonfinite", None, 4, None, ctypes.byref(n), None)
    assert e.value.code == _lib.GLIDE_E_NULL
    with pytest.raises(_lib.GlideError) as e:
        _lib.call("glide_count_nonfinite", 8, -1, 8, ctypes.byref(n), None)
    assert e.value.code == _lib.GLIDE_E_SIZE
    assert _lib.load().glide_version().startswith(b"glide_amd")


def test_last_error_copies_and_truncates():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    with pytest.raises(_lib.GlideError):
        _lib.call("glide_residual", None, None, None, None, None, None, None, None, None, None, None, None)
    f = _lib.fn("glide_last_error")
    n = f(None, 0)
    full = _lib.last_error()
    assert n == len(full) > 0 and "NULL" in full
    buf = ctypes.create_string_buffer(8)
    assert f(buf, 8) == n
    assert buf.value == full.encode()[:7]


def fake_levels(sizes):
    """Level descriptors with non-NULL dummy pointers (never dereferenced:
    the calls below fail validation before any device work)."""
    lv = (_lib.GlideLevel * len(sizes))()
    for k, n in enumerate(sizes):
        lv[k].g = _lib.GlideGeom(n, n, 1000.0 * (1 << k), 1.0)
        lv[k].p = _lib.GlideParams(3.0, 1e-5, 1e-3, 1.0, 0.1, 0)
        for name in _lib._LEVEL_PTRS:
            setattr(lv[k], name, 8)
    return lv


def test_vcycle_schedule_validation_without_a_gpu():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lv = fake_levels((16, 8))
    sc = _lib.schedule(local_skip=True)
    with pytest.raises(_lib.GlideError) as e:     # the idle-pass check needs a workspace
        _lib.call("glide_fascd_vcycle", lv, 2, 0.1, ctypes.byref(sc), None, None)
    assert e.value.code == _lib.GLIDE_E_NULL and "GlideVcycleWork" in str(e.value)
    sc = _lib.schedule(pre=-1, local_skip=False)
    with pytest.raises(_lib.GlideError) as e:
        _lib.call("glide_fascd_vcycle", lv, 2, 0.1, ctypes.byref(sc), None, None)
    assert e.value.code == _lib.GLIDE_E_ARG
    sc = _lib.schedule(boundary_weight=float("nan"), local_skip=False)
    with pytest.raises(_lib.GlideError) as e:
        _lib.call("glide_fascd_vcycle", lv, 2, 0.1, ctypes.byref(sc), None, None)
    assert e.value.code == _lib.GLIDE_E_ARG
    with pytest.raises(_lib.GlideError) as e:
        _lib.call("glide_vcycle_work_counts", None, None, None, 0)
    assert e.value.code == _lib.GLIDE_E_NULL


def test_residual_norm_rejects_null_buffers_before_launching():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lv = fake_levels((16,))
    out = ctypes.c_double()
    for name in ("r", "U", "f", "gamma", "scratch"):
        lv[0].__setattr__(name, None)
        with pytest.raises(_lib.GlideError) as e:
            _lib.call("glide_residual_norm", ctypes.byref(lv[0]), ctypes.byref(out), None)
        assert e.value.code == _lib.GLIDE_E_NULL
        lv[0].__setattr__(name, 8)


def test_batch_errors_are_reported_without_a_gpu():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    out = ctypes.c_void_p()
    with pytest.raises(_lib.GlideError) as e:
        _lib.call("glide_batch_create", None, None, 1, ctypes.byref(out))
    assert e.value.code == _lib.GLIDE_E_NULL
    # a member whose level 1 is not a 2x coarsening of level 0
    lv = (_lib.GlideLevel * 2)()
    for k, n in enumerate((16, 6)):
        lv[k].g = _lib.GlideGeom(n, n, 1.0, 1.0)
        for name in ("U", "f", "r", "w", "bed", "B", "beta", "gamma", "chi", "mask", "scratch", "H_prev", "smb"):
            setattr(lv[k], name, 8)   # never dereferenced: validation fails first
        lv[k].active = 8
    ptrs = (ctypes.POINTER(_lib.GlideLevel) * 1)(ctypes.cast(lv, ctypes.POINTER(_lib.GlideLevel)))
    nlev = (ctypes.c_int * 1)(2)
    with pytest.raises(_lib.GlideError) as e:
        _lib.call("glide_batch_create", ptrs, nlev, 1, ctypes.byref(out))
    assert e.value.code == _lib.GLIDE_E_SIZE and "coarsening" in str(e.value)
    nlev[0] = 0
    with pytest.raises(_lib.GlideError) as e:
        _lib.call("glide_batch_create", ptrs, nlev, 1, ctypes.byref(out))
    assert e.value.code == _lib.GLIDE_E_SIZE
    for name, args in (("glide_batch_vcycle", (None, 0.1, None, None)),
                       ("glide_batch_set_tiles", (None, 0, 0, 4, 0)),
                       ("glide_batch_forward_setup", (None, 1.0, None))):
        with pytest.raises(_lib.GlideError) as e:
            _lib.call(name, *args)
        assert e.value.code == _lib.GLIDE_E_NULL
