"""ctypes binding of ``libglide_hip.so`` (the C ABI declared in include/glide_hip.h).

The library is required: importing anything that launches work raises if it
has not been built (``python -m glide_amd.build``) - there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libglide_hip.so")

GLIDE_E_NULL, GLIDE_E_SIZE, GLIDE_E_ARG, GLIDE_E_NONFINITE = -1, -2, -3, -4


class GlideError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


class GlideParams(ctypes.Structure):
    _fields_ = [("n", c_double), ("eps_reg", c_double), ("water_drag", c_double),
                ("calving_rate", c_double), ("gl_sigmoid_c", c_double), ("gl_derivatives", c_int)]


class GlideGeom(ctypes.Structure):
    _fields_ = [("ny", c_int), ("nx", c_int), ("dx", c_double), ("dt", c_double)]


DP = c_void_p   # device pointers travel as plain addresses

_LEVEL_PTRS = ["U", "f", "r", "w", "z", "Lambda", "Lambda_out", "f_adj", "r_adj",
               "H_prev", "bed", "B", "beta", "smb", "gamma", "chi", "mask", "scratch", "active"]


class GlideLevel(ctypes.Structure):
    _fields_ = [("g", GlideGeom), ("p", GlideParams)] + [(n, c_void_p) for n in _LEVEL_PTRS]


class GlideSchedule(ctypes.Structure):
    _fields_ = [("pre", c_int), ("post", c_int), ("coarse", c_int), ("local_passes", c_int),
                ("local_sweeps", c_int), ("n_newton", c_int), ("adj_sweeps", c_int), ("adj_coarse", c_int),
                ("boundary_weight", c_double), ("local_skip", c_int)]


def schedule(pre=10, post=10, coarse=200, local_passes=10, local_sweeps=10, n_newton=10, adj_sweeps=10,
             adj_coarse=100, boundary_weight=1.0, local_skip=True) -> GlideSchedule:
    """A GlideSchedule; the defaults are the reference's sweep counts and
    residual restriction, with the idle-local-pass check on."""
    return GlideSchedule(int(pre), int(post), int(coarse), int(local_passes), int(local_sweeps), int(n_newton),
                         int(adj_sweeps), int(adj_coarse), float(boundary_weight), int(bool(local_skip)))


class GlideVtiArray(ctypes.Structure):
    _fields_ = [("comp", c_void_p * 3), ("n_comp", c_int)]


# name -> (restype, argtypes); the list mirrors include/glide_hip.h, glide_unet.h and glide_io.h
SIGNATURES = {
    "glide_residual": (c_int, [POINTER(GlideGeom), POINTER(GlideParams), DP, DP, DP, DP, DP, DP, DP, DP, DP, c_void_p]),
    "glide_jvp": (c_int, [POINTER(GlideGeom), POINTER(GlideParams), DP, DP, DP, DP, DP, DP, DP, DP, c_void_p]),
    "glide_vjp": (c_int, [POINTER(GlideGeom), POINTER(GlideParams), DP, DP, DP, DP, DP, DP, DP, DP, c_void_p]),
    "glide_vanka_sweep": (c_int, [POINTER(GlideGeom), POINTER(GlideParams), DP, DP, DP, DP, DP, DP, DP,
                                  c_int, c_int, DP, c_void_p]),
    "glide_vanka_sweep_noalloc": (c_int, [POINTER(GlideGeom), POINTER(GlideParams), DP, DP, DP, DP, DP, DP, DP,
                                          c_int, c_int, c_void_p]),
    "glide_vanka_block": (c_int, [POINTER(GlideGeom), POINTER(GlideParams), DP, DP, DP, DP, DP, DP, c_int, DP,
                                  c_void_p]),
    "glide_local_activation": (c_int, [POINTER(GlideGeom), DP, DP, DP, DP, c_void_p]),
    "glide_vanka_adjoint_pass": (c_int, [POINTER(GlideGeom), POINTER(GlideParams), DP, DP, DP, DP, DP, DP, DP,
                                         DP, DP, c_int, c_double, c_void_p]),
    "glide_grad_beta": (c_int, [POINTER(GlideGeom), POINTER(GlideParams), DP, DP, DP, DP, DP, c_int, c_void_p]),
    "glide_grad_B": (c_int, [POINTER(GlideGeom), POINTER(GlideParams), DP, DP, DP, DP, c_void_p]),
    "glide_restrict_u": (c_int, [DP, DP, c_int, c_int, c_void_p]),
    "glide_restrict_v": (c_int, [DP, DP, c_int, c_int, c_void_p]),
    "glide_restrict_cell": (c_int, [DP, DP, c_int, c_int, c_void_p]),
    "glide_restrict_max": (c_int, [DP, DP, c_int, c_int, c_void_p]),
    "glide_prolong_inject": (c_int, [DP, DP, c_int, c_int, c_double, c_void_p]),
    "glide_prolong_u": (c_int, [DP, DP, c_int, c_int, c_void_p]),
    "glide_prolong_v": (c_int, [DP, DP, c_int, c_int, c_void_p]),
    "glide_prolong_H": (c_int, [DP, DP, c_int, c_int, c_void_p]),
    "glide_restrict_parameters": (c_int, [POINTER(GlideLevel), c_int, c_void_p]),
    "glide_propagate_geometry": (c_int, [POINTER(GlideLevel), c_int, c_double, c_void_p]),
    "glide_fascd_vcycle": (c_int, [POINTER(GlideLevel), c_int, c_double, POINTER(GlideSchedule), c_void_p,
                                   c_void_p]),
    "glide_vcycle_work_create": (c_int, [POINTER(c_void_p)]),
    "glide_vcycle_work_destroy": (c_int, [c_void_p]),
    "glide_vcycle_work_counts": (c_int, [c_void_p, POINTER(ctypes.c_long), POINTER(ctypes.c_long), c_int]),
    "glide_forward_setup": (c_int, [POINTER(GlideLevel), c_int, c_double, c_void_p]),
    "glide_residual_norm": (c_int, [POINTER(GlideLevel), POINTER(c_double), c_void_p]),
    "glide_adjoint_vcycle": (c_int, [POINTER(GlideLevel), c_int, POINTER(GlideSchedule), c_void_p]),
    "glide_compute_mask": (c_int, [POINTER(GlideLevel), c_void_p]),
    "glide_adjoint_residual": (c_int, [POINTER(GlideLevel), c_void_p]),
    "glide_batch_create": (c_int, [POINTER(POINTER(GlideLevel)), POINTER(c_int), c_int, POINTER(c_void_p)]),
    "glide_batch_destroy": (c_int, [c_void_p]),
    "glide_batch_info": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
    "glide_batch_forward_setup": (c_int, [c_void_p, c_double, c_void_p]),
    "glide_batch_vcycle": (c_int, [c_void_p, c_double, POINTER(GlideSchedule), c_void_p]),
    "glide_batch_residual_norms": (c_int, [c_void_p, POINTER(c_double), c_void_p]),
    "glide_batch_residual": (c_int, [c_void_p, c_int, c_void_p]),
    "glide_batch_set_enabled": (c_int, [c_void_p, POINTER(ctypes.c_ubyte), c_void_p]),
    "glide_batch_set_tiles": (c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    "glide_batch_time_kernel": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, POINTER(c_double), c_void_p]),
    "glide_batch_time_restriction": (c_int, [c_void_p, c_int, c_int, POINTER(c_double), POINTER(c_double), c_void_p]),
    "glide_unet_conv3x3": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_int, c_void_p,
                                   c_int, c_void_p, c_void_p]),
    "glide_unet_conv3x3_dgrad": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_int,
                                         c_int, c_void_p, c_int, c_void_p]),
    "glide_unet_wgrad_chunks": (ctypes.c_long, [c_int, c_int, c_int]),
    "glide_unet_conv3x3_wgrad": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_int,
                                         c_void_p, c_void_p, c_void_p]),
    "glide_unet_gemm_bf16": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "glide_unet_prep_weights": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "glide_unet_relu_bwd": (c_int, [c_void_p, c_void_p, ctypes.c_long, c_void_p, c_void_p]),
    "glide_unet_pool2": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p]),
    "glide_unet_pool2_bwd": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_int, c_void_p]),
    "glide_unet_up2": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p]),
    "glide_unet_up2_bwd": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p]),
    "glide_unet_bias_chunks": (ctypes.c_long, [ctypes.c_long]),
    "glide_unet_bias_grad": (c_int, [c_void_p, ctypes.c_long, c_int, c_void_p, c_void_p, c_void_p]),
    "glide_unet_head": (c_int, [c_void_p, ctypes.c_long, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "glide_unet_head_bwd": (c_int, [c_void_p, c_void_p, ctypes.c_long, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p]),
    "glide_unet_mse_blocks": (ctypes.c_long, [ctypes.c_long]),
    "glide_unet_mse": (c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_long, ctypes.c_float, c_void_p, c_void_p,
                               c_void_p]),
    "glide_unet_adam": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_long, ctypes.c_float,
                                ctypes.c_float, ctypes.c_float, ctypes.c_float, c_int, ctypes.c_float, c_void_p]),
    "glide_unet_features": (c_int, [DP, DP, DP, DP, c_int, c_int, c_double, c_void_p, c_void_p]),
    "glide_unet_targets": (c_int, [DP, DP, DP, DP, c_int, c_int, c_double, c_double, c_void_p, c_void_p]),
    "glide_unet_apply_prediction": (c_int, [c_void_p, c_int, c_int, c_double, c_double, DP, DP, DP, DP, DP,
                                            c_void_p]),
    "glide_set_persistence": (c_int, [c_int, c_int]),
    "glide_set_small_levels": (c_int, [ctypes.c_long, c_int]),
    "glide_set_tile_height": (c_int, [c_int, c_int]),
    "glide_count_nonfinite": (c_int, [DP, ctypes.c_long, c_void_p, POINTER(ctypes.c_ulonglong), c_void_p]),
    "glide_vti_block_bytes": (ctypes.c_size_t, [POINTER(GlideVtiArray), c_int, c_int, c_int]),
    "glide_vti_pack": (c_int, [POINTER(GlideVtiArray), c_int, c_int, c_int, c_int, c_void_p, ctypes.c_size_t,
                               c_void_p]),
    "glide_interp_grid": (c_int, [DP, c_int, c_int, DP, c_int, DP, c_int, DP, c_int, DP, c_int, c_double, DP,
                                  c_void_p]),
    "glide_bedmachine_thickness": (c_int, [DP, DP, ctypes.c_long, c_double, DP, c_void_p]),
    "glide_last_error": (ctypes.c_size_t, [ctypes.c_char_p, ctypes.c_size_t]),
    "glide_version": (c_char_p, []),
}

_lib = None


def load():
    """Load the library (once).  Raises if it is missing: no fallback path."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing; build it with `python -m glide_amd.build`")
    # torch carries its own HIP runtime with the same soname (libamdhip64.so.7);
    # loading torch first makes this library bind to that one instance, so
    # device pointers and streams from torch are valid here.
    import torch  # noqa: F401
    lib = ctypes.CDLL(LIB_PATH)
    for name in ("glide_last_error", "glide_version"):
        f = getattr(lib, name)
        f.restype, f.argtypes = SIGNATURES[name]
    _lib = lib
    return lib


_bound: dict = {}


def fn(name: str):
    """The typed ctypes function for entry point ``name``."""
    f = _bound.get(name)
    if f is None:
        f = getattr(load(), name)
        f.restype, f.argtypes = SIGNATURES[name]
        _bound[name] = f
    return f


def last_error() -> str:
    """Message of the last failed call on this thread ("" if none)."""
    f = fn("glide_last_error")
    n = f(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    f(buf, n + 1)
    return buf.value.decode(errors="replace")


def call(name: str, *args) -> None:
    """Call a status-returning entry point and raise GlideError on failure."""
    rc = fn(name)(*args)
    if rc != 0:
        raise GlideError(name, rc, last_error())


class VcycleWork:
    """Owner of a GlideVcycleWork (the single-domain V-cycle's idle-pass
    check: pinned flag slots, events, pass counters), created for the
    current device and destroyed with this object."""

    def __init__(self):
        h = c_void_p()
        call("glide_vcycle_work_create", ctypes.byref(h))
        self.handle = h

    def counts(self, reset: bool = False):
        """(local passes run, local passes skipped) since the last reset."""
        run, skipped = ctypes.c_long(0), ctypes.c_long(0)
        call("glide_vcycle_work_counts", self.handle, ctypes.byref(run), ctypes.byref(skipped), int(bool(reset)))
        return run.value, skipped.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and _lib is not None:
            fn("glide_vcycle_work_destroy")(h)
            self.handle = None
