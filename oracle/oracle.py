"""ORACLE - test infrastructure only.

ctypes wrapper around ``oracle/_build/liboracle.so``, the FP64 CPU
restatement of GLIDE's SSA/FASCD solver (see ``glide_oracle.c``).  Only the
test suite, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
load this module, and only as the reference results / timed CPU baseline; the product
package ``glide_amd`` never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")
_LIBS: dict = {}

DP = POINTER(c_double)
UCP = POINTER(ctypes.c_ubyte)


class OParams(ctypes.Structure):
    _fields_ = [("n", c_double), ("eps_reg", c_double), ("water_drag", c_double),
                ("calving_rate", c_double), ("gl_sigmoid_c", c_double), ("gl_derivatives", c_int)]


class OGeom(ctypes.Structure):
    _fields_ = [("ny", c_int), ("nx", c_int), ("dx", c_double), ("dt", c_double)]


def build() -> None:
    """Compile the oracle libraries with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib(libm: bool = False):
    key = "libm" if libm else "det"
    if key in _LIBS:
        return _LIBS[key]
    name = "liboracle_libm.so" if libm else "liboracle.so"
    path = os.path.join(_BUILD, name)
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    P, G = POINTER(OParams), POINTER(OGeom)
    sig = {
        "or_eta": (None, [G, P, DP, DP, DP]),
        "or_residual": (None, [G, P, DP, DP, DP, DP, DP, DP, DP, DP, DP]),
        "or_jvp": (None, [G, P, DP, DP, DP, DP, DP, DP, DP, DP, c_void_p]),
        "or_vjp": (None, [G, P, DP, DP, DP, DP, DP, DP, DP, DP, c_void_p]),
        "or_lu5": (None, [DP, DP, DP]),
        "or_vanka_block": (None, [G, P, DP, DP, DP, DP, DP, DP, c_int, c_int, c_int, c_int, DP, DP]),
        "or_vanka_sweep": (None, [G, P, DP, DP, DP, DP, DP, DP, UCP, c_int, DP, DP]),
        "or_local_activation": (None, [c_int, c_int, DP, UCP]),
        "or_error_mask": (None, [G, DP, DP, DP, DP]),
        "or_vanka_adjoint_pass": (None, [G, P, DP, DP, DP, DP, DP, DP, DP, DP, DP, c_int, c_double, DP]),
        "or_grad_beta": (None, [G, P, DP, DP, DP, DP, DP, c_int]),
        "or_grad_B": (None, [G, P, DP, DP, DP, DP]),
        "or_restrict_u": (None, [DP, DP, c_int, c_int]),
        "or_restrict_v": (None, [DP, DP, c_int, c_int]),
        "or_restrict_cell": (None, [DP, DP, c_int, c_int]),
        "or_restrict_max": (None, [DP, DP, c_int, c_int]),
        "or_prolong_inject": (None, [DP, DP, c_int, c_int, c_double]),
        "or_prolong_u": (None, [DP, DP, c_int, c_int]),
        "or_prolong_v": (None, [DP, DP, c_int, c_int]),
        "or_prolong_H": (None, [DP, DP, c_int, c_int]),
        "oh_create": (c_void_p, [c_int, c_int, c_double, c_int, P, c_double]),
        "oh_destroy": (None, [c_void_p]),
        "oh_set_schedule": (None, [c_void_p] + [c_int] * 8),
        "oh_set_boundary_weight": (None, [c_void_p, c_double]),
        "oh_field": (DP, [c_void_p, c_int, c_int]),
        "oh_set_geometry": (None, [c_void_p, DP, DP]),
        "oh_set_parameters": (None, [c_void_p, DP, DP, DP]),
        "oh_restrict_parameters": (None, [c_void_p]),
        "oh_vcycle": (None, [c_void_p]),
        "oh_residual_norm": (c_double, [c_void_p]),
        "oh_forward": (c_int, [c_void_p, c_double, c_int, c_double, DP, c_int]),
        "oh_adjoint_vcycle": (None, [c_void_p]),
        "oh_adjoint": (None, [c_void_p, DP, DP, DP, c_int, DP, c_int]),
        "oh_adjoint_residual_norm": (c_double, [c_void_p]),
        "oh_num_threads": (c_int, []),
        "oh_set_threads": (None, [c_int]),
        "or_exp": (c_double, [c_double]),
        "or_log": (c_double, [c_double]),
        "or_pow": (c_double, [c_double, c_double]),
    }
    for name_, (res, args) in sig.items():
        fn = getattr(lib, name_)
        fn.restype = res
        fn.argtypes = args
    _LIBS[key] = lib
    return lib


def _p(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(DP)


def params(n=3.0, eps_reg=1e-5, water_drag=1e-3, calving_rate=1.0, gl_sigmoid_c=0.1,
           gl_derivatives=False) -> OParams:
    return OParams(n, eps_reg, water_drag, calving_rate, gl_sigmoid_c, int(bool(gl_derivatives)))


def sizes(ny: int, nx: int):
    nu, nv, nc = ny * (nx + 1), (ny + 1) * nx, ny * nx
    return nu, nv, nc, nu + nv + nc


def split(vec: np.ndarray, ny: int, nx: int):
    """Views (u, v, H) of a state vector [u | v | H] (glide/grid.py:160-165)."""
    nu, nv, nc, _ = sizes(ny, nx)
    return (vec[:nu].reshape(ny, nx + 1), vec[nu:nu + nv].reshape(ny + 1, nx),
            vec[nu + nv:].reshape(ny, nx))


def join(u, v, H) -> np.ndarray:
    return np.ascontiguousarray(np.concatenate([np.ravel(u), np.ravel(v), np.ravel(H)]), dtype=np.float64)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ------------------------------------------------------------------ kernels
def eta(ny, nx, dx, prm, U, B, libm=False):
    out = np.zeros(ny * nx)
    _lib(libm).or_eta(ctypes.byref(OGeom(ny, nx, dx, 1.0)), ctypes.byref(prm), _p(out), _p(_c(U)), _p(_c(B).ravel()))
    return out.reshape(ny, nx)


def residual(ny, nx, dx, dt, prm, U, f, bed, B, beta, mask, gamma, libm=False):
    """r = F(U) - f on one level (residuals.cu:6-292)."""
    _, _, nc, nt = sizes(ny, nx)
    r = np.zeros(nt)
    work = np.zeros(nc)
    _lib(libm).or_residual(ctypes.byref(OGeom(ny, nx, dx, dt)), ctypes.byref(prm), _p(r), _p(_c(U)), _p(_c(f)),
                           _p(_c(bed).ravel()), _p(_c(B).ravel()), _p(_c(beta).ravel()), _p(_c(mask).ravel()),
                           _p(_c(gamma).ravel()), _p(work))
    return r


def jvp(ny, nx, dx, dt, prm, U, dU, bed, B, beta, mask, gamma, libm=False):
    _, _, nc, nt = sizes(ny, nx)
    out = np.zeros(nt)
    work = np.zeros(2 * nc)
    _lib(libm).or_jvp(ctypes.byref(OGeom(ny, nx, dx, dt)), ctypes.byref(prm), _p(out), _p(_c(U)), _p(_c(dU)),
                      _p(_c(bed).ravel()), _p(_c(B).ravel()), _p(_c(beta).ravel()), _p(_c(mask).ravel()),
                      _p(_c(gamma).ravel()), work.ctypes.data_as(c_void_p))
    return out


def vjp(ny, nx, dx, dt, prm, U, lam, bed, B, beta, mask, gamma, libm=False):
    _, _, nc, nt = sizes(ny, nx)
    out = np.zeros(nt)
    work = np.zeros(2 * nc)
    _lib(libm).or_vjp(ctypes.byref(OGeom(ny, nx, dx, dt)), ctypes.byref(prm), _p(out), _p(_c(U)), _p(_c(lam)),
                      _p(_c(bed).ravel()), _p(_c(B).ravel()), _p(_c(beta).ravel()), _p(_c(mask).ravel()),
                      _p(_c(gamma).ravel()), work.ctypes.data_as(c_void_p))
    return out


def lu5(A, b):
    A = _c(A).reshape(25)
    b = _c(b).reshape(5)
    x = np.zeros(5)
    _lib().or_lu5(_p(A), _p(b), _p(x))
    return x


def vanka_block(ny, nx, dx, dt, prm, U, f, bed, B, beta, gamma, i, j, mode=0, adj_masked=0, eta_field=None):
    """Local 5x5 (J, r) of cell (i, j) at the current state; eta is frozen at
    ``eta_field`` (default: the viscosity of U itself)."""
    e = eta(ny, nx, dx, prm, U, B) if eta_field is None else _c(eta_field)
    J = np.zeros(25)
    r = np.zeros(5)
    _lib().or_vanka_block(ctypes.byref(OGeom(ny, nx, dx, dt)), ctypes.byref(prm), _p(_c(U)), _p(_c(f)),
                          _p(_c(bed).ravel()), _p(_c(beta).ravel()), _p(_c(gamma).ravel()), _p(_c(e).ravel()),
                          i, j, mode, adj_masked, _p(J), _p(r))
    return J.reshape(5, 5), r


def vanka_sweep(ny, nx, dx, dt, prm, U, f, bed, B, beta, gamma, active=None, n_newton=10, libm=False):
    """One forward Vanka sweep; returns the new state vector."""
    _, _, nc, _ = sizes(ny, nx)
    Uo = _c(U).copy()
    act = None
    if active is not None:
        act_arr = np.ascontiguousarray(active, dtype=np.uint8).ravel()
        act = act_arr.ctypes.data_as(UCP)
    _lib(libm).or_vanka_sweep(ctypes.byref(OGeom(ny, nx, dx, dt)), ctypes.byref(prm), _p(Uo), _p(_c(f)),
                              _p(_c(bed).ravel()), _p(_c(B).ravel()), _p(_c(beta).ravel()), _p(_c(gamma).ravel()),
                              act, n_newton, _p(np.zeros(nc)), _p(np.zeros(5 * nc)))
    return Uo


def local_activation(error_mask):
    em = _c(error_mask)
    ny, nx = em.shape
    act = np.zeros(ny * nx, dtype=np.uint8)
    _lib().or_local_activation(ny, nx, _p(em.ravel()), act.ctypes.data_as(UCP))
    return act.reshape(ny, nx)


def error_mask(ny, nx, r, U, gamma):
    out = np.zeros(ny * nx)
    _lib().or_error_mask(ctypes.byref(OGeom(ny, nx, 1.0, 1.0)), _p(_c(r)), _p(_c(U)), _p(_c(gamma).ravel()), _p(out))
    return out.reshape(ny, nx)


def vanka_adjoint_pass(ny, nx, dx, dt, prm, lam, mask, r_adj, U, bed, B, beta, gamma, color, omega=1.0):
    _, _, nc, _ = sizes(ny, nx)
    lam = _c(lam)
    out = lam.copy()
    _lib().or_vanka_adjoint_pass(ctypes.byref(OGeom(ny, nx, dx, dt)), ctypes.byref(prm), _p(out), _p(lam),
                                 _p(_c(mask).ravel()), _p(_c(r_adj)), _p(_c(U)), _p(_c(bed).ravel()),
                                 _p(_c(B).ravel()), _p(_c(beta).ravel()), _p(_c(gamma).ravel()), color, omega,
                                 _p(np.zeros(nc)))
    return out


def grad_beta(ny, nx, dx, dt, prm, U, lam, bed, beta, quirk=True):
    out = np.zeros(ny * nx)
    _lib().or_grad_beta(ctypes.byref(OGeom(ny, nx, dx, dt)), ctypes.byref(prm), _p(out), _p(_c(U)), _p(_c(lam)),
                        _p(_c(bed).ravel()), _p(_c(beta).ravel()), int(bool(quirk)))
    return out.reshape(ny, nx)


def grad_B(ny, nx, dx, prm, U, lam, B):
    """lambda . dR/dB (B enters through the viscosity only)."""
    out = np.zeros(ny * nx)
    _lib().or_grad_B(ctypes.byref(OGeom(ny, nx, dx, 1.0)), ctypes.byref(prm), _p(out), _p(_c(U)), _p(_c(lam)),
                     _p(_c(B).ravel()))
    return out.reshape(ny, nx)


def restrict_u(uf):
    uf = _c(uf)
    ny, nx1 = uf.shape
    nyc, nxc = ny // 2, (nx1 - 1) // 2
    out = np.zeros((nyc, nxc + 1))
    _lib().or_restrict_u(_p(uf), _p(out), nyc, nxc)
    return out


def restrict_v(vf):
    vf = _c(vf)
    ny1, nx = vf.shape
    nyc, nxc = (ny1 - 1) // 2, nx // 2
    out = np.zeros((nyc + 1, nxc))
    _lib().or_restrict_v(_p(vf), _p(out), nyc, nxc)
    return out


def restrict_cell(hf):
    hf = _c(hf)
    out = np.zeros((hf.shape[0] // 2, hf.shape[1] // 2))
    _lib().or_restrict_cell(_p(hf), _p(out), out.shape[0], out.shape[1])
    return out


def restrict_max(hf):
    hf = _c(hf)
    out = np.zeros((hf.shape[0] // 2, hf.shape[1] // 2))
    _lib().or_restrict_max(_p(hf), _p(out), out.shape[0], out.shape[1])
    return out


def prolong_inject(hc, sign=1.0):
    hc = _c(hc)
    out = np.zeros((hc.shape[0] * 2, hc.shape[1] * 2))
    _lib().or_prolong_inject(_p(hc), _p(out), out.shape[0], out.shape[1], sign)
    return out


def prolong_u(uc):
    uc = _c(uc)
    nyf, nxf = uc.shape[0] * 2, (uc.shape[1] - 1) * 2
    out = np.zeros((nyf, nxf + 1))
    _lib().or_prolong_u(_p(uc), _p(out), nyf, nxf)
    return out


def prolong_v(vc):
    vc = _c(vc)
    nyf, nxf = (vc.shape[0] - 1) * 2, vc.shape[1] * 2
    out = np.zeros((nyf + 1, nxf))
    _lib().or_prolong_v(_p(vc), _p(out), nyf, nxf)
    return out


def prolong_H(hc):
    hc = _c(hc)
    out = np.zeros((hc.shape[0] * 2, hc.shape[1] * 2))
    _lib().or_prolong_H(_p(hc), _p(out), out.shape[0], out.shape[1])
    return out


def exp(x, libm=False):
    return _lib(libm).or_exp(float(x))


def log(x, libm=False):
    return _lib(libm).or_log(float(x))


def pow(x, p, libm=False):  # noqa: A001 - mirrors the C name
    return _lib(libm).or_pow(float(x), float(p))


def set_threads(n: int) -> None:
    _lib().oh_set_threads(int(n))
    _lib(True).oh_set_threads(int(n))


def num_threads() -> int:
    return _lib().oh_num_threads()


# ---------------------------------------------------------------- hierarchy
FIELDS = {"U": 0, "f": 1, "r": 2, "F": 3, "w": 4, "y": 5, "z": 6, "Lambda": 7, "f_adj": 8,
          "r_adj": 9, "l": 10, "H_prev": 11, "bed": 12, "B": 13, "beta": 14, "smb": 15,
          "gamma": 16, "chi": 17, "phi": 18, "mask": 19, "error_mask": 20}


class Hierarchy:
    """Oracle counterpart of the reference's IcePhysics (glide/physics.py:19-259)."""

    def __init__(self, ny, nx, dx, n_levels=5, n=3.0, eps_reg=1e-5, thklim=0.1, water_drag=1e-3,
                 calving_rate=1.0, gl_sigmoid_c=0.1, gl_derivatives=False, libm=False):
        self.lib = _lib(libm)
        self.ny, self.nx, self.dx, self.n_levels, self.thklim = ny, nx, dx, n_levels, thklim
        self.prm = params(n, eps_reg, water_drag, calving_rate, gl_sigmoid_c, gl_derivatives)
        self.h = self.lib.oh_create(ny, nx, dx, n_levels, ctypes.byref(self.prm), thklim)
        if not self.h:
            raise ValueError("bad hierarchy")

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.oh_destroy(h)
            self.h = None

    def set_schedule(self, pre=10, post=10, coarse=200, local_passes=10, local_sweeps=10, n_newton=10,
                     adj_sweeps=10, adj_coarse=100, boundary_weight=1.0):
        self.lib.oh_set_schedule(self.h, pre, post, coarse, local_passes, local_sweeps, n_newton,
                                 adj_sweeps, adj_coarse)
        self.lib.oh_set_boundary_weight(self.h, float(boundary_weight))

    def field(self, name: str, level: int = 0) -> np.ndarray:
        ny, nx = self.ny >> level, self.nx >> level
        _, _, nc, nt = sizes(ny, nx)
        n = nt if FIELDS[name] <= 10 else nc
        ptr = self.lib.oh_field(self.h, level, FIELDS[name])
        arr = np.ctypeslib.as_array(ptr, shape=(n,))
        return arr if n == nt else arr.reshape(ny, nx)

    def state(self, level: int = 0):
        U = self.field("U", level)
        return split(U, self.ny >> level, self.nx >> level)

    def set_geometry(self, bed, H):
        self.lib.oh_set_geometry(self.h, _p(_c(bed).ravel()), _p(_c(H).ravel()))

    def set_parameters(self, B=None, beta=None, smb=None):
        def arr(x):
            if x is None:
                return None
            a = np.broadcast_to(np.asarray(x, dtype=np.float64), (self.ny, self.nx))
            return _c(a).ravel()
        b_, be_, s_ = arr(B), arr(beta), arr(smb)
        self.lib.oh_set_parameters(self.h, _p(b_), _p(be_), _p(s_))

    def restrict_parameters(self):
        self.lib.oh_restrict_parameters(self.h)

    def vcycle(self):
        self.lib.oh_vcycle(self.h)

    def residual_norm(self) -> float:
        return self.lib.oh_residual_norm(self.h)

    def forward(self, dt, n_vcycles=3, tol=0.0, history=False, update_geometry=True):
        hist = np.zeros(max(n_vcycles, 1)) if history else None
        done = self.lib.oh_forward(self.h, float(dt), int(n_vcycles), float(tol), _p(hist), int(update_geometry))
        u, v, H = self.state(0)
        out = (u.copy(), v.copy(), H.copy())
        if history:
            return out, hist[:done].copy()
        return out

    def adjoint(self, dL_du, dL_dv, dL_dH, n_vcycles=2, quirk=True):
        g = np.zeros(self.ny * self.nx)
        self.lib.oh_adjoint(self.h, _p(_c(dL_du).ravel()), _p(_c(dL_dv).ravel()), _p(_c(dL_dH).ravel()),
                            int(n_vcycles), _p(g), int(bool(quirk)))
        return g.reshape(self.ny, self.nx)

    def adjoint_vcycle(self):
        self.lib.oh_adjoint_vcycle(self.h)

    def adjoint_residual_norm(self) -> float:
        return self.lib.oh_adjoint_residual_norm(self.h)
