"""Independent vectorised NumPy statement of the SSA + mass residual, written
from the discrete equations (not from the oracle's loop structure), used to
pin the oracle's residual.

Grid: u[i, j] on the left face of cell (i, j) (ny x nx+1), v[i, j] on its top
face ((ny+1) x nx), cells (ny x nx); out-of-range reads are clamped
(glide/cuda/common.cu:78-125).
"""
import numpy as np


def sigmoid(z, c):
    s = np.clip(c * z, -20.0, 20.0)
    return 1.0 / (1.0 + np.exp(-s))


def residual(u, v, H, f_u, f_v, f_H, bed, B, beta, mask, gamma, dx, dt,
             n=3.0, eps_reg=1e-5, water_drag=1e-3, calving_rate=1.0, gl_c=0.1):
    ny, nx = H.shape
    I = np.arange(-2, ny + 3)[:, None]
    Jx = np.arange(-2, nx + 3)[None, :]

    # padded (clamped) copies, index offset 2
    def padc(a):
        return a[np.clip(I, 0, a.shape[0] - 1), np.clip(Jx, 0, a.shape[1] - 1)]
    up, vp, Hp, bp, betap = padc(u), padc(v), padc(H), padc(bed), padc(beta)

    def U(i, j):
        return up[i + 2, j + 2]

    def V(i, j):
        return vp[i + 2, j + 2]

    def C(a, i, j):
        return a[i + 2, j + 2]

    ii, jj = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    # --- viscosity at every cell
    ux = (U(ii, jj + 1) - U(ii, jj)) / dx
    vy = (V(ii, jj) - V(ii + 1, jj)) / dx

    def exy(a, b):  # shear strain at vertex (a, b) from the clamped field
        return 0.5 * ((U(a - 1, b) - U(a, b)) / dx + (V(a, b) - V(a, b - 1)) / dx)
    m_tl = (ii > 0) & (jj > 0)
    m_tr = (ii > 0) & (jj < nx - 1)
    m_bl = (ii < ny - 1) & (jj > 0)
    m_br = (ii < ny - 1) & (jj < nx - 1)
    e2 = 0.25 * ((exy(ii, jj) * m_tl) ** 2 + (exy(ii, jj + 1) * m_tr) ** 2 +
                 (exy(ii + 1, jj) * m_bl) ** 2 + (exy(ii + 1, jj + 1) * m_br) ** 2)
    eps2 = ux ** 2 + vy ** 2 + ux * vy + e2 + eps_reg
    eta = 0.5 * B * eps2 ** ((1.0 - n) / (2.0 * n))
    etaH = eta * H
    sxx = 2 * etaH * (2 * ux + vy)
    syy = 2 * etaH * (ux + 2 * vy)
    # vertex (a, b), interior only: a in 1..ny-1, b in 1..nx-1
    sxy = np.zeros((ny + 1, nx + 1))
    a, b = np.meshgrid(np.arange(1, ny), np.arange(1, nx), indexing="ij")
    eHv = 0.25 * (etaH[a - 1, b - 1] + etaH[a - 1, b] + etaH[a, b - 1] + etaH[a, b])
    sxy[1:ny, 1:nx] = 2 * eHv * exy(a, b)

    def SXX(i, j):
        ok = (j >= 0) & (j < nx)
        return np.where(ok, sxx[np.clip(i, 0, ny - 1), np.clip(j, 0, nx - 1)], 0.0)

    def SYY(i, j):
        ok = (i >= 0) & (i < ny)
        return np.where(ok, syy[np.clip(i, 0, ny - 1), np.clip(j, 0, nx - 1)], 0.0)

    g = sigmoid(bp + 0.917 * Hp, gl_c)
    be = g * betap + (1 - g) * water_drag
    base = np.maximum(bp, -0.917 * Hp)
    S = base + Hp

    # --- u faces
    iu, ju = np.meshgrid(np.arange(ny), np.arange(nx + 1), indexing="ij")
    r_u = (-f_u + (SXX(iu, ju) - SXX(iu, ju - 1)) / dx + (sxy[iu, ju] - sxy[iu + 1, ju]) / dx
           - 0.5 * (C(be, iu, ju - 1) + C(be, iu, ju)) * U(iu, ju))
    tdx = 0.5 * (C(Hp, iu, ju - 1) + C(Hp, iu, ju)) * (C(S, iu, ju) - C(S, iu, ju - 1)) / dx
    r_u = r_u - np.where((ju > 0) & (ju < nx), tdx, 0.0)
    # --- v faces
    iv, jv = np.meshgrid(np.arange(ny + 1), np.arange(nx), indexing="ij")
    r_v = (-f_v + (SYY(iv - 1, jv) - SYY(iv, jv)) / dx + (sxy[iv, jv + 1] - sxy[iv, jv]) / dx
           - 0.5 * (C(be, iv - 1, jv) + C(be, iv, jv)) * V(iv, jv))
    tdy = 0.5 * (C(Hp, iv - 1, jv) + C(Hp, iv, jv)) * (C(S, iv - 1, jv) - C(S, iv, jv)) / dx
    r_v = r_v - np.where((iv > 0) & (iv < ny), tdy, 0.0)
    # --- mass
    uf = U(iu, ju)
    qx = 0.5 * (C(Hp, iu, ju - 1) + C(Hp, iu, ju)) * uf - 0.5 * np.abs(uf) * (C(Hp, iu, ju) - C(Hp, iu, ju - 1))
    qx = np.where((ju > 0) & (ju < nx), qx, 0.0)
    vf = V(iv, jv)
    qy = 0.5 * (C(Hp, iv - 1, jv) + C(Hp, iv, jv)) * vf - 0.5 * np.abs(vf) * (C(Hp, iv - 1, jv) - C(Hp, iv, jv))
    qy = np.where((iv > 0) & (iv < ny), qy, 0.0)
    gc = sigmoid(bed + 0.917 * H, gl_c)
    r_H = (H / dt - f_H + calving_rate * (1 - gc) * H + (qx[:, 1:] - qx[:, :-1]) / dx
           + (qy[:-1, :] - qy[1:, :]) / dx)
    r_H = (1 - mask) * r_H + mask * (H - gamma)
    return r_u, r_v, r_H
