"""Misfit and regularisation helpers of the inversion (glide/physics.py:262-411)
on the CPU, against closed forms and torch autograd."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from glide_amd.physics import abs_grad, abs_loss, huber_grad, huber_loss, tikhonov_regularization  # noqa: E402


def t(a):
    return torch.as_tensor(np.asarray(a, dtype=np.float64))


def test_tikhonov_linear_field_closed_form():
    """f = a j + b i: every x difference is a and every y difference b, so
    the loss is (a^2 ny (nx-1) + b^2 (ny-1) nx) / 2 and the gradient is -a on
    the first column, +a on the last (and -b, +b on the first and last rows)."""
    ny, nx, a, b = 5, 7, 0.3, -1.25
    i, j = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    loss, grad = tikhonov_regularization(t(a * j + b * i))
    assert loss == pytest.approx(0.5 * (a * a * ny * (nx - 1) + b * b * (ny - 1) * nx), rel=1e-14)
    exp = np.zeros((ny, nx))
    exp[:, 0] -= a
    exp[:, -1] += a
    exp[0, :] -= b
    exp[-1, :] += b
    np.testing.assert_allclose(grad.numpy(), exp, atol=1e-14)


def test_tikhonov_gradient_is_the_derivative_of_the_loss():
    rng = np.random.default_rng(0)
    f = t(rng.normal(size=(6, 9))).requires_grad_(True)
    gx = f[:, 1:] - f[:, :-1]
    gy = f[1:, :] - f[:-1, :]
    (0.5 * ((gx ** 2).sum() + (gy ** 2).sum())).backward()
    loss, grad = tikhonov_regularization(f.detach())
    np.testing.assert_allclose(grad.numpy(), f.grad.numpy(), rtol=0, atol=1e-13)
    const_loss, const_grad = tikhonov_regularization(torch.full((4, 4), 3.7, dtype=torch.float64))
    assert const_loss == 0.0 and float(const_grad.abs().max()) == 0.0


def test_huber_values_and_nan_masking():
    """sqrt(d^2 + eps) with (d, eps) = (3, 16) is 5 and with d = 0 (an
    unobserved entry: NaN observation) sqrt(eps); the means run over all
    entries."""
    u, uo = t([[3.0, 1.0], [2.0, 7.0]]), t([[0.0, 1.0], [float("nan"), 7.0]])
    v, vo = t([[1.0, 4.0], [0.0, 0.0]]), t([[1.0, 0.0], [0.0, 3.0]])
    L = float(huber_loss(u, v, uo, vo, eps=16.0))
    assert L == pytest.approx((5 + 4 + 4 + 4) / 4 + (4 + math.sqrt(32) + 4 + 5) / 4, rel=1e-14)
    gu, gv = huber_grad(u, v, uo, vo, eps=16.0)
    np.testing.assert_allclose(gu.numpy(), [[3 / 5 / 4, 0], [0, 0]], atol=1e-15)
    np.testing.assert_allclose(gv.numpy(), [[0, 4 / math.sqrt(32) / 4], [0, -3 / 5 / 4]], atol=1e-15)


def test_huber_grad_is_the_derivative_of_the_loss_on_square_grids():
    rng = np.random.default_rng(1)
    ny = nx = 5
    u, v = t(rng.normal(size=(ny, nx + 1))), t(rng.normal(size=(ny + 1, nx)))
    uo, vo = t(rng.normal(size=u.shape)), t(rng.normal(size=v.shape))
    uo[1, 2] = float("nan")
    ur, vr = u.clone().requires_grad_(True), v.clone().requires_grad_(True)
    huber_loss(ur, vr, uo, vo, eps=2.0).backward()
    gu, gv = huber_grad(u, v, uo, vo, eps=2.0)
    np.testing.assert_allclose(gu.numpy(), ur.grad.numpy(), atol=1e-15)
    np.testing.assert_allclose(gv.numpy(), vr.grad.numpy(), atol=1e-15)
    assert float(gu[1, 2]) == 0.0


def test_abs_loss_and_grad():
    u, uo = t([[1.0, -2.0, 0.5]]), t([[0.0, float("nan"), 2.5]])
    v, vo = t([[3.0], [-1.0]]), t([[4.0], [-4.0]])
    assert float(abs_loss(u, v, uo, vo)) == pytest.approx((1 + 0 + 2) / 3 + (1 + 3) / 2, rel=1e-15)
    gu, gv = abs_grad(u, v, uo, vo)
    np.testing.assert_array_equal(gu.numpy(), [[1 / 3, 0.0, -1 / 3]])
    np.testing.assert_array_equal(gv.numpy(), [[-1 / 3], [1 / 3]])
