"""CPU pins of the U-Net oracle: forward == an independent torch float64
statement (strided/transposed convolutions instead of loops), finite
differences of the oracle's loss == torch autograd, Adam closed form."""
import numpy as np
import pytest

from oracle import unet_ref as ur

torch = pytest.importorskip("torch")
from unet_torch_ref import torch_forward, torch_loss_grad  # noqa: E402


def _rand_case(seed, n=2, nx=17, ny=9):
    rng = np.random.default_rng(seed)
    theta = rng.standard_normal(ur.num_params()) * 0.2
    feats = rng.standard_normal((n, nx * ny, 4))
    target = rng.standard_normal((n, nx * ny, 2))
    return theta, feats, target


def test_num_params():
    import helmholtz_amd as ha
    assert ur.num_params() == int(ha.lib().hh_unet_num_params())


@pytest.mark.parametrize("nx,ny", [(17, 9), (9, 9), (33, 17)])
def test_oracle_forward_equals_torch(nx, ny):
    theta, feats, _ = _rand_case(nx + ny, nx=nx, ny=ny)
    y = ur.forward(theta, feats, nx, ny)
    yt, _ = torch_forward(theta, feats, nx, ny)
    np.testing.assert_allclose(y, yt.detach().numpy(), rtol=1e-11, atol=1e-11)


def test_down_up_adjoint_pair():
    # <down(x), z> == <x, up(z)> / 4 on nodal grids (exact adjoints up to 4)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1, 13, 17, 2))
    z = rng.standard_normal((1, 7, 9, 2))
    lhs = np.sum(ur.down(x) * z)
    rhs = np.sum(x * ur.up(z, 13, 17)) / 4.0
    assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs))


def test_finite_differences_match_autograd():
    theta, feats, target = _rand_case(7, n=1, nx=9, ny=9)
    loss, g = torch_loss_grad(theta, feats, target, 9, 9)
    assert loss == pytest.approx(ur.mse(theta, feats, target, 9, 9), rel=1e-12)
    rng = np.random.default_rng(11)
    idx = rng.choice(theta.size, 24, replace=False)
    idx = np.concatenate([idx, [theta.size - 1, theta.size - 34, 0]])
    h = 1e-6
    for k in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (ur.mse(tp, feats, target, 9, 9) - ur.mse(tm, feats, target, 9, 9)) / (2 * h)
        assert fd == pytest.approx(g[k], rel=1e-5, abs=1e-9)


def test_adam_closed_form():
    g = np.array([0.5, -2.0, 1e-3])
    th = np.zeros(3)
    t1, m, v = ur.adam_step(th, g, np.zeros(3), np.zeros(3), 1, lr=0.1)
    # first step: m_hat = g, v_hat = g^2 -> update = lr * g / (|g| + eps)
    np.testing.assert_allclose(t1, -0.1 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    t2, _, _ = ur.adam_step(t1, g, m, v, 2, lr=0.1)
    np.testing.assert_allclose(t2, t1 - 0.1 * g / (np.abs(g) + 1e-8), rtol=1e-9)
