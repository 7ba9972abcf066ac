"""Pins the CPU oracle of the learned block-Jacobi correction (there is no
reference counterpart): the convolution adjoint identity, the weight
gradient against central finite differences of the loss (float64, no
rounding), linearity of the corrected preconditioner in its input, and the
BF16 rounding helper against hand-worked values."""
import numpy as np
import pytest

from oracle import learned_pc_oracle as lo


def _case(seed=0, ng=16, ny=8):
    rng = np.random.default_rng(seed)
    n = ng * ny
    active = (rng.uniform(size=n) > 0.2).astype(np.uint8)
    y = rng.standard_normal(3 * n)
    y = (y.reshape(3, n) / np.linalg.norm(y.reshape(3, n), axis=0)).ravel()
    u = rng.standard_normal(3 * n)
    x = rng.standard_normal(3 * n)
    w = rng.standard_normal(9 * (16 * 32 + 32 * 32 + 32 * 16)) * 0.1
    return rng, active, y, u, x, w, ng, ny


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.140625, 1e-3], np.float32)
    got = lo.bf16(vals)
    # 1+2^-8 is a tie -> even (1.0); 1+3*2^-9 rounds up to 1+2^-7
    assert list(got[:4]) == [1.0, 1.0, 1.0 + 2 ** -7, -2.5]
    assert got[4] == 3.140625
    assert abs(got[5] - 1e-3) / 1e-3 < 2 ** -8


@pytest.mark.parametrize("periodic", [False, True])
def test_conv_adjoint_identity(periodic):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((8, 16, 16))
    g = rng.standard_normal((8, 16, 32))
    W = rng.standard_normal((9, 16, 32))
    lhs = (lo.conv(a, W, periodic) * g).sum()
    rhs = (a * lo.conv_adjoint(g, W, periodic)).sum()
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    # wgrad is the adjoint in W
    dW = rng.standard_normal(W.shape)
    lhs2 = (lo.conv(a, dW, periodic) * g).sum()
    rhs2 = (lo.wgrad(a, g, periodic) * dW).sum()
    assert abs(lhs2 - rhs2) <= 1e-10 * abs(lhs2)


@pytest.mark.parametrize("periodic", [False, True])
def test_weight_gradient_matches_finite_differences(periodic):
    rng, active, y, u, x, w, ng, ny = _case(2)
    gamma, scale = 0.03, 1.7
    loss, g, _ = lo.loss_grad(w, u, y, x, active, gamma, scale, ng, ny, periodic, rounding=False)
    idx = rng.choice(w.size, 25, replace=False)
    eps = 1e-5
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        lp = lo.loss_grad(wp, u, y, x, active, gamma, scale, ng, ny, periodic, rounding=False)[0]
        lm = lo.loss_grad(wm, u, y, x, active, gamma, scale, ng, ny, periodic, rounding=False)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_correction_is_linear_in_the_residual():
    """z = u + N(u) must be linear in u (a fixed operator for GMRES)"""
    rng, active, y, u, x, w, ng, ny = _case(3)
    f = lambda uu: lo.forward(w, uu, y, active, 0.02, 1.0, ng, ny, False, rounding=False)["o"]
    u2 = rng.standard_normal(u.size)
    a, b = 0.7, -1.3
    lhs = f(a * u + b * u2)
    rhs = a * f(u) + b * f(u2)
    assert np.max(np.abs(lhs - rhs)) <= 1e-10 * np.max(np.abs(rhs))
