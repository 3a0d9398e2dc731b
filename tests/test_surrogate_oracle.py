"""Pins the surrogate's FP64 oracle: its analytic gradient equals central
finite differences of its own FP64 loss, and its forward obeys the
specification's closed forms."""
import numpy as np
import pytest

from oracle import surrogate_oracle as S


@pytest.fixture(scope="module")
def small_case():
    rng = np.random.default_rng(0)
    theta = S.init_params(7) * 3.0  # larger weights: nonlinear regime, some clipping
    x = rng.random((2, 8, 8))
    target = np.clip(x + 0.05 * rng.standard_normal(x.shape), 0, 1)
    return theta, x, target


def test_param_count():
    assert S.N_PARAMS == 104129


def test_gradient_matches_finite_differences(small_case):
    theta, x, target = small_case
    _, g = S.loss_and_grad(theta, x, target)
    rng = np.random.default_rng(1)
    # Sample coordinates from every parameter block.
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in S.SHAPES])
    idx = np.concatenate([rng.integers(offs[i], offs[i + 1], size=6) for i in range(len(S.SHAPES))])
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (S.loss_and_grad(tp, x, target)[0] - S.loss_and_grad(tm, x, target)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_conv5_is_a_circular_cross_correlation():
    rng = np.random.default_rng(2)
    X = rng.random((1, 3, 6, 7))
    W = rng.random((2, 25, 3))
    out = S.conv5(X, W)
    n, co, y, x = 0, 1, 0, 6
    want = sum(W[co, (dy + 2) * 5 + dx + 2, ci] * X[n, ci, (y + dy) % 6, (x + dx) % 7]
               for ci in range(3) for dy in range(-2, 3) for dx in range(-2, 3))
    assert out[n, co, y, x] == pytest.approx(want, rel=1e-12)


def test_zero_last_layer_is_identity():
    theta = S.init_params(3)
    P = S.unpack(theta)
    P["W4"][:] = 0.0
    P["b4"][:] = 0.0
    x = np.random.default_rng(3).random((1, 8, 8))
    y, _ = S.forward(S.pack(P), x)
    assert np.array_equal(y, x)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 3.0e-3])
    r = S.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # ties to even
    assert r[2] == 1.0 + 2 ** -7
    assert abs(r[3] - 3.0e-3) < 3.0e-3 * 2 ** -8
