"""Pins the NumPy MLP reference itself: its analytic gradient equals central
finite differences of its forward pass (FP64, no rounding), and the loss
gradient equals finite differences of the loss."""

from __future__ import annotations

import numpy as np

from tests import mlp_reference as ref


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = rng.normal(scale=0.2, size=ref.F * ref.H + ref.H + ref.H * ref.H + ref.H + ref.H + 1)
    X = rng.normal(size=(7, ref.F))
    dy = rng.normal(size=7)
    g = ref.gradient(p, X, dy, rounding=False)
    idx = np.concatenate([rng.choice(p.size, size=40, replace=False), [p.size - 1]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (dy @ ref.forward(pp, X, False)[0] - dy @ ref.forward(pm, X, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    D = 30
    u = rng.normal(size=D)
    u /= np.linalg.norm(u)
    e = rng.normal(size=D)
    e -= u * (u @ e)
    r = rng.normal(size=D)
    d = rng.uniform(1, 3, size=D)
    y = rng.normal(scale=0.3, size=D)
    loss, grad = ref.precond_loss(y, r, e, u, d, -1.0)
    assert 0.0 <= loss <= 1.0
    h = 1e-6
    for i in range(D):
        yp, ym = y.copy(), y.copy()
        yp[i] += h
        ym[i] -= h
        fd = (ref.precond_loss(yp, r, e, u, d, -1.0)[0] - ref.precond_loss(ym, r, e, u, d, -1.0)[0]) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6


def test_bf16_rounding_known_values():
    # bf16 spacing at 1 is 2^-7: 1 + 2^-8 ties to even (down to 1), 1 + 3 * 2^-8
    # ties to even (up to 1 + 2^-6), 1 + 2^-9 rounds down, 1 + 3 * 2^-9 rounds up
    vals = np.asarray([1.0, 1.00390625, 1.01171875, 1.001953125, 1.005859375, -2.5], dtype=np.float32)
    np.testing.assert_array_equal(ref.bf16(vals), [1.0, 1.0, 1.015625, 1.0, 1.0078125, -2.5])
