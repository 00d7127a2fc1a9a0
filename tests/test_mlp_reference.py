"""Pins the NumPy MLP reference itself: its analytic gradient equals central
finite differences of its forward pass (FP64, no rounding); the Ritz-gain
loss equals an independent 2 x 2 Rayleigh-Ritz solve and its gradient equals
finite differences of the loss."""

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


def _hermitian_state(rng, D, complex_=True):
    A = rng.normal(size=(D, D)) + (1j * rng.normal(size=(D, D)) if complex_ else 0)
    H = 0.5 * (A + A.conj().T)
    u = rng.normal(size=D) + (1j * rng.normal(size=D) if complex_ else 0)
    u /= np.linalg.norm(u)
    hu = H @ u
    theta = np.vdot(u, hu).real
    return H, u, hu - theta * u, theta


def test_ritz_gain_equals_two_by_two_rayleigh_ritz():
    # gain = theta - lowest eigenvalue of H projected on span{u, q}, computed
    # here from an orthonormal basis and numpy's Hermitian eigensolver
    rng = np.random.default_rng(4)
    for complex_ in (False, True):
        H, u, r, theta = _hermitian_state(rng, 40, complex_)
        scale = np.exp(rng.normal(scale=0.5, size=40))
        _, gain, _ = ref.ritz_loss(scale, r, u, lambda x: H @ x, theta)
        t = scale * r
        q = t - u * np.vdot(u, t)
        B = np.stack([u, q / np.linalg.norm(q)], axis=1)
        lowest = np.linalg.eigvalsh(B.conj().T @ H @ B)[0]
        assert abs((theta - lowest) - gain) <= 1e-12 * max(1.0, abs(theta)), complex_
        assert gain > 0


def test_ritz_gain_of_the_exact_correction_is_the_full_error():
    # for a two-level H the exact correction reaches the eigenvector: the gain
    # is the whole energy error theta - E0
    H = np.diag([-1.0, 2.0])
    u = np.array([np.cos(0.3), np.sin(0.3)])
    theta = u @ H @ u
    r = H @ u - theta * u
    _, gain, _ = ref.ritz_loss(np.ones(2), r, u, lambda x: H @ x, theta)
    assert abs(gain - (theta + 1.0)) <= 1e-14


def test_ritz_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    D = 30
    for complex_ in (False, True):
        H, u, r, theta = _hermitian_state(rng, D, complex_)
        m = rng.normal(scale=0.3, size=D)
        loss, _, grad = ref.ritz_loss(np.exp(m), r, u, lambda x: H @ x, theta)
        h = 1e-6
        for i in range(D):
            mp, mm = m.copy(), m.copy()
            mp[i] += h
            mm[i] -= h
            fd = (ref.ritz_loss(np.exp(mp), r, u, lambda x: H @ x, theta)[0] -
                  ref.ritz_loss(np.exp(mm), r, u, lambda x: H @ x, theta)[0]) / (2 * h)
            assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)), (complex_, i)


def test_bf16_rounding_known_values():
    # bf16 spacing at 1 is 2^-7: 1 + 2^-8 ties to even (down to 1), 1 + 3 * 2^-8
    # ties to even (up to 1 + 2^-6), 1 + 2^-9 rounds down, 1 + 3 * 2^-9 rounds up
    vals = np.asarray([1.0, 1.00390625, 1.01171875, 1.001953125, 1.005859375, -2.5], dtype=np.float32)
    np.testing.assert_array_equal(ref.bf16(vals), [1.0, 1.0, 1.015625, 1.0, 1.0078125, -2.5])
