"""NumPy reference of the preconditioner MLP and its loss (test infrastructure).

Mirrors the precision choices of the HIP kernels: bf16 weights, bf16 inputs
to every matrix product (features, the first hidden layer, the backward
pre-activation gradients), FP32/FP64 accumulation, tanh activations, and the
correction-angle loss of the learned preconditioner.  Computed in FP64.
"""

from __future__ import annotations

import numpy as np

F, H = 64, 128


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    out = (u.astype(np.uint32) << 16).view(np.float32)
    return out.astype(np.float64)


def unpack(p: np.ndarray):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + F * H].reshape(F, H); o += F * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    return W1, b1, W2, b2, w3, p[o]


def forward(p: np.ndarray, X: np.ndarray, rounding: bool = True):
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, w3, b3 = unpack(p)
    h1 = np.tanh(r(X) @ r(W1) + b1)
    h2 = np.tanh(r(h1) @ r(W2) + b2)
    return h2 @ w3 + b3, h1, h2


def gradient(p: np.ndarray, X: np.ndarray, dy: np.ndarray, rounding: bool = True) -> np.ndarray:
    """d/dp of sum_b dy[b] * forward(p, X)[b]."""
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, w3, b3 = unpack(p)
    _, h1, h2 = forward(p, X, rounding)
    dy = np.asarray(dy, dtype=np.float64)
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    da2 = dy[:, None] * w3[None, :] * (1.0 - h2 ** 2)
    gb2 = da2.sum(axis=0)
    gW2 = r(h1).T @ r(da2)
    dh1 = r(da2) @ r(W2).T
    da1 = dh1 * (1.0 - h1 ** 2)
    gb1 = da1.sum(axis=0)
    gW1 = r(X).T @ r(da1)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])


def precond_loss(y, r, e, u, d, theta):
    """L = 1 - <e,t>^2 / (|e|^2 |t - u<u,t>|^2), t = r e^y / (theta - d); returns (L, dL/dy)."""
    den = theta - d
    den = np.where(np.abs(den) < 1e-8, np.where(den > 0, 1e-8, -1e-8), den)
    t = r * np.exp(y) / den
    a = e @ t
    ut = u @ t
    c = t @ t - ut ** 2
    b = e @ e
    loss = 1.0 - a * a / (b * c)
    tperp = t - ut * u
    dt = -2.0 * a * e / (b * c) + 2.0 * a * a * tperp / (b * c * c)
    return loss, dt * t
