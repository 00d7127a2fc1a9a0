"""NumPy reference of the preconditioner MLP and its loss (test infrastructure).

Mirrors the precision choices of the HIP kernels: bf16 weights, bf16 inputs
to every matrix product (features, the first hidden layer, the backward
pre-activation gradients), FP32/FP64 accumulation, tanh activations, and the
one-step Ritz-gain loss of the learned preconditioner.  Computed in FP64.
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


def ritz_loss(scale, r, u, matvec, theta):
    """One-step Ritz-gain loss of the correction t = scale * r for a Davidson
    state (normalised u, r = H u - theta u).  q = t - u <u, t>; the lowest Ritz
    value of span{u, q} is theta - gain; loss = -log(gain).  Returns
    (loss, gain, dloss/dlog(scale)) -- the last is the gradient with respect to
    the learned log diagonal m when scale = e^m."""
    scale = np.asarray(scale, dtype=np.float64)
    t = scale * r
    q = t - u * np.vdot(u, t)
    n2 = np.vdot(q, q).real
    c = np.vdot(q, r)
    hq = matvec(q)
    gamma = np.vdot(q, hq).real / n2
    c2 = abs(c) ** 2
    rho2 = c2 / n2
    a = 0.5 * (gamma - theta)
    R = np.sqrt(a * a + rho2)
    gain = rho2 / (R + a) if a > 0 else R - a
    A1 = (1.0 - a / R) / (n2 * gain)
    A2 = -1.0 / (n2 * R * gain)
    A3 = c2 / (n2 * n2 * R * gain)
    gq = A1 * (hq - gamma * q) + A2 * np.conj(c) * r + A3 * q
    gt = gq - u * np.vdot(u, gq)
    dm = (np.conj(gt) * t).real
    return -np.log(gain), gain, dm
