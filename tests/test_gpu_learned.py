"""GPU checks of the learned preconditioner: MFMA forward and backward against
the NumPy reference (tests/mlp_reference.py, itself pinned by finite
differences), the loss kernel, training, and the gain over plain Jacobi."""

from __future__ import annotations

import numpy as np
import pytest

import tencirpauli_amd as tcp
from tencirpauli_amd import davidson, lattice, learned
from tests import mlp_reference as ref

pytestmark = pytest.mark.gpu


def _bf16_tensor(x: np.ndarray):
    import torch

    return torch.from_numpy(np.asarray(x, dtype=np.float32)).to(torch.bfloat16).cuda()


def _random_params(rng, scale=0.15):
    return rng.normal(scale=scale, size=learned.MLP.parameter_count()).astype(np.float32)


@pytest.mark.parametrize("batch", [1, 37, 64, 1000])
def test_mlp_forward_matches_reference(batch):
    rng = np.random.default_rng(batch)
    mlp = learned.MLP()
    p = _random_params(rng)
    mlp.set_parameters(p)
    np.testing.assert_array_equal(mlp.get_parameters(), p)
    X = rng.normal(size=(batch, 64))
    Xd = _bf16_tensor(X)
    got = mlp.forward(Xd).cpu().numpy()
    expected = ref.forward(p, Xd.float().cpu().numpy())[0]
    np.testing.assert_allclose(got, expected, atol=2e-3 * max(1.0, np.abs(expected).max()))


def test_mlp_gradient_matches_reference():
    import torch

    rng = np.random.default_rng(7)
    mlp = learned.MLP()
    p = _random_params(rng)
    mlp.set_parameters(p)
    batch = 3000
    X = rng.normal(size=(batch, 64))
    Xd = _bf16_tensor(X)
    dy = rng.normal(size=batch).astype(np.float32)
    got = mlp.gradient(Xd, torch.from_numpy(dy).cuda())
    expected = ref.gradient(p, Xd.float().cpu().numpy(), dy)
    offsets = np.cumsum([0, 64 * 128, 128, 128 * 128, 128, 128, 1])
    for name, lo, hi in zip(("W1", "b1", "W2", "b2", "w3", "b3"), offsets[:-1], offsets[1:]):
        err = np.linalg.norm(got[lo:hi] - expected[lo:hi]) / max(np.linalg.norm(expected[lo:hi]), 1e-12)
        assert err <= 1e-2, (name, err)


def test_loss_kernel_matches_reference():
    import ctypes

    import torch

    from tencirpauli_amd import _lib

    rng = np.random.default_rng(3)
    D = 5000
    u = rng.normal(size=D)
    u /= np.linalg.norm(u)
    e = rng.normal(size=D)
    e -= u * (u @ e)
    r = rng.normal(size=D)
    d = rng.uniform(0.5, 4.0, size=D)
    y = rng.normal(scale=0.3, size=D).astype(np.float32)
    theta = -1.3
    to_c = lambda a: torch.from_numpy(a.astype(np.complex128)).cuda()  # noqa: E731
    yd, rd, ed, ud = torch.from_numpy(y).cuda(), to_c(r), to_c(e), to_c(u)
    dd = torch.from_numpy(d).cuda()
    dy = torch.empty(D, dtype=torch.float32, device="cuda")
    loss = ctypes.c_double()
    _lib.check(_lib.lib().tcp_precond_loss_gradient(yd.data_ptr(), rd.data_ptr(), ed.data_ptr(), ud.data_ptr(),
                                                    dd.data_ptr(), theta, D, 1.0, dy.data_ptr(),
                                                    ctypes.byref(loss), None))
    exp_loss, exp_dy = ref.precond_loss(y.astype(np.float64), r, e, u, d, theta)
    assert abs(loss.value - exp_loss) <= 1e-10
    np.testing.assert_allclose(dy.cpu().numpy(), exp_dy, rtol=1e-5, atol=1e-6 * np.abs(exp_dy).max())


def _lattice_sector(rows, cols, seed, k=None):
    codes, coef = lattice.heisenberg_periodic(rows, cols, seed=seed)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    n = rows * cols
    return op.restrict_charge(tcp.U1Sector(n, n // 2 if k is None else k))


def test_training_lowers_the_correction_loss_and_generalises():
    mlp = learned.MLP(seed=11)
    train_snaps = []
    for rows, cols, seed in ((3, 4, 1), (4, 4, 2), (3, 5, 3)):
        train_snaps += learned.jacobi_snapshots(_lattice_sector(rows, cols, seed), mlp)
    held_out = learned.jacobi_snapshots(_lattice_sector(4, 4, 99), mlp)
    assert len(train_snaps) >= 10 and len(held_out) >= 3
    jacobi = np.mean([learned.jacobi_loss(s) for s in held_out])
    history = learned.train(mlp, train_snaps, epochs=25, lr=2e-3)
    assert history[-1] < history[0]
    trained = np.mean([learned.snapshot_loss(mlp, s) for s in held_out])
    print(f"held-out correction loss: jacobi {jacobi:.4f} learned {trained:.4f}", flush=True)
    assert trained < jacobi


def test_learned_preconditioner_in_davidson():
    mlp = learned.MLP(seed=12)
    snaps = []
    for rows, cols, seed in ((3, 4, 1), (4, 4, 2), (3, 5, 3)):
        snaps += learned.jacobi_snapshots(_lattice_sector(rows, cols, seed), mlp)
    learned.train(mlp, snaps, epochs=25, lr=2e-3)
    target = _lattice_sector(4, 5, 77)
    ref_run = davidson.lowest_eigenpair(target, tol=1e-10)
    plain = davidson.lowest_eigenpair(target, tol=1e-7)
    hook = learned.LearnedPreconditioner(mlp, target)
    smart = davidson.lowest_eigenpair(target, tol=1e-7, preconditioner=hook)
    assert smart.converged and abs(smart.eigenvalue - ref_run.eigenvalue) <= 1e-8
    print(f"4x5 Davidson iterations: jacobi {plain.iterations}, learned {smart.iterations}", flush=True)
    assert smart.iterations <= plain.iterations + 2
