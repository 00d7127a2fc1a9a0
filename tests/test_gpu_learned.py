"""GPU checks of the learned preconditioner: MFMA forward and backward against
the NumPy reference (tests/mlp_reference.py, itself pinned by finite
differences and an independent Rayleigh-Ritz solve), the Ritz-gain loss
kernel, training, and the gain over plain Jacobi in Davidson."""

from __future__ import annotations

import numpy as np
import pytest

import tencirpauli_amd as tcp
from oracle import dense
from tencirpauli_amd import davidson, lattice, learned
from tests import mlp_reference as ref
from tests.helpers import conserving_u1_terms

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
    # The reference rounds every product input to bf16 as the kernel does; what
    # remains is FP32 accumulation order, the kernel's exp2/rcp tanh (~1e-7) and
    # rare bf16 rounding-boundary flips of a first-layer activation (measured
    # max |error| 1.5e-4 of the output scale).
    np.testing.assert_allclose(got, expected, atol=5e-4 * max(1.0, np.abs(expected).max()))


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
        # the same rounding model as the forward check; measured worst block
        # (W1) 1.8e-4 over three seeds
        assert err <= 1e-3, (name, err)


@pytest.mark.parametrize("mode", ["learned", "jacobi"])
def test_ritz_loss_kernel_matches_reference(mode):
    import ctypes

    import torch

    from tencirpauli_amd import _lib

    rng = np.random.default_rng(3)
    n, k = 10, 5
    codes, coef = conserving_u1_terms(rng, n, 24)  # XY - YX hoppings: complex Hermitian matrix
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    restricted = op.restrict_charge(tcp.U1Sector(n, k))
    cc, cf = op.code_arrays()
    H = dense.restricted_dense(n, cc, cf, dense.u1_basis(n, k))
    assert np.abs(H.imag).max() > 0 and np.allclose(H, H.conj().T)
    D = H.shape[0]
    u = rng.normal(size=D) + 1j * rng.normal(size=D)
    u /= np.linalg.norm(u)
    theta = np.vdot(u, H @ u).real
    r = H @ u - theta * u
    precond = learned.LearnedPreconditioner(learned.MLP(), restricted)
    ud, rd = torch.from_numpy(u).cuda(), torch.from_numpy(r).cuda()
    logm = rng.normal(scale=0.4, size=D).astype(np.float32)
    if mode == "learned":
        md = torch.from_numpy(logm).cuda()
        dm = torch.empty(D, dtype=torch.float32, device="cuda")
        scale = np.exp(logm.astype(np.float64))
    else:
        md = dm = None
        den = theta - np.diag(H).real
        den = np.where(np.abs(den) < 1e-8, np.where(den < 0, -1e-8, 1e-8), den)
        scale = 1.0 / den
    loss, gain = ctypes.c_double(), ctypes.c_double()
    _lib.check(_lib.lib().tcp_learned_precond_ritz_loss(
        precond._ptr, theta, ud.data_ptr(), rd.data_ptr(), None if md is None else md.data_ptr(),
        None if dm is None else dm.data_ptr(), ctypes.byref(loss), ctypes.byref(gain), None))
    exp_loss, exp_gain, exp_dm = ref.ritz_loss(scale, r, u, lambda x: H @ x, theta)
    assert abs(gain.value - exp_gain) <= 1e-10 * abs(exp_gain)
    assert abs(loss.value - exp_loss) <= 1e-10 * max(1.0, abs(exp_loss))
    if mode == "learned":
        np.testing.assert_allclose(dm.cpu().numpy(), exp_dm, rtol=1e-5, atol=1e-6 * np.abs(exp_dm).max())


SNAPSHOTS = (1, 2, 3, 5, 8, 12, 18, 26, 36)


def _lattice_sector(rows, cols, seed, k=None):
    codes, coef = lattice.heisenberg_periodic(rows, cols, seed=seed)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    n = rows * cols
    return op.restrict_charge(tcp.U1Sector(n, n // 2 if k is None else k))


def test_training_lowers_the_correction_loss_and_generalises():
    mlp = learned.MLP(seed=11)
    train_snaps = []
    for rows, cols, seed in ((3, 4, 1), (4, 4, 2), (3, 5, 3)):
        train_snaps += learned.jacobi_snapshots(_lattice_sector(rows, cols, seed), mlp, SNAPSHOTS)
    held_out = learned.jacobi_snapshots(_lattice_sector(4, 4, 99), mlp, SNAPSHOTS)
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
        snaps += learned.jacobi_snapshots(_lattice_sector(rows, cols, seed), mlp, SNAPSHOTS)
    learned.train(mlp, snaps, epochs=25, lr=2e-3)
    target = _lattice_sector(4, 5, 77)
    ref_run = davidson.lowest_eigenpair(target, tol=1e-10)
    plain = davidson.lowest_eigenpair(target, tol=1e-7)
    hook = learned.LearnedPreconditioner(mlp, target)
    smart = davidson.lowest_eigenpair(target, tol=1e-7, preconditioner=hook)
    assert smart.converged and abs(smart.eigenvalue - ref_run.eigenvalue) <= 1e-8
    print(f"4x5 Davidson iterations: jacobi {plain.iterations}, learned {smart.iterations}", flush=True)
    assert smart.iterations < plain.iterations


class _DeviceFloats:
    """A float32 device array at a raw address, readable by torch.as_tensor."""

    def __init__(self, ptr, n):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "<f4", "data": (ptr, False), "version": 2}


def _device_floats(ptr, n):
    import torch

    return torch.as_tensor(_DeviceFloats(ptr, n), device="cuda").cpu().numpy().copy()


def test_refresh_threshold_controls_network_reevaluation():
    target = _lattice_sector(3, 4, 5)
    mlp = learned.MLP(seed=3)
    dim = target.dimension
    always = learned.LearnedPreconditioner(mlp, target, refresh=0.0)
    never = learned.LearnedPreconditioner(mlp, target, refresh=1e9)
    theta = -5.0
    a0 = _device_floats(always.log_multipliers(theta), dim)
    n0 = _device_floats(never.log_multipliers(theta), dim)
    np.testing.assert_array_equal(a0, n0)
    shifted = theta - 0.05 * always.scale
    a1 = _device_floats(always.log_multipliers(shifted), dim)
    n1 = _device_floats(never.log_multipliers(shifted), dim)
    assert not np.array_equal(a1, a0)  # feature 0 moved with theta: a new evaluation
    np.testing.assert_array_equal(n1, n0)  # within the threshold: the cached evaluation
    with pytest.raises(ValueError, match="refresh"):
        learned.LearnedPreconditioner(mlp, target, refresh=-1.0)


class _DeviceArray:
    """A device array of any dtype at a raw address, readable by torch.as_tensor."""

    def __init__(self, ptr, shape, typestr):
        self.__cuda_array_interface__ = {"shape": shape, "typestr": typestr, "data": (ptr, False), "version": 2}


def test_preconditioner_log_multipliers_match_the_reference_network():
    # The Davidson callback evaluates the network with feature 0 replaced by
    # bf16((H_ii - theta) / scale) from the diagonal, in the inference variant
    # of the forward kernel: it must equal the NumPy network on those inputs.
    import torch

    target = _lattice_sector(3, 4, 8)
    rng = np.random.default_rng(21)
    mlp = learned.MLP()
    p = _random_params(rng)
    mlp.set_parameters(p)
    hook = learned.LearnedPreconditioner(mlp, target, refresh=0.0)
    dim = hook.dimension
    theta = -4.5
    got = _device_floats(hook.log_multipliers(theta), dim)
    raw = torch.as_tensor(_DeviceArray(hook.features_ptr, (dim, 64), "<i2"), device="cuda").cpu().numpy()
    diag = torch.as_tensor(_DeviceArray(hook.diag_ptr, (dim,), "<f8"), device="cuda").cpu().numpy()
    X = (raw.astype(np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    X[:, 0] = ref.bf16(((diag - theta) * (1.0 / hook.scale)).astype(np.float32))
    expected = ref.forward(p, X)[0]
    np.testing.assert_allclose(got, expected, atol=2e-3 * max(1.0, np.abs(expected).max()))
