"""Learned diagonal preconditioner for the GPU Davidson solver.

No reference counterpart (the reference has no eigensolver of its own).  A
64-128-128-1 tanh MLP (bf16 weights, FP32 accumulation on MFMA, C ABI
``tcp_mlp_*``) maps per-state features of a restricted operator to the log of
a diagonal preconditioner: the correction is t_i = r_i e^{m_i}, in place of
Jacobi's r_i / (theta - H_ii).

Training data are snapshots of plain Jacobi Davidson runs on small lattices:
the Ritz vector u, its Ritz value theta and residual r = H u - theta u after a
few iterations, and the lattice's ground energy E0.  The loss of a snapshot is
the one-step Ritz gain of t: adding t to span{u} lowers the Ritz value by
gain, and the loss is -log(gain / (theta - E0)), the log of the fraction of
the energy error one step removes (``tcp_learned_precond_ritz_loss``, which
applies H to the projected correction).  Forward pass, loss gradient,
backward pass and Adam update all run as HIP kernels.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any, List, Sequence

import numpy as np

from . import _lib
from .davidson import lowest_eigenpair

FEATURES = 64
HIDDEN = 128


class MLP:
    """The preconditioner network (device-resident FP32 master weights)."""

    def __init__(self, seed: int = 20260805) -> None:
        _lib.require_device()
        out = ctypes.c_void_p()
        _lib.check(_lib.lib().tcp_mlp_create(seed, ctypes.byref(out)))
        self._ptr = out

    def __del__(self) -> None:
        if getattr(self, "_ptr", None) and getattr(_lib, "_lib", None) is not None:
            _lib.lib().tcp_mlp_destroy(self._ptr)
            self._ptr = None

    @staticmethod
    def parameter_count() -> int:
        return int(_lib.lib().tcp_mlp_parameter_count())

    def get_parameters(self) -> np.ndarray:
        p = np.zeros(self.parameter_count(), dtype=np.float32)
        _lib.check(_lib.lib().tcp_mlp_get_parameters(self._ptr, p.ctypes.data))
        return p

    def set_parameters(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.shape != (self.parameter_count(),):
            raise ValueError("parameter vector has the wrong length")
        _lib.check(_lib.lib().tcp_mlp_set_parameters(self._ptr, p.ctypes.data))

    @staticmethod
    def unpack(params: np.ndarray):
        """(W1 [64,128], b1, W2 [128,128], b2, w3 [128], b3) views of a flat vector."""
        o = 0
        W1 = params[o:o + FEATURES * HIDDEN].reshape(FEATURES, HIDDEN)
        o += FEATURES * HIDDEN
        b1 = params[o:o + HIDDEN]
        o += HIDDEN
        W2 = params[o:o + HIDDEN * HIDDEN].reshape(HIDDEN, HIDDEN)
        o += HIDDEN * HIDDEN
        b2 = params[o:o + HIDDEN]
        o += HIDDEN
        w3 = params[o:o + HIDDEN]
        o += HIDDEN
        return W1, b1, W2, b2, w3, float(params[o])

    def forward(self, features: Any) -> Any:
        """MLP outputs for a device bf16 tensor [batch, 64] (float32 tensor)."""
        import torch

        from ._plan import current_stream_handle

        if features.dtype != torch.bfloat16 or features.dim() != 2 or features.shape[1] != FEATURES:
            raise ValueError("features must be a bfloat16 tensor of shape (batch, 64)")
        features = features.contiguous()
        out = torch.empty(features.shape[0], dtype=torch.float32, device=features.device)
        _lib.check(_lib.lib().tcp_mlp_forward(self._ptr, features.data_ptr(), features.shape[0], out.data_ptr(),
                                              current_stream_handle(features)))
        return out

    def gradient(self, features: Any, dloss_dout: Any) -> np.ndarray:
        """Parameter gradient of sum_b dloss_dout[b] * MLP(features[b])."""
        from ._plan import current_stream_handle

        grad = np.zeros(self.parameter_count(), dtype=np.float32)
        _lib.check(_lib.lib().tcp_mlp_gradient(self._ptr, features.data_ptr(), features.shape[0],
                                               dloss_dout.data_ptr(), grad.ctypes.data,
                                               current_stream_handle(features)))
        return grad

    def _gradient_raw(self, features_ptr: int, batch: int, dy: Any) -> None:
        _lib.check(_lib.lib().tcp_mlp_gradient(self._ptr, features_ptr, batch, dy.data_ptr(), None, 0))

    def adam_step(self, lr: float, weight_decay: float = 0.0) -> None:
        _lib.check(_lib.lib().tcp_mlp_adam_step(self._ptr, ctypes.c_float(lr), ctypes.c_float(weight_decay), None))


class LearnedPreconditioner:
    """Features of one restricted operator plus the trained MLP, as a Davidson hook.

    The hook re-evaluates the network only when theta has moved by more than
    ``refresh`` times the operator's coupling scale since its last evaluation
    (0: on every call).
    """

    def __init__(self, mlp: MLP, restricted: Any, refresh: float = 4.0) -> None:
        from .symmetry import SectorMvpPlan, lazy_plan_of

        plan = lazy_plan_of(restricted)
        if not isinstance(plan, SectorMvpPlan):
            raise TypeError("the learned preconditioner needs a restricted (sector) operator")
        out = ctypes.c_void_p()
        _lib.check(_lib.lib().tcp_learned_precond_create(mlp._ptr, plan._ptr, ctypes.byref(out)))
        self._ptr = out
        _lib.check(_lib.lib().tcp_learned_precond_set_refresh(out, ctypes.c_double(refresh)))
        self.mlp = mlp
        self._plan = plan  # the loss applies H through it: keep it alive
        self.dimension = int(plan.dimension)
        feats = ctypes.c_void_p()
        diag = ctypes.c_void_p()
        scale = ctypes.c_double()
        _lib.check(_lib.lib().tcp_learned_precond_buffers(out, ctypes.byref(feats), ctypes.byref(diag),
                                                           ctypes.byref(scale)))
        self.features_ptr = int(feats.value)
        self.diag_ptr = int(diag.value)
        self.scale = float(scale.value)

    def __del__(self) -> None:
        if getattr(self, "_ptr", None) and getattr(_lib, "_lib", None) is not None:
            _lib.lib().tcp_learned_precond_destroy(self._ptr)
            self._ptr = None

    def callback(self):
        fn = ctypes.cast(_lib.lib().tcp_learned_precond_log_multiplier, ctypes.c_void_p).value
        return fn, self._ptr.value

    def log_multipliers(self, theta: float) -> int:
        """Device pointer to float[dimension] log diagonal for `theta` (default stream)."""
        ptr = _lib.lib().tcp_learned_precond_log_multiplier(self._ptr, theta, None)
        if not ptr:
            raise _lib.DeviceError(_lib.last_error())
        return int(ptr)


@dataclass
class Snapshot:
    """One training example: a Jacobi Davidson state of a small lattice."""

    precond: LearnedPreconditioner
    theta: float
    u: Any  # device complex128 [D], normalised Ritz vector
    r: Any  # H u - theta u
    e0: float  # the lattice's ground energy


def jacobi_snapshots(restricted: Any, mlp: MLP, iterations: Sequence[int] = (1, 2, 3, 5, 8, 12),
                     seed: int = 20260801) -> List[Snapshot]:
    """Snapshots after `iterations` steps of plain Jacobi Davidson on `restricted`."""
    import torch

    e0 = lowest_eigenpair(restricted, tol=1e-10, seed=seed).eigenvalue
    precond = LearnedPreconditioner(mlp, restricted)
    out = []
    for k in iterations:
        state = lowest_eigenpair(restricted, tol=1e-14, max_iterations=k, seed=seed, return_vector=True)
        u = state.eigenvector
        hu = restricted.apply(u)
        theta = torch.vdot(u, hu).real.item()
        if theta - e0 <= 1e-9 * max(1.0, abs(e0)):
            continue
        r = hu - theta * u
        out.append(Snapshot(precond, theta, u.contiguous(), r.contiguous(), e0))
    return out


def _ritz(snap: Snapshot, logm: Any, dlogm: Any) -> float:
    loss = ctypes.c_double()
    gain = ctypes.c_double()
    _lib.check(_lib.lib().tcp_learned_precond_ritz_loss(
        snap.precond._ptr, snap.theta, snap.u.data_ptr(), snap.r.data_ptr(),
        None if logm is None else logm.data_ptr(), None if dlogm is None else dlogm.data_ptr(),
        ctypes.byref(loss), ctypes.byref(gain), None))
    return gain.value / (snap.theta - snap.e0)


def snapshot_loss(mlp: MLP, snap: Snapshot, *, train: bool = False, lr: float = 1e-3) -> float:
    """-log of the fraction of the energy error one learned correction removes;
    with `train`, one Adam step on it."""
    import torch

    D = snap.precond.dimension
    lib = _lib.lib()
    _lib.check(lib.tcp_learned_precond_prepare(snap.precond._ptr, snap.theta, None))
    y = torch.empty(D, dtype=torch.float32, device="cuda")
    _lib.check(lib.tcp_mlp_forward(mlp._ptr, snap.precond.features_ptr, D, y.data_ptr(), None))
    dy = torch.empty(D, dtype=torch.float32, device="cuda")
    fraction = _ritz(snap, y, dy)
    if train:
        _lib.check(lib.tcp_mlp_gradient(mlp._ptr, snap.precond.features_ptr, D, dy.data_ptr(), None, None))
        mlp.adam_step(lr)
    return -float(np.log(fraction))


def jacobi_loss(snap: Snapshot) -> float:
    """The same loss for the plain Jacobi correction r_i / (theta - H_ii)."""
    return -float(np.log(_ritz(snap, None, None)))


def train(mlp: MLP, snapshots: Sequence[Snapshot], *, epochs: int = 30, lr: float = 2e-3,
          seed: int = 0) -> List[float]:
    """Adam over the snapshots (one step per snapshot, shuffled each epoch);
    returns the mean loss of every epoch."""
    rng = np.random.default_rng(seed)
    history = []
    order = np.arange(len(snapshots))
    for _ in range(epochs):
        rng.shuffle(order)
        losses = [snapshot_loss(mlp, snapshots[i], train=True, lr=lr) for i in order]
        history.append(float(np.mean(losses)))
    return history
