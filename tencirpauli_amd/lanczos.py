"""``eigsh`` on the GPU: thick-restart Lanczos over an MVP plan (``tcp_lanczos_*``).

The reference computes eigenpairs by handing its plan to SciPy:
``eigsh(plan.to_scipy_linear_operator(), k=1, which="SA", v0=..., tol=...,
maxiter=..., ncv=...)`` (examples/research/fermi_hubbard/run_tencirpauli.py:
194-211 over py/integrations/scipy.py:13-47), i.e. ARPACK's restarted Lanczos
on the host with one native MVP per step.  ``eigsh`` here takes the same
arguments and returns the same ``(w, v)`` (eigenvalues ascending, eigenvectors
as columns), but the whole iteration runs on the device: basis, products,
reorthogonalisation and residual estimates; the host solves only the
ncv x ncv projected problem at each restart.  Convergence uses ARPACK's test
(residual estimate <= tol * max(|theta|, eps^(2/3)), tol = 0 meaning machine
precision) and a run out of restarts raises SciPy's ``ArpackNoConvergence``.
Shift-invert and generalised problems (``sigma``, ``M``) are not provided.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any, Optional

import numpy as np

from . import _lib
from ._plan import is_device_tensor

MAX_NCV = 32
_WHICH = {"SA": 0, "LA": 1, "LM": 2}


class LanczosOptions(ctypes.Structure):
    _fields_ = [
        ("nev", ctypes.c_uint32),
        ("ncv", ctypes.c_uint32),
        ("which", ctypes.c_uint32),
        ("max_restarts", ctypes.c_uint32),
        ("tol", ctypes.c_double),
        ("v0", ctypes.c_void_p),
        ("seed", ctypes.c_uint64),
        ("stream", ctypes.c_void_p),
        ("keep", ctypes.c_uint32),
    ]


class LanczosResultStruct(ctypes.Structure):
    _fields_ = [
        ("converged", ctypes.c_uint32),
        ("restarts", ctypes.c_uint32),
        ("mvps", ctypes.c_uint64),
        ("seconds", ctypes.c_double),
    ]


@dataclass
class LanczosResult:
    eigenvalues: np.ndarray  # ascending
    eigenvectors: Any        # device complex128 tensor [dimension, k], or None
    residual_norms: Optional[np.ndarray]
    converged: int
    restarts: int
    mvps: int
    seconds: float


def _entry(plan: Any):
    """(handle, native entry point, dimension) of a plan, restricted operator or PauliOperator."""
    from .davidson import _plan_handle
    from .pauli import PauliOperator

    if isinstance(plan, PauliOperator):
        plan = plan.native_mvp_plan()
    handle, davidson_fn, _, inner = _plan_handle(plan)
    solve = getattr(_lib.lib(), davidson_fn.__name__.replace("davidson", "lanczos"))
    return handle, solve, int(inner.dimension), inner


def lanczos(plan: Any, k: int = 1, *, which: str = "SA", v0: Any = None, ncv: Optional[int] = None,
            max_restarts: Optional[int] = None, tol: float = 0.0, seed: int = 20260801,
            return_eigenvectors: bool = True, residuals: bool = False,
            keep: Optional[int] = None) -> LanczosResult:
    """Lowest / highest / largest-magnitude eigenpairs with every step on the GPU."""
    import torch

    handle, solve, dim, inner = _entry(plan)
    if which not in _WHICH:
        raise ValueError(f"which must be one of {sorted(_WHICH)} (shift-invert modes are not provided)")
    if not isinstance(k, int) or isinstance(k, bool) or not 0 < k < dim:
        raise ValueError(f"k must be between 1 and ndim(A)-1 ({dim - 1})")
    if ncv is None:
        ncv = min(dim, max(2 * k + 1, 20), MAX_NCV)
    if not k < ncv <= min(dim, MAX_NCV):
        raise ValueError(f"ncv must satisfy k < ncv <= min(n, {MAX_NCV}), got ncv={ncv}")
    if tol < 0:
        raise ValueError("tol must be non-negative")
    opt = LanczosOptions()
    _lib.lib().tcp_lanczos_default_options(ctypes.byref(opt))
    opt.nev, opt.ncv, opt.which, opt.tol, opt.seed = k, ncv, _WHICH[which], float(tol), seed
    opt.max_restarts = min(int(max_restarts if max_restarts is not None else 10 * dim), 2 ** 32 - 1)
    if keep is not None:
        if not k <= keep < ncv:
            raise ValueError(f"keep must satisfy k <= keep < ncv, got keep={keep}")
        opt.keep = keep
    start = None
    if v0 is not None:
        if is_device_tensor(v0):
            start = v0.reshape(-1).to(torch.complex128).contiguous()
        else:
            start = torch.from_numpy(np.ascontiguousarray(np.asarray(v0, dtype=np.complex128).reshape(-1))).cuda()
        if start.shape[0] != dim:
            raise ValueError(f"v0 must have {dim} entries, got {start.shape[0]}")
        opt.v0 = start.data_ptr()
    vectors = torch.empty((k, dim), dtype=torch.complex128, device="cuda") if return_eigenvectors else None
    opt.stream = int(torch.cuda.current_stream().cuda_stream)
    values = np.zeros(k, dtype=np.float64)
    norms = np.zeros(k, dtype=np.float64) if residuals else None
    res = LanczosResultStruct()
    dp = ctypes.POINTER(ctypes.c_double)
    _lib.check(solve(handle, ctypes.byref(opt), values.ctypes.data_as(dp),
                     vectors.data_ptr() if vectors is not None else None,
                     norms.ctypes.data_as(dp) if norms is not None else None, ctypes.byref(res)))
    return LanczosResult(values, vectors.T if vectors is not None else None, norms, int(res.converged),
                         int(res.restarts), int(res.mvps), float(res.seconds))


def eigsh(A: Any, k: int = 6, M: Any = None, sigma: Any = None, which: str = "LM", v0: Any = None,
          ncv: Optional[int] = None, maxiter: Optional[int] = None, tol: float = 0, return_eigenvectors: bool = True,
          Minv: Any = None, OPinv: Any = None, mode: str = "normal") -> Any:
    """scipy.sparse.linalg.eigsh's interface for a plan, restricted operator or PauliOperator.

    Returns ``w`` (ascending) and, with ``return_eigenvectors``, ``v`` whose
    column i belongs to ``w[i]``: a NumPy array, or a device tensor when
    ``v0`` is one.
    """
    if M is not None or sigma is not None or Minv is not None or OPinv is not None or mode != "normal":
        raise NotImplementedError("generalised and shift-invert eigenproblems are not provided on the GPU")
    result = lanczos(A, k, which=which, v0=v0, ncv=ncv, max_restarts=maxiter, tol=tol,
                     return_eigenvectors=return_eigenvectors)
    vectors = result.eigenvectors
    if vectors is not None and not is_device_tensor(v0):
        vectors = vectors.cpu().numpy()
    if result.converged < k:
        from scipy.sparse.linalg import ArpackNoConvergence

        raise ArpackNoConvergence(
            f"ARPACK error -1: No convergence ({result.restarts} iterations, {result.converged}/{k} "
            "eigenvectors converged)", result.eigenvalues, vectors)
    return (result.eigenvalues, vectors) if return_eigenvectors else result.eigenvalues


__all__ = ["LanczosResult", "eigsh", "lanczos"]
