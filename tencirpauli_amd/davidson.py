"""Davidson ground-state solver on the GPU (``tcp_davidson_*`` in the C ABI).

No reference counterpart: the reference computes eigenvalues with SciPy
``eigsh`` over a ``LinearOperator`` (py/integrations/scipy.py:13-47).  The
basis and its products stay on the device; orthogonalisation coefficients,
Rayleigh-quotient entries and residual norms accumulate in FP64.  With
``mvp_dtype="complex64"`` a first phase stores the basis in complex64 and runs
complex64 MVPs; at ``switch_tol`` its Ritz vector is formed in FP64 and checked
with one FP64 MVP, so a converged result belongs to the FP64 operator.  The
correction uses the Jacobi (diagonal) preconditioner or a learned diagonal
(``learned.LearnedPreconditioner``).  Full-space plans, lazy restricted plans
and eager (CSR) restricted operators are accepted.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any, Optional

import numpy as np

from . import _lib

LOG_MULTIPLIER_FN = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p)


class DavidsonOptions(ctypes.Structure):
    _fields_ = [
        ("max_subspace", ctypes.c_uint32),
        ("restart_keep", ctypes.c_uint32),
        ("max_iterations", ctypes.c_uint32),
        ("residual_tol", ctypes.c_double),
        ("mvp_dtype", ctypes.c_int),
        ("switch_tol", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("log_multiplier_fn", ctypes.c_void_p),
        ("log_multiplier_ctx", ctypes.c_void_p),
        ("log_multiplier_len", ctypes.c_uint64),
        ("trace", ctypes.POINTER(ctypes.c_double)),
        ("trace_capacity", ctypes.c_uint32),
        ("stream", ctypes.c_void_p),
    ]


class DavidsonResultStruct(ctypes.Structure):
    _fields_ = [
        ("eigenvalue", ctypes.c_double),
        ("residual_norm", ctypes.c_double),
        ("seconds", ctypes.c_double),
        ("iterations", ctypes.c_uint32),
        ("mvps_c64", ctypes.c_uint32),
        ("mvps_c128", ctypes.c_uint32),
        ("restarts", ctypes.c_uint32),
        ("trace_len", ctypes.c_uint32),
        ("converged", ctypes.c_uint32),
        ("final_fp64", ctypes.c_uint32),
    ]


@dataclass
class DavidsonResult:
    eigenvalue: float
    residual_norm: float
    seconds: float
    iterations: int
    mvps_c64: int
    mvps_c128: int
    restarts: int
    converged: bool
    final_fp64: bool
    trace: np.ndarray  # [iterations, 4]: seconds, theta, ||r||, MVP bits
    eigenvector: Any = None

    def seconds_to(self, reference: float, tolerance: float = 1e-8) -> Optional[float]:
        """Wall time after which every later FP64-phase Ritz value stays within
        `tolerance` of `reference` (None if never)."""
        ok = (np.abs(self.trace[:, 1] - reference) <= tolerance) & (self.trace[:, 3] == 64)
        if not ok.any() or not ok[-1]:
            return None
        last_bad = np.where(~ok)[0]
        first = 0 if last_bad.size == 0 else int(last_bad[-1]) + 1
        return float(self.trace[first, 0])


def _plan_handle(plan: Any):
    """(handle, solver entry point, diagonal entry point, plan object)."""
    from .eager import CsrMvpPlan
    from .pauli import NativeMVPPlan
    from .symmetry import lazy_plan_of

    L = _lib.lib()
    if getattr(plan, "storage", None) == "eager" and hasattr(plan, "_ensure_eager"):
        plan = plan.plan
    if isinstance(plan, CsrMvpPlan):
        return plan._ptr, L.tcp_davidson_csr, L.tcp_csr_plan_diagonal, plan
    inner = lazy_plan_of(plan)
    handle = getattr(inner, "_ptr", None)
    if handle is None:
        raise TypeError("expected a NativeMVPPlan, SectorMvpPlan, CsrMvpPlan or restricted operator")
    if isinstance(inner, NativeMVPPlan):
        return handle, L.tcp_davidson_full, L.tcp_mvp_plan_diagonal, inner
    return handle, L.tcp_davidson_sector, L.tcp_sector_plan_diagonal, inner


def lowest_eigenpair(plan: Any, *, tol: float = 1e-8, mvp_dtype: str = "complex128", switch_tol: float = 1e-4,
                     max_subspace: int = 24, restart_keep: int = 4, max_iterations: int = 1000,
                     seed: int = 20260801, preconditioner: Any = None, return_vector: bool = False,
                     trace_capacity: int = 4096) -> DavidsonResult:
    """Lowest eigenpair of a full-space or restricted plan by Davidson iteration."""
    handle, solve, _, inner = _plan_handle(plan)
    opt = DavidsonOptions()
    _lib.lib().tcp_davidson_default_options(ctypes.byref(opt))
    opt.max_subspace = max_subspace
    opt.restart_keep = restart_keep
    opt.max_iterations = max_iterations
    opt.residual_tol = tol
    opt.mvp_dtype = _lib.TCP_C64 if mvp_dtype in ("complex64", "c64") else _lib.TCP_C128
    opt.switch_tol = switch_tol
    opt.seed = seed
    trace = np.zeros((trace_capacity, 4), dtype=np.float64)
    opt.trace = trace.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    opt.trace_capacity = trace_capacity
    if preconditioner is not None:
        if int(preconditioner.dimension) != int(inner.dimension):
            raise ValueError(f"expected structure length {inner.dimension}, got {preconditioner.dimension} "
                             "(the preconditioner was built for another plan)")
        opt.log_multiplier_fn, opt.log_multiplier_ctx = preconditioner.callback()
        opt.log_multiplier_len = int(preconditioner.dimension)
    vector = None
    vec_ptr = None
    if return_vector:
        import torch

        vector = torch.empty(inner.dimension, dtype=torch.complex128, device="cuda")
        vec_ptr = vector.data_ptr()
    res = DavidsonResultStruct()
    _lib.check(solve(handle, ctypes.byref(opt), vec_ptr, ctypes.byref(res)))
    return DavidsonResult(
        eigenvalue=res.eigenvalue, residual_norm=res.residual_norm, seconds=res.seconds,
        iterations=res.iterations, mvps_c64=res.mvps_c64, mvps_c128=res.mvps_c128, restarts=res.restarts,
        converged=bool(res.converged), final_fp64=bool(res.final_fp64),
        trace=trace[: res.trace_len].copy(), eigenvector=vector)


def diagonal(plan: Any) -> Any:
    """Real diagonal H_ii of a plan as a device float64 tensor."""
    import torch

    from ._plan import current_stream_handle

    handle, _, diag_fn, inner = _plan_handle(plan)
    out = torch.empty(inner.dimension, dtype=torch.float64, device="cuda")
    _lib.check(diag_fn(handle, out.data_ptr(), current_stream_handle(out)))
    return out
