"""Shared plan facade: the public-call contract of the reference's native MVP plans.

Mirrors py/_native_mvp.py: ``DEFAULT_MAX_BYTES`` (:19), the advisory 1 GiB
floor of ``max_bytes`` (:28-58), the strict ``apply_into`` buffer protocol
(:61-81) and ``apply``/``apply_into``/``__call__`` (:272-308).  Numerical
application is one call into the gfx950 library.

Two buffer kinds are accepted:

* NumPy ``complex128`` host arrays, exactly as the reference accepts them;
  they are staged through device memory by the library;
* ``torch`` tensors on a ROCm device (``complex128`` -> FP64 path,
  ``complex64`` -> FP32 storage with FP64 accumulation), applied in place on
  torch's current stream without a host round trip.
"""

from __future__ import annotations

import ctypes
import sys
from dataclasses import dataclass
from typing import Any, Optional, Sequence

import numpy as np

from . import _lib

DEFAULT_MAX_BYTES = 16 * 1024 * 1024 * 1024
MIN_ENFORCED_MAX_BYTES = 1024 * 1024 * 1024


def validate_max_bytes(max_bytes: Optional[int]) -> None:
    if max_bytes is not None and (
        not isinstance(max_bytes, int) or isinstance(max_bytes, bool) or max_bytes < 0
    ):
        raise ValueError("max_bytes must be a non-negative integer or None")


def enforced_max_bytes(max_bytes: Optional[int]) -> Optional[int]:
    validate_max_bytes(max_bytes)
    if max_bytes is None or max_bytes < MIN_ENFORCED_MAX_BYTES:
        return None
    return max_bytes


def effective_max_bytes(max_bytes: Optional[int]) -> int:
    limit = enforced_max_bytes(max_bytes)
    return min(sys.maxsize if limit is None else limit, (1 << 64) - 1)


def check_allocation(requested: int, limit: Optional[int], context: str) -> None:
    enforced = enforced_max_bytes(limit)
    if enforced is not None and requested > enforced:
        raise MemoryError(
            f"{context} requires approximately {requested} bytes, exceeding max_bytes={enforced}"
        )


def is_device_tensor(value: object) -> bool:
    mod = type(value).__module__
    if not mod.startswith("torch"):
        return False
    return bool(getattr(value, "is_cuda", False))


def validate_apply_into_buffers(input_state: object, output_state: object, dimension: int) -> None:
    """The strict zero-copy buffer protocol of py/_native_mvp.py:61-81."""
    if not isinstance(input_state, np.ndarray) or not isinstance(output_state, np.ndarray):
        raise TypeError("apply_into input and output must be NumPy arrays")
    expected = np.dtype(np.complex128)
    if input_state.dtype != expected or output_state.dtype != expected:
        raise TypeError("apply_into input and output must have dtype complex128")
    for name, value in (("input", input_state), ("output", output_state)):
        if value.ndim != 1 or value.shape != (dimension,):
            raise ValueError(f"apply_into {name} must have shape ({dimension},)")
        if not value.flags.c_contiguous:
            raise ValueError(f"apply_into {name} must be C-contiguous")
    if not output_state.flags.writeable:
        raise ValueError("apply_into output must be writable")
    if np.shares_memory(input_state, output_state):
        raise ValueError("apply_into input and output must not overlap")


def validate_device_buffers(input_state: Any, output_state: Any, dimension: int) -> int:
    """Device-tensor counterpart of the buffer protocol; returns the tcp dtype."""
    import torch

    if not (is_device_tensor(input_state) and is_device_tensor(output_state)):
        raise TypeError("apply_into input and output must both be device tensors")
    if input_state.dtype != output_state.dtype:
        raise TypeError("apply_into input and output must have the same dtype")
    if input_state.dtype == torch.complex128:
        dtype = _lib.TCP_C128
    elif input_state.dtype == torch.complex64:
        dtype = _lib.TCP_C64
    else:
        raise TypeError("device tensors must be complex128 or complex64")
    for name, value in (("input", input_state), ("output", output_state)):
        if value.dim() != 1 or tuple(value.shape) != (dimension,):
            raise ValueError(f"apply_into {name} must have shape ({dimension},)")
        if not value.is_contiguous():
            raise ValueError(f"apply_into {name} must be C-contiguous")
    if input_state.device != output_state.device:
        raise ValueError("apply_into input and output must live on the same device")
    a0, a1 = input_state.data_ptr(), input_state.data_ptr() + input_state.numel() * input_state.element_size()
    b0, b1 = output_state.data_ptr(), output_state.data_ptr() + output_state.numel() * output_state.element_size()
    if a0 < b1 and b0 < a1:
        raise ValueError("apply_into input and output must not overlap")
    return dtype


def current_stream_handle(tensor: Any) -> int:
    import torch

    return int(torch.cuda.current_stream(tensor.device).cuda_stream)


def _dptr(array: np.ndarray) -> Any:
    return array.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass(frozen=True)
class NativePlanMetadata:
    """The metadata every native MVP plan exposes (py/_native_mvp.py:84-150)."""

    dimension: int
    term_count: int
    storage: str
    basis_ordering: str
    target: str
    execution_family: str
    native_variant: str
    retained_format: str
    source_term_count: int
    schedule_count: int
    estimated_retained_bytes: int
    estimated_construction_peak_bytes: int
    transition_count: Optional[int]

    def __post_init__(self) -> None:
        for name in ("dimension", "term_count", "source_term_count", "schedule_count", "estimated_retained_bytes",
                     "estimated_construction_peak_bytes"):
            value = getattr(self, name)
            if not isinstance(value, int) or isinstance(value, bool) or value < 0:
                raise ValueError(f"{name} must be a non-negative integer")
        if self.storage not in {"lazy", "eager"}:
            raise ValueError("storage must be either 'eager' or 'lazy'")
        if self.target != "native_mvp":
            raise ValueError("native MVP plans must have target='native_mvp'")
        for name in ("basis_ordering", "execution_family", "native_variant", "retained_format"):
            if not isinstance(getattr(self, name), str) or not getattr(self, name):
                raise ValueError(f"{name} must be a non-empty string")
        if self.transition_count is not None and (not isinstance(self.transition_count, int)
                                                  or self.transition_count < 0):
            raise ValueError("transition_count must be None or a non-negative integer")


class NativePlanFacade:
    """apply / apply_into / __call__ shared by full-space and sector plans.

    Subclasses provide ``dimension`` and the three native calls
    ``_native_apply_host``, ``_native_apply_device`` and the plan handle.
    """

    dimension: int

    @property
    def metadata(self) -> NativePlanMetadata:
        """The plan's metadata record (the fields of the reference's facades)."""
        term_count = int(getattr(self, "term_count", 0))
        retained = int(getattr(self, "estimated_retained_bytes", getattr(self, "estimated_bytes", 0)))
        transition = getattr(self, "transition_count", None)
        return NativePlanMetadata(
            dimension=int(self.dimension), term_count=term_count, storage=str(getattr(self, "storage", "lazy")),
            basis_ordering=str(getattr(self, "basis_ordering", "qubit0_msb_matrix")),
            target=str(getattr(self, "target", "native_mvp")),
            execution_family=str(getattr(self, "execution_family", "")),
            native_variant=str(getattr(self, "native_variant", "")),
            retained_format=str(getattr(self, "retained_format", "compiled_actions")),
            source_term_count=int(getattr(self, "source_term_count", term_count)),
            schedule_count=int(getattr(self, "schedule_count", term_count)),
            estimated_retained_bytes=retained,
            estimated_construction_peak_bytes=int(getattr(self, "estimated_construction_peak_bytes", retained)),
            transition_count=None if transition is None else int(transition))

    def _native_apply_host(self, inp: np.ndarray, out: np.ndarray) -> int:  # pragma: no cover
        raise NotImplementedError

    def _native_apply_device(self, inp: Any, out: Any, dtype: int, stream: int) -> int:  # pragma: no cover
        raise NotImplementedError

    def apply(self, state: Sequence[complex] | Any, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        """Apply the plan through exactly one coarse native call."""
        validate_max_bytes(max_bytes)
        if is_device_tensor(state):
            import torch

            if state.dim() != 1 or state.shape[0] != self.dimension:
                raise ValueError(f"state must have shape ({self.dimension},), got {tuple(state.shape)}")
            values = state.contiguous()
            out = torch.empty_like(values)
            self.apply_into(values, out, max_bytes=max_bytes)
            return out
        values = np.asarray(state, dtype=np.complex128)
        if values.ndim != 1 or values.shape[0] != self.dimension:
            raise ValueError(f"state must have shape ({self.dimension},), got {values.shape}")
        check_allocation(self.dimension * 16, max_bytes, "requested native MVP output")
        values = np.ascontiguousarray(values)
        out = np.empty(self.dimension, dtype=np.complex128)
        _lib.check(self._native_apply_host(values, out))
        return out

    def apply_into(self, input_state: Any, output_state: Any, *,
                   max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> None:
        """Apply into strict caller-owned buffers (output overwritten)."""
        validate_max_bytes(max_bytes)
        if is_device_tensor(input_state) or is_device_tensor(output_state):
            dtype = validate_device_buffers(input_state, output_state, self.dimension)
            _lib.check(self._native_apply_device(input_state, output_state, dtype,
                                                 current_stream_handle(input_state)))
            return
        validate_apply_into_buffers(input_state, output_state, self.dimension)
        _lib.check(self._native_apply_host(input_state, output_state))

    def __call__(self, state: Sequence[complex] | Any) -> Any:
        return self.apply(state)

    def to_scipy_linear_operator(self, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        """Expose this plan as a SciPy LinearOperator (py/integrations/scipy.py:13-47)."""
        from .integrations.scipy import to_scipy_linear_operator

        return to_scipy_linear_operator(self, max_bytes=max_bytes)
