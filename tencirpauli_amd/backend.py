"""BackendMVPPlan: the schema-v1 pure-array MVP plan as a GPU input.

Mirrors py/hamiltonian.py ``BackendMVPPlan`` (346-537) for Pauli plans: the
same fields, factory-only construction, read-only C-contiguous buffers and
validation messages (``__post_init__`` :383-454), ``dimension``,
``term_count``, ``__call__`` and ``apply(state, max_bytes)`` with the
``dimension * 80`` working-memory estimate (:156-207).  Where the reference
executes the arrays with NumPy, ``apply`` here compiles them once into the
full-space gfx950 plan (``tcp_mvp_plan_create_from_words``: the arrays are
the builder's input format, core/hamiltonian.rs:297-309, 501-524) and runs
every product on the GPU; ``native_mvp_plan`` hands that plan to callers.
Direct-Weyl qudit plans are outside the qubit MVP path and are rejected.
"""

from __future__ import annotations

import ctypes
import threading
from dataclasses import dataclass, field
from typing import Any, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._plan import DEFAULT_MAX_BYTES, check_allocation, is_device_tensor, validate_max_bytes

PLAN_FACTORY_TOKEN = object()

_MAPPINGS = (None, "jordan_wigner", "parity", "bravyi_kitaev")


@dataclass(frozen=True)
class BackendMVPPlan:
    """Versioned pure-array Pauli MVP plan (packed X/Z words, qubit q at bit q)."""

    schema_version: int
    nqubits: int
    word_count: int
    x_words: np.ndarray
    z_words: np.ndarray
    coefficients: np.ndarray
    ordering: str = "qubit0_msb_matrix"
    integer_width: int = 64
    required_operations: Tuple[str, ...] = ("xor", "phase", "scatter_add")
    local_dimensions: Tuple[int, ...] = ()
    basis_ordering: str = "qubit0_msb_matrix"
    estimated_bytes: int = 0
    plan_kind: str = "pauli"
    target: str = "backend_mvp"
    source_term_count: int = -1
    plan_term_count: int = -1
    mapping: Optional[str] = None
    _factory_token: object = field(default=None, repr=False, compare=False)

    def __post_init__(self) -> None:
        if self._factory_token is not PLAN_FACTORY_TOKEN:
            raise TypeError("BackendMVPPlan instances must be created by a plan factory")
        for name, dtype in (("x_words", np.uint64), ("z_words", np.uint64), ("coefficients", np.complex128)):
            values = np.array(getattr(self, name), dtype=dtype, order="C", copy=True)
            values.setflags(write=False)
            object.__setattr__(self, name, values)
        if self.plan_kind != "pauli":
            if self.plan_kind == "direct_weyl":
                raise ValueError("direct Weyl backend plans are not supported: only Pauli plans run on this library")
            raise ValueError("unknown backend MVP plan kind")
        for name, value in (("schema_version", self.schema_version), ("nqubits", self.nqubits),
                            ("word_count", self.word_count), ("integer_width", self.integer_width),
                            ("estimated_bytes", self.estimated_bytes)):
            if not isinstance(value, int) or isinstance(value, bool) or value < 0:
                raise ValueError(f"{name} must be a non-negative integer")
        if self.schema_version <= 0:
            raise ValueError("schema_version must be positive")
        if self.target != "backend_mvp":
            raise ValueError("backend MVP plans must have target='backend_mvp'")
        object.__setattr__(self, "local_dimensions", tuple(self.local_dimensions))
        object.__setattr__(self, "required_operations", tuple(self.required_operations))
        for name in ("source_term_count", "plan_term_count"):
            value = getattr(self, name)
            if not isinstance(value, int) or isinstance(value, bool):
                raise ValueError(f"{name} must be an integer")
            if value < 0:
                object.__setattr__(self, name, len(self.coefficients))
        if self.mapping not in _MAPPINGS:
            raise ValueError("mapping must be None, 'jordan_wigner', 'parity', or 'bravyi_kitaev'")
        if not np.isfinite(self.coefficients).all():
            raise ValueError("backend plan coefficients must be finite")
        if self.word_count != (self.nqubits + 63) // 64:
            raise ValueError("Pauli word_count does not match nqubits")
        if self.x_words.ndim != 2 or self.z_words.ndim != 2:
            raise ValueError("Pauli word arrays must be two-dimensional")
        expected = (len(self.coefficients), self.word_count)
        if self.x_words.shape != expected or self.z_words.shape != expected:
            raise ValueError("Pauli word arrays must match plan term and word counts")
        if self.local_dimensions and self.local_dimensions != (2,) * self.nqubits:
            raise ValueError("Pauli local dimensions must be binary")
        object.__setattr__(self, "_lock", threading.Lock())
        object.__setattr__(self, "_compiled", None)

    @property
    def dimension(self) -> int:
        if self.nqubits >= 63:
            raise OverflowError("finite basis dimension cannot be represented by platform indices")
        return 1 << self.nqubits

    @property
    def term_count(self) -> int:
        return len(self.coefficients)

    def to_operator(self) -> Any:
        """The PauliOperator these arrays encode (one native construction from the words)."""
        from .pauli import PauliOperator

        return PauliOperator.from_words(self.nqubits, self.x_words, self.z_words, self.coefficients)

    def native_mvp_plan(self, *, storage: str = "lazy", max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        """A full-space gfx950 plan compiled from these arrays."""
        return self.to_operator().native_mvp_plan(storage=storage, max_bytes=max_bytes)

    def _plan(self) -> Any:
        with self._lock:  # type: ignore[attr-defined]
            if self._compiled is None:  # type: ignore[attr-defined]
                object.__setattr__(self, "_compiled", self.native_mvp_plan(max_bytes=None))
            return self._compiled  # type: ignore[attr-defined]

    def apply(self, state: Sequence[complex] | Any, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        """H @ state on the GPU; NumPy input gives an owned complex128 array.

        Raises ValueError on a shape mismatch and MemoryError when the
        ``dimension * 80``-byte working estimate exceeds an enforced max_bytes.
        """
        validate_max_bytes(max_bytes)
        dimension = self.dimension
        check_allocation(dimension * 80, max_bytes, "backend MVP working memory")
        if is_device_tensor(state):
            if state.dim() != 1 or state.shape[0] != dimension:
                raise ValueError(f"state must have shape ({dimension},), got {tuple(state.shape)}")
            return self._plan().apply(state, max_bytes=None)
        values = np.asarray(state, dtype=np.complex128)
        if values.ndim != 1 or values.shape[0] != dimension:
            raise ValueError(f"state must have shape ({dimension},), got {values.shape}")
        return np.ascontiguousarray(self._plan().apply(values, max_bytes=None), dtype=np.complex128)

    def __call__(self, state: Sequence[complex] | Any) -> Any:
        return self.apply(state)


def plan_from_operator(operator: Any, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> BackendMVPPlan:
    """The factory behind ``PauliOperator.backend_mvp_plan`` (py/pauli.py:954-983)."""
    validate_max_bytes(max_bytes)
    T, n = operator.term_count, operator.nqubits
    W = (n + 63) // 64
    x = np.zeros((T, W), dtype=np.uint64)
    z = np.zeros((T, W), dtype=np.uint64)
    coef = np.zeros(T, dtype=np.complex128)
    u64 = ctypes.POINTER(ctypes.c_uint64)
    _lib.check(_lib.lib().tcp_operator_export_words(operator._handle.ptr, x.ctypes.data_as(u64),
                                                    z.ctypes.data_as(u64),
                                                    coef.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return BackendMVPPlan(1, n, W, x, z, coef, local_dimensions=(2,) * n, basis_ordering="qubit0_msb_matrix",
                          estimated_bytes=int(T * (W * 16 + 16)), _factory_token=PLAN_FACTORY_TOKEN)


__all__ = ["BackendMVPPlan", "plan_from_operator"]
