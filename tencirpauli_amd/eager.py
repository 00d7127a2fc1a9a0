"""Eager restricted operators: the sector-restricted matrix as CSR on the GPU.

Replaces ``U1RestrictedOperator::new`` + ``CsrGraph`` (core/sector.rs:345-432,
core/csr.rs:196-240) and the charge eager build (core/charge/eager.rs) behind
``restrict_charge(..., storage="eager")`` and ``mvp_plan(storage="eager")``.
Rows are destination ranks with ascending columns; the ``dense``/``coo``/
``csr`` materialisations (py/symmetry.py:477-512) read the device CSR back.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._plan import DEFAULT_MAX_BYTES, NativePlanFacade, check_allocation, effective_max_bytes, validate_max_bytes


@dataclass(frozen=True)
class COOMatrix:
    """Coordinate arrays, row-major with ascending columns (py/hamiltonian.py:110-133)."""

    row: np.ndarray
    column: np.ndarray
    data: np.ndarray
    shape: Tuple[int, int]

    @property
    def value(self) -> np.ndarray:
        return self.data

    def to_scipy(self) -> Any:
        from scipy.sparse import coo_matrix

        return coo_matrix((self.data, (self.row, self.column)), shape=self.shape)


@dataclass(frozen=True)
class CSRMatrix:
    """CSR arrays (py/hamiltonian.py:136-159)."""

    indptr: np.ndarray
    indices: np.ndarray
    data: np.ndarray
    shape: Tuple[int, int]

    @property
    def value(self) -> np.ndarray:
        return self.data

    def to_scipy(self) -> Any:
        from scipy.sparse import csr_matrix

        return csr_matrix((self.data, self.indices, self.indptr), shape=self.shape)


def _frozen(array: np.ndarray) -> np.ndarray:
    array.flags.writeable = False
    return array


def materialization_bytes(dimension: int, transitions: int, target: str) -> int:
    """py/symmetry.py:646-655."""
    if target == "dense":
        return dimension * dimension * 16
    if target == "coo":
        return transitions * 32
    if target == "csr":
        return (dimension + 1) * 8 + transitions * 24
    raise ValueError(f"unsupported U1 materialization target: {target}")


class CsrMvpPlan(NativePlanFacade):
    """Eager restricted plan: device CSR, row-parallel gather."""

    storage = "eager"
    retained_format = "csr"
    basis_ordering = "qubit0_msb_matrix"
    target = "native_mvp"

    def __init__(self, ptr: ctypes.c_void_p, native_variant: str, execution_family: str, term_count: int) -> None:
        self._ptr = ptr
        L = _lib.lib()
        self.dimension = int(L.tcp_csr_plan_dimension(ptr))
        self.transition_count = int(L.tcp_csr_plan_transition_count(ptr))
        self.device_bytes = int(L.tcp_csr_plan_estimated_bytes(ptr))
        # "compact": 24-bit column + 8-bit index into a table of the distinct
        # off-diagonal values per entry, diagonal in a dense array; "slab": the
        # compact entries regrouped by x-mask, one word per (x-group, row);
        # "plain": a value per entry.  The SpMV reads whichever form the plan
        # holds.
        self.storage_format = {0: "plain", 1: "compact", 2: "slab"}[L.tcp_csr_plan_storage_form(ptr)]
        self.slab_groups = int(L.tcp_csr_plan_slab_groups(ptr))  # x-group slots of the slab form (else 0)
        self.native_variant = native_variant
        self.execution_family = execution_family
        self.term_count = term_count
        self.schedule_count = term_count
        # the reference's retained-bytes formula for eager U1 plans (py/symmetry.py:565-569)
        self.estimated_bytes = self.dimension * 16 + (self.dimension + 1) * 8 + self.transition_count * 32
        self.estimated_retained_bytes = self.estimated_bytes
        self.estimated_construction_peak_bytes = self.estimated_bytes

    def __del__(self) -> None:
        if getattr(self, "_ptr", None) and getattr(_lib, "_lib", None) is not None:
            _lib.lib().tcp_csr_plan_destroy(self._ptr)
            self._ptr = None

    @property
    def handle(self) -> int:
        return int(self._ptr.value or 0)

    def _native_apply_host(self, inp: np.ndarray, out: np.ndarray) -> int:
        dp = ctypes.POINTER(ctypes.c_double)
        return _lib.lib().tcp_csr_plan_apply_host(self._ptr, inp.ctypes.data_as(dp), out.ctypes.data_as(dp),
                                                  inp.shape[0], out.shape[0])

    def _native_apply_device(self, inp: Any, out: Any, dtype: int, stream: int) -> int:
        return _lib.lib().tcp_csr_plan_apply_into(self._ptr, inp.data_ptr(), out.data_ptr(), inp.numel(), out.numel(),
                                                  dtype, stream)

    def csr_arrays(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(indptr u64, indices u64, values complex128) read back from the device."""
        indptr = np.empty(self.dimension + 1, dtype=np.uint64)
        indices = np.empty(self.transition_count, dtype=np.uint64)
        values = np.empty(self.transition_count, dtype=np.complex128)
        u64 = ctypes.POINTER(ctypes.c_uint64)
        _lib.check(_lib.lib().tcp_csr_plan_export(self._ptr, indptr.ctypes.data_as(u64), indices.ctypes.data_as(u64),
                                                  values.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return indptr, indices, values

    def materialize(self, target: str, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        validate_max_bytes(max_bytes)
        n = self.dimension
        check_allocation(materialization_bytes(n, self.transition_count, target), max_bytes,
                         f"U1 {target} materialization")
        indptr, indices, values = self.csr_arrays()
        if target == "csr":
            return CSRMatrix(_frozen(indptr), _frozen(indices), _frozen(values), (n, n))
        rows = np.repeat(np.arange(n, dtype=np.uint64), np.diff(indptr).astype(np.int64))
        if target == "coo":
            return COOMatrix(_frozen(rows), _frozen(indices), _frozen(values), (n, n))
        dense = np.zeros((n, n), dtype=np.complex128)
        dense[rows.astype(np.int64), indices.astype(np.int64)] = values
        return dense

    # the retained arrays, read back on demand and read-only (py/symmetry.py
    # eager plans expose indptr / columns / coefficients this way)
    @property
    def indptr(self) -> np.ndarray:
        return _frozen(self.csr_arrays()[0])

    @property
    def columns(self) -> np.ndarray:
        return _frozen(self.csr_arrays()[1])

    @property
    def coefficients(self) -> np.ndarray:
        return _frozen(self.csr_arrays()[2])


def csr_plan_u1(operator: Any, particle_number: int, max_bytes: Optional[int]) -> CsrMvpPlan:
    _lib.require_device()
    out = ctypes.c_void_p()
    _lib.check(_lib.lib().tcp_csr_plan_create_u1(operator._handle.ptr, particle_number,
                                                  effective_max_bytes(max_bytes), ctypes.byref(out)))
    return CsrMvpPlan(out, "qubit", "product", operator.term_count)


def csr_plan_blocks(operator: Any, sizes: Sequence[int], pops: Sequence[int],
                    max_bytes: Optional[int]) -> CsrMvpPlan:
    _lib.require_device()
    s = np.ascontiguousarray(sizes, dtype=np.uint32)
    p = np.ascontiguousarray(pops, dtype=np.uint32)
    u32 = ctypes.POINTER(ctypes.c_uint32)
    out = ctypes.c_void_p()
    _lib.check(_lib.lib().tcp_csr_plan_create_blocks(operator._handle.ptr, len(s), s.ctypes.data_as(u32),
                                                      p.ctypes.data_as(u32), effective_max_bytes(max_bytes),
                                                      ctypes.byref(out)))
    return CsrMvpPlan(out, "factorized", "factorized", operator.term_count)
