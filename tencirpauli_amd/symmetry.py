"""U(1) fixed-particle-number sectors (py/symmetry.py:144-241, 244-538, 658-706).

``U1Sector`` ranks and unranks in ascending integer (TensorCircuit) order with
qubit 0 as the most significant bit; the rank is the combinatorial number
system, computed by the native library (core/sector.rs:127-238).  A restricted
operator's lazy plan runs the gfx950 gather kernel.
"""

from __future__ import annotations

import ctypes
import math
import threading
from dataclasses import dataclass
from typing import Any, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._plan import (
    DEFAULT_MAX_BYTES,
    NativePlanFacade,
    check_allocation,
    effective_max_bytes,
    validate_max_bytes,
)


def _coerce_bitstring(nqubits: int, bitstring: Any) -> int:
    if isinstance(bitstring, bool):
        raise TypeError("bitstring must be an integer or a sequence of bits")
    if isinstance(bitstring, (int, np.integer)):
        value = int(bitstring)
    else:
        bits = tuple(bitstring)
        if len(bits) != nqubits or any(bit not in (0, 1) for bit in bits):
            raise ValueError("bitstring must contain exactly nqubits binary values")
        value = 0
        for bit in bits:
            value = (value << 1) | int(bit)
    if value < 0 or nqubits < value.bit_length():
        raise ValueError("bitstring is outside the computational basis")
    return value


def _word_count(nqubits: int) -> int:
    return max(1, (nqubits + 63) // 64)


def _packed_words(nqubits: int, value: int) -> np.ndarray:
    """Integer bitstring (qubit 0 = MSB) -> packed words, qubit q in word q // 64,
    bit q % 64 (core/word.rs:53-110)."""
    words = np.zeros(_word_count(nqubits), dtype=np.uint64)
    for q in range(nqubits):
        if (value >> (nqubits - 1 - q)) & 1:
            words[q >> 6] |= np.uint64(1 << (q & 63))
    return words


def _bits_of_words(nqubits: int, words: np.ndarray) -> Tuple[int, ...]:
    return tuple(int((int(words[q >> 6]) >> (q & 63)) & 1) for q in range(nqubits))


@dataclass(frozen=True)
class U1Sector:
    """Fixed-Hamming-weight basis with TensorCircuit integer ordering."""

    nqubits: int
    particle_number: int

    def __post_init__(self) -> None:
        if not isinstance(self.nqubits, int) or isinstance(self.nqubits, bool) or self.nqubits < 0:
            raise ValueError("nqubits must be a non-negative integer")
        if (not isinstance(self.particle_number, int) or isinstance(self.particle_number, bool)
                or not 0 <= self.particle_number <= self.nqubits):
            raise ValueError("particle_number must be between 0 and nqubits")
        if math.comb(self.nqubits, self.particle_number) > np.iinfo(np.uint64).max:
            raise OverflowError("U1 sector dimension exceeds uint64 restricted indices")

    @property
    def dimension(self) -> int:
        return math.comb(self.nqubits, self.particle_number)

    def rank(self, bitstring: Any) -> int:
        value = _coerce_bitstring(self.nqubits, bitstring)
        if bin(value).count("1") != self.particle_number:
            raise ValueError("bitstring has the wrong Hamming weight")
        out = ctypes.c_uint64()
        if self.nqubits > 64:
            words = _packed_words(self.nqubits, value)
            _lib.check(_lib.lib().tcp_u1_rank_words(self.nqubits, self.particle_number,
                                                    words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                                    ctypes.byref(out)))
        else:
            _lib.check(_lib.lib().tcp_u1_rank(self.nqubits, self.particle_number, value, ctypes.byref(out)))
        return int(out.value)

    def rank_words(self, words: Any) -> int:
        """Rank of packed words (qubit q in word q // 64, bit q % 64), the form
        of U1Sector::rank_words (core/sector.rs:127-238)."""
        w = np.ascontiguousarray(words, dtype=np.uint64)
        if w.shape != (_word_count(self.nqubits),):
            raise ValueError(f"expected {_word_count(self.nqubits)} packed words, got shape {w.shape}")
        out = ctypes.c_uint64()
        _lib.check(_lib.lib().tcp_u1_rank_words(self.nqubits, self.particle_number,
                                                w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(out)))
        return int(out.value)

    def unrank_words(self, index: int) -> np.ndarray:
        """Packed words of a restricted basis index (U1Sector::unrank_into)."""
        if not isinstance(index, (int, np.integer)) or isinstance(index, bool) or not 0 <= index < self.dimension:
            raise IndexError("restricted basis index is out of range")
        words = np.zeros(_word_count(self.nqubits), dtype=np.uint64)
        _lib.check(_lib.lib().tcp_u1_unrank_words(self.nqubits, self.particle_number, int(index),
                                                  words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return words

    def unrank(self, index: int) -> Tuple[int, ...]:
        if not isinstance(index, (int, np.integer)) or isinstance(index, bool) or not 0 <= index < self.dimension:
            raise IndexError("restricted basis index is out of range")
        if self.nqubits > 64:
            return _bits_of_words(self.nqubits, self.unrank_words(index))
        out = ctypes.c_uint64()
        _lib.check(_lib.lib().tcp_u1_unrank(self.nqubits, self.particle_number, int(index), ctypes.byref(out)))
        value = int(out.value)
        return tuple((value >> (self.nqubits - 1 - q)) & 1 for q in range(self.nqubits))

    def unrank_int(self, index: int) -> int:
        """Integer bitstring (qubit 0 = MSB) of a restricted basis index."""
        if self.nqubits > 64:
            value = 0
            for bit in self.unrank(index):
                value = (value << 1) | bit
            return value
        out = ctypes.c_uint64()
        _lib.check(_lib.lib().tcp_u1_unrank(self.nqubits, self.particle_number, int(index), ctypes.byref(out)))
        return int(out.value)

    def basis_words_packed(self, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> np.ndarray:
        """Read-only uint64 rows in restricted order: the integer bitstring for
        n <= 64, else the packed words (py/symmetry.py:230-240)."""
        validate_max_bytes(max_bytes)
        width = _word_count(self.nqubits)
        check_allocation(self.dimension * width * 8, max_bytes, "U1 packed basis")
        out = np.zeros((self.dimension, width), dtype=np.uint64)
        _lib.check(_lib.lib().tcp_u1_basis_words(self.nqubits, self.particle_number,
                                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                                 self.dimension))
        out.flags.writeable = False
        return out

    def basis_states(self, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> np.ndarray:
        """Read-only uint8 rows (qubit 0 first) in restricted order."""
        validate_max_bytes(max_bytes)
        check_allocation(self.dimension * self.nqubits, max_bytes, "U1 basis states")
        rows = np.asarray([self.unrank(i) for i in range(self.dimension)], dtype=np.uint8)
        rows = rows.reshape(self.dimension, self.nqubits)
        rows.flags.writeable = False
        return rows


class SectorMvpPlan(NativePlanFacade):
    """Lazy restricted plan on the GPU (U1LazyMvpPlan / PreparedChargeMvpPlan)."""

    def __init__(self, ptr: ctypes.c_void_p, native_variant: str, execution_family: str) -> None:
        self._ptr = ptr
        L = _lib.lib()
        self.dimension = int(L.tcp_sector_plan_dimension(ptr))
        self.nqubits = int(L.tcp_sector_plan_nqubits(ptr))
        self.term_count = int(L.tcp_sector_plan_term_count(ptr))
        self.group_count = int(L.tcp_sector_plan_group_count(ptr))
        self.estimated_bytes = int(L.tcp_sector_plan_estimated_bytes(ptr))
        self.native_variant = native_variant
        self.execution_family = execution_family
        self.storage = "lazy"
        self.retained_format = "compiled_actions"
        self.basis_ordering = "qubit0_msb_matrix"
        self.target = "native_mvp"
        self.schedule_count = self.term_count
        self.estimated_retained_bytes = self.estimated_bytes
        self.estimated_construction_peak_bytes = self.estimated_bytes
        self.transition_count: Optional[int] = None

    def __del__(self) -> None:
        if getattr(self, "_ptr", None) and getattr(_lib, "_lib", None) is not None:
            _lib.lib().tcp_sector_plan_destroy(self._ptr)
            self._ptr = None

    @property
    def handle(self) -> int:
        return int(self._ptr.value or 0)

    def _native_apply_host(self, inp: np.ndarray, out: np.ndarray) -> int:
        dp = ctypes.POINTER(ctypes.c_double)
        return _lib.lib().tcp_sector_plan_apply_host(self._ptr, inp.ctypes.data_as(dp), out.ctypes.data_as(dp),
                                                     inp.shape[0], out.shape[0])

    def _native_apply_device(self, inp: Any, out: Any, dtype: int, stream: int) -> int:
        return _lib.lib().tcp_sector_plan_apply_into(self._ptr, inp.data_ptr(), out.data_ptr(), inp.numel(),
                                                     out.numel(), dtype, stream)

    def rank_device(self, bitstrings: Any) -> Any:
        """Ranks of a device uint64 tensor of integer bitstrings (UINT64_MAX outside)."""
        import torch

        from ._plan import current_stream_handle

        out = torch.empty_like(bitstrings)
        _lib.check(_lib.lib().tcp_sector_plan_rank_device(self._ptr, bitstrings.data_ptr(), out.data_ptr(),
                                                          bitstrings.numel(), current_stream_handle(bitstrings)))
        return out

    def unrank_device(self, ranks: Any) -> Any:
        import torch

        from ._plan import current_stream_handle

        out = torch.empty_like(ranks)
        _lib.check(_lib.lib().tcp_sector_plan_unrank_device(self._ptr, ranks.data_ptr(), out.data_ptr(),
                                                            ranks.numel(), current_stream_handle(ranks)))
        return out

    def walk_device(self, first: int, count: int, run: int, device: Any = "cuda") -> Any:
        import torch

        from ._plan import current_stream_handle

        out = torch.empty(count, dtype=torch.int64, device=device)  # bit pattern of uint64
        _lib.check(_lib.lib().tcp_sector_plan_walk_device(self._ptr, first, out.data_ptr(), count, run,
                                                          current_stream_handle(out)))
        return out


class U1RestrictedOperator(NativePlanFacade):
    """A Pauli operator restricted to one U(1) sector (py/symmetry.py:244-538).

    ``storage="lazy"`` keeps the matrix-free GPU plan (exact conservation
    required up front); ``storage="eager"`` materialises the restricted
    operator as CSR on the device (no conservation precondition: leakage of any
    aggregated transition fails the build, as core/sector.rs:345-432 does).
    """

    def __init__(self, sector: U1Sector, operator: Any, plan: Optional[SectorMvpPlan], storage: str,
                 csr_plan: Any = None) -> None:
        self.sector = sector
        self.operator = operator
        self.dimension = sector.dimension
        self.storage = storage
        self.term_count = operator.term_count
        self._plan = plan
        self._csr = csr_plan
        self._lock = threading.Lock()
        self.native_variant = "qubit"

    @property
    def plan(self) -> Any:
        return self._csr if self.storage == "eager" else self._plan

    @property
    def transition_count(self) -> Optional[int]:
        return None if self._csr is None else self._csr.transition_count

    @property
    def estimated_bytes(self) -> int:
        lazy = self.dimension * 16 + (self.dimension + 1) * 8 + self.term_count * (self.sector.nqubits // 8 + 32)
        return lazy + (0 if self._csr is None else self._csr.transition_count * 32)

    def _ensure_lazy(self, max_bytes: Optional[int]) -> SectorMvpPlan:
        with self._lock:  # a concurrent first use builds one plan
            if self._plan is None:
                self._plan = _lazy_u1_plan(self.operator, self.sector, max_bytes)
            return self._plan

    def _ensure_eager(self, max_bytes: Optional[int]) -> Any:
        with self._lock:
            if self._csr is None:
                from .eager import csr_plan_u1

                self._csr = csr_plan_u1(self.operator, self.sector.particle_number, max_bytes)
            return self._csr

    def mvp_plan(self, *, storage: str = "lazy", max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        """Reusable restricted plan (py/symmetry.py:344-382)."""
        validate_max_bytes(max_bytes)
        if storage not in ("lazy", "eager"):
            raise ValueError("storage must be either 'eager' or 'lazy'")
        return self._ensure_eager(max_bytes) if storage == "eager" else self._ensure_lazy(max_bytes)

    def dense(self, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> np.ndarray:
        return self._ensure_eager(max_bytes).materialize("dense", max_bytes)

    def coo(self, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        return self._ensure_eager(max_bytes).materialize("coo", max_bytes)

    def csr(self, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        return self._ensure_eager(max_bytes).materialize("csr", max_bytes)

    def compile(self, target: str, *, storage: str = "lazy", max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        """py/symmetry.py:514-538."""
        if target in ("dense", "coo", "csr"):
            return getattr(self, target)(max_bytes=max_bytes)
        if target == "native_mvp":
            return self.mvp_plan(storage=storage, max_bytes=max_bytes)
        raise ValueError("target must be one of 'dense', 'coo', 'csr', or 'native_mvp'")

    def _native_apply_host(self, inp: np.ndarray, out: np.ndarray) -> int:
        return self.plan._native_apply_host(inp, out)

    def _native_apply_device(self, inp: Any, out: Any, dtype: int, stream: int) -> int:
        return self.plan._native_apply_device(inp, out, dtype, stream)


def lazy_plan_of(obj: Any) -> Any:
    """The matrix-free plan behind a plan or restricted operator (the learned
    preconditioner's features come from it); device CSR plans are rejected."""
    from .eager import CsrMvpPlan

    if hasattr(obj, "_ensure_lazy"):
        return obj._ensure_lazy(DEFAULT_MAX_BYTES)
    if isinstance(obj, CsrMvpPlan):
        raise TypeError("expected a full-space plan, a lazy restricted plan or a restricted operator, "
                        "not an eager CSR plan")
    return obj


def canonical_u1_sector(sector: Any) -> Optional[U1Sector]:
    """A single number charge over all qubits is the U(1) sector (py/symmetry.py:620-643)."""
    from .charge import ChargeSector

    if not isinstance(sector, ChargeSector) or len(sector.constraints) != 1:
        return None
    charge, requested = sector.constraints[0]
    if charge.offset != 0 or any(levels != (0, 1) for levels in charge.qubit_levels):
        return None
    if requested < 0 or requested > charge.nqubits:
        return None
    return U1Sector(charge.nqubits, int(requested))


def _lazy_u1_plan(operator: Any, sector: U1Sector, max_bytes: Optional[int]) -> SectorMvpPlan:
    _lib.require_device()
    out = ctypes.c_void_p()
    _lib.check(_lib.lib().tcp_sector_plan_create_u1(operator._handle.ptr, sector.particle_number,
                                                     effective_max_bytes(max_bytes), ctypes.byref(out)))
    return SectorMvpPlan(out, "qubit", "product")


def restrict_u1(operator: Any, sector: U1Sector, max_bytes: Optional[int], *,
                storage: str = "lazy") -> U1RestrictedOperator:
    """py/symmetry.py:658-706 on the GPU; lazy restriction checks conservation
    natively, eager restriction builds the device CSR."""
    if storage not in ("lazy", "eager"):
        raise ValueError("storage must be either 'eager' or 'lazy'")
    validate_max_bytes(max_bytes)
    if operator.nqubits != sector.nqubits:
        raise ValueError(f"incompatible qubit counts: {operator.nqubits} and {sector.nqubits}")
    if storage == "eager":
        from .eager import csr_plan_u1

        return U1RestrictedOperator(sector, operator, None, storage,
                                    csr_plan_u1(operator, sector.particle_number, max_bytes))
    return U1RestrictedOperator(sector, operator, _lazy_u1_plan(operator, sector, max_bytes), storage)
