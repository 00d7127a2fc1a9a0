"""Additive qubit charges and charge sectors (py/charge.py:162-483, 608-794, 1075-1502).

Only qubit charges are modelled (the Pauli path); a sector whose constraints
are particle numbers of contiguous, disjoint qubit blocks covering qubits
0..n-1 is the factorized layout (core/charge_sector.rs:150-217, 1786-1868) and
runs on the GPU block kernel.  Rank order is lexicographic over qubits with
qubit 0 most significant (``operator_space_axis0_msb_mixed_radix``).
"""

from __future__ import annotations

import ctypes
import math
import threading
from dataclasses import dataclass, field
from typing import Any, Dict, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._plan import DEFAULT_MAX_BYTES, NativePlanFacade, effective_max_bytes, validate_max_bytes


@dataclass(frozen=True, eq=True)
class AdditiveCharge:
    """Q = offset + sum_q level_q(bit_q) over qubits; levels default to (0, 0)."""

    nqubits: int
    qubits: Tuple[Tuple[int, Tuple[float, float]], ...] = ()
    offset: float = 0
    name: str = field(default="charge", compare=False)

    def __init__(self, nqubits: int, qubits: Optional[Dict[int, Tuple[float, float]]] = None, *,
                 offset: float = 0, name: str = "charge") -> None:
        if not isinstance(nqubits, int) or nqubits < 0:
            raise ValueError("nqubits must be a non-negative integer")
        items = []
        for index, levels in sorted((qubits or {}).items()):
            if not 0 <= int(index) < nqubits:
                raise ValueError(f"qubit index {index} is outside 0..{nqubits - 1}")
            l0, l1 = levels
            items.append((int(index), (float(l0), float(l1))))
        object.__setattr__(self, "nqubits", nqubits)
        object.__setattr__(self, "qubits", tuple(items))
        object.__setattr__(self, "offset", offset)
        object.__setattr__(self, "name", name)

    @classmethod
    def number(cls, nqubits: int, qubits: Sequence[int], *, name: str = "N") -> "AdditiveCharge":
        return cls(nqubits, {q: (0, 1) for q in qubits}, name=name)

    @property
    def qubit_levels(self) -> Tuple[Tuple[float, float], ...]:
        table = dict(self.qubits)
        return tuple(table.get(q, (0.0, 0.0)) for q in range(self.nqubits))

    def level_array(self) -> np.ndarray:
        return np.asarray(self.qubit_levels, dtype=np.float64).reshape(self.nqubits, 2)

    def sector(self, target: int) -> "ChargeSector":
        return ChargeSector(((self, target),))


class ChargeSector:
    """Computational states satisfying every (charge, target) constraint."""

    def __init__(self, constraints: Sequence[Tuple[AdditiveCharge, int]], *,
                 max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> None:
        validate_max_bytes(max_bytes)
        constraints = tuple((c, int(t)) for c, t in constraints)
        if not constraints:
            raise ValueError("a charge sector needs at least one constraint")
        n = constraints[0][0].nqubits
        if any(c.nqubits != n for c, _ in constraints):
            raise ValueError("all charges must act on the same qubits")
        self.constraints = constraints
        self.nqubits = n
        self.blocks = _factorized_blocks(constraints)

    @property
    def dimension(self) -> int:
        if self.blocks is None:
            raise NotImplementedError("only contiguous fixed-population qubit blocks are implemented")
        return math.prod(math.comb(s, k) for s, k in self.blocks)

    def _arrays(self) -> Tuple[np.ndarray, np.ndarray]:
        if self.blocks is None:
            raise NotImplementedError("only contiguous fixed-population qubit blocks are implemented")
        sizes = np.ascontiguousarray([b[0] for b in self.blocks], dtype=np.uint32)
        pops = np.ascontiguousarray([b[1] for b in self.blocks], dtype=np.uint32)
        return sizes, pops

    def rank(self, occupations: Sequence[int]) -> int:
        bits = tuple(int(b) for b in occupations)
        if len(bits) != self.nqubits or any(b not in (0, 1) for b in bits):
            raise ValueError("occupations must contain exactly nqubits binary values")
        value = 0
        for b in bits:
            value = (value << 1) | b
        sizes, pops = self._arrays()
        out = ctypes.c_uint64()
        u32 = ctypes.POINTER(ctypes.c_uint32)
        _lib.check(_lib.lib().tcp_blocks_rank(len(sizes), sizes.ctypes.data_as(u32), pops.ctypes.data_as(u32),
                                              value, ctypes.byref(out)))
        return int(out.value)

    def unrank(self, index: int) -> Tuple[int, ...]:
        if not 0 <= int(index) < self.dimension:
            raise IndexError("restricted basis index is out of range")
        sizes, pops = self._arrays()
        out = ctypes.c_uint64()
        u32 = ctypes.POINTER(ctypes.c_uint32)
        _lib.check(_lib.lib().tcp_blocks_unrank(len(sizes), sizes.ctypes.data_as(u32),
                                                pops.ctypes.data_as(u32), int(index), ctypes.byref(out)))
        value = int(out.value)
        return tuple((value >> (self.nqubits - 1 - q)) & 1 for q in range(self.nqubits))

    def basis_states(self) -> np.ndarray:
        rows = np.asarray([self.unrank(i) for i in range(self.dimension)], dtype=np.uint8)
        rows = rows.reshape(self.dimension, self.nqubits)
        rows.flags.writeable = False
        return rows


def _factorized_blocks(constraints: Sequence[Tuple[AdditiveCharge, int]]) -> Optional[Tuple[Tuple[int, int], ...]]:
    """(size, population) per block when every constraint is a 0/1 number charge
    on a contiguous block and the blocks tile qubits 0..n-1; otherwise None."""
    n = constraints[0][0].nqubits
    spans = []
    for charge, target in constraints:
        if charge.offset != 0:
            return None
        support = [q for q, levels in enumerate(charge.qubit_levels) if levels != (0.0, 0.0)]
        if any(charge.qubit_levels[q] != (0.0, 1.0) for q in support):
            return None
        if not support or support != list(range(support[0], support[-1] + 1)):
            return None
        if not 0 <= target <= len(support):
            raise ValueError("charge target is unreachable")
        spans.append((support[0], len(support), target))
    spans.sort()
    position = 0
    for start, size, _ in spans:
        if start != position:
            return None
        position += size
    if position != n:
        return None
    return tuple((size, target) for _, size, target in spans)


class ChargeRestrictedOperator(NativePlanFacade):
    """Pauli operator restricted to a factorized charge sector (py/charge.py:1075-1412).

    ``storage="eager"`` materialises the restricted operator as device CSR
    (core/charge/eager.rs); ``"lazy"`` keeps the matrix-free plan.
    """

    def __init__(self, operator: Any, sector: ChargeSector, plan: Any, storage: str, csr_plan: Any = None) -> None:
        self.operator = operator
        self.sector = sector
        self.dimension = sector.dimension
        self.storage = storage
        self.term_count = operator.term_count
        self._plan = plan
        self._csr = csr_plan
        self._lock = threading.Lock()
        self.native_variant = "factorized"
        self.execution_family = "factorized"

    @property
    def plan(self) -> Any:
        return self._csr if self.storage == "eager" else self._plan

    @property
    def transition_count(self) -> Optional[int]:
        return None if self._csr is None else self._csr.transition_count

    def _ensure_lazy(self, max_bytes: Optional[int]) -> Any:
        with self._lock:  # a concurrent first use builds one plan
            if self._plan is None:
                self._plan = _lazy_blocks_plan(self.operator, self.sector, max_bytes)
            return self._plan

    def _ensure_eager(self, max_bytes: Optional[int]) -> Any:
        with self._lock:
            if self._csr is None:
                from .eager import csr_plan_blocks

                sizes, pops = self.sector._arrays()
                self._csr = csr_plan_blocks(self.operator, sizes, pops, max_bytes)
            return self._csr

    def mvp_plan(self, *, storage: str = "lazy", max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
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

    def _native_apply_host(self, inp: np.ndarray, out: np.ndarray) -> int:
        return self.plan._native_apply_host(inp, out)

    def _native_apply_device(self, inp: Any, out: Any, dtype: int, stream: int) -> int:
        return self.plan._native_apply_device(inp, out, dtype, stream)


def _lazy_blocks_plan(operator: Any, sector: ChargeSector, max_bytes: Optional[int]) -> Any:
    from .symmetry import SectorMvpPlan

    _lib.require_device()
    sizes, pops = sector._arrays()
    out = ctypes.c_void_p()
    u32 = ctypes.POINTER(ctypes.c_uint32)
    _lib.check(_lib.lib().tcp_sector_plan_create_blocks(operator._handle.ptr, len(sizes), sizes.ctypes.data_as(u32),
                                                         pops.ctypes.data_as(u32), effective_max_bytes(max_bytes),
                                                         ctypes.byref(out)))
    return SectorMvpPlan(out, "factorized", "factorized")


def restrict_charge(operator: Any, sector: ChargeSector, *, storage: str = "lazy",
                    max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> ChargeRestrictedOperator:
    if storage not in ("lazy", "eager"):
        raise ValueError("storage must be either 'eager' or 'lazy'")
    validate_max_bytes(max_bytes)
    if operator.nqubits != sector.nqubits:
        raise ValueError("Pauli operator and charge sector layouts are incompatible")
    if sector.blocks is None:
        raise NotImplementedError("only contiguous fixed-population qubit blocks are implemented")
    if storage == "eager":
        from .eager import csr_plan_blocks

        sizes, pops = sector._arrays()
        return ChargeRestrictedOperator(operator, sector, None, storage,
                                        csr_plan_blocks(operator, sizes, pops, max_bytes))
    return ChargeRestrictedOperator(operator, sector, _lazy_blocks_plan(operator, sector, max_bytes), storage)
