"""Additive charges and charge sectors (py/charge.py:162-483, 608-794, 1075-1502).

Charges act on binary axes: qubits, or fermion modes of an ``OperatorSpace``
(occupation n_m in {0, 1}).  A sector whose constraints fix the particle
number of every block of a partition of the axes into at most four
contiguous blocks -- e.g. total number plus 2 S_z fixing N_up and N_down --
is the factorized layout (core/charge_sector.rs:150-217, 1786-1868) and runs
on the GPU block kernel.  Every other sector of integer charges on at most
63 axes (interleaved spin orbitals, weights other than 0/1, free spectator
axes, unreachable targets) uses the suffix-count rank tables of
ChargeSectorPlan (core/charge_sector.rs:1211-1580) and the generic gather
kernel.  Rank order is lexicographic over axes with axis 0 most significant
(``operator_space_axis0_msb_mixed_radix``).
"""

from __future__ import annotations

import ctypes
import math
import threading
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._plan import (DEFAULT_MAX_BYTES, NativePlanFacade, check_allocation, effective_max_bytes,
                    validate_max_bytes)


@dataclass(frozen=True, eq=True)
class AdditiveCharge:
    """Q = offset + sum_a level_a(bit_a) over the binary axes (py/charge.py:162-260).

    Constructed as the reference's ``AdditiveCharge(space, name=...,
    fermions={mode: weight}, qubits={index: (level_0, level_1)}, offset=...)``
    on an ``OperatorSpace`` (a fermion weight w is the level pair (0, w)), or as
    ``AdditiveCharge(nqubits, {index: (level_0, level_1)})`` on qubits alone.
    Boson and qudit axes are not on the GPU path.
    """

    nqubits: int
    qubits: Tuple[Tuple[int, Tuple[float, float]], ...] = ()
    offset: float = 0
    name: str = field(default="charge", compare=False)
    space: Any = field(default=None, compare=False)

    def __init__(self, space: Any, qubits: Optional[Dict[int, Tuple[float, float]]] = None, *,
                 name: str = "", fermions: Optional[Dict[int, float]] = None, bosons: Any = None,
                 qudits: Any = None, offset: float = 0) -> None:
        from .fermion import OperatorSpace

        if not isinstance(name, str):
            raise TypeError("name must be a string")
        if isinstance(space, OperatorSpace):
            if bosons or qudits or not space.is_binary:
                raise NotImplementedError("boson and qudit charges are not on the GPU path")
            n_f, n = space.fermions, space.binary_axes
            levels: Dict[int, Tuple[float, float]] = {}
            for mode, weight in (fermions or {}).items():
                if not isinstance(mode, int) or isinstance(mode, bool) or not 0 <= mode < n_f:
                    raise ValueError(f"fermions index {mode} is outside 0..{n_f}")
                levels[mode] = (0.0, float(weight))
            for index, pair in (qubits or {}).items():
                if not isinstance(index, int) or isinstance(index, bool) or not 0 <= index < space.qubits:
                    raise ValueError(f"qubits index {index} is outside 0..{space.qubits}")
                levels[n_f + index] = pair
            nqubits, table = n, levels
        else:
            if fermions or bosons or qudits:
                raise TypeError("space must be an OperatorSpace")
            if not isinstance(space, int) or isinstance(space, bool) or space < 0:
                raise ValueError("nqubits must be a non-negative integer")
            nqubits, table, space = space, dict(qubits or {}), None
        items = []
        for index, pair in sorted(table.items()):
            if not 0 <= int(index) < nqubits:
                raise ValueError(f"qubit index {index} is outside 0..{nqubits - 1}")
            l0, l1 = pair
            items.append((int(index), (float(l0), float(l1))))
        object.__setattr__(self, "nqubits", nqubits)
        object.__setattr__(self, "qubits", tuple(items))
        object.__setattr__(self, "offset", offset)
        object.__setattr__(self, "name", name or "charge")
        object.__setattr__(self, "space", space)

    @classmethod
    def number(cls, nqubits: int, qubits: Sequence[int], *, name: str = "N") -> "AdditiveCharge":
        return cls(nqubits, {q: (0, 1) for q in qubits}, name=name)

    @property
    def qubit_levels(self) -> Tuple[Tuple[float, float], ...]:
        table = dict(self.qubits)
        return tuple(table.get(q, (0.0, 0.0)) for q in range(self.nqubits))

    def level_array(self) -> np.ndarray:
        return np.asarray(self.qubit_levels, dtype=np.float64).reshape(self.nqubits, 2)

    def sector(self, target: int, *, boson_cutoffs: Any = None,
               max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> "ChargeSector":
        return ChargeSector(((self, target),), boson_cutoffs=boson_cutoffs, max_bytes=max_bytes)


class ChargeSector:
    """Computational states satisfying every (charge, target) constraint."""

    def __init__(self, constraints: Sequence[Tuple[AdditiveCharge, int]], *,
                 boson_cutoffs: Any = None, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> None:
        validate_max_bytes(max_bytes)
        if boson_cutoffs:
            raise NotImplementedError("boson cutoffs: boson axes are not on the GPU path")
        constraints = tuple((c, int(t)) for c, t in constraints)
        if not constraints:
            raise ValueError("a charge sector needs at least one constraint")
        n = constraints[0][0].nqubits
        if any(c.nqubits != n for c, _ in constraints):
            raise ValueError("all charges must act on the same qubits")
        spaces = {c.space for c, _ in constraints if c.space is not None}
        if len(spaces) > 1:
            raise ValueError("all charges must share one operator space")
        self.constraints = constraints
        self.nqubits = n
        self.space = spaces.pop() if spaces else None
        self.blocks = _factorized_blocks(constraints)
        self._general: Optional[_GeneralSector] = None
        if self.blocks is None:
            self._general = _GeneralSector(constraints, max_bytes)
        self.execution_family = "factorized" if self.blocks is not None else "generic"

    @property
    def dimension(self) -> int:
        if self._general is not None:
            return self._general.dimension
        return math.prod(math.comb(s, k) for s, k in self.blocks)

    def _arrays(self) -> Tuple[np.ndarray, np.ndarray]:
        if self.blocks is None:
            raise ValueError("the sector is not a product of contiguous fixed-population blocks")
        sizes = np.ascontiguousarray([b[0] for b in self.blocks], dtype=np.uint32)
        pops = np.ascontiguousarray([b[1] for b in self.blocks], dtype=np.uint32)
        return sizes, pops

    def rank(self, occupations: Sequence[int]) -> int:
        """Rank of a selected occupation (ValueError when a constraint fails)."""
        bits = tuple(int(b) for b in occupations)
        if len(bits) != self.nqubits:
            raise ValueError("occupations must match OperatorSpace.axes length")
        if any(b not in (0, 1) for b in bits):
            raise ValueError("occupation is outside the finite sector layout")
        value = 0
        for b in bits:
            value = (value << 1) | b
        out = ctypes.c_uint64()
        if self._general is not None:
            _lib.check(_lib.lib().tcp_charge_sector_rank(self._general.ptr, value, ctypes.byref(out)))
            return int(out.value)
        sizes, pops = self._arrays()
        u32 = ctypes.POINTER(ctypes.c_uint32)
        _lib.check(_lib.lib().tcp_blocks_rank(len(sizes), sizes.ctypes.data_as(u32), pops.ctypes.data_as(u32),
                                              value, ctypes.byref(out)))
        return int(out.value)

    def unrank(self, index: int) -> Tuple[int, ...]:
        if not 0 <= int(index) < self.dimension:
            raise IndexError("sector index is out of range")
        out = ctypes.c_uint64()
        if self._general is not None:
            _lib.check(_lib.lib().tcp_charge_sector_unrank(self._general.ptr, int(index), ctypes.byref(out)))
        else:
            sizes, pops = self._arrays()
            u32 = ctypes.POINTER(ctypes.c_uint32)
            _lib.check(_lib.lib().tcp_blocks_unrank(len(sizes), sizes.ctypes.data_as(u32),
                                                    pops.ctypes.data_as(u32), int(index), ctypes.byref(out)))
        value = int(out.value)
        return tuple((value >> (self.nqubits - 1 - q)) & 1 for q in range(self.nqubits))

    def basis_states(self, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> np.ndarray:
        """Every selected occupation as read-only uint64 rows in rank order (py/charge.py:770-791)."""
        validate_max_bytes(max_bytes)
        dim, n = self.dimension, self.nqubits
        check_allocation(dim * max(n, 1) * 8, max_bytes, "charge-sector basis states")
        if self._general is not None:
            words = np.empty(dim, dtype=np.uint64)
            _lib.check(_lib.lib().tcp_charge_sector_basis_words(
                self._general.ptr, 0, dim, words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
            shifts = np.arange(n - 1, -1, -1, dtype=np.uint64)
            rows = (words[:, None] >> shifts[None, :]) & np.uint64(1)
        else:
            rows = np.asarray([self.unrank(i) for i in range(dim)], dtype=np.uint64)
        rows = np.ascontiguousarray(rows, dtype=np.uint64).reshape(dim, n)
        rows.flags.writeable = False
        return rows


class _GeneralSector:
    """Owner of a native general-sector rank authority (``tcp_charge_sector``)."""

    def __init__(self, constraints: Sequence[Tuple[AdditiveCharge, int]], max_bytes: Optional[int]) -> None:
        self.ptr = None
        n, m = constraints[0][0].nqubits, len(constraints)
        levels = np.zeros((m, n, 2), dtype=np.int64)
        targets = np.zeros(m, dtype=np.int64)
        for j, (charge, target) in enumerate(constraints):
            pairs = charge.qubit_levels
            values = [v for pair in pairs for v in pair] + [charge.offset]
            if any(float(v) != int(v) for v in values):
                raise ValueError("charges of sectors that are not contiguous fixed-population blocks need "
                                 "integer levels")
            levels[j] = np.asarray(pairs, dtype=np.float64).reshape(n, 2).astype(np.int64)
            targets[j] = int(target) - int(charge.offset)
        ptr = ctypes.c_void_p()
        i64 = ctypes.POINTER(ctypes.c_int64)
        _lib.check(_lib.lib().tcp_charge_sector_create(n, m, levels.ctypes.data_as(i64),
                                                       targets.ctypes.data_as(i64), effective_max_bytes(max_bytes),
                                                       ctypes.byref(ptr)))
        self.ptr = ptr
        self.dimension = int(_lib.lib().tcp_charge_sector_dimension(ptr))

    def __del__(self) -> None:
        if getattr(self, "ptr", None) and getattr(_lib, "_lib", None) is not None:
            _lib.lib().tcp_charge_sector_destroy(self.ptr)
            self.ptr = None


def _factorized_blocks(constraints: Sequence[Tuple[AdditiveCharge, int]]) -> Optional[Tuple[Tuple[int, int], ...]]:
    """(size, population) of each contiguous fixed-population block, or None.

    Every charge is const_c + sum_a w_{c,a} b_a with w = level_1 - level_0.
    Maximal runs of axes with equal weight columns form the blocks; the
    constraints then fix the block populations n_B through
    sum_B w_{c,B} n_B = target_c - const_c.  The sector is the factorized
    layout (core/charge_sector.rs:150-217, 1786-1868) when exactly one
    population vector within the block sizes solves them; a block that no
    charge weighs, more than four blocks, no solution (an empty sector) or
    several solutions leave it to the general tables (None).
    """
    n = constraints[0][0].nqubits
    weights = [[l1 - l0 for l0, l1 in c.qubit_levels] for c, _ in constraints]
    rhs = [t - (c.offset + sum(l0 for l0, _ in c.qubit_levels)) for c, t in constraints]
    runs: List[List[Any]] = []  # [start, size, column]
    for a in range(n):
        column = tuple(w[a] for w in weights)
        if runs and runs[-1][2] == column:
            runs[-1][1] += 1
        else:
            runs.append([a, 1, column])
    if any(all(x == 0 for x in run[2]) for run in runs):
        return None
    if len(runs) > 4:
        return None
    solutions = _block_populations([r[2] for r in runs], [r[1] for r in runs], rhs)
    if len(solutions) != 1:
        return None
    return tuple((run[1], pop) for run, pop in zip(runs, solutions[0]))


def _block_populations(columns: Sequence[Tuple[float, ...]], sizes: Sequence[int],
                       rhs: Sequence[float]) -> List[Tuple[int, ...]]:
    """Population vectors 0 <= n_B <= size_B with sum_B columns[B] n_B == rhs (at most two)."""
    found: List[Tuple[int, ...]] = []

    def search(b: int, partial: List[float], chosen: List[int]) -> None:
        if len(found) > 1:
            return
        if b == len(columns):
            if all(abs(p - r) <= 1e-9 * max(1.0, abs(r)) for p, r in zip(partial, rhs)):
                found.append(tuple(chosen))
            return
        if b == len(columns) - 1:  # the last population follows from any charge that weighs it
            c = next(i for i, x in enumerate(columns[b]) if x != 0)
            pop_f = (rhs[c] - partial[c]) / columns[b][c]
            pop = int(round(pop_f))
            if abs(pop - pop_f) <= 1e-9 * max(1.0, abs(pop_f)) and 0 <= pop <= sizes[b]:
                search(b + 1, [p + x * pop for p, x in zip(partial, columns[b])], chosen + [pop])
            return
        for pop in range(sizes[b] + 1):
            search(b + 1, [p + c * pop for p, c in zip(partial, columns[b])], chosen + [pop])

    search(0, [0.0] * len(rhs), [])
    return found


class ChargeRestrictedOperator(NativePlanFacade):
    """Pauli operator restricted to a charge sector (py/charge.py:1075-1412).

    ``storage="eager"`` materialises the restricted operator as device CSR
    (core/charge/eager.rs); ``"lazy"`` keeps the matrix-free plan.
    Factorized sectors run the block kernel, the others the generic one.
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
        self.native_variant = sector.execution_family
        self.execution_family = sector.execution_family

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
                self._csr = _eager_pauli_plan(self.operator, self.sector, max_bytes)
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


class FermionRestrictedOperator(ChargeRestrictedOperator):
    """FermionOperator restricted to a factorized fermion charge sector.

    The plans gather from the canonical ladder words directly
    (``tcp_sector_plan_create_fermion`` / ``tcp_csr_plan_create_fermion``): a
    word acts where its annihilated modes are occupied and its other created
    modes empty, with its Jordan-Wigner sign -- the reference's raw-fermion
    restricted path (core/fermion_mask.rs, core/charge/generic.rs:642-809).
    """

    def _ensure_lazy(self, max_bytes: Optional[int]) -> Any:
        with self._lock:
            if self._plan is None:
                self._plan = _fermion_plan(self.operator, self.sector, max_bytes, eager=False)
            return self._plan

    def _ensure_eager(self, max_bytes: Optional[int]) -> Any:
        with self._lock:
            if self._csr is None:
                self._csr = _fermion_plan(self.operator, self.sector, max_bytes, eager=True)
            return self._csr


def _fermion_plan(operator: Any, sector: ChargeSector, max_bytes: Optional[int], *, eager: bool) -> Any:
    from .eager import CsrMvpPlan
    from .symmetry import SectorMvpPlan

    _lib.require_device()
    out = ctypes.c_void_p()
    L = _lib.lib()
    if sector._general is not None:
        create = L.tcp_csr_plan_create_fermion_charge if eager else L.tcp_sector_plan_create_fermion_charge
        _lib.check(create(operator._handle.ptr, sector._general.ptr, effective_max_bytes(max_bytes),
                          ctypes.byref(out)))
    else:
        sizes, pops = sector._arrays()
        u32 = ctypes.POINTER(ctypes.c_uint32)
        create = L.tcp_csr_plan_create_fermion if eager else L.tcp_sector_plan_create_fermion
        _lib.check(create(operator._handle.ptr, len(sizes), sizes.ctypes.data_as(u32), pops.ctypes.data_as(u32),
                          effective_max_bytes(max_bytes), ctypes.byref(out)))
    if eager:
        return CsrMvpPlan(out, sector.execution_family, "fermion", operator.term_count)
    return SectorMvpPlan(out, sector.execution_family, "fermion")


def restrict_fermion(operator: Any, sector: ChargeSector, *, storage: str = "lazy",
                     max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> FermionRestrictedOperator:
    if storage not in ("lazy", "eager"):
        raise ValueError("storage must be either 'eager' or 'lazy'")
    validate_max_bytes(max_bytes)
    plan = _fermion_plan(operator, sector, max_bytes, eager=storage == "eager")
    restricted = FermionRestrictedOperator(operator, sector, None if storage == "eager" else plan, storage,
                                           plan if storage == "eager" else None)
    restricted.execution_family = "fermion"
    return restricted


def _lazy_blocks_plan(operator: Any, sector: ChargeSector, max_bytes: Optional[int]) -> Any:
    from .symmetry import SectorMvpPlan

    _lib.require_device()
    out = ctypes.c_void_p()
    if sector._general is not None:
        _lib.check(_lib.lib().tcp_sector_plan_create_charge(operator._handle.ptr, sector._general.ptr,
                                                             effective_max_bytes(max_bytes), ctypes.byref(out)))
        return SectorMvpPlan(out, "generic", "generic")
    sizes, pops = sector._arrays()
    u32 = ctypes.POINTER(ctypes.c_uint32)
    _lib.check(_lib.lib().tcp_sector_plan_create_blocks(operator._handle.ptr, len(sizes), sizes.ctypes.data_as(u32),
                                                         pops.ctypes.data_as(u32), effective_max_bytes(max_bytes),
                                                         ctypes.byref(out)))
    return SectorMvpPlan(out, "factorized", "factorized")


def _eager_pauli_plan(operator: Any, sector: ChargeSector, max_bytes: Optional[int]) -> Any:
    from .eager import CsrMvpPlan, csr_plan_blocks

    if sector._general is None:
        sizes, pops = sector._arrays()
        return csr_plan_blocks(operator, sizes, pops, max_bytes)
    _lib.require_device()
    out = ctypes.c_void_p()
    _lib.check(_lib.lib().tcp_csr_plan_create_charge(operator._handle.ptr, sector._general.ptr,
                                                      effective_max_bytes(max_bytes), ctypes.byref(out)))
    return CsrMvpPlan(out, "generic", "generic", operator.term_count)


def restrict_charge(operator: Any, sector: ChargeSector, *, storage: str = "lazy",
                    max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> ChargeRestrictedOperator:
    if storage not in ("lazy", "eager"):
        raise ValueError("storage must be either 'eager' or 'lazy'")
    validate_max_bytes(max_bytes)
    if operator.nqubits != sector.nqubits:
        raise ValueError("Pauli operator and charge sector layouts are incompatible")
    if storage == "eager":
        return ChargeRestrictedOperator(operator, sector, None, storage, _eager_pauli_plan(operator, sector, max_bytes))
    return ChargeRestrictedOperator(operator, sector, _lazy_blocks_plan(operator, sector, max_bytes), storage)
