"""FermionOperator and OperatorSpace: the reference's fermion facade on the GPU path.

Mirrors py/_structured/fermion.py (``FermionOperator`` 51-276),
py/_structured/words.py (``FermionWord`` 18-68, ``FermionTerm`` 280-284) and
py/_structured/space.py (``OperatorSpace`` 45-120) for fermion-only and
qubit-only spaces.  Construction is one native CAR canonicalisation
(``tcp_fermion_from_ladder``: creations with increasing modes before
annihilations with decreasing modes, bit-pattern-ordered sums, exact zeros
dropped).  ``restrict_charge`` gathers from the ladder words themselves on
the GPU (``tcp_sector_plan_create_fermion``: a word acts where its annihilated
modes are occupied and its other created modes empty, with its Jordan-Wigner
sign; mode m is bit n-1-m of the occupation index), the reference's
raw-fermion restricted path in its occupation-basis order
(py/_structured/fermion.py:230-244).  Full-space targets (``compile``,
``mvp``) use the Jordan-Wigner image (``map_fermions``).
"""

from __future__ import annotations

import ctypes
import math
import numbers
from dataclasses import dataclass
from typing import Any, Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._plan import DEFAULT_MAX_BYTES, effective_max_bytes, validate_max_bytes

RawTerm = Tuple[Sequence[Tuple[int, str]], complex]


def _nonnegative_int(value: object, name: str) -> int:
    if not isinstance(value, int) or isinstance(value, bool) or value < 0:
        raise ValueError(f"{name} must be a non-negative integer")
    return int(value)


class OperatorSpace:
    """Ordered axes: fermions, bosons, qubits, then qudits (py/_structured/space.py:45-120).

    Only spaces whose axes are all fermion modes or all qubits reach the GPU
    path; bosons and qudits are accepted here and rejected where a sector or
    an operator would need them.
    """

    __slots__ = ("fermions", "bosons", "qubits", "qudits", "_locked")

    def __init__(self, fermions: int = 0, bosons: int = 0, qubits: int = 0, qudits: Sequence[int] = ()) -> None:
        fermions = _nonnegative_int(fermions, "fermions")
        bosons = _nonnegative_int(bosons, "bosons")
        qubits = _nonnegative_int(qubits, "qubits")
        if not isinstance(qudits, (tuple, list)):
            raise ValueError("qudits must be a tuple of local dimensions")
        dims = tuple(_nonnegative_int(v, "qudit dimension") for v in qudits)
        if dims and (any(v < 3 or v > 2 ** 32 - 1 for v in dims) or len(set(dims)) != 1):
            raise ValueError("qudits must have one uniform dimension 3 <= d <= 2**32-1")
        object.__setattr__(self, "fermions", fermions)
        object.__setattr__(self, "bosons", bosons)
        object.__setattr__(self, "qubits", qubits)
        object.__setattr__(self, "qudits", dims)
        object.__setattr__(self, "_locked", True)

    def __setattr__(self, name: str, value: object) -> None:
        raise AttributeError("OperatorSpace is immutable")

    def _key(self) -> Tuple[Any, ...]:
        return (self.fermions, self.bosons, self.qubits, self.qudits)

    def __eq__(self, other: object) -> bool:
        return isinstance(other, OperatorSpace) and self._key() == other._key()

    def __hash__(self) -> int:
        return hash(self._key())

    @property
    def binary_axes(self) -> int:
        """Fermion modes plus qubits: the axes of the occupation/computational basis."""
        return self.fermions + self.qubits

    @property
    def is_binary(self) -> bool:
        return self.bosons == 0 and not self.qudits

    def __repr__(self) -> str:
        return (f"OperatorSpace(fermions={self.fermions}, bosons={self.bosons}, qubits={self.qubits}, "
                f"qudits={self.qudits})")


@dataclass(frozen=True)
class FermionWord:
    """Canonical monomial: creations (increasing modes) then annihilations (decreasing)."""

    n_modes: int
    creation_modes: Tuple[int, ...] = ()
    annihilation_modes: Tuple[int, ...] = ()

    def __post_init__(self) -> None:
        n = _nonnegative_int(self.n_modes, "n_modes")
        cre, ann = tuple(self.creation_modes), tuple(self.annihilation_modes)
        if any(not isinstance(m, int) or isinstance(m, bool) or not 0 <= m < n for m in (*cre, *ann)):
            raise ValueError("fermion mode is outside the operator space")
        if cre != tuple(sorted(cre)) or len(set(cre)) != len(cre):
            raise ValueError("creation_modes must be strictly increasing")
        if ann != tuple(sorted(ann, reverse=True)) or len(set(ann)) != len(ann):
            raise ValueError("annihilation_modes must be strictly decreasing")
        object.__setattr__(self, "creation_modes", cre)
        object.__setattr__(self, "annihilation_modes", ann)

    @property
    def is_identity(self) -> bool:
        return not self.creation_modes and not self.annihilation_modes

    @property
    def factors(self) -> Tuple[Tuple[int, str], ...]:
        return tuple((m, "create") for m in self.creation_modes) + tuple(
            (m, "annihilate") for m in self.annihilation_modes)


@dataclass(frozen=True)
class FermionTerm:
    word: FermionWord
    coefficient: complex


class _FermionHandle:
    __slots__ = ("ptr",)

    def __init__(self, ptr: ctypes.c_void_p) -> None:
        self.ptr = ptr

    def __del__(self) -> None:
        if self.ptr and getattr(_lib, "_lib", None) is not None:
            _lib.lib().tcp_fermion_destroy(self.ptr)
            self.ptr = None


def _finite_complex(value: object) -> complex:
    if isinstance(value, (bool, np.ndarray)) or not isinstance(value, (numbers.Real, numbers.Complex)):
        raise ValueError("coefficient must be a finite real or complex scalar")
    result = complex(value)
    if not math.isfinite(result.real) or not math.isfinite(result.imag):
        raise ValueError("coefficient must be finite")
    return result


def _ladder_arrays(n_modes: int, terms: Iterable[RawTerm]):
    offsets: List[int] = [0]
    modes: List[int] = []
    actions: List[int] = []
    coef: List[float] = []
    for raw_factors, coefficient in terms:
        for factor in raw_factors:
            if not isinstance(factor, (tuple, list)) or len(factor) != 2:
                raise ValueError("fermion and boson factors must be (mode, action) pairs")
            mode = _nonnegative_int(factor[0], "mode")
            if mode >= n_modes:
                raise ValueError(f"mode {mode} is outside 0..{n_modes}")
            if factor[1] not in ("create", "annihilate"):
                raise ValueError("action must be exactly 'create' or 'annihilate'")
            modes.append(mode)
            actions.append(0 if factor[1] == "create" else 1)
        offsets.append(len(modes))
        c = _finite_complex(coefficient)
        coef.extend((c.real, c.imag))
    return (np.asarray(offsets, dtype=np.uint64), np.asarray(modes, dtype=np.uint32),
            np.asarray(actions, dtype=np.uint8), np.asarray(coef, dtype=np.float64))


class FermionOperator:
    """Immutable canonical fermionic operator governed by CAR (py/_structured/fermion.py:51-58)."""

    def __init__(self, n_modes: int, terms: Iterable[RawTerm], *,
                 max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> None:
        validate_max_bytes(max_bytes)
        n_modes = _nonnegative_int(n_modes, "n_modes")
        offsets, modes, actions, coef = _ladder_arrays(n_modes, terms)
        out = ctypes.c_void_p()
        _lib.check(_lib.lib().tcp_fermion_from_ladder(
            n_modes, len(offsets) - 1, offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
            modes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
            actions.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
            coef.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), effective_max_bytes(max_bytes), ctypes.byref(out)))
        self._handle = _FermionHandle(out)
        self.n_modes = n_modes
        self.space = OperatorSpace(fermions=n_modes)
        self._terms: Optional[Tuple[FermionTerm, ...]] = None
        self._pauli: Optional[Any] = None

    @classmethod
    def from_terms(cls, n_modes: int, terms: Iterable[RawTerm], *,
                   max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> "FermionOperator":
        """CAR-canonicalise raw ordered ladder products ``((mode, "create"|"annihilate"), ...)``."""
        return cls(n_modes, terms, max_bytes=max_bytes)

    # -- canonical content -------------------------------------------------
    @property
    def term_count(self) -> int:
        return int(_lib.lib().tcp_fermion_term_count(self._handle.ptr))

    def __len__(self) -> int:
        return self.term_count

    def _raw(self) -> List[RawTerm]:
        L = _lib.lib()
        T = self.term_count
        F = int(L.tcp_fermion_factor_count(self._handle.ptr))
        offsets = np.zeros(T + 1, dtype=np.uint64)
        modes = np.zeros(max(F, 1), dtype=np.uint32)
        actions = np.zeros(max(F, 1), dtype=np.uint8)
        coef = np.zeros(2 * max(T, 1), dtype=np.float64)
        _lib.check(L.tcp_fermion_export(self._handle.ptr, offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                        modes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                        actions.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                        coef.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        out = []
        for t in range(T):
            factors = tuple((int(modes[k]), "create" if actions[k] == 0 else "annihilate")
                            for k in range(int(offsets[t]), int(offsets[t + 1])))
            out.append((factors, complex(coef[2 * t], coef[2 * t + 1])))
        return out

    @property
    def terms(self) -> Tuple[FermionTerm, ...]:
        if self._terms is None:
            terms = []
            for factors, c in self._raw():
                cre = tuple(m for m, a in factors if a == "create")
                ann = tuple(m for m, a in factors if a == "annihilate")
                terms.append(FermionTerm(FermionWord(self.n_modes, cre, ann), c))
            self._terms = tuple(terms)
        return self._terms

    def to_dict(self) -> Dict[Tuple[Tuple[int, str], ...], complex]:
        return {factors: c for factors, c in self._raw()}

    def __eq__(self, other: object) -> bool:
        if not isinstance(other, FermionOperator):
            return NotImplemented
        return self.n_modes == other.n_modes and self._raw() == other._raw()

    __hash__ = None  # type: ignore[assignment]

    def __repr__(self) -> str:
        return f"FermionOperator(n_modes={self.n_modes}, term_count={self.term_count})"

    # -- algebra (products and sums re-canonicalised natively) ---------------
    def _check_space(self, other: "FermionOperator") -> None:
        if other.n_modes != self.n_modes:
            raise ValueError("fermion operators have incompatible operator spaces")

    def add(self, other: "FermionOperator", *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> "FermionOperator":
        self._check_space(other)
        return FermionOperator(self.n_modes, self._raw() + other._raw(), max_bytes=max_bytes)

    def multiply(self, other: "FermionOperator", *,
                 max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> "FermionOperator":
        """Product with CAR contractions (core/structured/fermion.rs:1-46)."""
        self._check_space(other)
        pairs = [(lf + rf, lc * rc) for lf, lc in self._raw() for rf, rc in other._raw()]
        return FermionOperator(self.n_modes, pairs, max_bytes=max_bytes)

    def scale(self, factor: complex) -> "FermionOperator":
        f = _finite_complex(factor)
        return FermionOperator(self.n_modes, [(w, c * f) for w, c in self._raw()])

    def __add__(self, other: object) -> "FermionOperator":
        return self.add(other) if isinstance(other, FermionOperator) else NotImplemented

    def __sub__(self, other: object) -> "FermionOperator":
        return self.add(other.scale(-1.0)) if isinstance(other, FermionOperator) else NotImplemented

    def __neg__(self) -> "FermionOperator":
        return self.scale(-1.0)

    def __mul__(self, other: object) -> "FermionOperator":
        if isinstance(other, FermionOperator):
            return self.multiply(other)
        if isinstance(other, numbers.Number) and not isinstance(other, bool):
            return self.scale(other)  # type: ignore[arg-type]
        return NotImplemented

    def __rmul__(self, other: object) -> "FermionOperator":
        if isinstance(other, numbers.Number) and not isinstance(other, bool):
            return self.scale(other)  # type: ignore[arg-type]
        return NotImplemented

    def adjoint(self) -> "FermionOperator":
        """Hermitian conjugate: reversed factors with swapped actions, conjugated coefficients."""
        flip = {"create": "annihilate", "annihilate": "create"}
        return FermionOperator(self.n_modes, [(tuple((m, flip[a]) for m, a in reversed(w)), c.conjugate())
                                              for w, c in self._raw()])

    # -- Pauli image and GPU execution --------------------------------------
    def map_fermions(self, mapping: str = "jordan_wigner", *,
                     max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        """Jordan-Wigner image as a canonical PauliOperator (mode m = qubit m)."""
        from .pauli import PauliOperator, _OperatorHandle

        validate_max_bytes(max_bytes)
        if mapping != "jordan_wigner":
            raise NotImplementedError("only the Jordan-Wigner mapping is provided")
        if self._pauli is None:
            out = ctypes.c_void_p()
            _lib.check(_lib.lib().tcp_fermion_jordan_wigner(self._handle.ptr, effective_max_bytes(max_bytes),
                                                            ctypes.byref(out)))
            op = object.__new__(PauliOperator)
            op._handle = _OperatorHandle(out)
            op.nqubits = self.n_modes
            op._terms = None
            op._arrays = None
            self._pauli = op
        return self._pauli

    def analyze_charge(self, charge: Any, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        return self.map_fermions().analyze_charge(charge, max_bytes=max_bytes)

    def conserves(self, charge: Any, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> bool:
        return self.analyze_charge(charge, max_bytes=max_bytes).is_conserved

    def restrict_charge(self, sector: Any, *, storage: str = "lazy",
                        max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        """Restricted operator on a factorized fermion charge sector, gathered on the GPU
        from the ladder words (lazy) or from their device CSR (eager)."""
        from .charge import ChargeSector, restrict_fermion

        if not isinstance(sector, ChargeSector):
            raise TypeError("sector must be a ChargeSector")
        if sector.space is not None and sector.space != self.space or sector.nqubits != self.n_modes:
            raise ValueError("fermion operator and charge sector layouts are incompatible")
        return restrict_fermion(self, sector, storage=storage, max_bytes=max_bytes)

    def mvp(self, state: Any, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        return self.map_fermions(max_bytes=max_bytes).mvp(state, max_bytes=max_bytes)

    def compile(self, target: str, *, storage: str = "lazy", mapping: str = "jordan_wigner",
                max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        """"dense", "coo", "csr", "native_mvp" or "backend_mvp" of the Jordan-Wigner image."""
        return self.map_fermions(mapping, max_bytes=max_bytes).compile(target, storage=storage, max_bytes=max_bytes)


__all__ = ["FermionOperator", "FermionTerm", "FermionWord", "OperatorSpace"]
