"""PauliOperator and NativeMVPPlan: the reference's operator/plan API on the GPU.

Mirrors py/pauli.py (``PauliOperator.__init__`` 324-354, ``from_code_arrays``
397-421, ``restrict_charge`` 705-736, ``mvp`` 931-952, ``backend_mvp_plan``
954-983, ``native_mvp_plan`` 985-1036, ``to_scipy_linear_operator``
1038-1045) and py/hamiltonian.py ``NativeMVPPlan`` (162-344).  Construction
canonicalises in the native library (PauliOperator::from_terms,
core/operator.rs:117-155); every apply is one call into the gfx950 kernels.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any, Dict, Iterable, Optional, Sequence, Tuple, Union

import numpy as np

from . import _lib
from .backend import BackendMVPPlan
from ._plan import (
    DEFAULT_MAX_BYTES,
    NativePlanFacade,
    check_allocation,
    effective_max_bytes,
    validate_max_bytes,
)

# tile geometries of the full-space kernel (threads x amplitudes per thread)
TILINGS = {"512x8": _lib.TCP_TILE_512x8, "1024x8": _lib.TCP_TILE_1024x8, "512x16": _lib.TCP_TILE_512x16,
           "512x4": _lib.TCP_TILE_512x4, "1024x4": _lib.TCP_TILE_1024x4,
           "1024x8x2": _lib.TCP_TILE_1024x8x2, "512x16x2": _lib.TCP_TILE_512x16x2}

_LETTERS = {"I": 0, "X": 1, "Y": 2, "Z": 3}
PauliInput = Union[str, Sequence[int]]


def _codes_of(nqubits: int, word: PauliInput) -> Tuple[int, ...]:
    if isinstance(word, str):
        try:
            codes = tuple(_LETTERS[ch] for ch in word.upper())
        except KeyError as exc:
            raise ValueError(f"invalid Pauli letter {exc.args[0]!r}; expected I/X/Y/Z") from None
    else:
        codes = tuple(int(c) for c in word)
    if len(codes) != nqubits:
        raise ValueError(f"expected structure length {nqubits}, got {len(codes)}")
    return codes


class _OperatorHandle:
    """Owns one native canonical operator."""

    def __init__(self, ptr: ctypes.c_void_p) -> None:
        self.ptr = ptr

    def __del__(self) -> None:
        if getattr(self, "ptr", None) and getattr(_lib, "_lib", None) is not None:
            _lib.lib().tcp_operator_destroy(self.ptr)
            self.ptr = None


@dataclass(frozen=True)
class PauliTerm:
    """One canonical term: codes from qubit 0 and its complex coefficient."""

    codes: Tuple[int, ...]
    coefficient: complex

    def to_string(self) -> str:
        return "".join("IXYZ"[c] for c in self.codes)


class PauliOperator:
    """Deterministic sparse Pauli operator with exact-zero aggregation.

    Terms are canonicalised on construction: duplicate words are summed in
    IEEE-bit-pattern order, exact zeros removed and the survivors sorted by
    code tuple from qubit 0 (I < X < Y < Z).  Qubit 0 is the most significant
    bit of every computational index.
    """

    def __init__(self, nqubits: int, terms: Iterable[Tuple[PauliInput, complex]], *,
                 max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> None:
        validate_max_bytes(max_bytes)
        if not isinstance(nqubits, int) or isinstance(nqubits, bool) or nqubits < 0:
            raise ValueError("nqubits must be a non-negative integer")
        rows = []
        coefs = []
        for word, coefficient in terms:
            rows.append(_codes_of(nqubits, word))
            coefs.append(complex(coefficient))
        codes = np.asarray(rows, dtype=np.int64).reshape(len(rows), nqubits)
        self._init_from_arrays(nqubits, codes, np.asarray(coefs, dtype=np.complex128))

    def _init_from_arrays(self, nqubits: int, codes: np.ndarray, coefficients: np.ndarray) -> None:
        if codes.size and (codes.min() < 0 or codes.max() > 255):
            bad = np.argwhere((codes < 0) | (codes > 255))[0]
            raise ValueError(f"invalid Pauli code {codes[tuple(bad)]} at index {bad[1]}; expected 0..=3")
        codes8 = np.ascontiguousarray(codes, dtype=np.uint8)
        coef = np.ascontiguousarray(coefficients, dtype=np.complex128)
        out = ctypes.c_void_p()
        _lib.check(_lib.lib().tcp_operator_from_codes(
            nqubits, coef.shape[0], codes8.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
            coef.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(out)))
        self._handle = _OperatorHandle(out)
        self.nqubits = nqubits
        self._terms: Optional[Tuple[PauliTerm, ...]] = None
        self._arrays: Optional[Tuple[np.ndarray, np.ndarray]] = None

    # -- constructors ----------------------------------------------------
    @classmethod
    def from_terms(cls, nqubits: int, terms: Iterable[Tuple[PauliInput, complex]], *,
                   max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> "PauliOperator":
        return cls(nqubits, terms, max_bytes=max_bytes)

    @classmethod
    def from_code_arrays(cls, structures: Any, coefficients: Any, *,
                         max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> "PauliOperator":
        """One native canonicalisation of [T, n] code rows and T coefficients."""
        validate_max_bytes(max_bytes)
        codes = np.asarray(structures)
        coef = np.asarray(coefficients, dtype=np.complex128).reshape(-1)
        if codes.ndim != 2:
            raise ValueError("structures must be a two-dimensional code array")
        if codes.shape[0] != coef.shape[0]:
            raise ValueError(f"expected structure length {codes.shape[0]}, got {coef.shape[0]}")
        instance = object.__new__(cls)
        instance._init_from_arrays(int(codes.shape[1]), codes.astype(np.int64), coef)
        return instance

    @classmethod
    def from_words(cls, nqubits: int, x_words: Any, z_words: Any, coefficients: Any) -> "PauliOperator":
        """From packed [T, W] X/Z words (qubit q at word q // 64, bit q % 64), as
        in the BackendMVPPlan arrays; canonicalised like every constructor."""
        if not isinstance(nqubits, int) or isinstance(nqubits, bool) or nqubits < 0:
            raise ValueError("nqubits must be a non-negative integer")
        W = (nqubits + 63) // 64
        coef = np.ascontiguousarray(np.asarray(coefficients, dtype=np.complex128).reshape(-1))
        x = np.ascontiguousarray(np.asarray(x_words, dtype=np.uint64))
        z = np.ascontiguousarray(np.asarray(z_words, dtype=np.uint64))
        if x.shape != (coef.shape[0], W) or z.shape != (coef.shape[0], W):
            raise ValueError(f"word arrays must have shape ({coef.shape[0]}, {W})")
        out = ctypes.c_void_p()
        u64 = ctypes.POINTER(ctypes.c_uint64)
        _lib.check(_lib.lib().tcp_operator_from_words(nqubits, coef.shape[0], x.ctypes.data_as(u64),
                                                      z.ctypes.data_as(u64),
                                                      coef.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                      ctypes.byref(out)))
        instance = object.__new__(cls)
        instance._handle = _OperatorHandle(out)
        instance.nqubits = nqubits
        instance._terms = None
        instance._arrays = None
        return instance

    @classmethod
    def empty(cls, nqubits: int) -> "PauliOperator":
        return cls(nqubits, ())

    # -- canonical content -----------------------------------------------
    @property
    def term_count(self) -> int:
        return int(_lib.lib().tcp_operator_term_count(self._handle.ptr))

    def code_arrays(self) -> Tuple[np.ndarray, np.ndarray]:
        """Canonical (codes [T, n] uint8, coefficients [T] complex128), read-only."""
        if self._arrays is None:
            T, n = self.term_count, self.nqubits
            codes = np.zeros((T, n), dtype=np.uint8)
            coef = np.zeros(T, dtype=np.complex128)
            _lib.check(_lib.lib().tcp_operator_export_codes(
                self._handle.ptr, codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                coef.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
            codes.flags.writeable = False
            coef.flags.writeable = False
            self._arrays = (codes, coef)
        return self._arrays

    @property
    def terms(self) -> Tuple[PauliTerm, ...]:
        if self._terms is None:
            codes, coef = self.code_arrays()
            self._terms = tuple(PauliTerm(tuple(int(c) for c in row), complex(v))
                                for row, v in zip(codes, coef))
        return self._terms

    def distinct_x_mask_count(self) -> int:
        """Number of distinct X-flip masks (native/operator.rs:50-59)."""
        return int(_lib.lib().tcp_operator_distinct_x_mask_count(self._handle.ptr))

    def backend_mvp_plan(self, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> "BackendMVPPlan":
        """Schema-v1 flat arrays (core/hamiltonian.rs:501-524, py/pauli.py:954-983)."""
        from .backend import plan_from_operator

        return plan_from_operator(self, max_bytes)

    def to_dict(self) -> Dict[str, complex]:
        """Canonical Pauli strings and coefficients (py/pauli.py:585-595)."""
        codes, coef = self.code_arrays()
        letters = np.frombuffer(b"IXYZ", dtype=np.uint8)
        return {letters[row].tobytes().decode(): complex(c) for row, c in zip(codes, coef)}

    @classmethod
    def from_strings(cls, terms: Iterable[Tuple[str, complex]], *,
                     max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> "PauliOperator":
        """From ("IXYZ" string, coefficient) pairs, inferring the qubit count (py/pauli.py:514-540)."""
        items = [(str(w), c) for w, c in terms]
        widths = {len(w) for w, _ in items}
        if len(widths) > 1:
            raise ValueError("all Pauli strings must have the same length")
        return cls(widths.pop() if widths else 0, items, max_bytes=max_bytes)

    # -- algebra (core/operator.rs:217-393, py/pauli.py:597-704) ------------
    def _wrap(self, ptr: ctypes.c_void_p) -> "PauliOperator":
        out = object.__new__(PauliOperator)
        out._handle = _OperatorHandle(ptr)
        out.nqubits = self.nqubits
        out._terms = None
        out._arrays = None
        return out

    def _check_compatible(self, other: object) -> "PauliOperator":
        if not isinstance(other, PauliOperator):
            raise TypeError("expected a PauliOperator")
        return other

    def add(self, other: "PauliOperator", *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> "PauliOperator":
        """Sum; equal words add their coefficients and exact zeros are dropped."""
        validate_max_bytes(max_bytes)
        out = ctypes.c_void_p()
        _lib.check(_lib.lib().tcp_operator_add(self._handle.ptr, self._check_compatible(other)._handle.ptr,
                                               ctypes.byref(out)))
        return self._wrap(out)

    def scale(self, scalar: complex, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> "PauliOperator":
        validate_max_bytes(max_bytes)
        value = complex(scalar)
        if not (np.isfinite(value.real) and np.isfinite(value.imag)):
            raise ValueError("scale must be a finite complex128 value")
        out = ctypes.c_void_p()
        _lib.check(_lib.lib().tcp_operator_scale(self._handle.ptr, value.real, value.imag, ctypes.byref(out)))
        return self._wrap(out)

    def multiply(self, other: "PauliOperator", *,
                 max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> "PauliOperator":
        """Operator product with exact Pauli phases absorbed into the coefficients."""
        validate_max_bytes(max_bytes)
        out = ctypes.c_void_p()
        _lib.check(_lib.lib().tcp_operator_multiply(self._handle.ptr, self._check_compatible(other)._handle.ptr,
                                                    effective_max_bytes(max_bytes), ctypes.byref(out)))
        return self._wrap(out)

    def commutator(self, other: "PauliOperator", *,
                   max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> "PauliOperator":
        """self * other - other * self."""
        return self.multiply(other, max_bytes=max_bytes).add(
            other.multiply(self, max_bytes=max_bytes).scale(-1.0), max_bytes=max_bytes)

    def anticommutator(self, other: "PauliOperator", *,
                       max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> "PauliOperator":
        """self * other + other * self."""
        return self.multiply(other, max_bytes=max_bytes).add(other.multiply(self, max_bytes=max_bytes),
                                                             max_bytes=max_bytes)

    def adjoint(self, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> "PauliOperator":
        validate_max_bytes(max_bytes)
        out = ctypes.c_void_p()
        _lib.check(_lib.lib().tcp_operator_adjoint(self._handle.ptr, ctypes.byref(out)))
        return self._wrap(out)

    def is_hermitian(self, tolerance: float = 0.0) -> bool:
        """Every coefficient real within ``tolerance`` (|c - conj(c)| <= tolerance)."""
        if not np.isfinite(tolerance) or tolerance < 0.0:
            raise ValueError("Hermiticity tolerance must be finite and non-negative")
        _, coef = self.code_arrays()
        return bool(np.all(np.abs(2.0 * coef.imag) <= tolerance))

    def __add__(self, other: object) -> "PauliOperator":
        return self.add(other) if isinstance(other, PauliOperator) else NotImplemented

    def __sub__(self, other: object) -> "PauliOperator":
        return self.add(other.scale(-1.0)) if isinstance(other, PauliOperator) else NotImplemented

    def __neg__(self) -> "PauliOperator":
        return self.scale(-1.0)

    def __mul__(self, other: object) -> "PauliOperator":
        if isinstance(other, PauliOperator):
            return self.multiply(other)
        if isinstance(other, (int, float, complex, np.number)) and not isinstance(other, bool):
            return self.scale(complex(other))
        return NotImplemented

    def __rmul__(self, other: object) -> "PauliOperator":
        if isinstance(other, (int, float, complex, np.number)) and not isinstance(other, bool):
            return self.scale(complex(other))
        return NotImplemented

    def __eq__(self, other: object) -> bool:
        if not isinstance(other, PauliOperator):
            return NotImplemented
        a, b = self.code_arrays(), other.code_arrays()
        return (self.nqubits == other.nqubits and a[0].shape == b[0].shape and bool(np.array_equal(a[0], b[0]))
                and bool(np.array_equal(a[1], b[1])))

    __hash__ = None  # type: ignore[assignment]

    # -- charges ----------------------------------------------------------
    def analyze_charge(self, charge: Any, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> "ChargeAnalysis":
        """Exact conservation analysis of a diagonal qubit charge (core/operator.rs:396-445)."""
        from .charge import AdditiveCharge

        validate_max_bytes(max_bytes)
        if not isinstance(charge, AdditiveCharge):
            raise TypeError("charge must be an AdditiveCharge")
        if charge.nqubits != self.nqubits:
            raise ValueError("Pauli operator and charge layouts are incompatible")
        levels = np.ascontiguousarray(charge.level_array(), dtype=np.float64)
        conserved = ctypes.c_int(0)
        nonzero = ctypes.c_uint64(0)
        _lib.check(_lib.lib().tcp_operator_analyze_charge(
            self._handle.ptr, levels.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            ctypes.byref(conserved), ctypes.byref(nonzero)))
        return ChargeAnalysis(bool(conserved.value), int(nonzero.value))

    def conserves(self, charge: Any, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> bool:
        """Whether this operator exactly conserves ``charge`` (py/pauli.py:696-703)."""
        return self.conserves_charge(charge, max_bytes=max_bytes)

    def conserves_charge(self, charge: Any, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> bool:
        return self.analyze_charge(charge, max_bytes=max_bytes).is_conserved

    def restrict_charge(self, sector: Any, *, storage: str = "lazy",
                        max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        """Restrict an exactly conserved operator to an occupation sector.

        ``U1Sector`` and single full-width number charges use the U(1) path;
        a ``ChargeSector`` of contiguous fixed-population qubit blocks uses
        the factorized block path.
        """
        from .charge import ChargeSector

        if isinstance(sector, ChargeSector) and sector.space is not None and sector.space.fermions:
            raise ValueError("Pauli operator and charge sector layouts are incompatible")
        return self._restrict(sector, storage, max_bytes)

    def _restrict(self, sector: Any, storage: str, max_bytes: Optional[int]) -> Any:
        """Sector dispatch shared with the Jordan-Wigner image of a FermionOperator."""
        from .charge import ChargeSector, restrict_charge
        from .symmetry import U1Sector, canonical_u1_sector, restrict_u1

        if isinstance(sector, U1Sector):
            return restrict_u1(self, sector, max_bytes, storage=storage)
        if isinstance(sector, ChargeSector):
            u1 = canonical_u1_sector(sector)
            if u1 is not None:
                return restrict_u1(self, u1, max_bytes, storage=storage)
            return restrict_charge(self, sector, storage=storage, max_bytes=max_bytes)
        raise TypeError("sector must be a U1Sector or a ChargeSector")

    def restrict_u1(self, sector: Any, *, storage: str = "lazy",
                    max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        """Deprecated alias of ``restrict_charge`` for a ``U1Sector``
        (py/pauli.py:803-828): same checks, same DeprecationWarning."""
        import warnings

        from .symmetry import U1Sector

        warnings.warn("PauliOperator.restrict_u1() is deprecated; use restrict_charge(U1Sector(...)) instead",
                      DeprecationWarning, stacklevel=2)
        validate_max_bytes(max_bytes)
        if not isinstance(sector, U1Sector):
            raise TypeError(f"expected U1Sector, got {type(sector).__name__}")
        if sector.nqubits != self.nqubits:
            raise ValueError(f"expected sector for {self.nqubits} qubits, got {sector.nqubits}")
        return self.restrict_charge(sector, storage=storage, max_bytes=max_bytes)

    # -- MVP --------------------------------------------------------------
    def mvp(self, state: Any, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        """One-shot H @ state (core/hamiltonian.rs:481-483); output is new."""
        validate_max_bytes(max_bytes)
        if not hasattr(state, "is_cuda"):
            values = np.asarray(state, dtype=np.complex128)
            if values.ndim != 1:
                raise ValueError(f"state must be one-dimensional, got shape {values.shape}")
            dim = 1 << self.nqubits if self.nqubits < 64 else None
            if dim is None:
                raise OverflowError("integer overflow while computing matrix dimension")
            if values.shape[0] != dim:
                raise ValueError(f"expected structure length {dim}, got {values.shape[0]}")
        return self.native_mvp_plan(max_bytes=max_bytes).apply(state, max_bytes=max_bytes)

    def native_mvp_plan(self, *, storage: str = "lazy",
                        max_bytes: Optional[int] = DEFAULT_MAX_BYTES,
                        tiling: Optional[str] = None, xcd_order: Optional[bool] = None,
                        prefetch: Optional[int] = None, nontemporal: bool = False) -> "NativeMVPPlan":
        """Compile a reusable matrix-free plan on the current GPU.

        ``tiling`` picks the kernel's tile geometry ("512x8", "512x4",
        "1024x4", "1024x8", "512x16", and the twin-tile "1024x8x2" and
        "512x16x2", whose workgroups own two output tiles; None = automatic:
        complex128 1024x8x2 from 22 qubits and 512x8 below, complex64 storage
        512x8 from 24 qubits and 512x4 below, with the XCD order) and ``xcd_order``
        walks the tiles XCD by XCD; ``prefetch`` (1-3) is the number of source
        tiles each workgroup keeps in flight and ``nontemporal`` marks those
        loads non-temporal (512x8, 512x4 and 1024x4 only).  None of these
        changes the result beyond floating-point summation order.
        """
        validate_max_bytes(max_bytes)
        if storage not in ("lazy", "eager"):
            raise ValueError("storage must be either 'eager' or 'lazy'")
        if tiling is not None and tiling not in TILINGS:
            raise ValueError(f"tiling must be one of {sorted(TILINGS)} or None")
        _lib.require_device()
        out = ctypes.c_void_p()
        st = _lib.TCP_STORAGE_EAGER if storage == "eager" else _lib.TCP_STORAGE_LAZY
        code = TILINGS.get(tiling, _lib.TCP_TILE_AUTO)
        if xcd_order:
            code |= _lib.TCP_TILE_XCD_ORDER
        elif xcd_order is False:
            code |= _lib.TCP_TILE_LINEAR_ORDER
        if prefetch is not None:
            if prefetch not in (1, 2, 3):
                raise ValueError("prefetch must be 1, 2 or 3")
            code |= (prefetch - 1) << _lib.TCP_TILE_PREFETCH_SHIFT
        if nontemporal:
            code |= _lib.TCP_TILE_NT_LOADS
        _lib.check(_lib.lib().tcp_mvp_plan_create_tiled(self._handle.ptr, st, effective_max_bytes(max_bytes),
                                                         code, ctypes.byref(out)))
        return NativeMVPPlan(out, self, storage)

    def to_scipy_linear_operator(self, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        return self.native_mvp_plan(max_bytes=max_bytes).to_scipy_linear_operator(max_bytes=max_bytes)

    # -- materialisation (py/pauli.py:866-928, core/hamiltonian.rs:313-460) --
    def _matrix_dimension(self) -> int:
        if self.nqubits >= 64:
            raise OverflowError("integer overflow while computing matrix dimension")
        return 1 << self.nqubits

    def _sparse_estimate(self, target: str) -> int:
        """Working bytes the reference reserves for COO/CSR (core/hamiltonian.rs:360-460):
        G * 2^n candidate entries, with a candidate buffer and row counts unless
        every X group holds one term."""
        dim = self._matrix_dimension()
        groups = self.distinct_x_mask_count()
        upper = groups * dim
        direct = groups == self.term_count
        if target == "coo":
            return upper * 32 if direct else upper * (24 + 32) + dim * 8
        output = 8 * (dim + 1) + 24 * upper
        return output if direct else upper * 24 + output + dim * 8

    def _full_csr_plan(self, max_bytes: Optional[int]) -> Any:
        from .eager import CsrMvpPlan

        _lib.require_device()
        out = ctypes.c_void_p()
        _lib.check(_lib.lib().tcp_csr_plan_create_full(self._handle.ptr, effective_max_bytes(max_bytes),
                                                        ctypes.byref(out)))
        return CsrMvpPlan(out, "pauli", "pauli", self.term_count)

    def dense(self, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> np.ndarray:
        """Dense complex128 matrix, qubit 0 the most significant index bit.

        Built from the device CSR of the operator; the budget is the dense
        array's 16 * 4^n bytes (core/hamiltonian.rs:313-320)."""
        validate_max_bytes(max_bytes)
        dim = self._matrix_dimension()
        check_allocation(dim * dim * 16, max_bytes, "dense matrix")
        return self._full_csr_plan(max_bytes).materialize("dense", None)

    def coo(self, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        """Deterministic aggregated COO: row-major, columns ascending, exact zeros dropped."""
        validate_max_bytes(max_bytes)
        check_allocation(self._sparse_estimate("coo"), max_bytes, "COO working memory")
        return self._full_csr_plan(max_bytes).materialize("coo", None)

    def csr(self, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        """Deterministic CSR (uint64 indptr/indices, complex128 data)."""
        validate_max_bytes(max_bytes)
        check_allocation(self._sparse_estimate("csr"), max_bytes, "CSR working memory")
        return self._full_csr_plan(max_bytes).materialize("csr", None)

    def csr_mvp_plan(self, *, max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        """The full-space matrix retained as device CSR, applied by the row-parallel
        SpMV (the same plan object as the eager restricted operators)."""
        validate_max_bytes(max_bytes)
        check_allocation(self._sparse_estimate("csr"), max_bytes, "CSR working memory")
        return self._full_csr_plan(max_bytes)

    def compile(self, target: str, *, storage: str = "lazy", max_bytes: Optional[int] = DEFAULT_MAX_BYTES) -> Any:
        """One named target (py/pauli.py:1047-1071): "dense", "coo", "csr",
        "native_mvp" or "backend_mvp"."""
        if target == "dense":
            return self.dense(max_bytes=max_bytes)
        if target == "coo":
            return self.coo(max_bytes=max_bytes)
        if target == "csr":
            return self.csr(max_bytes=max_bytes)
        if target == "backend_mvp":
            return self.backend_mvp_plan(max_bytes=max_bytes)
        if target == "native_mvp":
            return self.native_mvp_plan(storage=storage, max_bytes=max_bytes)
        raise ValueError("target must be one of 'dense', 'coo', 'csr', 'native_mvp', or 'backend_mvp'")

    def __repr__(self) -> str:
        return f"PauliOperator(nqubits={self.nqubits}, term_count={self.term_count})"


@dataclass(frozen=True)
class ChargeAnalysis:
    is_conserved: bool
    nonzero_commutator_terms: int


# the earlier name of the schema-v1 plan class
BackendPlanArrays = BackendMVPPlan


class NativeMVPPlan(NativePlanFacade):
    """Reusable full-space plan (py/hamiltonian.py:162-344) executed by gfx950 kernels."""

    execution_family = "pauli"
    native_variant = "pauli"
    basis_ordering = "qubit0_msb_matrix"
    target = "native_mvp"

    def __init__(self, ptr: ctypes.c_void_p, operator: PauliOperator, storage: str) -> None:
        self._ptr = ptr
        self._operator = operator
        L = _lib.lib()
        self.nqubits = int(L.tcp_mvp_plan_nqubits(ptr))
        self.term_count = int(L.tcp_mvp_plan_term_count(ptr))
        self.group_count = int(L.tcp_mvp_plan_group_count(ptr))
        self.class_count = int(L.tcp_mvp_plan_class_count(ptr))
        self.pass_count = int(L.tcp_mvp_plan_pass_count(ptr))
        self.class_count_c64 = int(L.tcp_mvp_plan_class_count_for(ptr, _lib.TCP_C64))
        names = {v: k for k, v in TILINGS.items()}
        self.tile_bits = int(L.tcp_mvp_plan_tile_bits(ptr, _lib.TCP_C128))
        self.tile_bits_c64 = int(L.tcp_mvp_plan_tile_bits(ptr, _lib.TCP_C64))
        code, code64 = int(L.tcp_mvp_plan_tiling(ptr, _lib.TCP_C128)), int(L.tcp_mvp_plan_tiling(ptr, _lib.TCP_C64))
        self.tiling = names.get(code & 0xFF, "small")        # complex128 applies
        self.tiling_c64 = names.get(code64 & 0xFF, "small")  # complex64 applies
        self.xcd_order = bool(code & _lib.TCP_TILE_XCD_ORDER)
        self.dimension = int(L.tcp_mvp_plan_dimension(ptr))
        self.storage = storage
        self.retained_format = "compact_graph" if storage == "eager" else "compiled_actions"
        self.schedule_count = self.term_count
        word_count = (self.nqubits + 63) // 64
        term_bytes = self.term_count * (word_count * 16 + 16)
        if storage == "eager":  # py/pauli.py:1011-1015
            self.estimated_bytes = term_bytes + operator.distinct_x_mask_count() * self.dimension * 16
            count = int(L.tcp_mvp_plan_transition_count(ptr))
            self.transition_count: Optional[int] = count
        else:
            self.estimated_bytes = term_bytes
            self.transition_count = None
        self.estimated_retained_bytes = self.estimated_bytes
        self.estimated_construction_peak_bytes = self.estimated_bytes
        self.device_bytes = int(L.tcp_mvp_plan_estimated_bytes(ptr))

    def __del__(self) -> None:
        if getattr(self, "_ptr", None) and getattr(_lib, "_lib", None) is not None:
            _lib.lib().tcp_mvp_plan_destroy(self._ptr)
            self._ptr = None

    @property
    def handle(self) -> int:
        """Raw ``tcp_mvp_plan*`` for callers of the C ABI."""
        return int(self._ptr.value or 0)

    def bytes_alg(self, dtype: str = "complex128") -> int:
        """Grouped-gather algorithmic bytes of one apply."""
        code = _lib.TCP_C64 if dtype in ("complex64", "c64") else _lib.TCP_C128
        return int(_lib.lib().tcp_mvp_plan_bytes_alg(self._ptr, code))

    def _native_apply_host(self, inp: np.ndarray, out: np.ndarray) -> int:
        dp = ctypes.POINTER(ctypes.c_double)
        return _lib.lib().tcp_mvp_plan_apply_host(self._ptr, inp.ctypes.data_as(dp), out.ctypes.data_as(dp),
                                                  inp.shape[0], out.shape[0])

    def _native_apply_device(self, inp: Any, out: Any, dtype: int, stream: int) -> int:
        return _lib.lib().tcp_mvp_plan_apply_into(self._ptr, inp.data_ptr(), out.data_ptr(), inp.numel(),
                                                  out.numel(), dtype, stream)

    def apply_batch(self, inputs: Any, outputs: Any) -> None:
        """outputs[r] = H inputs[r] for device tensors of shape (nrhs, 2^n)."""
        import torch

        if inputs.shape != outputs.shape or inputs.dim() != 2 or inputs.shape[1] != self.dimension:
            raise ValueError(f"batched buffers must have shape (nrhs, {self.dimension})")
        if not (inputs.is_contiguous() and outputs.is_contiguous()):
            raise ValueError("batched buffers must be C-contiguous")
        dtype = {torch.complex128: _lib.TCP_C128, torch.complex64: _lib.TCP_C64}.get(inputs.dtype)
        if dtype is None or outputs.dtype != inputs.dtype:
            raise TypeError("batched buffers must both be complex128 or complex64")
        from ._plan import current_stream_handle

        _lib.check(_lib.lib().tcp_mvp_plan_apply_batch(self._ptr, inputs.data_ptr(), outputs.data_ptr(),
                                                       inputs.shape[0], dtype, current_stream_handle(inputs)))

    def __repr__(self) -> str:
        return (f"NativeMVPPlan(nqubits={self.nqubits}, term_count={self.term_count}, "
                f"groups={self.group_count}, classes={self.class_count}, storage={self.storage!r})")


__all__ = [
    "DEFAULT_MAX_BYTES",
    "BackendMVPPlan",
    "BackendPlanArrays",
    "ChargeAnalysis",
    "NativeMVPPlan",
    "PauliOperator",
    "PauliTerm",
    "check_allocation",
]
