"""CPU oracle for the Pauli-sum MVP path (test infrastructure only).

Only the test suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, and only as the checker or as the timed CPU
baseline.  The product library never imports, links or calls it.

Two independent oracles live here:

* ``tcp_oracle.c`` (loaded through ctypes below): a plain-C restatement of the
  reference algorithms -- canonicalisation (core/operator.rs:117-155),
  TermDirect apply (core/hamiltonian.rs:204-270), U(1) rank/unrank and the lazy
  restricted scatter (core/sector.rs:79-342, 579-642, 1051-1127), contiguous
  fixed-population blocks and general sectors (the suffix-count plan,
  core/charge_sector.rs:1426-1578, 1620-1722) with the generic charge loop
  (core/charge_sector.rs:973-1038, core/charge/generic.rs:949-1045), the eager
  CSR apply (core/csr.rs:194-238) and ``analyze_charge``
  (core/operator.rs:396-445).
* ``dense.py``: a dense Kronecker-sum oracle and brute-force sector bases built
  from textbook definitions only (the pattern of the reference's
  tests/reference.py:15-115 and tests/test_charge.py:63-83).
* ``wide.py``: the U(1) restricted operator over Python big integers for
  sectors of more than 64 qubits.
* ``fermion.py``: ladder operators on the occupation basis from their
  definition (tests/test_structured_algebra.py:256-277), the checker of the
  fermion canonicalisation, the Jordan-Wigner map and the fermion sector
  plans.

Parity of each function is pinned in ``tests/test_oracle.py`` against the
reference's known-answer vectors (``tests/golden/``) and the dense oracle.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libtcp_oracle.so")
_lib: Optional[ctypes.CDLL] = None

STATUS_NAMES = {
    0: "OK",
    1: "INVALID_STRUCTURE_LENGTH",
    2: "INVALID_CODE",
    3: "NON_FINITE_COEFFICIENT",
    4: "MEMORY_LIMIT",
    5: "OVERFLOW",
    6: "INVALID_SECTOR",
    7: "SECTOR_LEAKAGE",
    8: "INVALID_INDEX",
    9: "INVALID_ARGUMENT",
}


class OracleError(RuntimeError):
    def __init__(self, status: int, detail: str = "") -> None:
        self.status = status
        super().__init__(f"oracle status {STATUS_NAMES.get(status, status)} {detail}".strip())


def build() -> str:
    """Compile the C oracle with gcc (oracle/Makefile) and return the .so path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_dp = ctypes.POINTER(ctypes.c_double)


def _declare(L: ctypes.CDLL) -> None:
    c = ctypes
    L.orc_canonicalize.argtypes = [c.c_uint32, c.c_uint64, _u8p, _dp, _u8p, _dp, _u64p, _u64p]
    L.orc_matrix_terms.argtypes = [c.c_uint32, c.c_uint64, _u8p, _dp, _u64p, _u64p, _dp]
    L.orc_matrix_terms.restype = None
    L.orc_termdirect_apply.argtypes = [c.c_uint32, c.c_uint64, _u64p, _u64p, _dp, _dp, _dp, c.c_int]
    L.orc_termdirect_rows.argtypes = [
        c.c_uint32, c.c_uint64, _u64p, _u64p, _dp, _dp, _u64p, c.c_uint64, _dp, c.c_int]
    L.orc_u1_dimension.argtypes = [c.c_uint32, c.c_uint32, _u64p]
    L.orc_u1_rank.argtypes = [c.c_uint32, c.c_uint32, c.c_uint64, _u64p]
    L.orc_u1_unrank.argtypes = [c.c_uint32, c.c_uint32, c.c_uint64, _u64p]
    L.orc_u1_rank_many.argtypes = [c.c_uint32, c.c_uint32, _u64p, c.c_uint64, _u64p]
    L.orc_u1_unrank_many.argtypes = [c.c_uint32, c.c_uint32, _u64p, c.c_uint64, _u64p]
    L.orc_u1_lazy_apply.argtypes = [
        c.c_uint32, c.c_uint32, c.c_uint64, _u8p, _dp, _dp, _dp, _u64p, _u32p]
    L.orc_u1_lazy_apply_range.argtypes = [
        c.c_uint32, c.c_uint32, c.c_uint64, _u8p, _dp, _dp, _dp, c.c_uint64, c.c_uint64, _u64p, _u32p]
    L.orc_blocks_dimension.argtypes = [c.c_uint32, _u32p, _u32p, _u64p]
    L.orc_blocks_rank_many.argtypes = [c.c_uint32, _u32p, _u32p, _u64p, c.c_uint64, _u64p]
    L.orc_blocks_unrank_many.argtypes = [c.c_uint32, _u32p, _u32p, _u64p, c.c_uint64, _u64p]
    L.orc_blocks_generic_apply.argtypes = [
        c.c_uint32, _u32p, _u32p, c.c_uint64, _u8p, _dp, _dp, _dp, _u64p]
    L.orc_blocks_generic_apply_range.argtypes = [
        c.c_uint32, _u32p, _u32p, c.c_uint64, _u8p, _dp, _dp, _dp, c.c_uint64, c.c_uint64, _u64p]
    i64p = c.POINTER(c.c_int64)
    L.orc_charge_dimension.argtypes = [c.c_uint32, c.c_uint32, i64p, i64p, _u64p]
    L.orc_charge_rank_many.argtypes = [c.c_uint32, c.c_uint32, i64p, i64p, _u64p, c.c_uint64, _u64p]
    L.orc_charge_unrank_many.argtypes = [c.c_uint32, c.c_uint32, i64p, i64p, _u64p, c.c_uint64, _u64p]
    L.orc_charge_generic_apply_range.argtypes = [
        c.c_uint32, c.c_uint32, i64p, i64p, c.c_uint64, _u8p, _dp, _dp, _dp, c.c_uint64, c.c_uint64, _u64p]
    L.orc_csr_apply.argtypes = [c.c_uint64, _u64p, _u64p, _dp, _dp, _dp, c.c_uint64, c.c_uint64, c.c_int]
    L.orc_analyze_charge.argtypes = [c.c_uint32, c.c_uint64, _u8p, _dp, _dp, _u64p]


def _p(array: np.ndarray, ptype):  # type: ignore[no-untyped-def]
    return array.ctypes.data_as(ptype)


def _check(status: int, detail: str = "") -> None:
    if status != 0:
        raise OracleError(status, detail)


def _terms(nqubits: int, codes, coef) -> Tuple[np.ndarray, np.ndarray]:  # type: ignore[no-untyped-def]
    """Return (codes [T, n] u8, coefficients [T] c128), both C-contiguous."""
    cf = np.ascontiguousarray(np.asarray(coef, dtype=np.complex128).reshape(-1))
    codes = np.asarray(codes, dtype=np.uint8).reshape(len(cf), nqubits)
    return np.ascontiguousarray(codes), cf


def canonicalize(nqubits: int, codes: np.ndarray, coef: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """PauliOperator::from_terms: returns (canonical codes [T,n] u8, coefficients c128)."""
    codes, cf = _terms(nqubits, codes, coef)
    T = codes.shape[0]
    out_codes = np.zeros((max(T, 1), max(nqubits, 1)), dtype=np.uint8)
    out_coef = np.zeros(max(T, 1), dtype=np.complex128)
    out_T = ctypes.c_uint64(0)
    err = ctypes.c_uint64(0)
    st = lib().orc_canonicalize(
        nqubits, T, _p(codes, _u8p), _p(cf.view(np.float64), _dp),
        _p(out_codes, _u8p), _p(out_coef.view(np.float64), _dp), ctypes.byref(out_T), ctypes.byref(err))
    _check(st, f"index={err.value}")
    t = out_T.value
    return out_codes[:t, :nqubits].copy(), out_coef[:t].copy()


def matrix_terms(nqubits: int, codes: np.ndarray, coef: np.ndarray) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Index-space (x_mask, z_mask, weighted phase w = c * i^#Y) per term."""
    codes, cf = _terms(nqubits, codes, coef)
    T = codes.shape[0]
    x = np.zeros(T, dtype=np.uint64)
    z = np.zeros(T, dtype=np.uint64)
    w = np.zeros(T, dtype=np.complex128)
    lib().orc_matrix_terms(nqubits, T, _p(codes, _u8p), _p(cf.view(np.float64), _dp),
                           _p(x, _u64p), _p(z, _u64p), _p(w.view(np.float64), _dp))
    return x, z, w


def termdirect_apply(nqubits: int, codes: np.ndarray, coef: np.ndarray, state: np.ndarray,
                     nthreads: int = 1) -> np.ndarray:
    """MvpPlan::apply_into (TermDirect) on canonical terms."""
    x, z, w = matrix_terms(nqubits, codes, coef)
    v = np.ascontiguousarray(state, dtype=np.complex128)
    if v.shape != (1 << nqubits,):
        raise ValueError("state has the wrong length")
    y = np.empty_like(v)
    _check(lib().orc_termdirect_apply(nqubits, len(x), _p(x, _u64p), _p(z, _u64p),
                                      _p(w.view(np.float64), _dp), _p(v.view(np.float64), _dp),
                                      _p(y.view(np.float64), _dp), nthreads))
    return y


def termdirect_rows(nqubits: int, codes: np.ndarray, coef: np.ndarray, state: np.ndarray,
                    rows: np.ndarray, nthreads: int = 1,
                    masks: Optional[Tuple[np.ndarray, np.ndarray, np.ndarray]] = None) -> np.ndarray:
    """(H v)[rows] by the TermDirect row gather, without a full output vector."""
    x, z, w = masks if masks is not None else matrix_terms(nqubits, codes, coef)
    rows = np.ascontiguousarray(rows, dtype=np.uint64)
    v = np.ascontiguousarray(state)
    if v.dtype != np.complex128:
        v = v.astype(np.complex128)
    y = np.empty(len(rows), dtype=np.complex128)
    _check(lib().orc_termdirect_rows(nqubits, len(x), _p(x, _u64p), _p(z, _u64p),
                                     _p(w.view(np.float64), _dp), _p(v.view(np.float64), _dp),
                                     _p(rows, _u64p), len(rows), _p(y.view(np.float64), _dp), nthreads))
    return y


def u1_dimension(nqubits: int, k: int) -> int:
    d = ctypes.c_uint64(0)
    _check(lib().orc_u1_dimension(nqubits, k, ctypes.byref(d)))
    return d.value


def u1_rank_many(nqubits: int, k: int, bitstrings: np.ndarray) -> np.ndarray:
    b = np.ascontiguousarray(bitstrings, dtype=np.uint64)
    r = np.zeros(len(b), dtype=np.uint64)
    _check(lib().orc_u1_rank_many(nqubits, k, _p(b, _u64p), len(b), _p(r, _u64p)))
    return r


def u1_unrank_many(nqubits: int, k: int, indices: np.ndarray) -> np.ndarray:
    i = np.ascontiguousarray(indices, dtype=np.uint64)
    b = np.zeros(len(i), dtype=np.uint64)
    _check(lib().orc_u1_unrank_many(nqubits, k, _p(i, _u64p), len(i), _p(b, _u64p)))
    return b


def u1_lazy_apply(nqubits: int, k: int, codes: np.ndarray, coef: np.ndarray,
                  state: np.ndarray, sources: Optional[Tuple[int, int]] = None) -> np.ndarray:
    """U1LazyMvpPlan::apply_into on canonical terms (raises OracleError on leakage).
    With ``sources=(begin, end)`` only those source ranks scatter (a bounded
    timing sample; the full apply is every source)."""
    codes, cf = _terms(nqubits, codes, coef)
    T = codes.shape[0]
    v = np.ascontiguousarray(state, dtype=np.complex128)
    y = np.empty_like(v)
    leak = ctypes.c_uint64(0)
    weight = ctypes.c_uint32(0)
    lo, hi = sources if sources is not None else (0, 2**64 - 1)
    st = lib().orc_u1_lazy_apply_range(nqubits, k, T, _p(codes, _u8p), _p(cf.view(np.float64), _dp),
                                       _p(v.view(np.float64), _dp), _p(y.view(np.float64), _dp), lo, hi,
                                       ctypes.byref(leak), ctypes.byref(weight))
    _check(st, f"source_index={leak.value} actual={weight.value}")
    return y


def _blocks_arrays(blocks: Sequence[Tuple[int, int]]) -> Tuple[np.ndarray, np.ndarray]:
    sizes = np.ascontiguousarray([b[0] for b in blocks], dtype=np.uint32)
    pops = np.ascontiguousarray([b[1] for b in blocks], dtype=np.uint32)
    return sizes, pops


def blocks_dimension(blocks: Sequence[Tuple[int, int]]) -> int:
    s, p = _blocks_arrays(blocks)
    d = ctypes.c_uint64(0)
    _check(lib().orc_blocks_dimension(len(s), _p(s, _u32p), _p(p, _u32p), ctypes.byref(d)))
    return d.value


def blocks_rank_many(blocks: Sequence[Tuple[int, int]], bitstrings: np.ndarray) -> np.ndarray:
    s, p = _blocks_arrays(blocks)
    b = np.ascontiguousarray(bitstrings, dtype=np.uint64)
    r = np.zeros(len(b), dtype=np.uint64)
    _check(lib().orc_blocks_rank_many(len(s), _p(s, _u32p), _p(p, _u32p), _p(b, _u64p), len(b), _p(r, _u64p)))
    return r


def blocks_unrank_many(blocks: Sequence[Tuple[int, int]], indices: np.ndarray) -> np.ndarray:
    s, p = _blocks_arrays(blocks)
    i = np.ascontiguousarray(indices, dtype=np.uint64)
    b = np.zeros(len(i), dtype=np.uint64)
    _check(lib().orc_blocks_unrank_many(len(s), _p(s, _u32p), _p(p, _u32p), _p(i, _u64p), len(i), _p(b, _u64p)))
    return b


def blocks_generic_apply(blocks: Sequence[Tuple[int, int]], codes: np.ndarray, coef: np.ndarray,
                         state: np.ndarray, sources: Optional[Tuple[int, int]] = None) -> np.ndarray:
    """Generic charge loop over a contiguous fixed-population block sector
    (``sources=(begin, end)``: those source ranks only)."""
    s, p = _blocks_arrays(blocks)
    n = int(s.sum())
    codes, cf = _terms(n, codes, coef)
    T = codes.shape[0]
    v = np.ascontiguousarray(state, dtype=np.complex128)
    y = np.empty_like(v)
    leak = ctypes.c_uint64(0)
    lo, hi = sources if sources is not None else (0, 2**64 - 1)
    st = lib().orc_blocks_generic_apply_range(len(s), _p(s, _u32p), _p(p, _u32p), T, _p(codes, _u8p),
                                              _p(cf.view(np.float64), _dp), _p(v.view(np.float64), _dp),
                                              _p(y.view(np.float64), _dp), lo, hi, ctypes.byref(leak))
    _check(st, f"source_index={leak.value}")
    return y


def _charge_arrays(levels, targets) -> Tuple[np.ndarray, np.ndarray]:  # type: ignore[no-untyped-def]
    """levels [m][n][2] and targets [m] as contiguous int64 arrays."""
    lv = np.ascontiguousarray(np.asarray(levels, dtype=np.int64))
    if lv.ndim != 3 or lv.shape[2] != 2:
        raise ValueError("levels must have shape [constraints, axes, 2]")
    tg = np.ascontiguousarray(np.asarray(targets, dtype=np.int64).reshape(-1))
    if tg.shape[0] != lv.shape[0]:
        raise ValueError("one target per constraint")
    return lv, tg


_i64p = ctypes.POINTER(ctypes.c_int64)


def charge_dimension(levels, targets) -> int:  # type: ignore[no-untyped-def]
    """Dimension of the general charge sector (build_charge_sector_plan)."""
    lv, tg = _charge_arrays(levels, targets)
    d = ctypes.c_uint64(0)
    _check(lib().orc_charge_dimension(lv.shape[1], lv.shape[0], _p(lv, _i64p), _p(tg, _i64p), ctypes.byref(d)))
    return int(d.value)


def charge_rank_many(levels, targets, bitstrings: np.ndarray) -> np.ndarray:  # type: ignore[no-untyped-def]
    """ChargeSectorPlan::rank of every bitstring (UINT64_MAX outside the sector)."""
    lv, tg = _charge_arrays(levels, targets)
    b = np.ascontiguousarray(bitstrings, dtype=np.uint64)
    r = np.zeros(len(b), dtype=np.uint64)
    _check(lib().orc_charge_rank_many(lv.shape[1], lv.shape[0], _p(lv, _i64p), _p(tg, _i64p), _p(b, _u64p), len(b),
                                      _p(r, _u64p)))
    return r


def charge_unrank_many(levels, targets, indices: np.ndarray) -> np.ndarray:  # type: ignore[no-untyped-def]
    """ChargeSectorPlan::unrank of every index."""
    lv, tg = _charge_arrays(levels, targets)
    i = np.ascontiguousarray(indices, dtype=np.uint64)
    b = np.zeros(len(i), dtype=np.uint64)
    _check(lib().orc_charge_unrank_many(lv.shape[1], lv.shape[0], _p(lv, _i64p), _p(tg, _i64p), _p(i, _u64p),
                                        len(i), _p(b, _u64p)))
    return b


def charge_generic_apply(levels, targets, codes: np.ndarray, coef: np.ndarray,  # type: ignore[no-untyped-def]
                         state: np.ndarray, sources: Optional[Tuple[int, int]] = None) -> np.ndarray:
    """The generic charge loop of a Pauli operator on a general charge sector
    (``sources=(begin, end)``: those source ranks only)."""
    lv, tg = _charge_arrays(levels, targets)
    n = lv.shape[1]
    codes, cf = _terms(n, codes, coef)
    T = codes.shape[0]
    v = np.ascontiguousarray(state, dtype=np.complex128)
    y = np.empty_like(v)
    leak = ctypes.c_uint64(0)
    lo, hi = sources if sources is not None else (0, 2**64 - 1)
    st = lib().orc_charge_generic_apply_range(n, lv.shape[0], _p(lv, _i64p), _p(tg, _i64p), T, _p(codes, _u8p),
                                              _p(cf.view(np.float64), _dp), _p(v.view(np.float64), _dp),
                                              _p(y.view(np.float64), _dp), lo, hi, ctypes.byref(leak))
    _check(st, f"source_index={leak.value}")
    return y


def csr_apply(indptr: np.ndarray, indices: np.ndarray, values: np.ndarray, state: np.ndarray,
              rows: Optional[Tuple[int, int]] = None, nthreads: int = 1) -> np.ndarray:
    """CsrGraph::apply_into (core/csr.rs:194-238): row-parallel gather in stored
    column order.  With ``rows=(begin, end)`` only those rows are computed (the
    others are zero)."""
    ip = np.ascontiguousarray(indptr, dtype=np.uint64)
    ix = np.ascontiguousarray(indices, dtype=np.uint64)
    va = np.ascontiguousarray(values, dtype=np.complex128)
    v = np.ascontiguousarray(state, dtype=np.complex128)
    dim = len(ip) - 1
    y = np.zeros(dim, dtype=np.complex128)
    lo, hi = rows if rows is not None else (0, dim)
    _check(lib().orc_csr_apply(dim, _p(ip, _u64p), _p(ix, _u64p), _p(va.view(np.float64), _dp),
                               _p(v.view(np.float64), _dp), _p(y.view(np.float64), _dp), lo, hi, nthreads))
    return y


def analyze_charge(nqubits: int, codes: np.ndarray, coef: np.ndarray,
                   levels: Sequence[Tuple[float, float]]) -> Tuple[bool, int]:
    """PauliOperator::analyze_charge: (is_conserved, nonzero commutator words)."""
    codes, cf = _terms(nqubits, codes, coef)
    T = codes.shape[0]
    lv = np.ascontiguousarray(np.asarray(levels, dtype=np.float64).reshape(nqubits, 2))
    nz = ctypes.c_uint64(0)
    _check(lib().orc_analyze_charge(nqubits, T, _p(codes, _u8p), _p(cf.view(np.float64), _dp),
                                    _p(lv, _dp), ctypes.byref(nz)))
    return nz.value == 0, nz.value
