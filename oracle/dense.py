"""Independent dense oracles built from textbook definitions (test infrastructure).

Nothing here follows the reference's code paths: the operator is an explicit
sum of Kronecker products of the 2x2 Pauli matrices with qubit 0 as the most
significant factor (the definition the reference's own tests/reference.py:15-115
uses), sectors are enumerated by brute force over all 2^n integers, and ranks
come from ``math.comb``.  Usable for n <= ~11 (dense) and n <= ~20 (sectors).
"""

from __future__ import annotations

import itertools
import math
from typing import Iterable, List, Sequence, Tuple

import numpy as np

PAULI = (
    np.array([[1, 0], [0, 1]], dtype=np.complex128),
    np.array([[0, 1], [1, 0]], dtype=np.complex128),
    np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    np.array([[1, 0], [0, -1]], dtype=np.complex128),
)


def word_matrix(codes: Sequence[int]) -> np.ndarray:
    out = np.ones((1, 1), dtype=np.complex128)
    for c in codes:
        out = np.kron(out, PAULI[int(c)])
    return out


def dense_operator(nqubits: int, codes: Iterable[Sequence[int]], coef: Sequence[complex]) -> np.ndarray:
    dim = 1 << nqubits
    out = np.zeros((dim, dim), dtype=np.complex128)
    for word, c in zip(codes, coef):
        out += complex(c) * word_matrix(word)
    return out


def popcount(value: int) -> int:
    return bin(value).count("1")


def u1_basis(nqubits: int, k: int) -> List[int]:
    """All n-bit integers of Hamming weight k, ascending (TensorCircuit order)."""
    return [b for b in range(1 << nqubits) if popcount(b) == k]


def blocks_basis(blocks: Sequence[Tuple[int, int]]) -> List[int]:
    """Brute-force basis of contiguous fixed-population blocks (qubit 0 = MSB).

    Block j covers the next ``blocks[j][0]`` qubits and must hold
    ``blocks[j][1]`` ones; the order is ascending integer order, which equals
    the lexicographic mixed-radix order over qubits 0..n-1 with qubit 0 most
    significant (the pattern of tests/test_charge.py:63-83).
    """
    n = sum(s for s, _ in blocks)
    out = []
    for b in range(1 << n):
        shift = n
        ok = True
        for size, pop in blocks:
            shift -= size
            if popcount((b >> shift) & ((1 << size) - 1)) != pop:
                ok = False
                break
        if ok:
            out.append(b)
    return out


def charge_basis(levels, targets) -> List[int]:  # type: ignore[no-untyped-def]
    """Brute-force basis of a general charge sector (qubit 0 = MSB).

    ``levels[j][a] = (value with bit a clear, value with bit a set)`` of
    constraint j; a state is selected when every charge, summed over the
    axes, equals its target.  Ascending integer order, as
    tests/test_charge.py:63-83 enumerates ``itertools.product`` over axes.
    Vectorised over all 2^n integers (n <= ~24).
    """
    lv = np.asarray(levels, dtype=np.int64)
    m, n = lv.shape[0], lv.shape[1]
    states = np.arange(1 << n, dtype=np.int64)
    keep = np.ones(states.shape[0], dtype=bool)
    for j in range(m):
        total = np.zeros(states.shape[0], dtype=np.int64)
        for a in range(n):
            bit = (states >> (n - 1 - a)) & 1
            total += np.where(bit == 1, lv[j, a, 1], lv[j, a, 0])
        keep &= total == int(targets[j])
    return [int(b) for b in states[keep]]


def comb_rank(bits: int, k: int) -> int:
    """Combinatorial number system: sum_j C(p_j, j) over ascending set bits."""
    rank = 0
    j = 0
    p = 0
    while bits:
        if bits & 1:
            j += 1
            rank += math.comb(p, j)
        bits >>= 1
        p += 1
    assert j == k
    return rank


def restricted_dense(nqubits: int, codes, coef, basis: Sequence[int]) -> np.ndarray:
    full = dense_operator(nqubits, codes, coef)
    idx = np.asarray(basis, dtype=np.int64)
    return full[np.ix_(idx, idx)]


def random_codes(rng: np.random.Generator, nqubits: int, count: int, p_identity: float = 0.7) -> np.ndarray:
    """Per qubit: I with probability p_identity, else X/Y/Z uniformly."""
    letters = rng.integers(1, 4, size=(count, nqubits))
    identity = rng.random(size=(count, nqubits)) < p_identity
    return np.where(identity, 0, letters).astype(np.uint8)


def all_words(nqubits: int) -> Iterable[Tuple[int, ...]]:
    return itertools.product(range(4), repeat=nqubits)
