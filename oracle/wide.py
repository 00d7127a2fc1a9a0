"""Big-integer oracle for wide U(1) sectors (test infrastructure only).

Restates, with Python integers of any width, the U(1) restricted operator of
the reference for sectors beyond 64 qubits:

* the basis is every n-bit integer of weight k in ascending order, qubit 0
  the most significant bit (U1Sector public order, core/sector.rs:79-238);
* a Pauli term acts on a source integer s as w (-1)^{popcount(z & s)} with
  destination s ^ x, w = c i^{#Y mod 4} (core/pauli_action.rs:24-109);
* per source, terms with the same destination are aggregated, exact zeros
  dropped, and a nonzero destination outside the sector is leakage
  (aggregate_source, core/sector.rs:1051-1127).

Pinned in tests/test_wide_sector.py by the closed form of the combinatorial
number system, brute-force enumeration and the reference's known answers.
"""

from __future__ import annotations

from itertools import combinations
from typing import Dict, List, Sequence, Tuple

import numpy as np


class LeakageError(ValueError):
    pass


def basis_values(nqubits: int, k: int) -> List[int]:
    return sorted(sum(1 << (nqubits - 1 - q) for q in occ) for occ in combinations(range(nqubits), k))


def packed_words(value: int, nqubits: int) -> Tuple[int, ...]:
    """Integer bitstring -> words with qubit q in word q // 64, bit q % 64."""
    words = [0] * max(1, (nqubits + 63) // 64)
    for q in range(nqubits):
        if (value >> (nqubits - 1 - q)) & 1:
            words[q // 64] |= 1 << (q % 64)
    return tuple(words)


def term_masks(codes: Sequence[int], nqubits: int) -> Tuple[int, int, complex]:
    """(x mask, z mask, i^{#Y}) of one code row, qubit q at bit n - 1 - q."""
    x = z = 0
    ny = 0
    for q, c in enumerate(codes):
        bit = 1 << (nqubits - 1 - q)
        if c in (1, 2):
            x |= bit
        if c in (2, 3):
            z |= bit
        ny += c == 2
    return x, z, 1j ** (ny % 4)


def restricted_sparse(nqubits: int, k: int, codes: np.ndarray, coef: np.ndarray) -> Dict[Tuple[int, int], complex]:
    """{(row, column): value} of the restricted operator (rows = destinations)."""
    basis = basis_values(nqubits, k)
    lookup = {v: i for i, v in enumerate(basis)}
    terms = [term_masks(row, nqubits) + (complex(c),) for row, c in zip(np.asarray(codes), np.asarray(coef))]
    out: Dict[Tuple[int, int], complex] = {}
    for col, s in enumerate(basis):
        agg: Dict[int, complex] = {}
        for x, z, phase, c in terms:
            sign = -1 if bin(z & s).count("1") % 2 else 1
            d = s ^ x
            agg[d] = agg.get(d, 0j) + c * phase * sign
        for d, value in agg.items():
            if value == 0:
                continue
            if bin(d).count("1") != k:
                raise LeakageError(f"leakage from restricted source index {col}")
            out[(lookup[d], col)] = out.get((lookup[d], col), 0j) + value
    return out


def restricted_apply(nqubits: int, k: int, codes: np.ndarray, coef: np.ndarray, state: np.ndarray) -> np.ndarray:
    y = np.zeros(len(state), dtype=np.complex128)
    for (row, col), value in restricted_sparse(nqubits, k, codes, coef).items():
        y[row] += value * state[col]
    return y
