"""Helpers shared by the test modules (no GPU work at import time)."""

from __future__ import annotations

import json
import os
from typing import Dict, Tuple

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def kats() -> Dict:
    with open(os.path.join(GOLDEN, "kats.json")) as fh:
        return json.load(fh)


def seeded() -> Dict[str, np.ndarray]:
    with np.load(os.path.join(GOLDEN, "seeded.npz"), allow_pickle=False) as data:
        return {k: data[k] for k in data.files}


def word_codes(word: str) -> Tuple[int, ...]:
    return tuple("IXYZ".index(ch) for ch in word)


def cplx(pair) -> complex:  # type: ignore[no-untyped-def]
    return complex(pair[0], pair[1])


def rel_err(actual: np.ndarray, expected: np.ndarray) -> float:
    denom = max(float(np.linalg.norm(expected)), 1e-300)
    return float(np.linalg.norm(np.asarray(actual) - np.asarray(expected)) / denom)


def conserving_u1_terms(rng: np.random.Generator, n: int, count: int):
    """Random number-conserving Pauli sums: hopping c (XX + YY) and c (XY - YX),
    with the Jordan-Wigner Z string between the two sites when a + b is even,
    plus ZZ and Z terms.  Each (kind, a, b) hopping pair is drawn at most once,
    so every x-mask group cancels exactly on the sources it would move out of
    the sector (no floating-point leakage)."""
    rows, coefs = [], []
    used = set()
    for _ in range(count):
        kind = int(rng.integers(0, 4))
        if kind <= 1 and n >= 2:
            a, b = sorted(int(q) for q in rng.choice(n, size=2, replace=False))
            string = (a + b) % 2 == 0
            key = (kind, a, b)
            if key in used:
                continue
            used.add(key)
            c = complex(rng.normal(), 0.0)
            pairs = ((1, 1, c), (2, 2, c)) if kind == 0 else ((1, 2, c), (2, 1, -c))
            for la, lb, cc in pairs:
                w = [0] * n
                w[a], w[b] = la, lb
                if string:
                    for q in range(a + 1, b):
                        w[q] = 3
                rows.append(w)
                coefs.append(cc)
        elif kind == 2 and n >= 2:
            a, b = rng.choice(n, size=2, replace=False)
            w = [0] * n
            w[a] = w[b] = 3
            rows.append(w)
            coefs.append(complex(rng.normal(), 0.0))
        else:
            w = [0] * n
            if n:
                w[int(rng.integers(0, n))] = 3
            rows.append(w)
            coefs.append(complex(rng.normal(), 0.0))
    return np.asarray(rows, dtype=np.uint8).reshape(-1, n), np.asarray(coefs, dtype=np.complex128)


def interleave_spins(codes: np.ndarray) -> np.ndarray:
    """Relabel the qubits of Pauli words from the block layout (spin up on
    qubits 0..L-1, spin down on L..2L-1) to interleaved spin orbitals
    (site i up on 2i, down on 2i + 1).  A qubit relabeling is exact for Pauli
    operators (no Jordan-Wigner string is re-derived)."""
    codes = np.asarray(codes, dtype=np.uint8)
    n = codes.shape[1]
    sites = n // 2
    out = np.zeros_like(codes)
    for i in range(sites):
        out[:, 2 * i] = codes[:, i]
        out[:, 2 * i + 1] = codes[:, sites + i]
    return out


def spin_levels(n: int, interleaved: bool = True) -> np.ndarray:
    """levels [2][n][2] of the total number N and of 2 S_z (+1 up, -1 down)."""
    levels = np.zeros((2, n, 2), dtype=np.int64)
    levels[0, :, 1] = 1
    for a in range(n):
        up = (a % 2 == 0) if interleaved else (a < n // 2)
        levels[1, a, 1] = 1 if up else -1
    return levels


def words_of(bits: np.ndarray, n: int) -> np.ndarray:
    """uint64 words (qubit 0 = MSB) of occupation rows."""
    rows = np.asarray(bits, dtype=np.uint64).reshape(-1, n)
    shifts = np.arange(n - 1, -1, -1, dtype=np.uint64)
    return (rows << shifts[None, :]).sum(axis=1, dtype=np.uint64)
