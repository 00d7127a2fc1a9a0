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
