"""Dense fermion oracle from the occupation-basis definition (test infrastructure only).

Mode m is the bit n-1-m of a basis index (mode 0 most significant); the
ladder operators act on occupation strings with the Jordan-Wigner sign of
the occupied modes below m:

    a_m |..1_m..> = (-1)^{sum_{l<m} n_l} |..0_m..>,
    a+_m |..0_m..> = (-1)^{sum_{l<m} n_l} |..1_m..>,

and a raw product is the matrix product of its factors in the order written
(the definition of the reference's tests/test_structured_algebra.py:256-277).
Nothing here uses Pauli algebra, normal ordering or the library: it pins both
the CAR canonicalisation and the Jordan-Wigner map of ``FermionOperator``.
"""

from __future__ import annotations

from typing import Iterable, Sequence, Tuple

import numpy as np

Factor = Tuple[int, str]


def ladder_matrix(n_modes: int, mode: int, action: str) -> np.ndarray:
    dim = 1 << n_modes
    out = np.zeros((dim, dim), dtype=np.complex128)
    bit = 1 << (n_modes - 1 - mode)
    lower = sum(1 << (n_modes - 1 - m) for m in range(mode))
    for col in range(dim):
        occupied = bool(col & bit)
        if occupied == (action == "create"):
            continue
        row = col ^ bit
        out[row, col] = -1.0 if bin(col & lower).count("1") % 2 else 1.0
    return out


def product_matrix(n_modes: int, factors: Sequence[Factor]) -> np.ndarray:
    out = np.eye(1 << n_modes, dtype=np.complex128)
    for mode, action in factors:
        out = out @ ladder_matrix(n_modes, mode, action)
    return out


def operator_matrix(n_modes: int, terms: Iterable[Tuple[Sequence[Factor], complex]]) -> np.ndarray:
    out = np.zeros((1 << n_modes, 1 << n_modes), dtype=np.complex128)
    for factors, coefficient in terms:
        out += complex(coefficient) * product_matrix(n_modes, factors)
    return out


def spinful_terms(sites: int):
    """Complex hopping, on-site and pair-hopping terms with exact cancellations
    (the shape of the reference's tests/test_charge.py:652-690)."""
    h = 0.7 + 0.2j
    return [
        (((0, "create"), (0, "annihilate"), (sites, "create"), (sites, "annihilate")), 1.3),
        (((0, "create"), (1, "annihilate")), h),
        (((1, "create"), (0, "annihilate")), h.conjugate()),
        (((sites, "create"), (sites + 1, "annihilate")), -h),
        (((sites + 1, "create"), (sites, "annihilate")), -h.conjugate()),
        (((0, "create"), (sites + 1, "create"), (1, "annihilate"), (sites, "annihilate")), 0.35j),
        (((sites, "create"), (1, "create"), (sites + 1, "annihilate"), (0, "annihilate")), -0.35j),
        (((0, "create"), (0, "annihilate")), 0.25),
        (((0, "create"), (0, "annihilate")), -0.25),
        (((1, "create"), (1, "annihilate")), 0.0),
    ]
