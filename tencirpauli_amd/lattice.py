"""Seeded Pauli-sum workloads: random Pauli sums, XXZ lattices, Jordan-Wigner Hubbard.

These build the synthetic operators the MVP is measured and tested on.  All
randomness comes from ``numpy.random.default_rng(seed)`` (PCG64) in the order
documented per function, so the same seed gives the same operator anywhere.
Codes are 0=I 1=X 2=Y 3=Z with qubit 0 first.
"""

from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

import numpy as np

BASE_SEED = 20260801

Terms = Tuple[np.ndarray, np.ndarray]  # (codes [T, n] uint8, coefficients [T] complex128)


def random_pauli_sum(nqubits: int, nterms: int, seed: int, p_identity: float = 0.7) -> Terms:
    """Random strings: per qubit I with probability p_identity, else X/Y/Z uniformly;
    real N(0, 1) coefficients.  Draw order: letters [T, n] in 1..3, then
    uniforms [T, n] (identity where < p_identity), then T normals."""
    rng = np.random.default_rng(seed)
    letters = rng.integers(1, 4, size=(nterms, nqubits))
    identity = rng.random(size=(nterms, nqubits)) < p_identity
    codes = np.where(identity, 0, letters).astype(np.uint8)
    coef = rng.standard_normal(nterms).astype(np.complex128)
    return codes, coef


def open_lattice_bonds(rows: int, cols: int) -> List[Tuple[int, int]]:
    """Nearest-neighbour bonds of an open rows x cols lattice, site = row*cols + col,
    in row-major order with the horizontal bond of a site before its vertical one."""
    bonds = []
    for r in range(rows):
        for c in range(cols):
            s = r * cols + c
            if c + 1 < cols:
                bonds.append((s, s + 1))
            if r + 1 < rows:
                bonds.append((s, s + cols))
    return bonds


def periodic_lattice_bonds(rows: int, cols: int) -> List[Tuple[int, int]]:
    """Bonds of a periodic rows x cols torus (each site to its right and lower neighbour)."""
    bonds = []
    for r in range(rows):
        for c in range(cols):
            s = r * cols + c
            bonds.append((s, r * cols + (c + 1) % cols))
            bonds.append((s, ((r + 1) % rows) * cols + c))
    return bonds


def _two_site(n: int, a: int, b: int, code: int) -> np.ndarray:
    w = np.zeros(n, dtype=np.uint8)
    w[a] = code
    w[b] = code
    return w


def xxz_terms(n: int, bonds: Sequence[Tuple[int, int]], jxy: Sequence[float], jz: Sequence[float],
              field: Sequence[float]) -> Terms:
    """sum_b jxy_b (X X + Y Y) + jz_b Z Z over bonds, plus sum_i field_i Z_i."""
    rows, coefs = [], []
    for (a, b), j, z in zip(bonds, jxy, jz):
        for code, c in ((1, j), (2, j), (3, z)):
            rows.append(_two_site(n, a, b, code))
            coefs.append(c)
    for i, h in enumerate(field):
        w = np.zeros(n, dtype=np.uint8)
        w[i] = 3
        rows.append(w)
        coefs.append(h)
    return np.asarray(rows, dtype=np.uint8).reshape(-1, n), np.asarray(coefs, dtype=np.complex128)


def heisenberg_4x7_with_random_terms(seed: int = BASE_SEED, random_terms: int = 1885) -> Terms:
    """28-qubit 4 x 7 open XXZ lattice (Jz = 1, Jxy ~ U[0.5, 1.5] per bond), field
    h_i ~ U[-1, 1] on Z_i, plus `random_terms` random weight-1..4 strings with
    N(0, 0.1) coefficients.  Draw order: 45 Jxy, 28 h, then per random term:
    weight in 1..4, distinct sites, letters in 1..3, coefficient.  With the
    default 1,885 random terms the raw term count is 2,048."""
    rng = np.random.default_rng(seed)
    n, rows, cols = 28, 4, 7
    bonds = open_lattice_bonds(rows, cols)
    jxy = rng.uniform(0.5, 1.5, size=len(bonds))
    h = rng.uniform(-1.0, 1.0, size=n)
    codes, coef = xxz_terms(n, bonds, jxy, np.ones(len(bonds)), h)
    extra = np.zeros((random_terms, n), dtype=np.uint8)
    extra_coef = np.zeros(random_terms, dtype=np.complex128)
    for t in range(random_terms):
        weight = int(rng.integers(1, 5))
        sites = rng.choice(n, size=weight, replace=False)
        letters = rng.integers(1, 4, size=weight)
        extra[t, sites] = letters
        extra_coef[t] = rng.normal(0.0, 0.1)
    return np.concatenate([codes, extra]), np.concatenate([coef, extra_coef])


def heisenberg_periodic(rows: int, cols: int, seed: int, jz_range=(0.8, 1.2), field_range=(-0.2, 0.2)) -> Terms:
    """Periodic XXZ: Jxy = 1, Jz ~ U[jz_range] per bond, h_i ~ U[field_range] on Z_i.
    Draw order: Jz per bond, then h per site."""
    rng = np.random.default_rng(seed)
    n = rows * cols
    bonds = periodic_lattice_bonds(rows, cols)
    jz = rng.uniform(*jz_range, size=len(bonds))
    h = rng.uniform(*field_range, size=n)
    return xxz_terms(n, bonds, np.ones(len(bonds)), jz, h)


# Single-qubit Pauli products: (a, b) -> (code, phase) with a * b = phase * code.
_PRODUCT = {
    (0, 0): (0, 1), (0, 1): (1, 1), (0, 2): (2, 1), (0, 3): (3, 1),
    (1, 0): (1, 1), (1, 1): (0, 1), (1, 2): (3, 1j), (1, 3): (2, -1j),
    (2, 0): (2, 1), (2, 1): (3, -1j), (2, 2): (0, 1), (2, 3): (1, 1j),
    (3, 0): (3, 1), (3, 1): (2, 1j), (3, 2): (1, -1j), (3, 3): (0, 1),
}

PauliDict = Dict[Tuple[int, ...], complex]


def _multiply(a: PauliDict, b: PauliDict) -> PauliDict:
    out: PauliDict = {}
    for wa, ca in a.items():
        for wb, cb in b.items():
            phase = 1 + 0j
            word = []
            for x, y in zip(wa, wb):
                code, ph = _PRODUCT[(x, y)]
                word.append(code)
                phase *= ph
            key = tuple(word)
            out[key] = out.get(key, 0) + ca * cb * phase
    return {k: v for k, v in out.items() if v != 0}


def jw_ladder(n: int, mode: int, creation: bool) -> PauliDict:
    """Jordan-Wigner c_j = Z_{<j} (X_j + i Y_j) / 2, c_j^dag = Z_{<j} (X_j - i Y_j) / 2,
    occupied = |1> (the convention of the reference's core/structured/fermion.rs:245-283)."""
    base = [3] * mode + [0] * (n - mode)
    x = list(base)
    x[mode] = 1
    y = list(base)
    y[mode] = 2
    return {tuple(x): 0.5, tuple(y): (-0.5j if creation else 0.5j)}


def fermi_hubbard_jw(rows: int, cols: int, t: float, u: float, seed: int,
                     eps_range=(-0.5, 0.5)) -> Terms:
    """Open rows x cols Fermi-Hubbard, modes up = site, down = sites + site:
    -t sum_<ij>,s (c_is^dag c_js + h.c.) + U sum_i n_i^up n_i^dn + sum_i eps_i (n_i^up + n_i^dn),
    eps_i ~ U[eps_range] (one draw per site), mapped by Jordan-Wigner."""
    rng = np.random.default_rng(seed)
    sites = rows * cols
    n = 2 * sites
    eps = rng.uniform(*eps_range, size=sites)
    total: PauliDict = {}

    def add(d: PauliDict, scale: complex) -> None:
        for k, v in d.items():
            total[k] = total.get(k, 0) + scale * v

    def number(mode: int) -> PauliDict:
        return _multiply(jw_ladder(n, mode, True), jw_ladder(n, mode, False))

    for a, b in open_lattice_bonds(rows, cols):
        for spin in (0, 1):
            i, j = a + spin * sites, b + spin * sites
            add(_multiply(jw_ladder(n, i, True), jw_ladder(n, j, False)), -t)
            add(_multiply(jw_ladder(n, j, True), jw_ladder(n, i, False)), -t)
    for s in range(sites):
        add(_multiply(number(s), number(sites + s)), u)
        add(number(s), eps[s])
        add(number(sites + s), eps[s])
    items = [(k, v) for k, v in sorted(total.items()) if v != 0]
    codes = np.asarray([k for k, _ in items], dtype=np.uint8).reshape(-1, n)
    coef = np.asarray([v for _, v in items], dtype=np.complex128)
    return codes, coef


def gaussian_state(dimension: int, seed: int, dtype=np.complex128) -> np.ndarray:
    """i.i.d. complex Gaussian amplitudes normalised to 1: real parts, then imaginary parts."""
    rng = np.random.default_rng(seed)
    re = rng.standard_normal(dimension)
    im = rng.standard_normal(dimension)
    v = (re + 1j * im).astype(dtype)
    v /= np.linalg.norm(v)
    return v


def batched_random_sums(seed: int = BASE_SEED + 2, count: int = 512, qubits=(10, 22),
                        terms=(64, 4096)) -> List[Terms]:
    """`count` independent random Pauli sums for one batched launch: qubit counts
    uniform in qubits[0]..qubits[1] and raw term counts uniform in
    terms[0]..terms[1] (inclusive), both drawn from ``default_rng(seed)`` in
    that order per sum; sum i is then ``random_pauli_sum(n, t, seed + 9998 + i)``
    (the default seed gives the per-sum seeds BASE_SEED + 10000 + i)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(qubits[0], qubits[1] + 1))
        t = int(rng.integers(terms[0], terms[1] + 1))
        out.append(random_pauli_sum(n, t, seed=seed + 9998 + i))
    return out


def hubbard_ladder_terms(rows: int, cols: int, hopping: float, interaction: float):
    """Raw ladder terms of the spinful Fermi-Hubbard model, up modes 0..S-1 and
    down modes S..2S-1, open bonds, U n_up n_dn on every site and -t on every
    bond and spin in both directions: the FermionOperator input of the
    reference's study (examples/research/fermi_hubbard/run_tencirpauli.py:22-80)."""
    sites = rows * cols
    bonds = []
    for r in range(rows):
        for c in range(cols):
            s = r * cols + c
            if c + 1 < cols:
                bonds.append((s, s + 1))
            if r + 1 < rows:
                bonds.append((s, s + cols))
    terms = []
    for s in range(sites):
        up, dn = s, sites + s
        terms.append((((up, "create"), (up, "annihilate"), (dn, "create"), (dn, "annihilate")), complex(interaction)))
    for left, right in bonds:
        for spin in (0, 1):
            a, b = left + spin * sites, right + spin * sites
            terms.append((((a, "create"), (b, "annihilate")), -hopping))
            terms.append((((b, "create"), (a, "annihilate")), -hopping))
    return terms
