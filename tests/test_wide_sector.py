"""Wide U(1) sectors (64..1024 qubits): host rank/unrank of packed words.

Pinned by the reference's known answers (core/tests.rs:441-449: 129 qubits,
k = 2, indices crossing three limbs), brute-force enumeration of the basis as
sorted Python integers (the public order of tests/test_u1.py:84-117), and the
closed form of the combinatorial number system (oracle/dense.py comb_rank).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import tencirpauli_amd as tcp
from oracle import dense, wide


def test_packed_rank_crosses_three_limb_boundaries():
    # core/tests.rs:441-449
    sector = tcp.U1Sector(129, 2)
    for index in (0, 1, 63, 64, 127, 128, 8255):
        words = sector.unrank_words(index)
        assert words.shape == (3,)
        assert sector.rank_words(words) == index
        assert int(words[2]) & ~1 == 0


@pytest.mark.parametrize("n,k", [(63, 1), (64, 2), (65, 1), (127, 2), (128, 1), (129, 2), (256, 2)])
def test_wide_basis_order_and_packing(n, k):
    sector = tcp.U1Sector(n, k)
    expected = wide.basis_values(n, k)
    assert sector.dimension == len(expected)
    assert sector.rank(expected[0]) == 0
    assert sector.rank(expected[-1]) == len(expected) - 1
    assert sector.unrank(0) == tuple((expected[0] >> (n - 1 - q)) & 1 for q in range(n))
    basis = sector.basis_words_packed()
    assert basis.flags.writeable is False
    if n <= 64:
        assert basis.shape == (len(expected), 1)
        np.testing.assert_array_equal(basis[:, 0], np.asarray(expected, dtype=np.uint64))
    else:
        assert basis.shape == (len(expected), (n + 63) // 64)
        packed = np.asarray([wide.packed_words(v, n) for v in expected], dtype=np.uint64)
        np.testing.assert_array_equal(basis, packed)
        if n % 64:
            assert np.all(basis[:, -1] < np.uint64(1 << (n % 64)))


@pytest.mark.parametrize("n,k", [(70, 2), (65, 63), (66, 64), (80, 78), (130, 1), (72, 3)])
def test_wide_round_trip_and_order_exhaustive(n, k):
    # includes the complement path (k > n / 2): ranks still ascend with the integer value
    sector = tcp.U1Sector(n, k)
    expected = wide.basis_values(n, k)
    basis = sector.basis_words_packed()
    np.testing.assert_array_equal(basis, np.asarray([wide.packed_words(v, n) for v in expected], dtype=np.uint64))
    for i in (0, 1, len(expected) // 3, len(expected) - 2, len(expected) - 1):
        assert sector.rank_words(basis[i]) == i
        assert sector.rank(expected[i]) == i


@pytest.mark.parametrize("n,k", [(300, 4), (1024, 3), (200, 197), (513, 2)])
def test_wide_rank_matches_the_closed_form(n, k):
    sector = tcp.U1Sector(n, k)
    rng = np.random.default_rng(n + k)
    dim = math.comb(n, k)
    picks = {0, dim - 1} | {int(x) for x in rng.integers(0, dim, size=40)}
    for index in sorted(picks):
        value = sector.unrank_int(index)
        assert bin(value).count("1") == k and value < (1 << n)
        assert dense.comb_rank(value, k) == index
        assert sector.rank(value) == index


def test_wide_rank_errors():
    sector = tcp.U1Sector(100, 2)
    with pytest.raises(ValueError, match="Hamming weight"):
        sector.rank((1 << 99) | (1 << 5) | 1)
    words = sector.unrank_words(7)
    words[1] |= np.uint64(1 << 40)  # qubit 104 >= 100: a padding bit
    with pytest.raises(ValueError):
        sector.rank_words(words)
    with pytest.raises(IndexError):
        sector.unrank(sector.dimension)
    with pytest.raises(tcp.UnsupportedOnDeviceError):
        tcp.U1Sector(1100, 1).rank(1)


def test_wide_oracle_sparse_matches_dense_on_a_narrow_case():
    # the big-integer oracle agrees with the dense Kronecker oracle where both run
    from tests.helpers import conserving_u1_terms

    rng = np.random.default_rng(3)
    codes, coef = conserving_u1_terms(rng, 9, 14)
    H = dense.restricted_dense(9, codes, coef, dense.u1_basis(9, 4))
    got = np.zeros_like(H)
    for (r, c), v in wide.restricted_sparse(9, 4, codes, coef).items():
        got[r, c] = v
    np.testing.assert_allclose(got, H, rtol=0, atol=1e-13)


def test_restrict_u1_is_a_deprecated_alias_with_checks():
    op = tcp.PauliOperator.from_terms(3, [("XXI", 1.0), ("YYI", 1.0)])
    with pytest.warns(DeprecationWarning, match="restrict_charge"):
        with pytest.raises(TypeError, match="expected U1Sector"):
            op.restrict_u1("not a sector")
    with pytest.warns(DeprecationWarning):
        with pytest.raises(ValueError, match="expected sector for 3 qubits, got 4"):
            op.restrict_u1(tcp.U1Sector(4, 1))
