"""Fermion operators on the host: CAR canonicalisation, the Jordan-Wigner map
and factorized fermion charge sectors.

Pinned by the occupation-basis dense oracle (``oracle/fermion.py``, the
definition of the reference's tests/test_structured_algebra.py:256-277): the
Jordan-Wigner image's dense Kronecker matrix must equal the matrix of the raw
ladder products, which checks canonicalisation (signs, contractions, dropped
nilpotent words) and the map together.  Known answers of the canonical form
follow py/_structured/words.py:18-68 (creations increasing, annihilations
decreasing).
"""

from __future__ import annotations

import numpy as np
import pytest

import tencirpauli_amd as tcp
from tencirpauli_amd import lattice
from oracle import dense
from oracle import fermion as foracle


def _jw_dense(op: tcp.FermionOperator) -> np.ndarray:
    codes, coef = op.map_fermions().code_arrays()
    return dense.dense_operator(op.n_modes, codes, coef)


def test_canonical_words_known_answers():
    # a a+ = 1 - a+ a
    op = tcp.FermionOperator.from_terms(1, [(((0, "annihilate"), (0, "create")), 1.0)])
    assert [(t.word.factors, t.coefficient) for t in op.terms] == [((), 1.0), (((0, "create"), (0, "annihilate")), -1.0)]
    # n_0 n_2 = a+_0 a+_2 a_2 a_0
    op = tcp.FermionOperator.from_terms(4, [(((0, "create"), (0, "annihilate"), (2, "create"), (2, "annihilate")), 1.0)])
    assert op.terms[0].word.creation_modes == (0, 2) and op.terms[0].word.annihilation_modes == (2, 0)
    assert op.terms[0].coefficient == 1.0
    # {a_0, a_1} = 0 and a+ a+ = 0 leave nothing; the number operator is one word
    assert tcp.FermionOperator.from_terms(2, [(((1, "annihilate"), (0, "annihilate")), 1.0),
                                              (((0, "annihilate"), (1, "annihilate")), 1.0)]).term_count == 0
    assert tcp.FermionOperator.from_terms(2, [(((1, "create"), (1, "create")), 1.0)]).term_count == 0
    assert tcp.FermionOperator.from_terms(1, [(((0, "create"), (0, "annihilate")), 1.0)]).term_count == 1
    # a+_1 a+_0 = - a+_0 a+_1; exact zeros are dropped, equal words summed
    op = tcp.FermionOperator.from_terms(2, [(((1, "create"), (0, "create")), 2.0),
                                            (((0, "create"), (1, "create")), 0.5),
                                            (((1, "create"), (1, "annihilate")), 0.0)])
    assert op.to_dict() == {((0, "create"), (1, "create")): -1.5}


def test_terms_are_sorted_by_creation_then_annihilation_modes():
    op = tcp.FermionOperator.from_terms(3, [(((2, "create"), (0, "annihilate")), 1.0),
                                            (((0, "create"), (2, "annihilate")), 2.0),
                                            (((1, "create"), (1, "annihilate")), 3.0),
                                            ((), 4.0)])
    keys = [(t.word.creation_modes, t.word.annihilation_modes) for t in op.terms]
    assert keys == sorted(keys) and keys[0] == ((), ())


@pytest.mark.parametrize("seed", range(6))
def test_jordan_wigner_image_matches_the_occupation_basis_oracle(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 6))
    terms = []
    for _ in range(8):
        length = int(rng.integers(0, 5))
        factors = tuple((int(rng.integers(0, n)), "create" if rng.random() < 0.5 else "annihilate")
                        for _ in range(length))
        terms.append((factors, complex(rng.normal(), rng.normal())))
    op = tcp.FermionOperator.from_terms(n, terms)
    np.testing.assert_allclose(_jw_dense(op), foracle.operator_matrix(n, terms), atol=1e-12)


def test_car_identities_across_modes():
    # tests/test_structured_algebra.py:483-510 of the reference
    n = 3
    a = [tcp.FermionOperator.from_terms(n, [(((m, "annihilate"),), 1.0)]) for m in range(n)]
    c = [tcp.FermionOperator.from_terms(n, [(((m, "create"),), 1.0)]) for m in range(n)]
    for i in range(n):
        for j in range(n):
            assert (a[i] * a[j] + a[j] * a[i]).term_count == 0
            assert (c[i] * c[j] + c[j] * c[i]).term_count == 0
            anti = a[i] * c[j] + c[j] * a[i]
            if i == j:
                assert anti.to_dict() == {(): 1.0}
            else:
                assert anti.term_count == 0
    np.testing.assert_allclose(_jw_dense(a[0]), foracle.ladder_matrix(n, 0, "annihilate"))
    np.testing.assert_allclose(_jw_dense(c[2]), foracle.ladder_matrix(n, 2, "create"))


def test_hubbard_quartic_terms_and_adjoint():
    # tests/test_structured_algebra.py:513-530: two-site spinful Hubbard
    terms = lattice.hubbard_ladder_terms(1, 2, 0.7, 1.25)
    op = tcp.FermionOperator.from_terms(4, terms)
    assert max(len(t.word.factors) for t in op.terms) == 4
    expected = foracle.operator_matrix(4, terms)
    np.testing.assert_allclose(_jw_dense(op), expected, atol=1e-12)
    assert op.adjoint() == op  # Hermitian
    np.testing.assert_allclose(_jw_dense((2.0 * op) - op), expected, atol=1e-12)


def test_validation_errors():
    with pytest.raises(ValueError, match="action"):
        tcp.FermionOperator.from_terms(2, [(((0, "destroy"),), 1.0)])
    with pytest.raises(ValueError, match="outside"):
        tcp.FermionOperator.from_terms(2, [(((2, "create"),), 1.0)])
    with pytest.raises(ValueError, match="finite"):
        tcp.FermionOperator.from_terms(2, [(((0, "create"),), float("nan"))])
    with pytest.raises(ValueError, match="pairs"):
        tcp.FermionOperator.from_terms(2, [((0, "create"), 1.0)])
    with pytest.raises(ValueError, match="strictly"):
        tcp.FermionWord(3, (2, 1), ())
    with pytest.raises(NotImplementedError):
        tcp.FermionOperator.from_terms(2, []).map_fermions("parity")
    op = tcp.FermionOperator.from_terms(2, [(((0, "create"), (1, "annihilate")), 1.0)])
    with pytest.raises(ValueError, match="incompatible"):
        op + tcp.FermionOperator.from_terms(3, [])


def test_spinful_charge_sector_factorizes_into_two_blocks():
    # examples/research/fermi_hubbard/run_tencirpauli.py:82-98: total number and 2 S_z
    sites = 16
    space = tcp.OperatorSpace(fermions=2 * sites)
    total = tcp.AdditiveCharge(space, name="particle_number", fermions={m: 1 for m in range(2 * sites)})
    spin = tcp.AdditiveCharge(space, name="two_Sz", fermions={m: (1 if m < sites else -1) for m in range(2 * sites)})
    sector = tcp.ChargeSector(((total, sites), (spin, 0)), max_bytes=16 * 1024 ** 3)
    assert sector.blocks == ((16, 8), (16, 8)) and sector.dimension == 165_636_900
    assert sector.space == space
    # one marked mode pins both blocks; a single total charge is one block
    space4 = tcp.OperatorSpace(fermions=4)
    number = tcp.AdditiveCharge(space4, fermions={m: 1 for m in range(4)})
    marked = tcp.AdditiveCharge(space4, fermions={0: 1})
    assert tcp.ChargeSector(((number, 1), (marked, 1))).blocks == ((1, 1), (3, 0))
    assert tcp.ChargeSector(((number, 2), (marked, 0))).blocks == ((1, 0), (3, 2))
    assert number.sector(2).blocks == ((4, 2),)
    # an unreachable target is an empty sector (no error), as in the reference
    assert tcp.ChargeSector(((number, 1), (marked, 2))).dimension == 0
    # the rank order is the ascending occupation integer (mode 0 most significant)
    small = tcp.ChargeSector(((number, 2), (marked, 0)))
    expected = dense.blocks_basis([(1, 0), (3, 2)])
    got = [int("".join(map(str, small.unrank(i))), 2) for i in range(small.dimension)]
    assert got == list(expected)
    # a weighted qubit charge on an OperatorSpace of qubits
    qspace = tcp.OperatorSpace(qubits=2)
    qcharge = tcp.AdditiveCharge(qspace, qubits={0: (0, 1), 1: (0, 1)})
    assert qcharge.sector(1).blocks == ((2, 1),)
    with pytest.raises(NotImplementedError):
        tcp.AdditiveCharge(tcp.OperatorSpace(bosons=1), bosons={0: 1})


def test_pauli_operator_rejects_a_fermion_sector():
    space = tcp.OperatorSpace(fermions=2)
    number = tcp.AdditiveCharge(space, fermions={0: 1, 1: 1})
    with pytest.raises(ValueError, match="incompatible"):
        tcp.PauliOperator.from_terms(2, [("ZZ", 1.0)]).restrict_charge(number.sector(1))


def test_occupation_oracle_is_pinned_by_the_car_and_known_matrices():
    # The oracle itself: {a_i, a+_j} = delta_ij, {a_i, a_j} = 0, a+ = a^H, the
    # single-mode a = |0><1|, and the sign of a_1 on |11> (mode 0 occupied below).
    n = 3
    eye = np.eye(1 << n)
    a = [foracle.ladder_matrix(n, m, "annihilate") for m in range(n)]
    c = [foracle.ladder_matrix(n, m, "create") for m in range(n)]
    for i in range(n):
        np.testing.assert_array_equal(c[i], a[i].conj().T)
        for j in range(n):
            np.testing.assert_array_equal(a[i] @ c[j] + c[j] @ a[i], eye if i == j else 0 * eye)
            np.testing.assert_array_equal(a[i] @ a[j] + a[j] @ a[i], 0 * eye)
    np.testing.assert_array_equal(foracle.ladder_matrix(1, 0, "annihilate"), [[0, 1], [0, 0]])
    a1 = foracle.ladder_matrix(2, 1, "annihilate")
    assert a1[0b10, 0b11] == -1.0 and a1[0b00, 0b01] == 1.0
    # products apply right to left: a+_0 a_0 is the number operator of mode 0
    np.testing.assert_array_equal(foracle.product_matrix(2, ((0, "create"), (0, "annihilate"))),
                                  np.diag([0, 0, 1, 1]))


def test_long_products_without_contractions_take_the_sign_of_their_sort():
    # 200 annihilations in increasing order reverse into canonical order:
    # 19,900 transpositions, an even permutation
    n = 200
    op = tcp.FermionOperator.from_terms(n, [(tuple((m, "annihilate") for m in range(n)), 1.0)])
    assert op.term_count == 1 and op.terms[0].coefficient == 1.0
    assert op.terms[0].word.annihilation_modes == tuple(range(n - 1, -1, -1))
    # a_0 a_1 a+_0 = -a_1 - a+_0 a_1 a_0 needs the contraction path
    op = tcp.FermionOperator.from_terms(2, [(((0, "annihilate"), (1, "annihilate"), (0, "create")), 1.0)])
    assert op.to_dict() == {((1, "annihilate"),): -1.0, ((0, "create"), (1, "annihilate"), (0, "annihilate")): -1.0}
    np.testing.assert_allclose(_jw_dense(op), foracle.product_matrix(2, ((0, "annihilate"), (1, "annihilate"),
                                                                         (0, "create"))), atol=1e-15)
