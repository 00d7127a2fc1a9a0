"""Pauli operator algebra on the host (core/operator.rs:217-393, py/pauli.py:514-704).

Pinned by the dense Kronecker oracle: every result's dense matrix equals the
same matrix expression of the operands' dense matrices (Pauli products and
phases are exact, so for dyadic coefficients the match is exact), plus the
reference's canonical-form rules (exact zeros dropped, zero scale -> empty).
"""

from __future__ import annotations

import numpy as np
import pytest

import tencirpauli_amd as tcp
from oracle import dense


def _dense(op):
    return dense.dense_operator(op.nqubits, *op.code_arrays())


def _random(rng, n, count):
    codes = dense.random_codes(rng, n, count, p_identity=0.4)
    coef = (np.round(rng.normal(size=count) * 64) + 1j * np.round(rng.normal(size=count) * 64)) / 64
    return tcp.PauliOperator.from_code_arrays(codes, coef)


@pytest.mark.parametrize("seed", range(5))
def test_sum_product_commutators_and_adjoint_match_dense_matrices(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 5))
    a, b = _random(rng, n, 7), _random(rng, n, 5)
    A, B = _dense(a), _dense(b)
    np.testing.assert_array_equal(_dense(a + b), A + B)
    np.testing.assert_array_equal(_dense(a - b), A - B)
    np.testing.assert_array_equal(_dense(a * b), A @ B)
    np.testing.assert_array_equal(_dense(a.commutator(b)), A @ B - B @ A)
    np.testing.assert_array_equal(_dense(a.anticommutator(b)), A @ B + B @ A)
    np.testing.assert_array_equal(_dense(a.adjoint()), A.conj().T)
    np.testing.assert_array_equal(_dense(0.5j * a), 0.5j * A)
    np.testing.assert_array_equal(_dense(-a), -A)


def test_canonical_form_rules():
    hop = tcp.PauliOperator.from_strings([("XX", 0.5), ("YY", 0.5)])
    zz = tcp.PauliOperator.from_strings([("ZZ", 1.0)])
    assert hop.commutator(zz).term_count == 0  # XX + YY commutes with ZZ: every word cancels exactly
    assert (hop * 0).term_count == 0 and hop.scale(0.0).nqubits == 2
    assert (hop - hop).term_count == 0
    assert hop.to_dict() == {"XX": 0.5, "YY": 0.5}
    assert hop.is_hermitian() and not (1j * hop).is_hermitian() and (1j * hop).is_hermitian(tolerance=1.0)
    assert hop == tcp.PauliOperator.from_terms(2, [("YY", 0.5), ("XX", 0.5)])
    with pytest.raises(ValueError, match="incompatible"):
        hop + tcp.PauliOperator.from_strings([("ZZZ", 1.0)])
    with pytest.raises(ValueError, match="finite"):
        hop.scale(float("inf"))
    with pytest.raises(ValueError, match="same length"):
        tcp.PauliOperator.from_strings([("XX", 1.0), ("X", 1.0)])
    with pytest.raises(ValueError, match="tolerance"):
        hop.is_hermitian(-1.0)
