"""GPU checks of the full-space materialisations and the schema-v1 backend plan.

``PauliOperator.dense/coo/csr/compile`` (py/pauli.py:866-928, 1047-1071;
core/hamiltonian.rs:313-460) are built from the device CSR.  The CSR structure
(indptr, ascending columns, exact zeros dropped) is compared bit-exactly with
the CSR of the Kronecker-product oracle; values within 1e-12 relative.  The
cases follow the reference's tests/test_hamiltonian.py:16-189.
"""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import tencirpauli_amd as tcp
from oracle import dense
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def make_operator(nqubits: int) -> tcp.PauliOperator:
    # tests/test_hamiltonian.py:16-25 of the reference
    structures = [tuple((index + qubit) % 4 for qubit in range(nqubits)) for index in range(min(10, 4 ** nqubits))]
    terms = tuple((structure, complex(index + 1, -index / 3)) for index, structure in enumerate(structures))
    return tcp.PauliOperator.from_terms(nqubits, terms)


def reconstruct_coo(target) -> np.ndarray:
    result = np.zeros(target.shape, dtype=np.complex128)
    result[target.row.astype(np.int64), target.column.astype(np.int64)] = target.data
    return result


def reconstruct_csr(target) -> np.ndarray:
    result = np.zeros(target.shape, dtype=np.complex128)
    for row in range(target.shape[0]):
        start, end = int(target.indptr[row]), int(target.indptr[row + 1])
        result[row, target.indices[start:end].astype(np.int64)] = target.data[start:end]
    return result


def oracle_dense(op: tcp.PauliOperator) -> np.ndarray:
    codes, coef = op.code_arrays()
    return dense.dense_operator(op.nqubits, codes, coef)


@pytest.mark.parametrize("nqubits", (0, 1, 2, 5))
def test_dense_coo_csr_and_mvp_match_the_dense_oracle(nqubits):
    op = make_operator(nqubits)
    expected = oracle_dense(op)
    np.testing.assert_allclose(op.dense(), expected, rtol=FP64_TOL, atol=FP64_TOL)
    np.testing.assert_allclose(reconstruct_coo(op.coo()), expected, rtol=FP64_TOL, atol=FP64_TOL)
    np.testing.assert_allclose(reconstruct_csr(op.csr()), expected, rtol=FP64_TOL, atol=FP64_TOL)
    state = np.arange(1 << nqubits, dtype=np.float64) + 1j * np.arange(1 << nqubits)
    np.testing.assert_allclose(op.mvp(state), expected @ state, rtol=FP64_TOL, atol=FP64_TOL)


@pytest.mark.parametrize("code", (1, 2, 3))
def test_first_and_last_qubit_matrix_ordering(code):
    first = tcp.PauliOperator.from_terms(3, ((tuple([code, 0, 0]), 1.0),)).dense()
    last = tcp.PauliOperator.from_terms(3, ((tuple([0, 0, code]), 1.0),)).dense()
    np.testing.assert_array_equal(first, dense.dense_operator(3, ((code, 0, 0),), (1.0,)))
    np.testing.assert_array_equal(last, dense.dense_operator(3, ((0, 0, code),), (1.0,)))
    assert not np.array_equal(first, last)


def test_csr_structure_bit_exact_against_the_oracle():
    from scipy.sparse import csr_matrix

    rng = np.random.default_rng(20260803)
    cases = [tcp.PauliOperator.from_terms(2, [("XX", 1.0), ("YY", 1.0), ("ZI", 0.5)])]  # exact cancellations
    for n in range(1, 10):
        codes = dense.random_codes(rng, n, 3 * n + 4, p_identity=0.5)
        coef = rng.normal(size=len(codes)) + 1j * rng.normal(size=len(codes))
        cases.append(tcp.PauliOperator.from_code_arrays(codes, coef))
    for op in cases:
        expected = csr_matrix(oracle_dense(op))
        expected.sort_indices()
        got = op.csr()
        assert got.indptr.dtype == np.uint64 and got.indices.dtype == np.uint64 and got.data.dtype == np.complex128
        np.testing.assert_array_equal(got.indptr.astype(np.int64), expected.indptr)
        np.testing.assert_array_equal(got.indices.astype(np.int64), expected.indices)
        if expected.nnz:
            assert rel_err(got.data, expected.data) <= FP64_TOL, op
        coo = op.coo()
        np.testing.assert_array_equal(coo.row, np.repeat(np.arange(got.shape[0], dtype=np.uint64),
                                                         np.diff(got.indptr.astype(np.int64))))
        np.testing.assert_array_equal(coo.column, got.indices)
        np.testing.assert_array_equal(coo.data, got.data)
        assert not coo.data.flags.writeable and not got.indptr.flags.writeable
    # XX + YY cancels on |00> and |11>: only the two flip-flop entries and the Z diagonal remain
    first = cases[0].csr().to_scipy().toarray()
    assert np.count_nonzero(first) == 6 and first[1, 2] == 2.0 and first[0, 3] == 0.0


def test_full_csr_rows_at_16_qubits_are_sorted_grouped_and_apply_like_the_lazy_plan():
    import torch

    rng = np.random.default_rng(7)
    n = 16
    codes = dense.random_codes(rng, n, 120, p_identity=0.75)
    coef = rng.normal(size=len(codes)) + 1j * rng.normal(size=len(codes))
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    plan = op.csr_mvp_plan()
    assert plan.dimension == 1 << n and plan.execution_family == "pauli"
    indptr, indices, values = plan.csr_arrays()
    counts = np.diff(indptr.astype(np.int64))
    groups = op.distinct_x_mask_count()
    assert counts.max() <= groups and plan.transition_count == int(indptr[-1])
    rows = np.repeat(np.arange(1 << n, dtype=np.uint64), counts)
    xs = np.sort(np.unique(rows ^ indices))
    assert len(xs) == groups  # every column is r ^ x for one of the operator's X masks
    step = np.diff(indices.astype(np.int64))
    starts = indptr[1:-1].astype(np.int64)
    inside = np.ones(len(step), dtype=bool)
    inside[starts[(starts > 0) & (starts <= len(step))] - 1] = False
    assert np.all(step[inside] > 0)  # columns strictly ascend inside each row
    assert np.all(values != 0)
    state = torch.from_numpy(rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).cuda()
    y_csr = plan.apply(state)
    y_lazy = op.native_mvp_plan().apply(state)
    torch.cuda.synchronize()
    assert rel_err(y_csr.cpu().numpy(), y_lazy.cpu().numpy()) <= FP64_TOL


def test_compile_target_dispatch_is_explicit():
    # tests/test_hamiltonian.py:176-189 of the reference
    op = tcp.PauliOperator.from_terms(1, (((1,), 2.0),))
    np.testing.assert_array_equal(op.compile("dense"), op.dense())
    assert isinstance(op.compile("backend_mvp"), tcp.BackendMVPPlan)
    native_plan = op.compile("native_mvp")
    np.testing.assert_array_equal(native_plan(np.ones(2, dtype=np.complex128)), op.mvp(np.ones(2)))
    np.testing.assert_array_equal(op.compile("csr").to_scipy().toarray(), [[0, 2], [2, 0]])
    np.testing.assert_array_equal(reconstruct_coo(op.compile("coo")), [[0, 2], [2, 0]])
    assert op.compile("native_mvp", storage="eager").storage == "eager"
    with pytest.raises(ValueError, match="target"):
        op.compile("unknown")


def test_empty_identity_and_allocation_guards():
    np.testing.assert_array_equal(tcp.PauliOperator.empty(0).dense(), np.zeros((1, 1), dtype=np.complex128))
    identity = tcp.PauliOperator.from_terms(0, (((), 1.0),))
    np.testing.assert_array_equal(identity.dense(), np.ones((1, 1), dtype=np.complex128))
    empty = tcp.PauliOperator.empty(3)
    assert empty.csr().indptr.tolist() == [0] * 9 and empty.coo().row.shape == (0,)
    assert make_operator(10).dense(max_bytes=1).shape == (1 << 10, 1 << 10)  # advisory below 1 GiB
    with pytest.raises(OverflowError, match="dimension"):
        tcp.PauliOperator.empty(64).dense()
    big = make_operator(24)  # 10 X groups x 2^24 rows: ~8 GB of CSR working memory
    with pytest.raises(MemoryError):
        big.csr(max_bytes=1 << 30)
    with pytest.raises(MemoryError):
        big.dense()
    with pytest.raises(tcp.UnsupportedOnDeviceError):
        tcp.PauliOperator.from_terms(33, [("Z" + "I" * 32, 1.0)]).csr(max_bytes=None)


def test_backend_plan_versioned_read_only_arrays_and_gpu_executor():
    # tests/test_hamiltonian.py:80-105 of the reference
    import torch

    op = make_operator(3)
    plan = op.backend_mvp_plan()
    assert isinstance(plan, tcp.BackendMVPPlan) and plan.target == "backend_mvp"
    assert plan.schema_version == 1 and plan.ordering == "qubit0_msb_matrix" and plan.estimated_bytes > 0
    assert plan.x_words.dtype == np.uint64 and plan.z_words.dtype == np.uint64
    assert plan.coefficients.dtype == np.complex128
    assert plan.local_dimensions == (2, 2, 2) and plan.required_operations == ("xor", "phase", "scatter_add")
    assert plan.dimension == 8 and plan.term_count == op.term_count == plan.source_term_count
    for array in (plan.x_words, plan.z_words, plan.coefficients):
        assert array.flags.c_contiguous and not array.flags.writeable
    with pytest.raises(ValueError, match="read-only"):
        plan.x_words[0, 0] = 0
    with pytest.raises(ValueError, match="read-only"):
        plan.coefficients[0] = 0
    rng = np.random.default_rng(20260801)
    state = rng.normal(size=8) + 1j * np.random.default_rng(7).normal(size=8)
    expected = oracle_dense(op) @ state
    np.testing.assert_allclose(plan.apply(state), expected, rtol=FP64_TOL, atol=FP64_TOL)
    np.testing.assert_allclose(plan.apply(state, max_bytes=1), expected, rtol=FP64_TOL, atol=FP64_TOL)
    np.testing.assert_allclose(plan(state), expected, rtol=FP64_TOL, atol=FP64_TOL)
    on_device = plan.apply(torch.from_numpy(state).cuda())
    assert on_device.is_cuda and rel_err(on_device.cpu().numpy(), expected) <= FP64_TOL
    with pytest.raises(ValueError, match="shape"):
        plan.apply(np.ones(4))
    back = plan.to_operator()
    np.testing.assert_array_equal(back.code_arrays()[0], op.code_arrays()[0])
    np.testing.assert_array_equal(back.code_arrays()[1], op.code_arrays()[1])
    with pytest.raises(TypeError, match="factory"):
        tcp.BackendMVPPlan(1, 3, 1, plan.x_words, plan.z_words, plan.coefficients)
    with pytest.raises(ValueError, match="Weyl"):
        dataclasses.replace(plan, plan_kind="direct_weyl")
    with pytest.raises(ValueError, match="word_count"):
        dataclasses.replace(plan, word_count=2)
    with pytest.raises(ValueError, match="finite"):
        dataclasses.replace(plan, coefficients=np.full(plan.term_count, np.nan + 0j))


def test_backend_plan_wide_words_run_on_the_gpu():
    # 70 qubits: two words per mask; a U(1)-conserving operator restricted through the plan's operator
    rng = np.random.default_rng(11)
    n = 70
    terms = [("I" * q + "XX" + "I" * (n - q - 2), 0.5) for q in range(0, n - 1, 7)]
    terms += [("I" * q + "YY" + "I" * (n - q - 2), 0.5) for q in range(0, n - 1, 7)]
    terms += [("I" * q + "Z" + "I" * (n - q - 1), float(rng.normal())) for q in range(0, n, 5)]
    op = tcp.PauliOperator.from_terms(n, terms)
    plan = op.backend_mvp_plan()
    assert plan.word_count == 2 and plan.x_words.shape == (op.term_count, 2)
    back = plan.to_operator()
    sector = tcp.U1Sector(n, 2)
    state = rng.normal(size=sector.dimension) + 1j * rng.normal(size=sector.dimension)
    a = back.restrict_charge(sector).apply(state)
    b = op.restrict_charge(sector).apply(state)
    np.testing.assert_array_equal(a, b)
    with pytest.raises(OverflowError):
        plan.apply(np.ones(4))
