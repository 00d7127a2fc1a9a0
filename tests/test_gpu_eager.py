"""GPU checks of the eager restricted operators (device CSR).

Structure (indptr, column order) is compared bit-exactly with the CSR of the
dense restricted-matrix oracle (rows = destination ranks, columns ascending,
exact zeros dropped: the layout of core/sector.rs:345-432); values and applies
within 1e-12 relative (FP64) / 2e-6 (FP32 storage) of the oracle.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
import tencirpauli_amd as tcp
from oracle import dense
from tencirpauli_amd import lattice
from tests.helpers import conserving_u1_terms, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-6


def _oracle_csr(n, k, op):
    from scipy.sparse import csr_matrix

    basis = dense.u1_basis(n, k)
    matrix = csr_matrix(dense.restricted_dense(n, *op.code_arrays(), basis))
    matrix.sort_indices()
    return matrix


def test_hopping_known_answer_dense_and_apply():
    # core/tests.rs:384-418: restricted XX+YY on (3, 1)
    op = tcp.PauliOperator.from_terms(3, [("XXI", 1.0), ("YYI", 1.0)])
    restricted = op.restrict_charge(tcp.U1Sector(3, 1), storage="eager")
    assert restricted.storage == "eager" and restricted.transition_count == 2
    flat = restricted.dense().ravel()
    assert flat[5] == 2.0 and flat[7] == 2.0 and np.count_nonzero(flat) == 2
    np.testing.assert_array_equal(restricted.apply(np.asarray([3.0, 5.0, 7.0])), [0.0, 14.0, 10.0])
    plan = restricted.mvp_plan(storage="eager")
    assert plan.retained_format == "csr" and plan.transition_count == 2


def test_csr_structure_bit_exact_and_values_match_oracle():
    rng = np.random.default_rng(20260802)
    for n in range(2, 11):
        codes, coef = conserving_u1_terms(rng, n, 16)
        # Dyadic coefficients keep every x-group sum exact: XX+YY and XY-YX on
        # one pair share a group, and with general complex weights their sum on
        # a leaving source rounds to ~1e-17 instead of zero, which the exact
        # leakage check (as core/sector.rs:1051-1127) reports.
        coef = np.round(coef.real * 256.0) / 256.0
        op = tcp.PauliOperator.from_code_arrays(codes, coef * (1 + 0.25j))
        for k in range(n + 1):
            expected = _oracle_csr(n, k, op)
            got = op.restrict_charge(tcp.U1Sector(n, k), storage="eager").csr()
            np.testing.assert_array_equal(got.indptr.astype(np.int64), expected.indptr)
            np.testing.assert_array_equal(got.indices.astype(np.int64), expected.indices)
            if expected.nnz:
                assert rel_err(got.data, expected.data) <= FP64_TOL, (n, k)
            coo = op.restrict_charge(tcp.U1Sector(n, k), storage="eager").coo()
            np.testing.assert_array_equal(coo.to_scipy().toarray(), got.to_scipy().toarray())


def test_eager_apply_matches_oracle_and_lazy():
    import torch

    for rows, cols, k in ((4, 4, 8), (4, 5, 10), (3, 5, 7)):
        codes, coef = lattice.heisenberg_periodic(rows, cols, seed=lattice.BASE_SEED)
        op = tcp.PauliOperator.from_code_arrays(codes, coef)
        n = rows * cols
        sector = tcp.U1Sector(n, k)
        eager = op.restrict_charge(sector, storage="eager")
        lazy = op.restrict_charge(sector)
        state = lattice.gaussian_state(sector.dimension, 11)
        cc, cf = op.code_arrays()
        expected = oracle.u1_lazy_apply(n, k, cc, cf, state)
        assert rel_err(eager.apply(state), expected) <= FP64_TOL
        v = torch.from_numpy(state).cuda()
        y = torch.empty_like(v)
        eager.apply_into(v, y)
        y_lazy = lazy.apply(v)
        y32 = eager.apply(v.to(torch.complex64))
        torch.cuda.synchronize()
        assert rel_err(y.cpu().numpy(), expected) <= FP64_TOL
        assert rel_err(y.cpu().numpy(), y_lazy.cpu().numpy()) <= FP64_TOL
        assert rel_err(y32.cpu().numpy().astype(np.complex128), expected) <= FP32_TOL
        # every stored entry is nonzero and the count matches the lazy gather's transitions
        assert eager.transition_count == int(np.count_nonzero(eager.csr().data))


def test_eager_accepts_nonconserving_operator_that_never_leaks():
    # XX - YY changes the particle number by two, but on N = 1 of two qubits
    # every aggregated transition is exactly zero: lazy restriction requires
    # conservation up front, the eager build (as core/sector.rs:345-432) only
    # checks the transitions it meets.
    op = tcp.PauliOperator.from_terms(2, [("XX", 1.0), ("YY", -1.0), ("ZI", 0.5)])
    sector = tcp.U1Sector(2, 1)
    with pytest.raises(ValueError, match="conserved"):
        op.restrict_charge(sector)
    eager = op.restrict_charge(sector, storage="eager")
    np.testing.assert_allclose(eager.dense(), np.diag([0.5, -0.5]))


def test_eager_leakage_error_matches_lazy_apply_error():
    op = tcp.PauliOperator.from_terms(3, [("XII", 1.0), ("ZZI", 0.5)])
    with pytest.raises(ValueError, match="U\\(1\\) sector leakage from restricted source index 0: "
                                         "expected particle number 1, got destination weight 2"):
        op.restrict_charge(tcp.U1Sector(3, 1), storage="eager")


def test_block_sector_eager_matches_lazy_and_oracle():
    import torch

    codes, coef = lattice.fermi_hubbard_jw(2, 3, 1.0, 4.0, seed=lattice.BASE_SEED)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    n = op.nqubits
    blocks = [(n // 2, 2), (n // 2, 3)]
    charges, start = [], 0
    for size, pop in blocks:
        charges.append((tcp.AdditiveCharge.number(n, range(start, start + size)), pop))
        start += size
    sector = tcp.ChargeSector(charges)
    eager = op.restrict_charge(sector, storage="eager")
    lazy = op.restrict_charge(sector)
    state = lattice.gaussian_state(sector.dimension, 13)
    cc, cf = op.code_arrays()
    expected = oracle.blocks_generic_apply(blocks, cc, cf, state)
    assert rel_err(eager.apply(state), expected) <= FP64_TOL
    assert rel_err(lazy.apply(state), expected) <= FP64_TOL
    dense_expected = dense.restricted_dense(n, cc, cf, dense.blocks_basis(blocks))
    np.testing.assert_allclose(eager.dense(), dense_expected, rtol=0, atol=1e-14)


def test_materialization_budget_and_errors():
    op = tcp.PauliOperator.from_terms(4, [("XXII", 1.0), ("YYII", 1.0), ("IIZZ", 0.3)])
    eager = op.restrict_charge(tcp.U1Sector(4, 2), storage="eager")
    assert eager.compile("csr").shape == (6, 6)
    assert eager.compile("native_mvp", storage="eager") is eager.mvp_plan(storage="eager")
    with pytest.raises(ValueError):
        eager.compile("graph")
    with pytest.raises(MemoryError):
        big = lattice.heisenberg_periodic(4, 4, seed=1)
        tcp.PauliOperator.from_code_arrays(*big).restrict_charge(tcp.U1Sector(16, 8), storage="eager").dense(
            max_bytes=1 << 30)


def test_plan_metadata_records():
    # the fields and values of the reference's plan facades (py/_native_mvp.py:84-150,
    # py/symmetry.py:541-590 for U(1): execution_family "product", variant "qubit")
    op = tcp.PauliOperator.from_terms(4, [("XXII", 1.0), ("YYII", 1.0), ("IIZZ", 0.3), ("ZIII", -0.2)])
    full = op.native_mvp_plan().metadata
    assert (full.execution_family, full.native_variant, full.retained_format, full.storage) == \
        ("pauli", "pauli", "compiled_actions", "lazy")
    assert full.target == "native_mvp" and full.basis_ordering == "qubit0_msb_matrix"
    assert full.dimension == 16 and full.term_count == 4 and full.transition_count is None
    sector = tcp.U1Sector(4, 2)
    lazy = op.restrict_charge(sector).mvp_plan().metadata
    assert (lazy.execution_family, lazy.native_variant, lazy.retained_format, lazy.storage) == \
        ("product", "qubit", "compiled_actions", "lazy")
    assert lazy.transition_count is None and lazy.schedule_count == lazy.term_count == 4
    eager = op.restrict_charge(sector, storage="eager").mvp_plan(storage="eager").metadata
    assert (eager.execution_family, eager.native_variant, eager.retained_format, eager.storage) == \
        ("product", "qubit", "csr", "eager")
    D, nnz = 6, eager.transition_count
    assert nnz == int(np.count_nonzero(op.restrict_charge(sector, storage="eager").dense()))
    assert eager.estimated_retained_bytes == D * 16 + (D + 1) * 8 + nnz * 32
    assert eager.estimated_construction_peak_bytes == eager.estimated_retained_bytes


def test_materialization_is_cached_read_only_and_leaves_lazy_plans_lazy():
    # tests/test_native_mvp_resources.py:47-88 of the reference, on a Pauli operator
    from concurrent.futures import ThreadPoolExecutor

    op = tcp.PauliOperator.from_terms(6, [("XXIIII", 1.0), ("YYIIII", 1.0), ("IIXXII", 0.5), ("IIYYII", 0.5),
                                          ("ZIIIIZ", 0.2)])
    restricted = op.restrict_charge(tcp.U1Sector(6, 3))
    lazy_plan = restricted.mvp_plan()
    before = restricted.estimated_bytes
    with ThreadPoolExecutor(max_workers=4) as pool:
        matrices = list(pool.map(lambda _: restricted.csr(max_bytes=None), range(4)))
    assert all(m.shape == (20, 20) for m in matrices)
    assert restricted.estimated_bytes > before
    assert lazy_plan.storage == "lazy" and restricted.mvp_plan() is lazy_plan
    eager = restricted.mvp_plan(storage="eager")
    assert restricted.mvp_plan(storage="eager") is eager
    first = restricted.csr()
    assert first.data.flags.writeable is False and first.indices.flags.writeable is False
    assert eager.indptr.flags.writeable is False
    assert eager.columns.flags.writeable is False and eager.coefficients.flags.writeable is False
    np.testing.assert_array_equal(eager.indptr, first.indptr)
    np.testing.assert_array_equal(eager.coefficients, first.data)
