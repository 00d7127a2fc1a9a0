"""GPU checks of the compact and slab CSR forms against the Kronecker-product
oracle.

An eager plan whose matrix has fewer than 2^24 rows and at most 255 distinct
off-diagonal values stores each entry as a 24-bit column plus an 8-bit index
into a value table, with the diagonal in a dense array; otherwise it keeps a
value per entry.  A restricted plan whose x-groups (distinct off-diagonal
x-masks holding entries) times rows are at most 8/3 of its off-diagonal
entries regroups the compact entries by x-mask (the slab form).  Every form
must give the oracle's product (FP64 within 1e-12 relative, complex64 within
FP32 rounding) and export the same CSR.  The cases pin the table boundary
(255 distinct values compact, 256 plain), rows without a diagonal entry,
complex table values, the slab choice computed independently from the
exported CSR and the sector basis, and sectors on either side of it.
"""

from __future__ import annotations

import numpy as np
import pytest

import tencirpauli_amd as tcp
from oracle import dense
from tencirpauli_amd import lattice
from tests.helpers import interleave_spins, rel_err, spin_levels

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-6


def _x_word(n, mask):
    return tuple(1 if (mask >> (n - 1 - q)) & 1 else 0 for q in range(n))


def _distinct_value_operator(n, distinct, phase=1.0):
    """`distinct` X-only words with coefficients phase * (k + 1): every row
    holds one entry per word with exactly that value, so the matrix has
    `distinct` distinct off-diagonal values; plus a diagonal that vanishes on
    half of the rows (0.5 Z0 Z1 + 0.5 Z0)."""
    terms = [(_x_word(n, m), phase * (k + 1)) for k, m in enumerate(range(1, distinct + 1))]
    terms.append(((3, 3) + (0,) * (n - 2), 0.5))
    terms.append(((3,) + (0,) * (n - 1), 0.5))
    return tcp.PauliOperator.from_terms(n, terms)


def _check(plan, op_dense, seed):
    import torch

    rng = np.random.default_rng(seed)
    state = rng.normal(size=plan.dimension) + 1j * rng.normal(size=plan.dimension)
    expected = op_dense @ state
    assert rel_err(plan.apply(state), expected) <= FP64_TOL
    got32 = torch.empty(plan.dimension, dtype=torch.complex64, device="cuda")
    plan.apply_into(torch.from_numpy(state.astype(np.complex64)).cuda(), got32)
    assert rel_err(got32.cpu().numpy(), expected) <= FP32_TOL


@pytest.mark.parametrize("distinct,form", [(1, "compact"), (255, "compact"), (256, "plain")])
def test_table_boundary_and_rows_without_diagonal(distinct, form):
    n = 9
    op = _distinct_value_operator(n, distinct, phase=(0.6 - 0.8j) if distinct == 255 else 1.0)
    plan = op.csr_mvp_plan()
    assert plan.storage_format == form
    full = dense.dense_operator(n, *op.code_arrays())
    _check(plan, full, distinct)
    # the exported CSR is the same in either form
    csr = op.csr()
    np.testing.assert_allclose(csr.to_scipy().toarray(), full, atol=1e-13)


def _expected_form(plan, basis):
    """The form the selection rule gives, from the exported CSR and the
    sector basis (integers in rank order)."""
    indptr, cols, vals = plan.csr_arrays()
    rows = np.repeat(np.arange(plan.dimension), np.diff(indptr).astype(np.int64))
    x = np.asarray(basis, dtype=np.int64)[rows] ^ np.asarray(basis, dtype=np.int64)[cols.astype(np.int64)]
    off = x != 0
    if len(np.unique(vals[off])) > 255 or plan.dimension > (1 << 24) or len(cols) == 0:
        return "plain"
    slots = len(np.unique(x[off]))
    return "slab" if 0 < slots and 3 * slots * plan.dimension <= 8 * int(off.sum()) else "compact"


def test_u1_lattice_sector_is_slab_and_matches_lazy_and_oracle():
    codes, coef = lattice.heisenberg_periodic(3, 4, seed=5)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    sector = tcp.U1Sector(12, 6)
    plan = op.restrict_charge(sector).mvp_plan(storage="eager")
    basis = dense.blocks_basis([(12, 6)])
    assert plan.storage_format == _expected_form(plan, basis) == "slab"
    basis = np.asarray(dense.blocks_basis([(12, 6)]), dtype=np.int64)
    full = dense.dense_operator(12, codes, coef)[np.ix_(basis, basis)]
    _check(plan, full, 11)
    state = np.random.default_rng(2).normal(size=plan.dimension) + 0j
    assert rel_err(plan.apply(state), op.restrict_charge(sector).apply(state)) <= FP64_TOL


def test_random_coefficients_keep_plain_csr():
    rng = np.random.default_rng(9)
    n = 10
    codes = dense.random_codes(rng, n, 600, p_identity=0.3)
    coef = rng.normal(size=600) + 1j * rng.normal(size=600)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    plan = op.csr_mvp_plan()
    assert plan.storage_format == "plain"
    _check(plan, dense.dense_operator(n, codes, coef), 4)


def test_degenerate_matrices():
    # diagonal only (no off-diagonal values: an empty table), a zero operator
    # (no entries: plain CSR) and a one-state sector
    n = 6
    diag_only = tcp.PauliOperator.from_terms(n, [((3,) + (0,) * (n - 1), 0.5), ((0, 3) + (0,) * (n - 2), -1.25)])
    plan = diag_only.csr_mvp_plan()
    assert plan.storage_format == "compact"
    _check(plan, dense.dense_operator(n, *diag_only.code_arrays()), 1)
    zero = tcp.PauliOperator.from_terms(n, [((1,) + (0,) * (n - 1), 0.0)])
    plan = zero.csr_mvp_plan()
    assert plan.transition_count == 0 and plan.storage_format == "plain"
    state = np.arange(1 << n, dtype=np.complex128)
    np.testing.assert_array_equal(plan.apply(state), np.zeros(1 << n))
    hop = tcp.PauliOperator.from_terms(2, [("XX", 0.5), ("YY", 0.5), ("ZZ", 0.25)])
    one = hop.restrict_charge(tcp.U1Sector(2, 2)).mvp_plan(storage="eager")
    assert one.dimension == 1
    np.testing.assert_allclose(one.apply(np.array([2.0 + 1.0j])), [0.25 * (2.0 + 1.0j)], rtol=1e-15)


def test_values_that_differ_only_in_the_sign_of_an_imaginary_part():
    # A Hermitian complex operator holds v and conj(v) as separate table
    # values: X0 X1 + 2 X0 Y1 maps |00> -> |11> with 1 + 2i and |01> -> |10>
    # with 1 - 2i.  Both must keep their own table entry (a hash or equality
    # test that reads only the real part merges them).
    n = 4
    op = tcp.PauliOperator.from_terms(n, [("XXII", 1.0), ("XYII", 2.0), ("IIZI", 0.3)])
    plan = op.csr_mvp_plan()
    assert plan.storage_format == "compact"
    _check(plan, dense.dense_operator(n, *op.code_arrays()), 3)
    # a complex-hopping U(1) lattice sector: every hop and its reverse are conjugates
    codes, coef = lattice.heisenberg_periodic(3, 3, seed=5)
    rows, vals = [], []
    for (a, b) in lattice.periodic_lattice_bonds(3, 3):
        for pa, pb, c in ((1, 1, 0.8), (2, 2, 0.8), (1, 2, 0.6), (2, 1, -0.6)):
            w = np.zeros(9, dtype=np.uint8)
            w[a], w[b] = pa, pb
            rows.append(w)
            vals.append(c)
    rows = np.concatenate([np.asarray(rows), codes])
    vals = np.concatenate([np.asarray(vals, dtype=np.complex128), coef])
    op = tcp.PauliOperator.from_code_arrays(rows, vals)
    sector = tcp.U1Sector(9, 4)
    plan = op.restrict_charge(sector).mvp_plan(storage="eager")
    basis = np.asarray(dense.blocks_basis([(9, 4)]), dtype=np.int64)
    assert plan.storage_format == _expected_form(plan, basis) == "slab"
    full = dense.dense_operator(9, rows, vals)[np.ix_(basis, basis)]
    assert np.any(np.abs(full.imag) > 0.1)
    _check(plan, full, 12)


def test_sparse_groups_stay_compact():
    # all-to-all hopping of one particle on 10 sites: 45 x-groups, each row
    # holds 9 transitions, so the slab would be 45 / 9 = 5 words per entry
    n = 10
    terms = []
    for a in range(n):
        for b in range(a + 1, n):
            for code in (1, 2):
                w = [0] * n
                w[a] = w[b] = code
                terms.append((tuple(w), 0.5))
    terms.append(((3,) + (0,) * (n - 1), 0.2))
    op = tcp.PauliOperator.from_terms(n, terms)
    sector = tcp.U1Sector(n, 1)
    plan = op.restrict_charge(sector).mvp_plan(storage="eager")
    basis = np.asarray(dense.blocks_basis([(n, 1)]), dtype=np.int64)
    assert plan.storage_format == _expected_form(plan, basis) == "compact"
    full = dense.dense_operator(n, *op.code_arrays())[np.ix_(basis, basis)]
    _check(plan, full, 21)


@pytest.mark.parametrize("filling", [(1, 1), (2, 2), (3, 2)])
def test_block_and_general_sector_forms_match_oracle(filling):
    # a 2x3 Fermi-Hubbard (Jordan-Wigner): N_up, N_down on spin-major modes
    # (block sector), then the same operator on interleaved spin orbitals with
    # N and 2 S_z (general sector); each eager form against the dense oracle
    codes, coef = lattice.fermi_hubbard_jw(2, 3, 1.0, 4.0, seed=7)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    blocks = [(6, filling[0]), (6, filling[1])]
    sector = tcp.ChargeSector(((tcp.AdditiveCharge.number(12, range(6)), filling[0]),
                               (tcp.AdditiveCharge.number(12, range(6, 12)), filling[1])))
    plan = op.restrict_charge(sector, storage="eager").mvp_plan(storage="eager")
    basis = np.asarray(dense.blocks_basis(blocks), dtype=np.int64)
    assert plan.storage_format == _expected_form(plan, basis)
    _check(plan, dense.dense_operator(12, codes, coef)[np.ix_(basis, basis)], 5)

    icodes = interleave_spins(codes)
    iop = tcp.PauliOperator.from_code_arrays(icodes, coef)
    levels = spin_levels(12)
    targets = (filling[0] + filling[1], filling[0] - filling[1])
    charges = [(tcp.AdditiveCharge(12, {a: (int(levels[j, a, 0]), int(levels[j, a, 1])) for a in range(12)}),
                targets[j]) for j in range(2)]
    iplan = iop.restrict_charge(tcp.ChargeSector(charges), storage="eager").mvp_plan(storage="eager")
    ibasis = np.asarray(dense.charge_basis(levels, targets), dtype=np.int64)
    assert iplan.storage_format == _expected_form(iplan, ibasis)
    _check(iplan, dense.dense_operator(12, icodes, coef)[np.ix_(ibasis, ibasis)], 6)
