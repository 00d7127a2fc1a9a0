"""Pins the CPU oracle against the reference's known answers and independent math.

Each oracle function is checked against (a) the known-answer vectors of the
reference's own tests (tests/golden/kats.json), (b) the dense Kronecker oracle
or a brute-force sector basis, and (c) closed forms -- chosen so a plausible
mistake (wrong qubit order, transposed operand, sign on the wrong index,
dropped Y phase, wrong aggregation order) fails at least one of them.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from oracle import dense
from tests.helpers import (conserving_u1_terms, cplx, interleave_spins, kats, rel_err, seeded, spin_levels,
                           word_codes)


def _terms(spec):
    codes = np.asarray([word_codes(w) for w, _ in spec], dtype=np.uint8)
    coef = np.asarray([cplx(c) if isinstance(c, list) else complex(c) for _, c in spec], dtype=np.complex128)
    return codes, coef


# ---------------------------------------------------------------- canonical form
def test_canonicalize_matches_dense_and_orders_terms():
    rng = np.random.default_rng(20260801)
    for n in range(0, 6):
        codes = rng.integers(0, 4, size=(40, n)).astype(np.uint8)
        codes[20:] = codes[:20]  # duplicates
        coef = rng.normal(size=40) + 1j * rng.normal(size=40)
        cc, cf = oracle.canonicalize(n, codes, coef)
        np.testing.assert_allclose(dense.dense_operator(n, cc, cf), dense.dense_operator(n, codes, coef),
                                   atol=1e-12)
        rows = [tuple(r) for r in cc]
        assert rows == sorted(rows) and len(set(rows)) == len(rows)
        assert np.all(cf != 0)


def test_canonicalize_sums_in_bit_pattern_order_and_drops_exact_zero():
    # Bit-pattern order is (1.0, 1e16, -1e16): 1 + 1e16 rounds to 1e16, then the
    # sum is exactly zero and the word is dropped.  Input order 1e16, -1e16, 1.0
    # would give 1.0 instead.
    codes = np.asarray([[1], [1], [1], [3]], dtype=np.uint8)
    coef = np.asarray([1e16, -1e16, 1.0, 2.0], dtype=np.complex128)
    cc, cf = oracle.canonicalize(1, codes, coef)
    assert cc.tolist() == [[3]] and cf.tolist() == [2.0]


def test_canonicalize_rejects_bad_input():
    with pytest.raises(oracle.OracleError) as err:
        oracle.canonicalize(2, np.asarray([[0, 4]], dtype=np.uint8), np.asarray([1.0]))
    assert err.value.status == 2
    with pytest.raises(oracle.OracleError) as err:
        oracle.canonicalize(1, np.asarray([[1]], dtype=np.uint8), np.asarray([np.nan]))
    assert err.value.status == 3


# ---------------------------------------------------------------- full-space MVP
def test_termdirect_known_answers():
    k = kats()
    case = k["two_z_plus_identity"]
    codes, coef = _terms(case["terms"])
    cc, cf = oracle.canonicalize(1, codes, coef)
    state = np.asarray([cplx(p) for p in case["state"]])
    np.testing.assert_allclose(oracle.termdirect_apply(1, cc, cf, state), [cplx(p) for p in case["output"]])
    np.testing.assert_allclose(dense.dense_operator(1, cc, cf), case["dense"])
    case = k["x_on_plus"]
    codes, coef = _terms(case["terms"])
    state = np.asarray([cplx(p) for p in case["state"]])
    out = oracle.termdirect_apply(1, codes, coef, state)
    np.testing.assert_allclose(out, [cplx(p) for p in case["output"]], atol=1e-15)
    assert abs(np.vdot(state, out) - 1.0) < 1e-15


@pytest.mark.parametrize("code", (1, 2, 3))
def test_termdirect_first_and_last_qubit_ordering(code):
    # tests/test_hamiltonian.py:71-77: qubit 0 is the most significant bit.
    state = np.arange(8) + 1j * np.arange(8)[::-1]
    first = oracle.termdirect_apply(3, np.asarray([[code, 0, 0]], np.uint8), [1.0], state)
    last = oracle.termdirect_apply(3, np.asarray([[0, 0, code]], np.uint8), [1.0], state)
    np.testing.assert_allclose(first, dense.word_matrix((code, 0, 0)) @ state)
    np.testing.assert_allclose(last, dense.word_matrix((0, 0, code)) @ state)
    assert not np.allclose(first, last)


def test_termdirect_seeded_fixtures():
    s = seeded()
    names = [f"make_operator_{n}" for n in (0, 1, 2, 5)] + ["mvp_test_operator"] + \
        [f"random_{n}" for n in range(1, 9)]
    for name in names:
        codes, coef = s[name + "_codes"], s[name + "_coef"]
        n = codes.shape[1]
        cc, cf = oracle.canonicalize(n, codes, coef)
        out = oracle.termdirect_apply(n, cc, cf, s[name + "_state"], nthreads=4)
        np.testing.assert_allclose(out, s[name + "_output"], rtol=1e-12, atol=1e-12 * max(1, np.abs(out).max()))


def test_termdirect_rows_equal_full_apply():
    rng = np.random.default_rng(5)
    codes = dense.random_codes(rng, 12, 80)
    coef = rng.normal(size=80) + 0j
    cc, cf = oracle.canonicalize(12, codes, coef)
    state = rng.normal(size=4096) + 1j * rng.normal(size=4096)
    full = oracle.termdirect_apply(12, cc, cf, state, nthreads=3)
    rows = rng.choice(4096, size=100, replace=False).astype(np.uint64)
    np.testing.assert_array_equal(oracle.termdirect_rows(12, cc, cf, state, rows), full[rows.astype(np.int64)])


def test_fixtures_reject_plausible_mistakes():
    """The seeded fixtures distinguish the right action from common mistakes."""
    s = seeded()
    codes, coef, state, expected = (s["random_4_codes"], s["random_4_coef"], s["random_4_state"],
                                    s["random_4_output"])
    n = 4
    x, z, w = oracle.matrix_terms(n, codes, coef)
    rows = np.arange(1 << n, dtype=np.uint64)

    def apply(sign_on_dest=False, drop_y=False, transpose=False):
        y = np.zeros(1 << n, dtype=np.complex128)
        for xt, zt, wt, ct in zip(x, z, w, coef):
            wt = ct if drop_y else wt
            col = rows ^ xt
            idx = col if not sign_on_dest else rows
            par = np.array([bin(int(i) & int(zt)).count("1") & 1 for i in idx])
            term = wt * np.where(par, -1, 1) * state[col.astype(np.int64)]
            if transpose:
                term = np.zeros_like(y)
                term[col.astype(np.int64)] = wt * np.where(par, -1, 1) * state
            y += term
        return y

    assert rel_err(apply(), expected) < 1e-12
    for mistake in ({"sign_on_dest": True}, {"drop_y": True}, {"transpose": True}):
        assert rel_err(apply(**mistake), expected) > 1e-3, mistake


# ---------------------------------------------------------------- U(1) sectors
def test_u1_known_answers():
    k = kats()["u1_basis"]
    for n, kk, expected in k["cases"]:
        assert oracle.u1_dimension(n, kk) == len(expected)
        np.testing.assert_array_equal(oracle.u1_unrank_many(n, kk, np.arange(len(expected))), expected)
        np.testing.assert_array_equal(oracle.u1_rank_many(n, kk, np.asarray(expected)), np.arange(len(expected)))
    assert oracle.u1_rank_many(3, 1, np.asarray([2]))[0] == k["rank_3_1_of_2"]
    assert oracle.u1_unrank_many(3, 1, np.asarray([2]))[0] == k["unrank_3_1_of_2"]


def test_u1_rank_exhaustive_against_brute_force():
    for n in range(0, 13):
        for k in range(n + 1):
            basis = np.asarray(dense.u1_basis(n, k), dtype=np.uint64)
            np.testing.assert_array_equal(oracle.u1_rank_many(n, k, basis), np.arange(len(basis)))
            np.testing.assert_array_equal(oracle.u1_unrank_many(n, k, np.arange(len(basis))), basis)


def test_u1_rank_wide_against_closed_form():
    rng = np.random.default_rng(11)
    for n in (25, 33, 48, 63, 64):
        for k in (1, 2, n // 2, n // 2 + 1, n - 2):
            dim = math.comb(n, k)
            idx = np.unique(np.concatenate([np.asarray([0, dim - 1], dtype=np.uint64),
                                            rng.integers(0, dim, size=50, dtype=np.uint64)]))
            bits = oracle.u1_unrank_many(n, k, idx.astype(np.uint64))
            for b, i in zip(bits, idx):
                assert bin(int(b)).count("1") == k
                assert dense.comb_rank(int(b), k) == int(i)
            np.testing.assert_array_equal(oracle.u1_rank_many(n, k, bits), idx)


def test_u1_lazy_apply_known_answer_and_leakage():
    k = kats()["u1_hopping"]
    codes, coef = _terms(k["terms"])
    cc, cf = oracle.canonicalize(3, codes, coef)
    out = oracle.u1_lazy_apply(3, 1, cc, cf, np.asarray(k["state"], dtype=np.complex128))
    np.testing.assert_array_equal(out, np.asarray(k["output"], dtype=np.complex128))
    restricted = dense.restricted_dense(3, cc, cf, dense.u1_basis(3, 1)).reshape(-1)
    for idx, val in k["dense_flat_nonzero"].items():
        assert restricted[int(idx)] == val
    leak = kats()["u1_leakage"]
    lc, lf = _terms([(w, [c, 0.0]) for w, c in leak["leaking"]])
    with pytest.raises(oracle.OracleError) as err:
        oracle.u1_lazy_apply(3, 1, lc, lf, np.ones(3, dtype=np.complex128))
    assert err.value.status == 7


def test_u1_lazy_apply_matches_dense_projection():
    rng = np.random.default_rng(20260802)
    for n in range(1, 9):
        codes, coef = conserving_u1_terms(rng, n, 12)
        cc, cf = oracle.canonicalize(n, codes, coef)
        for k in range(n + 1):
            basis = dense.u1_basis(n, k)
            state = rng.normal(size=len(basis)) + 1j * rng.normal(size=len(basis))
            expected = dense.restricted_dense(n, cc, cf, basis) @ state
            np.testing.assert_allclose(oracle.u1_lazy_apply(n, k, cc, cf, state), expected, atol=1e-12)


# ---------------------------------------------------------------- block sectors
def test_blocks_known_order():
    k = kats()["charge_three_modes_one_particle"]
    basis = oracle.blocks_unrank_many([(3, 1)], np.arange(3))
    assert [[(int(b) >> (2 - q)) & 1 for q in range(3)] for b in basis] == k["order"]
    two = oracle.blocks_unrank_many([(1, 0), (1, 1)], np.arange(1))
    assert int(two[0]) == 0b01


def test_blocks_rank_against_brute_force():
    for blocks in ([(3, 1), (3, 2)], [(4, 2), (4, 2)], [(2, 1), (3, 0), (3, 3)], [(5, 2), (1, 1), (4, 3)]):
        basis = np.asarray(dense.blocks_basis(blocks), dtype=np.uint64)
        assert oracle.blocks_dimension(blocks) == len(basis)
        np.testing.assert_array_equal(oracle.blocks_rank_many(blocks, basis), np.arange(len(basis)))
        np.testing.assert_array_equal(oracle.blocks_unrank_many(blocks, np.arange(len(basis))), basis)


def test_blocks_generic_apply_matches_dense_projection_hubbard():
    from tencirpauli_amd import lattice

    codes, coef = lattice.fermi_hubbard_jw(1, 3, 1.0, 4.0, seed=3)
    cc, cf = oracle.canonicalize(6, codes, coef)
    rng = np.random.default_rng(4)
    for up, dn in ((1, 1), (2, 1), (1, 2), (3, 0)):
        blocks = [(3, up), (3, dn)]
        basis = dense.blocks_basis(blocks)
        state = rng.normal(size=len(basis)) + 1j * rng.normal(size=len(basis))
        expected = dense.restricted_dense(6, cc, cf, basis) @ state
        np.testing.assert_allclose(oracle.blocks_generic_apply(blocks, cc, cf, state), expected, atol=1e-12)


def test_blocks_generic_apply_reports_leakage():
    codes = np.asarray([[1, 0, 0, 0]], dtype=np.uint8)
    with pytest.raises(oracle.OracleError) as err:
        oracle.blocks_generic_apply([(2, 1), (2, 1)], codes, [1.0], np.ones(4, dtype=np.complex128))
    assert err.value.status == 6


# ---------------------------------------------------------------- general charge sectors
GENERAL_CASES = [
    # (levels [m][n][2], targets): interleaved N and 2 S_z, weights other than
    # 0/1 with nonzero level_0, free spectator axes, three constraints, and an
    # unreachable target
    (spin_levels(8), [4, 0]),
    (spin_levels(10), [5, 1]),
    (spin_levels(10), [3, -3]),
    (np.asarray([[[0, 2], [1, -1], [0, 3], [0, 0], [2, 1], [0, 1], [-1, 1]]]), [2]),
    (np.asarray([[[0, 1]] * 3 + [[0, 0]] * 4 + [[0, 1]] * 2]), [2]),
    (np.concatenate([spin_levels(9), np.asarray([[[0, 1 if a % 3 == 0 else 0] for a in range(9)]])]), [4, 0, 1]),
    (spin_levels(6), [7, 1]),
]


def test_charge_known_answers():
    k = kats()
    for name in ("charge_spin_and_number", "charge_free_spectators"):
        case = k[name]
        assert oracle.charge_dimension(case["levels"], case["targets"]) == case["dimension"]
    case = k["charge_spin_and_number"]
    assert [int(b) for b in oracle.charge_unrank_many(case["levels"], case["targets"], np.arange(1))] == [0b10]
    for n in k["charge_wide_one_particle"]["modes"]:
        number = np.zeros((1, n, 2), dtype=np.int64)
        number[0, :, 1] = 1
        assert oracle.charge_dimension(number, [1]) == n
        last = oracle.charge_unrank_many(number, [1], np.asarray([n - 1]))
        assert int(oracle.charge_rank_many(number, [1], last)[0]) == n - 1
        marked = np.concatenate([number, np.zeros((1, n, 2), dtype=np.int64)])
        marked[1, 0, 1] = 1
        assert oracle.charge_dimension(marked, [1, 1]) == 1
        assert int(oracle.charge_unrank_many(marked, [1, 1], np.arange(1))[0]) == 1 << (n - 1)
    # the reference's restricted qubit hopping: dense matrix and its product
    for name in ("charge_qubit_hopping", "charge_aggregated_cancellation"):
        case = k[name]
        codes, coef = _terms(case["terms"])
        dim = oracle.charge_dimension(case["levels"], case["targets"])
        expected = np.asarray([[cplx(v) for v in row] for row in case["dense"]], dtype=np.complex128)
        columns = [oracle.charge_generic_apply(case["levels"], case["targets"], codes, coef, np.eye(dim)[j] + 0j)
                   for j in range(dim)]
        np.testing.assert_array_equal(np.stack(columns, axis=1), expected)
        assert int(np.count_nonzero(expected)) == case["transition_count"]


@pytest.mark.parametrize("case", range(len(GENERAL_CASES)))
def test_charge_rank_against_brute_force(case):
    levels, targets = GENERAL_CASES[case]
    n = levels.shape[1]
    basis = np.asarray(dense.charge_basis(levels, targets), dtype=np.uint64)
    assert oracle.charge_dimension(levels, targets) == len(basis)
    np.testing.assert_array_equal(oracle.charge_rank_many(levels, targets, basis), np.arange(len(basis)))
    np.testing.assert_array_equal(oracle.charge_unrank_many(levels, targets, np.arange(len(basis))), basis)
    outside = np.setdiff1d(np.arange(1 << n, dtype=np.uint64), basis)
    assert np.all(oracle.charge_rank_many(levels, targets, outside) == np.iinfo(np.uint64).max)


def test_charge_sector_agrees_with_the_block_restatement_on_factorized_sectors():
    # two independent restatements of one order: contiguous blocks (closed-form
    # binomial ranks) and the suffix-count tables
    for blocks in ([(3, 1), (3, 2)], [(4, 2), (4, 2)], [(2, 1), (3, 0), (3, 3)]):
        n = sum(s for s, _ in blocks)
        levels = np.zeros((len(blocks), n, 2), dtype=np.int64)
        start = 0
        for j, (size, _) in enumerate(blocks):
            levels[j, start:start + size, 1] = 1
            start += size
        targets = [p for _, p in blocks]
        dim = oracle.blocks_dimension(blocks)
        assert oracle.charge_dimension(levels, targets) == dim
        np.testing.assert_array_equal(oracle.charge_unrank_many(levels, targets, np.arange(dim)),
                                      oracle.blocks_unrank_many(blocks, np.arange(dim)))


def test_charge_generic_apply_matches_dense_projection_interleaved_hubbard():
    from tencirpauli_amd import lattice

    codes, coef = lattice.fermi_hubbard_jw(1, 3, 1.0, 4.0, seed=3)
    cc, cf = oracle.canonicalize(6, interleave_spins(codes), coef)
    rng = np.random.default_rng(5)
    for number, spin in ((2, 0), (3, 1), (4, 0), (3, -3)):
        levels, targets = spin_levels(6), [number, spin]
        basis = dense.charge_basis(levels, targets)
        state = rng.normal(size=len(basis)) + 1j * rng.normal(size=len(basis))
        expected = dense.restricted_dense(6, cc, cf, basis) @ state
        got = oracle.charge_generic_apply(levels, targets, cc, cf, state)
        np.testing.assert_allclose(got, expected, atol=1e-12)
    # the non-interleaved image equals the block-sector loop exactly
    bc, bf = oracle.canonicalize(6, codes, coef)
    state = rng.normal(size=9) + 0j
    np.testing.assert_array_equal(oracle.charge_generic_apply(spin_levels(6, interleaved=False), [2, 0], bc, bf, state),
                                  oracle.blocks_generic_apply([(3, 1), (3, 1)], bc, bf, state))


def test_charge_generic_apply_reports_leakage():
    codes = np.asarray([[1, 0, 0, 0]], dtype=np.uint8)
    with pytest.raises(oracle.OracleError) as err:
        oracle.charge_generic_apply(spin_levels(4), [2, 0], codes, [1.0], np.ones(4, dtype=np.complex128))
    assert err.value.status == 6


# ---------------------------------------------------------------- charge analysis
def test_analyze_charge():
    leak = kats()["u1_leakage"]
    hc, hf = _terms([(w, [c, 0.0]) for w, c in leak["conserving"]])
    lc, lf = _terms([(w, [c, 0.0]) for w, c in leak["leaking"]])
    levels = [(0.0, 1.0)] * 3
    assert oracle.analyze_charge(3, hc, hf, levels) == (True, 0)
    conserved, nonzero = oracle.analyze_charge(3, lc, lf, levels)
    assert not conserved and nonzero == 1
    # XX - YY creates pairs: not conserved
    pc, pf = _terms([("XXI", [1.0, 0.0]), ("YYI", [-1.0, 0.0])])
    assert not oracle.analyze_charge(3, pc, pf, levels)[0]
    # same letters on a qubit whose levels are equal are ignored
    assert oracle.analyze_charge(3, lc, lf, [(0.0, 0.0), (0.0, 1.0), (0.0, 1.0)])[0]


def test_sector_applies_by_source_range_sum_to_the_full_apply():
    # the bounded-sample forms used for CPU timing: disjoint source ranges are
    # separate scatters whose outputs add up to the full apply
    from tests.helpers import conserving_u1_terms

    rng = np.random.default_rng(11)
    n, k = 9, 4
    codes, coef = conserving_u1_terms(rng, n, 20)
    dim = oracle.u1_dimension(n, k)
    v = rng.normal(size=dim) + 1j * rng.normal(size=dim)
    full = oracle.u1_lazy_apply(n, k, codes, coef, v)
    parts = sum(oracle.u1_lazy_apply(n, k, codes, coef, v, sources=(lo, hi))
                for lo, hi in ((0, 17), (17, 90), (90, dim)))
    np.testing.assert_allclose(parts, full, rtol=0, atol=1e-12)
    from tencirpauli_amd import lattice

    blocks = [(4, 2), (4, 2)]
    bcodes, bcoef = lattice.fermi_hubbard_jw(2, 2, 1.0, 4.0, seed=5)  # conserves each spin block
    bdim = oracle.blocks_dimension(blocks)
    bv = rng.normal(size=bdim) + 1j * rng.normal(size=bdim)
    bfull = oracle.blocks_generic_apply(blocks, bcodes, bcoef, bv)
    bparts = sum(oracle.blocks_generic_apply(blocks, bcodes, bcoef, bv, sources=(lo, hi))
                 for lo, hi in ((0, 50), (50, bdim)))
    np.testing.assert_allclose(bparts, bfull, rtol=0, atol=1e-12)


def test_csr_apply_matches_dense_product_and_row_ranges():
    from scipy.sparse import csr_matrix

    from tests.helpers import conserving_u1_terms

    rng = np.random.default_rng(12)
    n, k = 8, 3
    codes, coef = conserving_u1_terms(rng, n, 18)
    H = dense.restricted_dense(n, codes, coef, dense.u1_basis(n, k))
    m = csr_matrix(H)
    m.sort_indices()
    v = rng.normal(size=H.shape[0]) + 1j * rng.normal(size=H.shape[0])
    for threads in (1, 3):
        y = oracle.csr_apply(m.indptr, m.indices, m.data, v, nthreads=threads)
        np.testing.assert_allclose(y, H @ v, rtol=0, atol=1e-12)
    part = oracle.csr_apply(m.indptr, m.indices, m.data, v, rows=(5, 20), nthreads=2)
    np.testing.assert_allclose(part[5:20], (H @ v)[5:20], rtol=0, atol=1e-12)
    assert not part[:5].any() and not part[20:].any()
