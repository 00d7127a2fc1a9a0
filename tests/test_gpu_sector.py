"""GPU parity of the restricted (U(1) and contiguous-block) MVP against the oracle.

Sector indices are bit-exact; FP64 outputs within 1e-12 relative (normwise);
FP32-storage outputs within 2e-6 relative.
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


def _u64_device(values: np.ndarray):
    import torch

    return torch.from_numpy(np.ascontiguousarray(values, dtype=np.uint64).view(np.int64)).cuda()


def _to_u64(tensor) -> np.ndarray:
    return tensor.cpu().numpy().view(np.uint64)


def _blocks_sector(blocks):
    n = sum(s for s, _ in blocks)
    charges, start = [], 0
    for size, pop in blocks:
        charges.append((tcp.AdditiveCharge.number(n, range(start, start + size)), pop))
        start += size
    return tcp.ChargeSector(charges)


def test_u1_hopping_known_answer():
    op = tcp.PauliOperator.from_terms(3, [("XXI", 1.0), ("YYI", 1.0)])
    restricted = op.restrict_charge(tcp.U1Sector(3, 1))
    assert restricted.dimension == 3
    np.testing.assert_array_equal(restricted.apply(np.asarray([3.0, 5.0, 7.0])), [0.0, 14.0, 10.0])


def test_u1_small_operator_matches_projection():
    # tests/test_symmetry.py:185-207
    op = tcp.PauliOperator.from_terms(3, [("XXI", 0.5), ("YYI", 0.5), ("IZZ", -0.25), ("ZII", 0.1)])
    sector = tcp.U1Sector(3, 1)
    restricted = op.restrict_charge(sector)
    basis = dense.u1_basis(3, 1)
    expected = dense.restricted_dense(3, *op.code_arrays(), basis)
    state = np.asarray([1.0, 2.0, 3.0], dtype=np.complex128)
    np.testing.assert_allclose(restricted.apply(state), expected @ state, atol=1e-15)
    np.testing.assert_allclose(restricted.mvp_plan().apply(state), expected @ state, atol=1e-15)


def test_u1_restriction_rejects_leakage():
    sector = tcp.U1Sector(3, 1)
    with pytest.raises(ValueError, match="leakage"):
        tcp.PauliOperator.from_terms(3, [("XII", 1.0)]).restrict_charge(sector)


def test_u1_floating_point_leakage_raises_like_the_reference():
    # Four terms of one x-group that cancel exactly per commutator word but
    # not in floating point on the all-zero source: the reference's lazy apply
    # raises SectorLeakage there; the oracle restates it.
    n = 6
    words = [([1, 0, 0, 2, 0, 0], 0.65115314), ([1, 3, 3, 2, 0, 0], 0.88191665),
             ([2, 0, 0, 1, 0, 0], -0.65115314), ([2, 3, 3, 1, 0, 0], -0.88191665)]
    codes = np.asarray([w for w, _ in words], dtype=np.uint8)
    coef = np.asarray([c for _, c in words], dtype=np.complex128)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    cc, cf = op.code_arrays()
    assert oracle.analyze_charge(n, cc, cf, [(0, 1)] * n)[0]
    for k in range(n + 1):
        state = np.ones(math.comb(n, k), dtype=np.complex128)
        try:
            expected = oracle.u1_lazy_apply(n, k, cc, cf, state)
        except oracle.OracleError as err:
            restricted = op.restrict_charge(tcp.U1Sector(n, k))
            source = str(err).split("source_index=")[1].split()[0]
            with pytest.raises(ValueError, match=f"source index {source}:"):
                restricted.apply(state)
        else:
            restricted = op.restrict_charge(tcp.U1Sector(n, k))
            assert rel_err(restricted.apply(state), expected) <= FP64_TOL


def test_u1_random_conserving_operators_all_sectors():
    rng = np.random.default_rng(20260802)
    for n in range(1, 11):
        codes, coef = conserving_u1_terms(rng, n, 14)
        op = tcp.PauliOperator.from_code_arrays(codes, coef)
        cc, cf = op.code_arrays()
        for k in range(n + 1):
            dim = math.comb(n, k)
            state = rng.normal(size=dim) + 1j * rng.normal(size=dim)
            out = op.restrict_charge(tcp.U1Sector(n, k)).apply(state)
            expected = oracle.u1_lazy_apply(n, k, cc, cf, state)
            assert rel_err(out, expected) <= FP64_TOL, (n, k)


def test_u1_heisenberg_lattices_against_oracle():
    import torch

    for rows, cols, k in ((4, 4, 8), (4, 5, 10), (4, 5, 9)):
        codes, coef = lattice.heisenberg_periodic(rows, cols, seed=lattice.BASE_SEED)
        op = tcp.PauliOperator.from_code_arrays(codes, coef)
        n = rows * cols
        sector = tcp.U1Sector(n, k)
        restricted = op.restrict_charge(sector)
        state = lattice.gaussian_state(sector.dimension, 5)
        cc, cf = op.code_arrays()
        expected = oracle.u1_lazy_apply(n, k, cc, cf, state)
        assert rel_err(restricted.apply(state), expected) <= FP64_TOL
        v32 = torch.from_numpy(state).cuda().to(torch.complex64)
        y32 = restricted.apply(v32)
        torch.cuda.synchronize()
        assert rel_err(y32.cpu().numpy().astype(np.complex128), expected) <= 2e-6


def test_u1_device_rank_unrank_exhaustive_bit_exact():
    for n in range(1, 17):
        for k in range(0, n + 1):
            op = tcp.PauliOperator.empty(n)
            plan = op.restrict_charge(tcp.U1Sector(n, k)).plan
            basis = np.asarray(dense.u1_basis(n, k), dtype=np.uint64) if n <= 12 else \
                oracle.u1_unrank_many(n, k, np.arange(math.comb(n, k), dtype=np.uint64))
            ranks = _to_u64(plan.rank_device(_u64_device(basis)))
            np.testing.assert_array_equal(ranks, np.arange(len(basis), dtype=np.uint64))
            bits = _to_u64(plan.unrank_device(_u64_device(np.arange(len(basis), dtype=np.uint64))))
            np.testing.assert_array_equal(bits, basis)
            walk = _to_u64(plan.walk_device(0, len(basis), 7))
            np.testing.assert_array_equal(walk, basis)


def test_u1_device_rank_wide_sectors():
    rng = np.random.default_rng(12)
    for n, k in ((25, 12), (25, 13), (32, 5), (40, 3), (63, 2)):
        dim = math.comb(n, k)
        plan = tcp.PauliOperator.empty(n).restrict_charge(tcp.U1Sector(n, k)).plan
        idx = np.unique(np.concatenate([np.asarray([0, dim - 1], dtype=np.uint64),
                                        rng.integers(0, dim, size=4000, dtype=np.uint64)]))
        expected = oracle.u1_unrank_many(n, k, idx)
        np.testing.assert_array_equal(_to_u64(plan.unrank_device(_u64_device(idx))), expected)
        np.testing.assert_array_equal(_to_u64(plan.rank_device(_u64_device(expected))), idx)
        outside = expected ^ np.uint64(1)  # flips one bit: wrong weight
        assert np.all(_to_u64(plan.rank_device(_u64_device(outside))) == np.uint64(2 ** 64 - 1))


@pytest.mark.parametrize("rows,cols", [(1, 3), (2, 2), (2, 3)])
def test_block_sector_hubbard_against_oracle(rows, cols):
    codes, coef = lattice.fermi_hubbard_jw(rows, cols, 1.0, 4.0, seed=lattice.BASE_SEED + 3)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    cc, cf = op.code_arrays()
    sites = rows * cols
    rng = np.random.default_rng(sites)
    for up in range(0, sites + 1):
        for dn in (0, sites // 2, sites):
            blocks = [(sites, up), (sites, dn)]
            restricted = op.restrict_charge(_blocks_sector(blocks))
            state = rng.normal(size=restricted.dimension) + 1j * rng.normal(size=restricted.dimension)
            expected = oracle.blocks_generic_apply(blocks, cc, cf, state)
            assert rel_err(restricted.apply(state), expected) <= FP64_TOL, blocks
            if sites <= 4:
                full = dense.restricted_dense(2 * sites, cc, cf, dense.blocks_basis(blocks)) @ state
                assert rel_err(restricted.apply(state), full) <= FP64_TOL


def test_block_sector_device_rank_bit_exact():
    for blocks in ([(3, 1), (3, 2)], [(8, 2), (8, 3)], [(2, 1), (5, 2), (4, 4)], [(16, 4), (16, 4)]):
        n = sum(s for s, _ in blocks)
        plan = tcp.PauliOperator.empty(n).restrict_charge(_blocks_sector(blocks)).plan
        dim = plan.dimension
        idx = np.unique(np.linspace(0, dim - 1, num=min(dim, 20000)).astype(np.uint64))
        expected = oracle.blocks_unrank_many(blocks, idx)
        np.testing.assert_array_equal(_to_u64(plan.unrank_device(_u64_device(idx))), expected)
        np.testing.assert_array_equal(_to_u64(plan.rank_device(_u64_device(expected))), idx)
        if dim <= 200000:
            walk = _to_u64(plan.walk_device(0, dim, 5))
            np.testing.assert_array_equal(walk, oracle.blocks_unrank_many(blocks, np.arange(dim, dtype=np.uint64)))


def test_block_sector_rejects_nonconserving_operator():
    op = tcp.PauliOperator.from_terms(4, [("XIXI", 1.0), ("YIYI", 1.0)])  # hops between blocks
    with pytest.raises(ValueError, match="exactly conserved"):
        op.restrict_charge(_blocks_sector([(2, 1), (2, 1)]))
