"""GPU checks of fermion operators restricted to factorized charge sectors.

The reference's raw-fermion path (examples/research/fermi_hubbard/
run_tencirpauli.py:82-160; tests/test_charge.py:652-771) runs here as the
block-sector gather over ladder words with occupancy conditions.  Pinned by
the occupation-basis dense oracle restricted to the sector basis (ascending
occupation integers, mode 0 most significant) and, at larger sizes, by the C
oracle's generic charge loop on the Jordan-Wigner image.  (On the
Jordan-Wigner image itself, the 16-term groups of the 0.35i pair hopping sum
to rounding residue instead of exact zeros on some out-of-sector sources,
which the Pauli sector path reports as leakage; ladder words are exact.)
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import tencirpauli_amd as tcp
from tencirpauli_amd import lattice
from oracle import dense
from oracle import fermion as foracle
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _spinful_sector(space, sites, particles):
    total = tcp.AdditiveCharge(space, fermions={m: 1 for m in range(2 * sites)})
    balance = tcp.AdditiveCharge(space, fermions={m: (1 if m < sites else -1) for m in range(2 * sites)})
    return tcp.ChargeSector(((total, 2 * particles), (balance, 0)))


def _restricted_oracle(n_modes, terms, blocks):
    basis = np.asarray(dense.blocks_basis(blocks), dtype=np.int64)
    full = foracle.operator_matrix(n_modes, terms)
    return full[np.ix_(basis, basis)]


@pytest.mark.parametrize("sites,particles", [(s, p) for s in range(2, 6) for p in range(1, s + 1)])
def test_spinful_lazy_and_eager_match_the_occupation_oracle(sites, particles):
    terms = foracle.spinful_terms(sites)
    op = tcp.FermionOperator.from_terms(2 * sites, terms)
    sector = _spinful_sector(tcp.OperatorSpace(fermions=2 * sites), sites, particles)
    assert sector.blocks == ((sites, particles), (sites, particles))
    expected = _restricted_oracle(2 * sites, terms, sector.blocks)
    lazy = op.restrict_charge(sector)
    eager = op.restrict_charge(sector, storage="eager")
    assert lazy.dimension == expected.shape[0]
    state = np.arange(lazy.dimension, dtype=np.float64).astype(np.complex128)
    state += 0.125j * state[::-1]
    want = expected @ state
    out = np.empty_like(state)
    plan = lazy.mvp_plan()
    assert plan.native_variant == "factorized"
    plan.apply_into(state, out, max_bytes=None)
    assert rel_err(out, want) <= FP64_TOL
    assert rel_err(eager.apply(state), want) <= FP64_TOL
    np.testing.assert_allclose(eager.dense(), expected, atol=1e-13)
    # the ladder words give exact values: the CSR holds exactly the oracle's nonzeros
    assert eager.transition_count == int(np.count_nonzero(expected))
    assert plan.execution_family == "fermion" and plan.term_count == op.term_count


def test_reference_example_api_on_a_2x2_hubbard_at_half_filling():
    # examples/research/fermi_hubbard/run_tencirpauli.py:82-160 at rows = cols = 2
    rows, cols, sites = 2, 2, 4
    terms = lattice.hubbard_ladder_terms(rows, cols, 1.0, 4.0)
    operator = tcp.FermionOperator.from_terms(2 * sites, terms)
    space = tcp.OperatorSpace(fermions=2 * sites)
    total = tcp.AdditiveCharge(space, name="particle_number", fermions={m: 1 for m in range(2 * sites)})
    spin = tcp.AdditiveCharge(space, name="two_Sz", fermions={m: (1 if m < sites else -1) for m in range(2 * sites)})
    sector = tcp.ChargeSector(((total, sites), (spin, 0)), max_bytes=16 * 1024 ** 3)
    restricted = operator.restrict_charge(sector, storage="lazy")
    plan = restricted.mvp_plan()
    assert restricted.dimension == 36 and plan.term_count > 0 and plan.estimated_bytes > 0
    expected = _restricted_oracle(2 * sites, terms, sector.blocks)
    rng = np.random.default_rng(20260804)
    state = rng.normal(size=36) + 1j * rng.normal(size=36)
    state /= np.linalg.norm(state)
    assert rel_err(restricted.apply(state), expected @ state) <= FP64_TOL
    # the example's eigen call
    w, v = tcp.eigsh(plan, k=1, which="SA", v0=state, tol=1e-10, maxiter=300, ncv=20)
    exact = np.linalg.eigvalsh(expected)[0]
    assert abs(w[0] - exact) <= 1e-9 * abs(exact)
    assert np.linalg.norm(restricted.apply(v[:, 0]) - w[0] * v[:, 0]) <= 1e-8


def test_compile_targets_follow_the_occupation_basis():
    # tests/test_structured_algebra.py:471-530 of the reference
    annihilate = tcp.FermionOperator.from_terms(2, [(((0, "annihilate"),), 1.0)])
    create = tcp.FermionOperator.from_terms(2, [(((0, "create"),), 1.0)])
    aa, ad = annihilate.compile("dense"), create.compile("dense")
    np.testing.assert_allclose(aa @ ad + ad @ aa, np.eye(4))
    np.testing.assert_allclose(aa, foracle.product_matrix(2, ((0, "annihilate"),)))
    terms = lattice.hubbard_ladder_terms(1, 2, 0.7, 1.25)
    op = tcp.FermionOperator.from_terms(4, terms)
    expected = foracle.operator_matrix(4, terms)
    np.testing.assert_allclose(op.compile("dense"), expected, atol=1e-12)
    state = np.arange(16, dtype=np.complex128)
    np.testing.assert_allclose(op.compile("native_mvp").apply(state), expected @ state, atol=1e-12)
    np.testing.assert_allclose(op.mvp(state), expected @ state, atol=1e-12)
    np.testing.assert_allclose(op.compile("csr").to_scipy().toarray(), expected, atol=1e-12)


def test_3x4_hubbard_quarter_filling_against_the_generic_charge_loop():
    # 24 modes, N_up = N_dn = 3: D = C(12,3)^2 = 48,400; the C oracle's serial
    # generic loop (core/charge/generic.rs:949-1045) over the same Pauli image
    terms = lattice.hubbard_ladder_terms(3, 4, 1.0, 4.0)
    op = tcp.FermionOperator.from_terms(24, terms)
    sector = _spinful_sector(tcp.OperatorSpace(fermions=24), 12, 3)
    restricted = op.restrict_charge(sector)
    codes, coef = op.map_fermions().code_arrays()
    state = np.random.default_rng(3).normal(size=restricted.dimension) + 0j
    expected = oracle.blocks_generic_apply([(12, 3), (12, 3)], codes, coef, state)
    assert rel_err(restricted.apply(state), expected) <= FP64_TOL
    # conservation of both charges, and a non-conserving term refused by the lazy path
    total, balance = (c for c, _ in sector.constraints)
    assert op.conserves(total) and op.conserves(balance)
    broken = tcp.FermionOperator.from_terms(24, terms + [(((0, "create"), (12, "annihilate")), 0.1)])
    assert not broken.conserves(balance)
    with pytest.raises(ValueError):
        broken.restrict_charge(sector)
    with pytest.raises(ValueError, match="leaks"):
        broken.restrict_charge(sector, storage="eager")
