"""GPU Davidson eigensolver against independent spectra.

Pinned by numpy ``eigvalsh`` on the dense restricted matrix (dimension up to a
few thousand) and by SciPy ``eigsh`` over the oracle's restricted MVP (the
pattern of the reference's tests/test_scipy_linear_operator.py:104-124).
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import tencirpauli_amd as tcp
from oracle import dense
from tencirpauli_amd import davidson, lattice

pytestmark = pytest.mark.gpu


def _heisenberg(rows, cols, seed=lattice.BASE_SEED):
    codes, coef = lattice.heisenberg_periodic(rows, cols, seed=seed)
    return tcp.PauliOperator.from_code_arrays(codes, coef)


def test_sector_diagonal_matches_dense():
    op = _heisenberg(2, 4)
    restricted = op.restrict_charge(tcp.U1Sector(8, 4))
    diag = davidson.diagonal(restricted).cpu().numpy()
    expected = np.real(np.diag(dense.restricted_dense(8, *op.code_arrays(), dense.u1_basis(8, 4))))
    np.testing.assert_allclose(diag, expected, atol=1e-13)


def test_full_space_diagonal_matches_dense():
    codes, coef = lattice.random_pauli_sum(10, 80, seed=4, p_identity=0.5)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    diag = davidson.diagonal(op.native_mvp_plan()).cpu().numpy()
    expected = np.real(np.diag(dense.dense_operator(10, *op.code_arrays())))
    np.testing.assert_allclose(diag, expected, atol=1e-12)


@pytest.mark.parametrize("rows,cols,k", [(2, 4, 4), (3, 4, 6), (3, 4, 5)])
def test_davidson_fp64_matches_eigvalsh(rows, cols, k):
    op = _heisenberg(rows, cols)
    n = rows * cols
    restricted = op.restrict_charge(tcp.U1Sector(n, k))
    res = davidson.lowest_eigenpair(restricted, tol=1e-9, return_vector=True)
    assert res.converged and res.final_fp64
    matrix = dense.restricted_dense(n, *op.code_arrays(), dense.u1_basis(n, k))
    exact = np.linalg.eigvalsh(matrix)[0]
    assert abs(res.eigenvalue - exact) <= 1e-9
    vec = res.eigenvector.cpu().numpy()
    assert np.linalg.norm(matrix @ vec - exact * vec) <= 1e-7


def test_davidson_fp32_phase_then_fp64_reaches_fp64_energy():
    op = _heisenberg(4, 4)
    restricted = op.restrict_charge(tcp.U1Sector(16, 8))
    ref = davidson.lowest_eigenpair(restricted, tol=1e-10)
    mixed = davidson.lowest_eigenpair(restricted, tol=1e-8, mvp_dtype="complex64", switch_tol=1e-4)
    assert mixed.converged and mixed.final_fp64
    assert mixed.mvps_c64 > 0 and mixed.mvps_c128 > 0
    assert abs(mixed.eigenvalue - ref.eigenvalue) <= 1e-8
    assert mixed.seconds_to(ref.eigenvalue, 1e-8) is not None


def test_davidson_matches_eigsh_on_oracle_operator():
    from scipy.sparse.linalg import LinearOperator, eigsh

    op = _heisenberg(4, 4, seed=7)
    cc, cf = op.code_arrays()
    dim = tcp.U1Sector(16, 7).dimension
    linear = LinearOperator((dim, dim), matvec=lambda x: oracle.u1_lazy_apply(16, 7, cc, cf, np.asarray(x).ravel()),
                            dtype=np.complex128)
    exact = eigsh(linear, k=1, which="SA", tol=1e-12, return_eigenvectors=False, v0=np.ones(dim))[0]
    res = davidson.lowest_eigenpair(op.restrict_charge(tcp.U1Sector(16, 7)), tol=1e-9)
    assert abs(res.eigenvalue - exact) <= 1e-9


def test_davidson_full_space_plan():
    codes, coef = lattice.heisenberg_periodic(2, 5, seed=3)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    res = davidson.lowest_eigenpair(op.native_mvp_plan(), tol=1e-9)
    exact = np.linalg.eigvalsh(dense.dense_operator(10, *op.code_arrays()))[0]
    assert abs(res.eigenvalue - exact) <= 1e-9


def test_davidson_block_sector_hubbard():
    codes, coef = lattice.fermi_hubbard_jw(2, 2, 1.0, 4.0, seed=5)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    n1 = tcp.AdditiveCharge.number(8, range(4))
    n2 = tcp.AdditiveCharge.number(8, range(4, 8))
    restricted = op.restrict_charge(tcp.ChargeSector(((n1, 2), (n2, 2))))
    res = davidson.lowest_eigenpair(restricted, tol=1e-9)
    exact = np.linalg.eigvalsh(dense.restricted_dense(8, *op.code_arrays(), dense.blocks_basis([(4, 2), (4, 2)])))[0]
    assert abs(res.eigenvalue - exact) <= 1e-9
