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


def test_davidson_fp32_phase_refinement_alone_meets_a_loose_tolerance():
    # complex64 phase to ||r|| <= 1e-5, then one FP64 MVP: the FP64 residual of
    # that Ritz vector already meets 1e-4, so no FP64 iterations follow, and
    # the FP64 Rayleigh quotient is within (1e-4)^2 / gap of the FP64 energy
    op = _heisenberg(4, 4)
    restricted = op.restrict_charge(tcp.U1Sector(16, 8))
    ref = davidson.lowest_eigenpair(restricted, tol=1e-10)
    mixed = davidson.lowest_eigenpair(restricted, tol=1e-4, mvp_dtype="complex64", switch_tol=1e-5)
    assert mixed.converged and mixed.final_fp64
    assert mixed.mvps_c128 == 1 and mixed.restarts >= 1
    assert abs(mixed.eigenvalue - ref.eigenvalue) <= 1e-8
    assert mixed.trace[-1, 3] == 64 and mixed.trace[-2, 3] == 32


def test_davidson_fp32_phase_stall_switches_to_fp64():
    # a switch tolerance below the complex64 rounding floor cannot be met: the
    # phase ends when its residual stops improving and FP64 finishes the solve
    op = _heisenberg(3, 4)
    restricted = op.restrict_charge(tcp.U1Sector(12, 6))
    ref = davidson.lowest_eigenpair(restricted, tol=1e-10)
    mixed = davidson.lowest_eigenpair(restricted, tol=1e-9, mvp_dtype="complex64", switch_tol=1e-13,
                                      max_iterations=400)
    assert mixed.converged and mixed.final_fp64 and mixed.mvps_c128 >= 1
    assert abs(mixed.eigenvalue - ref.eigenvalue) <= 1e-9


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_davidson_on_eager_csr_matches_lazy_and_eigvalsh(dtype):
    from tests.helpers import conserving_u1_terms

    rng = np.random.default_rng(5)
    n, k = 12, 6
    codes, coef = conserving_u1_terms(rng, n, 40)  # complex Hermitian restricted matrix
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    eager = op.restrict_charge(tcp.U1Sector(n, k), storage="eager")
    lazy = op.restrict_charge(tcp.U1Sector(n, k))
    exact = np.linalg.eigvalsh(dense.restricted_dense(n, *op.code_arrays(), dense.u1_basis(n, k)))[0]
    diag_e = davidson.diagonal(eager).cpu().numpy()
    np.testing.assert_allclose(diag_e, davidson.diagonal(lazy).cpu().numpy(), atol=1e-13)
    res = davidson.lowest_eigenpair(eager, tol=1e-9, mvp_dtype=dtype, switch_tol=1e-5, return_vector=True)
    res_lazy = davidson.lowest_eigenpair(lazy, tol=1e-9)
    assert res.converged and res.final_fp64
    assert abs(res.eigenvalue - exact) <= 1e-9 and abs(res_lazy.eigenvalue - exact) <= 1e-9
    vec = res.eigenvector
    resid = eager.apply(vec) - res.eigenvalue * vec
    import torch

    assert torch.linalg.vector_norm(resid).item() <= 1e-8


def test_davidson_small_subspace_restarts_and_limits():
    op = _heisenberg(3, 4)
    restricted = op.restrict_charge(tcp.U1Sector(12, 6))
    exact = np.linalg.eigvalsh(dense.restricted_dense(12, *op.code_arrays(), dense.u1_basis(12, 6)))[0]
    res = davidson.lowest_eigenpair(restricted, tol=1e-9, max_subspace=6, restart_keep=2)
    assert res.converged and res.restarts > 0 and abs(res.eigenvalue - exact) <= 1e-9
    res32 = davidson.lowest_eigenpair(restricted, tol=1e-9, max_subspace=32, restart_keep=8)
    assert res32.converged and abs(res32.eigenvalue - exact) <= 1e-9
    with pytest.raises(ValueError, match="max_subspace"):
        davidson.lowest_eigenpair(restricted, max_subspace=33)


def test_abi_rejects_multiplier_buffer_of_another_length():
    # tcp_davidson_* checks log_multiplier_len against the plan's dimension
    # before any launch; the callback is never invoked.
    import ctypes

    from tencirpauli_amd import _lib

    restricted = _heisenberg(2, 4).restrict_charge(tcp.U1Sector(8, 4))
    plan = restricted.mvp_plan()
    called = []
    fn = davidson.LOG_MULTIPLIER_FN(lambda ctx, theta, stream: called.append(theta) or None)
    opt = davidson.DavidsonOptions()
    _lib.lib().tcp_davidson_default_options(ctypes.byref(opt))
    opt.log_multiplier_fn = ctypes.cast(fn, ctypes.c_void_p).value
    opt.log_multiplier_len = plan.dimension - 1
    res = davidson.DavidsonResultStruct()
    status = _lib.lib().tcp_davidson_sector(plan._ptr, ctypes.byref(opt), None, ctypes.byref(res))
    assert status == _lib.TCP_ERR_INVALID_STRUCTURE_LENGTH
    assert not called
