"""GPU thick-restart Lanczos (``tcp.eigsh``) against independent spectra.

Pinned by numpy ``eigvalsh`` of the dense Kronecker-product / restricted
matrices (dimension up to ~1k), including degenerate multiplets of a uniform
Heisenberg ring, and by explicit residuals ||H x - lambda x|| of the returned
vectors.  The call pattern is the reference's: ``eigsh(linear, k=1,
which="SA", v0=state, tol=..., maxiter=..., ncv=...)``
(examples/research/fermi_hubbard/run_tencirpauli.py:194-211).
"""

from __future__ import annotations

import numpy as np
import pytest

import tencirpauli_amd as tcp
from oracle import dense
from tencirpauli_amd import lanczos, lattice

pytestmark = pytest.mark.gpu


def _random_operator(n, terms, seed):
    codes, coef = lattice.random_pauli_sum(n, terms, seed=seed, p_identity=0.5)
    return tcp.PauliOperator.from_code_arrays(codes, coef)


def _check_pairs(matrix, w, v, exact, tol_value=1e-10, tol_residual=1e-8):
    scale = max(1.0, np.abs(exact).max())
    np.testing.assert_allclose(w, exact, rtol=0, atol=tol_value * scale)
    assert np.all(np.diff(w) >= 0)
    for i in range(len(w)):
        x = v[:, i]
        assert abs(np.linalg.norm(x) - 1.0) <= 1e-10
        assert np.linalg.norm(matrix @ x - w[i] * x) <= tol_residual * scale
    gram = v.conj().T @ v
    np.testing.assert_allclose(gram, np.eye(len(w)), atol=1e-8)


@pytest.mark.parametrize("which", ["SA", "LA", "LM"])
def test_full_space_eigenpairs_match_eigvalsh(which):
    op = _random_operator(10, 60, seed=12)
    matrix = dense.dense_operator(10, *op.code_arrays())
    spectrum = np.linalg.eigvalsh(matrix)
    w, v = tcp.eigsh(op.native_mvp_plan(), k=4, which=which)
    if which == "SA":
        exact = spectrum[:4]
    elif which == "LA":
        exact = spectrum[-4:]
    else:
        exact = np.sort(spectrum[np.argsort(-np.abs(spectrum))[:4]])
    _check_pairs(matrix, w, v, exact)


def test_sector_plans_lazy_and_eager_match_the_restricted_spectrum():
    codes, coef = lattice.heisenberg_periodic(3, 4, seed=lattice.BASE_SEED)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    sector = tcp.U1Sector(12, 6)
    matrix = dense.restricted_dense(12, codes, coef, dense.u1_basis(12, 6))
    exact = np.linalg.eigvalsh(matrix)[:3]
    for storage in ("lazy", "eager"):
        restricted = op.restrict_charge(sector, storage=storage)
        w, v = tcp.eigsh(restricted.mvp_plan(storage=storage), k=3, which="SA", tol=1e-12)
        _check_pairs(matrix, w, v, exact)
    # the restricted operator itself and the full-space CSR plan are accepted too
    w1 = tcp.eigsh(op.restrict_charge(sector), k=1, which="SA", return_eigenvectors=False)
    assert abs(w1[0] - exact[0]) <= 1e-10 * abs(exact[0])
    small = _random_operator(8, 30, seed=5)
    ws = tcp.eigsh(small.csr_mvp_plan(), k=2, which="SA", return_eigenvectors=False)
    np.testing.assert_allclose(ws, np.linalg.eigvalsh(dense.dense_operator(8, *small.code_arrays()))[:2],
                               atol=1e-10)


def test_degenerate_multiplets_of_a_uniform_heisenberg_ring():
    # The first excited level of the ring is a triplet.  A single-vector Krylov
    # method (ARPACK's as well) sees one direction of a degenerate eigenspace
    # in exact arithmetic and further copies only through rounding, so k = 4
    # may return the triplet once, twice or three times; every returned pair
    # must still be an exact eigenpair, orthonormal to the others.
    n = 10
    bonds = [(i, (i + 1) % n) for i in range(n)]
    codes, coef = lattice.xxz_terms(n, bonds, [1.0] * n, [1.0] * n, [0.0] * n)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    matrix = dense.dense_operator(n, *op.code_arrays())
    spectrum = np.linalg.eigvalsh(matrix)
    assert abs(spectrum[1] - spectrum[3]) <= 1e-9 and spectrum[3] - spectrum[0] > 1e-3
    w, v = tcp.eigsh(op, k=4, which="SA", ncv=24)
    assert abs(w[0] - spectrum[0]) <= 1e-10 * abs(spectrum[0]) and abs(w[1] - spectrum[1]) <= 1e-9
    levels = np.unique(np.round(spectrum, 8))
    assert all(np.min(np.abs(levels - x)) <= 1e-8 for x in w)
    for i in range(4):
        assert np.linalg.norm(matrix @ v[:, i] - w[i] * v[:, i]) <= 1e-8 * abs(w[i])
    np.testing.assert_allclose(v.conj().T @ v, np.eye(4), atol=1e-8)


def test_reference_call_pattern_with_start_vector_and_outputs():
    codes, coef = lattice.heisenberg_periodic(4, 4, seed=lattice.BASE_SEED)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    restricted = op.restrict_charge(tcp.U1Sector(16, 8))
    plan = restricted.mvp_plan()
    state = lattice.gaussian_state(plan.dimension, 3)
    w, v = tcp.eigsh(plan, k=1, which="SA", v0=state, tol=1e-10, maxiter=300, ncv=20)
    assert w.shape == (1,) and v.shape == (plan.dimension, 1) and isinstance(v, np.ndarray)
    residual = np.linalg.norm(restricted.apply(v[:, 0]) - w[0] * v[:, 0])
    assert residual <= 1e-8 * abs(w[0])
    ref = tcp.davidson.lowest_eigenpair(restricted, tol=1e-10)
    assert abs(ref.eigenvalue - w[0]) <= 1e-9 * abs(w[0])
    # device start vector -> device eigenvectors; explicit residual norms
    import torch

    result = lanczos.lanczos(plan, 2, v0=torch.from_numpy(state).cuda(), tol=1e-10, residuals=True)
    assert result.converged == 2 and result.eigenvectors.is_cuda
    assert np.all(result.residual_norms <= 1e-8 * np.abs(result.eigenvalues))
    assert result.mvps > 0 and result.seconds > 0


def test_exhausted_krylov_space_of_a_small_sector():
    # D = 6: the basis spans the whole sector and every pair is exact
    op = tcp.PauliOperator.from_terms(4, [("XXII", 1.0), ("YYII", 1.0), ("IXXI", 0.7), ("IYYI", 0.7),
                                          ("IIXX", 0.4), ("IIYY", 0.4), ("ZIIZ", 0.3)])
    restricted = op.restrict_charge(tcp.U1Sector(4, 2))
    matrix = restricted.dense()
    w, v = tcp.eigsh(restricted.mvp_plan(), k=5, which="SA", ncv=6)
    _check_pairs(matrix, w, v, np.linalg.eigvalsh(matrix)[:5])


def test_argument_errors_and_no_convergence():
    from scipy.sparse.linalg import ArpackNoConvergence

    op = _random_operator(6, 20, seed=2)
    plan = op.native_mvp_plan()
    with pytest.raises(ValueError, match="k must be"):
        tcp.eigsh(plan, k=64)
    with pytest.raises(ValueError, match="ncv"):
        tcp.eigsh(plan, k=4, ncv=4)
    with pytest.raises(ValueError, match="ncv"):
        tcp.eigsh(plan, k=2, ncv=40)
    with pytest.raises(ValueError, match="which"):
        tcp.eigsh(plan, k=2, which="SM")
    with pytest.raises(NotImplementedError):
        tcp.eigsh(plan, k=2, sigma=0.5)
    with pytest.raises(ValueError, match="v0"):
        tcp.eigsh(plan, k=1, v0=np.ones(10))
    with pytest.raises(ValueError, match="keep"):
        lanczos.lanczos(plan, 2, keep=1)
    kept = lanczos.lanczos(plan, 2, ncv=12, keep=6, tol=1e-10)
    np.testing.assert_allclose(kept.eigenvalues, np.linalg.eigvalsh(dense.dense_operator(6, *op.code_arrays()))[:2],
                               atol=1e-9)
    big = _random_operator(12, 80, seed=9).native_mvp_plan()
    with pytest.raises(ArpackNoConvergence) as info:
        tcp.eigsh(big, k=3, which="SA", ncv=8, maxiter=0, tol=1e-14)
    assert info.value.eigenvalues.shape == (3,)
