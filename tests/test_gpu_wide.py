"""GPU checks of wide U(1) sectors (64..1024 qubits) against the big-integer
oracle (oracle/wide.py): eager CSR structure and values, lazy applies in both
precisions, leakage after aggregation, and Davidson -- the cases of the
reference's tests/test_u1.py:84-199 at the sizes a Python oracle finishes in
seconds."""

from __future__ import annotations

import math

import numpy as np
import pytest

import tencirpauli_amd as tcp
from oracle import wide
from tencirpauli_amd import davidson

pytestmark = pytest.mark.gpu


def _chain(n: int, coefficient: complex = 0.5):
    """coefficient (X_q X_{q+1} + Y_q Y_{q+1}) on every bond of an open chain."""
    codes = []
    for q in range(n - 1):
        for c in (1, 2):
            row = [0] * n
            row[q] = row[q + 1] = c
            codes.append(row)
    return np.asarray(codes, dtype=np.uint8), np.full(len(codes), coefficient, dtype=np.complex128)


def _sparse_csr(n, k, codes, coef):
    from scipy.sparse import csr_matrix

    ref = wide.restricted_sparse(n, k, codes, coef)
    dim = math.comb(n, k)
    rows, cols = zip(*ref.keys())
    m = csr_matrix((list(ref.values()), (rows, cols)), shape=(dim, dim))
    m.sort_indices()
    return m


@pytest.mark.parametrize("n,k", [(64, 1), (65, 2), (65, 63), (127, 1), (128, 2), (129, 2), (256, 1)])
def test_wide_restriction_matches_big_int_reference(n, k):
    codes, coef = _chain(n)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    expected = _sparse_csr(n, k, *op.code_arrays())
    eager = op.restrict_charge(tcp.U1Sector(n, k), storage="eager")
    got = eager.csr()
    np.testing.assert_array_equal(got.indptr.astype(np.int64), expected.indptr)
    np.testing.assert_array_equal(got.indices.astype(np.int64), expected.indices)
    np.testing.assert_allclose(got.data, expected.data, rtol=1e-14, atol=0)
    lazy = op.restrict_charge(tcp.U1Sector(n, k))
    rng = np.random.default_rng(n + k)
    v = rng.normal(size=expected.shape[0]) + 1j * rng.normal(size=expected.shape[0])
    y = expected @ v
    for restricted in (lazy, eager):
        out = restricted.apply(v)
        assert np.linalg.norm(out - y) <= 1e-12 * np.linalg.norm(y)


def test_wide_complex_directed_hopping_crosses_a_limb_boundary():
    n = 65
    coefficient = 0.7 + 0.2j
    codes = np.zeros((2, n), dtype=np.uint8)
    codes[0, 63] = codes[0, 64] = 1
    codes[1, 63] = codes[1, 64] = 2
    op = tcp.PauliOperator.from_code_arrays(codes, np.full(2, coefficient))
    expected = _sparse_csr(n, 1, *op.code_arrays())
    got = op.restrict_charge(tcp.U1Sector(n, 1), storage="eager").csr()
    np.testing.assert_array_equal(got.indices.astype(np.int64), expected.indices)
    np.testing.assert_allclose(got.data, expected.data, rtol=1e-14, atol=0)
    v = np.arange(n) + 1j * np.arange(n)[::-1]
    out = op.restrict_charge(tcp.U1Sector(n, 1), storage="eager").apply(v)
    np.testing.assert_allclose(out, expected @ v, rtol=1e-13, atol=1e-13)


def test_wide_leakage_is_checked_after_x_group_aggregation():
    n = 129
    codes = np.zeros((2, n), dtype=np.uint8)
    codes[0, 63] = codes[0, 64] = 1
    codes[1, 63] = codes[1, 64] = 2
    hopping = tcp.PauliOperator.from_code_arrays(codes, np.full(2, 0.5))
    assert hopping.restrict_charge(tcp.U1Sector(n, 1)).dimension == n
    leak = np.zeros((1, n), dtype=np.uint8)
    leak[0, 64] = 1
    leaking = tcp.PauliOperator.from_code_arrays(leak, np.ones(1))
    with pytest.raises(ValueError, match=r"U\(1\) sector leakage"):
        leaking.restrict_charge(tcp.U1Sector(n, 1))
    with pytest.raises(ValueError, match=r"U\(1\) sector leakage from restricted source index 0"):
        leaking.restrict_charge(tcp.U1Sector(n, 1), storage="eager")


def test_wide_fp32_storage_and_device_apply():
    import torch

    codes, coef = _chain(200, 0.5)
    codes = np.vstack([codes, np.eye(200, dtype=np.uint8)[:50] * 3])  # Z fields on the first 50 qubits
    coef = np.concatenate([coef, np.linspace(-1, 1, 50)])
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    lazy = op.restrict_charge(tcp.U1Sector(200, 2))
    eager = op.restrict_charge(tcp.U1Sector(200, 2), storage="eager")
    g = torch.Generator(device="cuda").manual_seed(5)
    v = torch.randn(lazy.dimension, dtype=torch.complex128, device="cuda", generator=g)
    y64 = lazy.apply(v)
    y_e = eager.apply(v)
    y32 = lazy.apply(v.to(torch.complex64)).to(torch.complex128)
    assert (torch.linalg.vector_norm(y_e - y64) / torch.linalg.vector_norm(y64)).item() <= 1e-13
    assert (torch.linalg.vector_norm(y32 - y64) / torch.linalg.vector_norm(y64)).item() <= 2e-6
    with pytest.raises(tcp.UnsupportedOnDeviceError):
        lazy.plan.rank_device(torch.zeros(4, dtype=torch.int64, device="cuda"))


def test_davidson_on_a_wide_sector_matches_sparse_eigsh():
    from scipy.sparse.linalg import eigsh

    n, k = 100, 2
    codes, coef = _chain(n, 0.5)
    codes = np.vstack([codes, np.eye(n, dtype=np.uint8) * 3])
    coef = np.concatenate([coef, np.random.default_rng(9).uniform(-0.3, 0.3, n)])
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    m = _sparse_csr(n, k, *op.code_arrays())
    exact = eigsh(m, k=1, which="SA", tol=1e-12, return_eigenvectors=False, v0=np.ones(m.shape[0]))[0]
    for storage in ("lazy", "eager"):
        res = davidson.lowest_eigenpair(op.restrict_charge(tcp.U1Sector(n, k), storage=storage), tol=1e-9)
        assert res.converged and abs(res.eigenvalue - exact) <= 1e-9, storage
