"""GPU parity of the full-space MVP (gfx950 kernels) against the CPU oracle.

Tolerance: FP64 results must match the oracle within 1e-12 relative (normwise,
the bar of the reference's own tests, tests/test_hamiltonian.py:58-68); the
FP32-storage path (complex64 amplitudes, FP32 coefficients, FP64 accumulation)
within 2e-6 relative of the FP64 oracle.
"""

from __future__ import annotations

import os
import threading

import numpy as np
import pytest

import oracle
import tencirpauli_amd as tcp
from oracle import dense
from tencirpauli_amd import lattice
from tests.helpers import rel_err, seeded

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-6
THREADS = max(1, min(16, os.cpu_count() or 1))


def _oracle_apply(op: tcp.PauliOperator, state: np.ndarray) -> np.ndarray:
    codes, coef = op.code_arrays()
    return oracle.termdirect_apply(op.nqubits, codes, coef, state, nthreads=THREADS)


def _random_state(n: int, seed: int) -> np.ndarray:
    return lattice.gaussian_state(1 << n, seed)


def test_seeded_fixtures_match_dense_oracle():
    s = seeded()
    names = [f"make_operator_{n}" for n in (0, 1, 2, 5)] + ["mvp_test_operator"] + \
        [f"random_{n}" for n in range(1, 9)]
    for name in names:
        op = tcp.PauliOperator.from_code_arrays(s[name + "_codes"], s[name + "_coef"])
        out = op.native_mvp_plan().apply(s[name + "_state"])
        assert rel_err(out, s[name + "_output"]) <= FP64_TOL, name
        np.testing.assert_allclose(op.mvp(s[name + "_state"]), out, rtol=0, atol=0)


def test_known_answer_two_z_plus_identity():
    op = tcp.PauliOperator.from_terms(1, [("Z", 2.0), ("I", 1.0)])
    plan = op.native_mvp_plan()
    np.testing.assert_allclose(plan.apply(np.asarray([1.0, 2.0])), [3.0, -2.0])
    out = np.empty(2, dtype=np.complex128)
    plan.apply_into(np.asarray([1.0, 2.0], dtype=np.complex128), out)
    np.testing.assert_allclose(out, [3.0, -2.0])


@pytest.mark.parametrize("n", [0, 1, 3, 5, 8, 9, 10, 11, 12, 13, 14, 16])
def test_random_operators_every_tile_shape(n):
    codes, coef = lattice.random_pauli_sum(n, 96, seed=1000 + n, p_identity=0.5)
    coef = coef * (1 + 0.3j)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    state = _random_state(n, 77 + n)
    plan = op.native_mvp_plan()
    out = plan.apply(state)
    expected = _oracle_apply(op, state)
    assert rel_err(out, expected) <= FP64_TOL
    if n <= 8:
        full = dense.dense_operator(n, *op.code_arrays()) @ state
        assert rel_err(out, full) <= FP64_TOL


@pytest.mark.parametrize("xcd_order", [False, True])
@pytest.mark.parametrize("tiling", [None, "512x8", "512x4", "1024x8", "512x16", "1024x8x2", "512x16x2"])
def test_every_tiling_against_oracle(tiling, xcd_order):
    import torch

    codes, coef = lattice.random_pauli_sum(17, 500, seed=91, p_identity=0.55)
    op = tcp.PauliOperator.from_code_arrays(codes, coef * (1 - 0.4j))
    plan = op.native_mvp_plan(tiling=tiling, xcd_order=xcd_order)
    assert plan.tiling == (tiling or "512x8") and plan.xcd_order == xcd_order
    assert plan.tiling_c64 == (tiling or "512x4")
    bits = {"512x4": 11, "512x8": 12, "1024x8": 13, "512x16": 13, "1024x8x2": 13, "512x16x2": 13}
    assert plan.tile_bits == bits[plan.tiling] and plan.tile_bits_c64 == bits[plan.tiling_c64]
    assert plan.pass_count >= plan.group_count
    state = _random_state(17, 92)
    expected = _oracle_apply(op, state)
    assert rel_err(plan.apply(state), expected) <= FP64_TOL
    y32 = plan.apply(torch.from_numpy(state).cuda().to(torch.complex64))
    torch.cuda.synchronize()
    assert rel_err(y32.cpu().numpy().astype(np.complex128), expected) <= FP32_TOL


def test_default_tiling_is_xcd_ordered():
    codes, coef = lattice.random_pauli_sum(14, 50, seed=95)
    plan = tcp.PauliOperator.from_code_arrays(codes, coef).native_mvp_plan()
    assert plan.tiling == "512x8" and plan.xcd_order and plan.tile_bits == 12
    assert plan.tiling_c64 == "512x4" and plan.tile_bits_c64 == 11


def test_tiling_falls_back_on_small_spaces():
    codes, coef = lattice.random_pauli_sum(12, 80, seed=93)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    for tiling in ("1024x8", "512x16", "1024x8x2", "512x16x2"):
        plan = op.native_mvp_plan(tiling=tiling, xcd_order=True)
        assert plan.tiling == "512x8" and plan.tile_bits == 12
        state = _random_state(12, 94)
        assert rel_err(plan.apply(state), _oracle_apply(op, state)) <= FP64_TOL
    with pytest.raises(ValueError):
        op.native_mvp_plan(tiling="64x64")


@pytest.mark.parametrize("tiling", ["1024x8x2", "512x16x2"])
def test_twin_tiles_with_bit_map_against_oracle(tiling):
    """Twin tiles on a space large enough for the operator-chosen bit map
    (window and XCD bits): every row against the oracle, in both storages,
    and against the single-tile geometry of the same staged tile size."""
    import torch

    n = 22
    codes, coef = lattice.random_pauli_sum(n, 300, seed=96, p_identity=0.6)
    op = tcp.PauliOperator.from_code_arrays(codes, coef * (0.7 + 0.2j))
    plan = op.native_mvp_plan(tiling=tiling, xcd_order=True)
    assert plan.tiling == tiling and plan.xcd_order and plan.tile_bits == 13
    single = op.native_mvp_plan(tiling="1024x8", xcd_order=True)
    print(f"{tiling}: classes {plan.class_count} (1024x8: {single.class_count})", flush=True)
    state = _random_state(n, 97)
    expected = _oracle_apply(op, state)
    out = plan.apply(state)
    assert rel_err(out, expected) <= FP64_TOL
    assert rel_err(out, single.apply(state)) <= FP64_TOL
    y32 = plan.apply(torch.from_numpy(state).cuda().to(torch.complex64))
    torch.cuda.synchronize()
    assert rel_err(y32.cpu().numpy().astype(np.complex128), expected) <= FP32_TOL


@pytest.mark.parametrize("tiling", ["512x8", "512x4", "1024x8", "512x16", "1024x8x2"])
def test_real_imaginary_and_complex_passes_against_oracle(tiling):
    """Passes whose terms' weights are all real, all imaginary, or mixed run on
    different loops of the single-tile kernels: an operator with real
    coefficients (real or imaginary weights by the parity of #Y) plus a few
    complex ones puts all three kinds into the same classes; every row
    against the oracle, both storages."""
    import torch

    n = 17
    codes, coef = lattice.random_pauli_sum(n, 600, seed=98, p_identity=0.6)
    coef = np.asarray(coef, dtype=np.complex128).copy()
    coef[::7] *= (0.6 + 0.8j)  # mixed passes where these meet real ones
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    plan = op.native_mvp_plan(tiling=tiling, xcd_order=True)
    state = _random_state(n, 99)
    expected = _oracle_apply(op, state)
    assert rel_err(plan.apply(state), expected) <= FP64_TOL
    y32 = plan.apply(torch.from_numpy(state).cuda().to(torch.complex64))
    torch.cuda.synchronize()
    assert rel_err(y32.cpu().numpy().astype(np.complex128), expected) <= FP32_TOL


def test_many_classes_and_large_groups():
    # full-width random strings: nearly every term is its own x-group and class
    codes, coef = lattice.random_pauli_sum(18, 700, seed=5, p_identity=0.3)
    diag = np.zeros((300, 18), dtype=np.uint8)
    rng = np.random.default_rng(6)
    diag[:] = np.where(rng.random((300, 18)) < 0.5, 3, 0)
    codes = np.concatenate([codes, diag])
    coef = np.concatenate([coef, rng.normal(size=300) + 0j])
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    plan = op.native_mvp_plan()
    # every 6-bit high part occurs; the 300-term diagonal group spans several class blocks
    assert plan.class_count >= 64 and plan.group_count > 500
    state = _random_state(18, 8)
    assert rel_err(plan.apply(state), _oracle_apply(op, state)) <= FP64_TOL


def test_device_path_fp64_and_fp32():
    import torch

    codes, coef = lattice.random_pauli_sum(16, 300, seed=21)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    plan = op.native_mvp_plan()
    state = _random_state(16, 22)
    expected = _oracle_apply(op, state)
    v = torch.from_numpy(state).cuda()
    y = torch.empty_like(v)
    plan.apply_into(v, y)
    torch.cuda.synchronize()
    assert rel_err(y.cpu().numpy(), expected) <= FP64_TOL
    v32 = v.to(torch.complex64)
    y32 = plan.apply(v32)
    torch.cuda.synchronize()
    err32 = rel_err(y32.cpu().numpy().astype(np.complex128), expected)
    assert err32 <= FP32_TOL
    assert err32 > 0.0  # really FP32 storage


def test_apply_into_overwrites_prefilled_output():
    codes, coef = lattice.random_pauli_sum(10, 64, seed=3)
    plan = tcp.PauliOperator.from_code_arrays(codes, coef).native_mvp_plan()
    state = _random_state(10, 4)
    zeroed = np.zeros(1024, dtype=np.complex128)
    prefilled = np.full(1024, 7.0 - 11.0j)
    plan.apply_into(state, zeroed)
    plan.apply_into(state, prefilled)
    np.testing.assert_array_equal(prefilled, zeroed)


def test_empty_identity_and_shape_errors():
    zero = tcp.PauliOperator.empty(3)
    np.testing.assert_array_equal(zero.native_mvp_plan().apply(np.ones(8)), np.zeros(8))
    ident = tcp.PauliOperator.from_terms(0, [((), 2.5)])
    np.testing.assert_array_equal(ident.mvp(np.asarray([1.0 + 1j])), [2.5 + 2.5j])
    with pytest.raises(ValueError, match="expected structure length"):
        tcp.PauliOperator.empty(2).mvp(np.ones(3, dtype=np.complex128))
    plan = tcp.PauliOperator.from_terms(2, [("XY", 1.0)]).native_mvp_plan()
    with pytest.raises(ValueError, match="shape"):
        plan.apply(np.ones(3))
    with pytest.raises(TypeError):
        plan.apply_into(np.ones(4, dtype=np.complex64), np.ones(4, dtype=np.complex64))
    # max_bytes is advisory below 1 GiB (tests/test_hamiltonian.py:150-175)
    big = tcp.PauliOperator.empty(20)
    np.testing.assert_array_equal(big.mvp(np.zeros(1 << 20, dtype=np.complex128), max_bytes=1), 0)


def test_config0_20q_1024_terms_against_oracle():
    codes, coef = lattice.random_pauli_sum(20, 1024, seed=lattice.BASE_SEED)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    state = lattice.gaussian_state(1 << 20, lattice.BASE_SEED + 1)
    plan = op.native_mvp_plan()
    out = plan.apply(state)
    expected = _oracle_apply(op, state)
    assert rel_err(out, expected) <= FP64_TOL
    print(f"config0: T={plan.term_count} G={plan.group_count} H={plan.class_count} "
          f"rel_err={rel_err(out, expected):.2e}", flush=True)


def test_batched_right_hand_sides_equal_single_applies():
    import torch

    codes, coef = lattice.random_pauli_sum(13, 128, seed=31)
    plan = tcp.PauliOperator.from_code_arrays(codes, coef).native_mvp_plan()
    rng = np.random.default_rng(32)
    states = rng.normal(size=(5, 1 << 13)) + 1j * rng.normal(size=(5, 1 << 13))
    v = torch.from_numpy(states).cuda()
    y = torch.empty_like(v)
    plan.apply_batch(v, y)
    torch.cuda.synchronize()
    for r in range(5):
        np.testing.assert_array_equal(y[r].cpu().numpy(), plan.apply(states[r]))


def test_eager_plan_counts_transitions():
    for n in (3, 6, 9):
        codes, coef = lattice.random_pauli_sum(n, 40, seed=50 + n, p_identity=0.4)
        codes = np.concatenate([codes, np.asarray([[3] * n, [0] * n], dtype=np.uint8)])
        coef = np.concatenate([coef, [1.0, -1.0]])  # ZZ..Z and I cancel on part of the diagonal
        op = tcp.PauliOperator.from_code_arrays(codes, coef)
        plan = op.native_mvp_plan(storage="eager")
        assert plan.storage == "eager" and plan.retained_format == "compact_graph"
        matrix = dense.dense_operator(n, *op.code_arrays())
        assert plan.transition_count == int(np.count_nonzero(matrix))
        state = _random_state(n, 3)
        assert rel_err(plan.apply(state), matrix @ state) <= FP64_TOL
    assert tcp.PauliOperator.from_terms(2, [("XX", 1.0)]).native_mvp_plan().transition_count is None


def test_large_state_properties_n24():
    import torch

    n = 24
    codes, coef = lattice.random_pauli_sum(n, 400, seed=61, p_identity=0.75)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)  # Hermitian: real coefficients
    plan = op.native_mvp_plan()
    u = _random_state(n, 62)
    w = _random_state(n, 63)
    du, dw = torch.from_numpy(u).cuda(), torch.from_numpy(w).cuda()
    hu, hw = plan.apply(du), plan.apply(dw)
    a, b = 0.3 - 1.1j, -0.7 + 0.2j
    hlin = plan.apply(a * du + b * dw)
    torch.cuda.synchronize()
    hu_n, hw_n, hlin_n = hu.cpu().numpy(), hw.cpu().numpy(), hlin.cpu().numpy()
    assert rel_err(hlin_n, a * hu_n + b * hw_n) <= 1e-13
    assert abs(np.vdot(w, hu_n) - np.vdot(hw_n, u)) <= 1e-12 * np.linalg.norm(hu_n)
    rows = np.random.default_rng(64).choice(1 << n, size=3000, replace=False).astype(np.uint64)
    codes_c, coef_c = op.code_arrays()
    expected = oracle.termdirect_rows(n, codes_c, coef_c, u, rows, nthreads=THREADS)
    assert rel_err(hu_n[rows.astype(np.int64)], expected) <= FP64_TOL


def test_config1_28q_full_size_rows_and_properties():
    # the 28-qubit workload itself (4x7 XXZ + field + 1,885 random terms, 2^28
    # complex128 amplitudes): linearity and Hermiticity on the device, and a
    # row sample against the oracle's TermDirect rows
    import torch

    n = 28
    codes, coef = lattice.heisenberg_4x7_with_random_terms(lattice.BASE_SEED, 1885)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)  # real coefficients: Hermitian
    plan = op.native_mvp_plan()
    g = torch.Generator(device="cuda").manual_seed(281)
    u = torch.randn(1 << n, dtype=torch.complex128, device="cuda", generator=g)
    w = torch.randn(1 << n, dtype=torch.complex128, device="cuda", generator=g)
    hu, hw = plan.apply(u), plan.apply(w)
    a, b = 0.3 - 1.1j, -0.7 + 0.2j
    hlin = plan.apply(a * u + b * w)
    lin_err = (torch.linalg.vector_norm(hlin - (a * hu + b * hw)) / torch.linalg.vector_norm(hlin)).item()
    assert lin_err <= 1e-13
    herm = abs((torch.vdot(w, hu) - torch.vdot(hw, u)).item())
    assert herm <= 1e-12 * torch.linalg.vector_norm(hu).item() * torch.linalg.vector_norm(w).item()
    del hw, hlin, w
    rows = np.random.default_rng(282).choice(1 << n, size=2000, replace=False).astype(np.uint64)
    u_host = u.cpu().numpy()
    codes_c, coef_c = op.code_arrays()
    expected = oracle.termdirect_rows(n, codes_c, coef_c, u_host, rows, nthreads=THREADS)
    got = hu[torch.from_numpy(rows.astype(np.int64)).cuda()].cpu().numpy()
    assert rel_err(got, expected) <= FP64_TOL


def test_concurrent_host_applies_from_threads():
    codes, coef = lattice.random_pauli_sum(12, 100, seed=71)
    plan = tcp.PauliOperator.from_code_arrays(codes, coef).native_mvp_plan()
    states = [_random_state(12, 80 + i) for i in range(6)]
    expected = [plan.apply(s) for s in states]
    results = [None] * 6

    def work(i):
        out = np.empty(4096, dtype=np.complex128)
        for _ in range(5):
            plan.apply_into(states[i], out)
        results[i] = out

    threads = [threading.Thread(target=work, args=(i,)) for i in range(6)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for r, e in zip(results, expected):
        np.testing.assert_array_equal(r, e)


def test_scipy_eigsh_matches_dense_spectrum():
    from scipy.sparse.linalg import eigsh

    codes, coef = lattice.heisenberg_periodic(2, 4, seed=3)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    linear = op.to_scipy_linear_operator()
    values = eigsh(linear, k=3, which="SA", return_eigenvectors=False, tol=1e-13, v0=np.ones(256))
    exact = np.linalg.eigvalsh(dense.dense_operator(8, *op.code_arrays()))[:3]
    np.testing.assert_allclose(np.sort(values), exact, atol=1e-10)


def test_scipy_linear_operator_contracts():
    # tests/test_scipy_linear_operator.py:26-101 of the reference, on this API
    op = tcp.PauliOperator.from_terms(2, (("XX", 0.5), ("ZI", -1.25j)))
    plan = op.native_mvp_plan()
    linear = plan.to_scipy_linear_operator()
    rng = np.random.default_rng(20260807)
    vector = rng.normal(size=4) + 1j * rng.normal(size=4)
    np.testing.assert_allclose(linear.matvec(vector), plan.apply(vector))
    np.testing.assert_allclose(linear.matvec(vector[:, None]), plan.apply(vector)[:, None])
    columns = np.column_stack((vector, 2.0 * vector))
    np.testing.assert_allclose(linear.matmat(columns), np.column_stack((plan.apply(vector), plan.apply(2 * vector))))
    assert linear.shape == (4, 4) and linear.dtype == np.dtype(np.complex128)
    for bad in (np.ones(3), np.ones((3, 1)), np.ones((1, 4)), np.ones((4, 1, 1))):
        with pytest.raises(ValueError, match=r"dimension mismatch|shape"):
            linear.matvec(bad)
    # the convenience form compiles one plan and reuses it
    calls = []
    original = tcp.PauliOperator.native_mvp_plan

    def counted(self, **kwargs):
        calls.append(1)
        return original(self, **kwargs)

    op2 = tcp.PauliOperator.from_terms(2, (("XX", 1.0), ("IZ", 0.25)))
    expected = op2.native_mvp_plan().apply(vector)
    tcp.PauliOperator.native_mvp_plan = counted
    try:
        lin2 = op2.to_scipy_linear_operator()
        np.testing.assert_allclose(lin2 @ vector, expected)
        np.testing.assert_allclose(lin2 @ (2.0 * vector), 2.0 * expected)
    finally:
        tcp.PauliOperator.native_mvp_plan = original
    assert len(calls) == 1
    with pytest.raises(TypeError):
        op2.to_scipy_linear_operator(mapping="jordan_wigner")
    np.testing.assert_allclose(op2.native_mvp_plan().to_scipy_linear_operator(max_bytes=1).matvec(vector), expected)
    assert not hasattr(op2.backend_mvp_plan(), "to_scipy_linear_operator")
    # restricted plans, both storages; no adjoint product
    hop = tcp.PauliOperator.from_terms(2, (("XX", 0.5), ("YY", 0.5), ("ZI", 0.3)))
    for storage in ("lazy", "eager"):
        rplan = hop.restrict_charge(tcp.U1Sector(2, 1), storage=storage).mvp_plan(storage=storage)
        lin = rplan.to_scipy_linear_operator()
        v2 = np.array([1.0 + 2.0j, -0.5 + 0.25j])
        np.testing.assert_allclose(lin.matvec(v2), rplan.apply(v2))
        with pytest.raises(NotImplementedError):
            lin.rmatvec(v2)
