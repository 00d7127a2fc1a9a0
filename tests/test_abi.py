"""CPU checks of the C ABI: the library loads, exports every declared symbol,
and its host-side logic (canonicalisation, packing, charge analysis, sector
rank/unrank, error taxonomy) matches the oracle.  No kernel is launched."""

from __future__ import annotations

import ctypes
import math
import os
import re

import numpy as np
import pytest

import oracle
import tencirpauli_amd as tcp
from oracle import dense
from tencirpauli_amd import _lib, _plan
from tests.helpers import conserving_u1_terms

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    include = os.path.join(ROOT, "include")
    for fname in os.listdir(include):
        if fname.endswith(".h"):
            text = open(os.path.join(include, fname)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names.update(re.findall(r"\b(tcp_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 40
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    assert set(_lib.SIGNATURES) == set(names)
    assert lib.tcp_version().decode().startswith("tencirpauli_amd")


def test_operator_canonicalisation_matches_oracle():
    rng = np.random.default_rng(20260801)
    for n in range(0, 13):
        codes = rng.integers(0, 4, size=(60, n)).astype(np.uint8)
        codes[30:] = codes[:30]
        coef = rng.normal(size=60) + 1j * rng.normal(size=60)
        coef[5] = 0.0
        op = tcp.PauliOperator.from_code_arrays(codes, coef)
        cc, cf = oracle.canonicalize(n, codes, coef)
        got_codes, got_coef = op.code_arrays()
        np.testing.assert_array_equal(got_codes, cc)
        np.testing.assert_array_equal(got_coef, cf)  # bit-exact sums


def test_bit_pattern_summation_drops_exact_zero():
    op = tcp.PauliOperator.from_terms(1, [("X", 1e16), ("X", -1e16), ("X", 1.0), ("Z", 2.0)])
    assert [t.codes for t in op.terms] == [(3,)]


def test_string_terms_and_backend_arrays():
    op = tcp.PauliOperator.from_terms(3, [("XYZ", 0.5), ("ZII", 1.0), ("XYZ", 0.25)])
    assert [t.to_string() for t in op.terms] == ["XYZ", "ZII"]
    plan = op.backend_mvp_plan()
    assert plan.x_words.dtype == np.uint64 and plan.x_words.shape == (2, 1)
    # qubit q at bit q (core/word.rs:94-110): XYZ -> x = 0b011, z = 0b110
    assert plan.x_words[0, 0] == 0b011 and plan.z_words[0, 0] == 0b110
    assert plan.z_words[1, 0] == 0b001
    np.testing.assert_array_equal(plan.coefficients, [0.75, 1.0])
    assert not plan.x_words.flags.writeable
    assert op.distinct_x_mask_count() == 2


def test_operator_from_words_round_trip():
    rng = np.random.default_rng(3)
    for n in (1, 7, 63, 64, 65, 130):
        codes = rng.integers(0, 4, size=(12, n)).astype(np.uint8)
        coef = rng.normal(size=12) + 0j
        op = tcp.PauliOperator.from_code_arrays(codes, coef)
        arrays = op.backend_mvp_plan()
        out = ctypes.c_void_p()
        u64 = ctypes.POINTER(ctypes.c_uint64)
        cf = np.ascontiguousarray(arrays.coefficients)
        xw = np.ascontiguousarray(arrays.x_words)
        zw = np.ascontiguousarray(arrays.z_words)
        _lib.check(_lib.lib().tcp_operator_from_words(n, len(cf), xw.ctypes.data_as(u64), zw.ctypes.data_as(u64),
                                                      cf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                      ctypes.byref(out)))
        assert _lib.lib().tcp_operator_term_count(out) == op.term_count
        codes2 = np.zeros((op.term_count, n), dtype=np.uint8)
        _lib.check(_lib.lib().tcp_operator_export_codes(out, codes2.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                                        None))
        np.testing.assert_array_equal(codes2, op.code_arrays()[0])
        _lib.lib().tcp_operator_destroy(out)


def test_error_taxonomy():
    with pytest.raises(ValueError, match="invalid Pauli code"):
        tcp.PauliOperator.from_code_arrays([[0, 4]], [1.0])
    with pytest.raises(ValueError, match="not finite"):
        tcp.PauliOperator.from_terms(1, [("X", float("inf"))])
    with pytest.raises(ValueError, match="expected structure length"):
        tcp.PauliOperator.from_terms(2, [("XYZ", 1.0)])
    # n >= 64 is an overflow before any device work (core/hamiltonian.rs:544-555)
    op = tcp.PauliOperator.empty(64)
    out = ctypes.c_void_p()
    status = _lib.lib().tcp_mvp_plan_create(op._handle.ptr, 0, 0, ctypes.byref(out))
    assert status == _lib.TCP_ERR_OVERFLOW
    with pytest.raises(OverflowError, match="matrix dimension"):
        _lib.check(status)
    status = _lib.lib().tcp_mvp_plan_create(op._handle.ptr, 7, 0, ctypes.byref(out))
    assert status == _lib.TCP_ERR_INVALID_ARGUMENT


def test_no_cpu_fallback_without_device():
    if _lib.device_count() > 0:
        pytest.skip("a HIP device is visible; this checks the device-less failure path")
    op = tcp.PauliOperator.from_terms(1, [("X", 1.0)])
    with pytest.raises(tcp.DeviceError):
        op.native_mvp_plan()
    with pytest.raises(tcp.DeviceError):
        op.mvp(np.ones(2, dtype=np.complex128))
    with pytest.raises(tcp.DeviceError):
        op.restrict_charge(tcp.U1Sector(1, 1))


def test_analyze_charge_matches_oracle():
    rng = np.random.default_rng(20260802)
    for n in range(1, 9):
        codes, coef = conserving_u1_terms(rng, n, 10)
        extra = rng.integers(0, 4, size=(3, n)).astype(np.uint8)
        for c, f in ((codes, coef), (np.concatenate([codes, extra]), np.concatenate([coef, np.ones(3)]))):
            op = tcp.PauliOperator.from_code_arrays(c, f)
            cc, cf = op.code_arrays()
            levels = [(0.0, 1.0)] * n
            got = op.analyze_charge(tcp.AdditiveCharge.number(n, range(n)))
            assert (got.is_conserved, got.nonzero_commutator_terms) == oracle.analyze_charge(n, cc, cf, levels)


def test_u1_host_rank_unrank_bit_exact():
    for n in range(0, 13):
        for k in range(n + 1):
            sector = tcp.U1Sector(n, k)
            basis = dense.u1_basis(n, k)
            assert [sector.rank(b) for b in basis] == list(range(len(basis)))
            assert [sector.unrank_int(i) for i in range(len(basis))] == basis
    rng = np.random.default_rng(9)
    for n in (25, 40, 63, 64):
        for k in (1, n // 2, n // 2 + 1, n - 1):
            dim = math.comb(n, k)
            idx = np.asarray(sorted({0, dim - 1, *(int(i) for i in rng.integers(0, dim, size=40, dtype=np.uint64))}),
                             dtype=np.uint64)
            expected = oracle.u1_unrank_many(n, k, idx)
            sector = tcp.U1Sector(n, k)
            assert [sector.unrank_int(int(i)) for i in idx] == [int(b) for b in expected]
            assert [sector.rank(int(b)) for b in expected] == [int(i) for i in idx]


def test_u1_known_basis_orders():
    for n, k, expected in [(0, 0, [0]), (3, 0, [0]), (3, 1, [1, 2, 4]), (3, 2, [3, 5, 6]), (3, 3, [7])]:
        sector = tcp.U1Sector(n, k)
        assert sector.dimension == len(expected)
        bits = [tuple((v >> (n - 1 - q)) & 1 for q in range(n)) for v in expected]
        assert [sector.unrank(i) for i in range(sector.dimension)] == bits
        assert [sector.rank(v) for v in expected] == list(range(len(expected)))
        basis = sector.basis_states()
        np.testing.assert_array_equal(basis, np.asarray(bits, dtype=np.uint8).reshape(len(bits), n))
        assert basis.flags.writeable is False
    with pytest.raises(ValueError, match="wrong Hamming weight"):
        tcp.U1Sector(3, 1).rank(3)
    with pytest.raises(IndexError):
        tcp.U1Sector(3, 1).unrank(3)


def test_block_sector_rank_matches_oracle():
    for blocks in ([(3, 1), (3, 2)], [(4, 2), (4, 2)], [(2, 1), (3, 0), (3, 3)], [(16, 4), (16, 4)]):
        n = sum(s for s, _ in blocks)
        charges = []
        start = 0
        for size, pop in blocks:
            charges.append((tcp.AdditiveCharge.number(n, range(start, start + size)), pop))
            start += size
        sector = tcp.ChargeSector(charges[::-1])  # constraint order does not matter
        assert sector.blocks == tuple(blocks)
        assert sector.dimension == oracle.blocks_dimension(blocks)
        idx = np.unique(np.linspace(0, sector.dimension - 1, num=min(sector.dimension, 300)).astype(np.uint64))
        bits = oracle.blocks_unrank_many(blocks, idx)
        for i, b in zip(idx, bits):
            occ = tuple((int(b) >> (n - 1 - q)) & 1 for q in range(n))
            assert sector.unrank(int(i)) == occ
            assert sector.rank(occ) == int(i)
    three = tcp.ChargeSector(((tcp.AdditiveCharge.number(3, range(3)), 1),))
    assert [three.unrank(i) for i in range(3)] == [(0, 0, 1), (0, 1, 0), (1, 0, 0)]


def _fermion_charges(n, levels):
    space = tcp.OperatorSpace(fermions=n)
    return [tcp.AdditiveCharge(space, fermions={a: int(levels[j, a, 1]) for a in range(n)})
            for j in range(levels.shape[0])]


def test_general_charge_sector_rank_matches_oracle():
    # interleaved spin orbitals (N, 2 S_z) and weighted qubit charges: the
    # sectors that are not contiguous blocks use the native suffix-count tables
    from tests.helpers import spin_levels

    for n, targets in ((10, (4, 0)), (12, (5, -1)), (32, (8, 0))):
        levels = spin_levels(n)
        sector = tcp.ChargeSector(list(zip(_fermion_charges(n, levels), targets)))
        assert sector.blocks is None and sector.execution_family == "generic"
        assert sector.dimension == oracle.charge_dimension(levels, targets)
        idx = np.unique(np.linspace(0, sector.dimension - 1, num=min(sector.dimension, 300)).astype(np.uint64))
        words = oracle.charge_unrank_many(levels, targets, idx)
        for i, w in zip(idx, words):
            occ = tuple((int(w) >> (n - 1 - q)) & 1 for q in range(n))
            assert sector.unrank(int(i)) == occ
            assert sector.rank(occ) == int(i)
        if sector.dimension <= 5000:
            basis = sector.basis_states()
            assert basis.dtype == np.uint64 and basis.flags.writeable is False
            np.testing.assert_array_equal(basis, np.asarray(
                [[(int(w) >> (n - 1 - q)) & 1 for q in range(n)] for w in dense.charge_basis(levels, targets)]))
    weighted = tcp.AdditiveCharge(7, {0: (0, 2), 1: (1, -1), 2: (0, 3), 4: (2, 1), 5: (0, 1), 6: (-1, 1)})
    sector = weighted.sector(2)
    lv = np.asarray([weighted.qubit_levels], dtype=np.int64)
    assert [int("".join(map(str, r)), 2) for r in sector.basis_states()] == dense.charge_basis(lv, [2])
    with pytest.raises(ValueError, match="constraint"):
        sector.rank((1, 1, 1, 1, 1, 1, 1))
    with pytest.raises(IndexError):
        sector.unrank(sector.dimension)


def test_general_charge_sector_edges():
    space = tcp.OperatorSpace(fermions=4)
    number = tcp.AdditiveCharge(space, fermions={m: 1 for m in range(4)})
    marked = tcp.AdditiveCharge(space, fermions={0: 1})
    # an unreachable target is an empty sector, as build_charge_sector_plan gives
    empty = tcp.ChargeSector(((number, 1), (marked, 2)))
    assert empty.dimension == 0 and empty.basis_states().shape == (0, 4)
    # unweighted axes are free spectators (tests/test_charge.py:855-861)
    assert marked.sector(0, max_bytes=1 << 20).dimension == 8
    with pytest.raises(ValueError, match="integer"):
        tcp.ChargeSector(((tcp.AdditiveCharge(3, {0: (0, 0.5), 2: (0, 1)}), 1),))
    # 64 axes exceed the general tables; 2^24 table entries bound the levels
    wide = tcp.OperatorSpace(fermions=64)
    interleaved = tcp.AdditiveCharge(wide, fermions={m: (1 if m % 2 else -1) for m in range(64)})
    with pytest.raises(tcp.UnsupportedOnDeviceError, match="63"):
        tcp.ChargeSector(((interleaved, 0),))
    # the remainder box holds only what a prefix and a suffix can both reach:
    # powers of two with a small target stay small, a middle target does not
    heavy = tcp.AdditiveCharge(24, {a: (0, 1 << a) for a in range(24)})
    five = heavy.sector(5)  # distinct powers of two: only the binary digits of 5
    assert five.dimension == 1 and five.unrank(0) == tuple(1 if q in (0, 2) else 0 for q in range(24))
    with pytest.raises(tcp.UnsupportedOnDeviceError, match="2\\^24"):
        heavy.sector(1 << 23)
    # two population solutions: a union of block layouts; budgets below the
    # enforced floor are advisory (tests/test_charge.py:882-897)
    union = tcp.ChargeSector(((number, 2), (marked, 1), (tcp.AdditiveCharge(space, fermions={1: 1, 3: 1}), 1)),
                             max_bytes=16)
    assert union.blocks is None and [int("".join(map(str, union.unrank(i))), 2) for i in range(2)] == [0b1001, 0b1100]


def test_apply_into_buffer_protocol():
    good = np.zeros(4, dtype=np.complex128)
    with pytest.raises(TypeError, match="NumPy arrays"):
        _plan.validate_apply_into_buffers([0, 0, 0, 0], good, 4)
    with pytest.raises(TypeError, match="complex128"):
        _plan.validate_apply_into_buffers(np.zeros(4, dtype=np.complex64), good, 4)
    with pytest.raises(ValueError, match="shape"):
        _plan.validate_apply_into_buffers(np.zeros(3, dtype=np.complex128), good, 4)
    with pytest.raises(ValueError, match="C-contiguous"):
        _plan.validate_apply_into_buffers(np.zeros(8, dtype=np.complex128)[::2], good, 4)
    with pytest.raises(ValueError, match="overlap"):
        _plan.validate_apply_into_buffers(good, good, 4)
    ro = np.zeros(4, dtype=np.complex128)
    ro.flags.writeable = False
    with pytest.raises(ValueError, match="writable"):
        _plan.validate_apply_into_buffers(good, ro, 4)
    with pytest.raises(ValueError, match="max_bytes"):
        _plan.validate_max_bytes(-1)
    assert _plan.enforced_max_bytes(1) is None
    with pytest.raises(MemoryError):
        _plan.check_allocation(2 << 30, 1 << 30, "output")


def test_bench_refuses_more_than_one_gpu():
    # the measured path is one GPU's share; --gpus above 1 exits with a message and status 2
    import json
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    proc = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                          text=True, cwd=root, timeout=120)
    assert proc.returncode == 2
    assert "one GPU" in json.loads(proc.stdout.strip().splitlines()[-1])["error"]


def test_davidson_rejects_a_preconditioner_of_another_dimension(monkeypatch):
    # The learned preconditioner's callback returns a bare float buffer; a hook
    # built for another plan must be refused before the solver launches
    # anything (and the ABI carries the buffer length for the same check).
    from tencirpauli_amd import davidson

    class Inner:
        dimension = 924

    calls = []
    monkeypatch.setattr(davidson, "_plan_handle",
                        lambda plan: (None, lambda *a: calls.append(a) or 0, None, Inner()))

    class Hook:
        dimension = 495

        def callback(self):  # pragma: no cover - must not be reached
            raise AssertionError("callback requested for a mismatched plan")

    with pytest.raises(ValueError, match="expected structure length 924, got 495"):
        davidson.lowest_eigenpair(object(), preconditioner=Hook())
    assert not calls
    names = [f[0] for f in davidson.DavidsonOptions._fields_]
    assert names.index("log_multiplier_len") == names.index("log_multiplier_ctx") + 1
    header = open(os.path.join(ROOT, "include", "tencirpauli_amd.h")).read()
    assert re.search(r"void \*log_multiplier_ctx;\s*uint64_t log_multiplier_len;", header)
