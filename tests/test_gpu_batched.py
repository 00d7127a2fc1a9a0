"""GPU checks of the batched persistent MVP: every operator/right-hand side of
the batch equals the single-operator plan (itself pinned to the oracle), for
every tile geometry, in both precisions, eagerly and as replayed hipGraphs."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
import tencirpauli_amd as tcp
from tencirpauli_amd import lattice
from tencirpauli_amd.batched import BatchedMVP
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _batch(seed, count=10, nmin=9, nmax=14, tmin=16, tmax=300):
    rng = np.random.default_rng(seed)
    ops = []
    for i in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        t = int(rng.integers(tmin, tmax + 1))
        codes, coef = lattice.random_pauli_sum(n, t, seed=seed * 1000 + i)
        ops.append(tcp.PauliOperator.from_code_arrays(codes, coef))
    ops.append(tcp.PauliOperator.empty(nmin))  # an empty operator in the batch
    return ops


def _inputs(batch, dtype, seed=5):
    import torch

    gen = torch.Generator(device="cuda").manual_seed(seed)
    v = torch.randn(batch.total_amplitudes, dtype=torch.complex128, device="cuda", generator=gen)
    return v.to(dtype)


def _check_against_single(batch, ops, v, y, tol):
    import torch

    for p, op in enumerate(ops):
        plan = op.native_mvp_plan()
        vin = batch.block(v, p)
        expected = torch.empty_like(vin.to(torch.complex128))
        plan.apply_batch(vin.to(torch.complex128).contiguous(), expected)
        got = batch.block(y, p).to(torch.complex128)
        torch.cuda.synchronize()
        e = expected.cpu().numpy()
        g = got.cpu().numpy()
        if np.linalg.norm(e) == 0:
            assert np.all(g == 0)
        else:
            assert rel_err(g, e) <= tol, (p, op.nqubits, op.term_count)


def test_batched_matches_single_plans_every_geometry():
    import torch

    ops = _batch(1)
    batch = BatchedMVP(ops, nrhs=3)
    v = _inputs(batch, torch.complex128)
    for cfg in range(len(batch.configs())):
        batch.select(cfg)
        y = torch.full_like(v, 7.0 - 3.0j)
        batch.apply(v, y)
        torch.cuda.synchronize()
        _check_against_single(batch, ops, v, y, 1e-12)


def test_batched_against_oracle():
    import torch

    ops = _batch(2, count=4)
    batch = BatchedMVP(ops, nrhs=2)
    v = _inputs(batch, torch.complex128, seed=9)
    y = torch.empty_like(v)
    batch.apply(v, y)
    torch.cuda.synchronize()
    for p, op in enumerate(ops[:-1]):
        codes, coef = op.code_arrays()
        for r in range(2):
            vin = batch.block(v, p)[r].cpu().numpy()
            expected = oracle.termdirect_apply(op.nqubits, codes, coef, vin, nthreads=8)
            assert rel_err(batch.block(y, p)[r].cpu().numpy(), expected) <= 1e-12


def test_batched_fp32_storage():
    import torch

    ops = _batch(3)
    batch = BatchedMVP(ops, nrhs=2)
    v = _inputs(batch, torch.complex64, seed=4)
    y = torch.empty_like(v)
    batch.apply(v, y)
    torch.cuda.synchronize()
    _check_against_single(batch, ops, v.to(torch.complex128), y, 2e-6)


def test_autotune_and_graph_replay():
    import torch

    ops = _batch(4, count=16, nmax=16)
    batch = BatchedMVP(ops, nrhs=4)
    v = _inputs(batch, torch.complex64, seed=6)
    y = torch.empty_like(v)
    report = batch.autotune(v, y, repeats=2)
    assert 0 <= report["chosen"] < len(report["configs"])
    assert report["times_ms"][report["chosen"]] > 0
    batch.apply(v, y)
    torch.cuda.synchronize()
    eager = y.clone()
    stream = torch.cuda.Stream()
    batch.capture(v, y, stream)
    exact = report["chosen_split"] == 0  # split items add with atomics in varying order
    for i in range(4):
        y.fill_(0)
        torch.cuda.synchronize()
        _ = (v * 2).sum()  # eager work between replays
        batch.replay(stream)
        stream.synchronize()
        torch.testing.assert_close(y, eager, rtol=0 if exact else 1e-5, atol=0 if exact else 1e-5)


def test_split_items_match_single_plans_and_replay():
    # a few operators with many classes and few tiles: split-K over their
    # classes gives the device more items; every split against the single
    # plans in both precisions, then a captured split apply replayed over a
    # NaN-filled output (the graph clears it)
    import torch

    ops = []
    for i, n in enumerate((16, 17, 18)):
        codes, coef = lattice.random_pauli_sum(n, 1500, seed=70 + i, p_identity=0.3)
        ops.append(tcp.PauliOperator.from_code_arrays(codes, coef))
    batch = BatchedMVP(ops, nrhs=2)
    splits = batch.splits()
    assert splits[0] == 0 and all(s > 0 for s in splits[1:])
    v = _inputs(batch, torch.complex128, seed=11)
    for k in range(len(splits)):
        batch.select_split(k)
        assert batch.selected_split == k
        y = torch.full_like(v, float("nan"))
        batch.apply(v, y)
        torch.cuda.synchronize()
        _check_against_single(batch, ops, v, y, 1e-12)
        y32 = torch.full_like(v.to(torch.complex64), float("nan"))
        batch.apply(v.to(torch.complex64), y32)
        torch.cuda.synchronize()
        _check_against_single(batch, ops, v, y32, 2e-6)
    report = batch.autotune(v, torch.empty_like(v), repeats=1)
    assert report["split_times_ms"][0] > 0 and 0 <= report["chosen_split"] < len(splits)
    batch.select_split(len(splits) - 1)
    y = torch.empty_like(v)
    stream = torch.cuda.Stream()
    batch.capture(v, y, stream)
    for _ in range(3):
        y.fill_(float("nan"))
        torch.cuda.synchronize()
        batch.replay(stream)
        stream.synchronize()
        _check_against_single(batch, ops, v, y, 1e-12)
