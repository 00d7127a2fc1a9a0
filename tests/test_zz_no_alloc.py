"""GPU checks of tcp_mvp_apply_static, the full-space MVP entry point that
allocates no device memory (its tables live in a static per-device
workspace).  Called through the C ABI, it must equal
tcp_mvp_plan_apply_into on a plan of the same operator bit for bit -- the
same tables, geometry and kernel -- on consecutive calls whose second one is
smaller (the workspace keeps the previous call's contents), in both storage
precisions, and it must refuse operators whose tables exceed the workspace.
"""

from __future__ import annotations

import numpy as np
import pytest

import tencirpauli_amd as tcp
from tencirpauli_amd import _lib, lattice

pytestmark = pytest.mark.gpu


def _static_apply(op, v, y) -> int:
    import torch

    code = _lib.TCP_C128 if v.dtype == torch.complex128 else _lib.TCP_C64
    stream = torch.cuda.current_stream(v.device).cuda_stream
    return int(_lib.lib().tcp_mvp_apply_static(op._handle.ptr, v.data_ptr(), y.data_ptr(), v.numel(), y.numel(),
                                               code, stream or None))


def _state(n, dtype, seed):
    import torch

    g = torch.Generator(device="cuda").manual_seed(seed)
    return torch.randn(1 << n, dtype=torch.complex128, device="cuda", generator=g).to(dtype)


@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_static_entry_equals_plan_apply_on_consecutive_calls(precision):
    import torch

    dtype = getattr(torch, precision)
    cases = [(24, 1500, 31), (13, 300, 32), (20, 1024, 33)]  # second call smaller than the first
    for n, terms, seed in cases:
        codes, coef = lattice.random_pauli_sum(n, terms, seed=seed)
        op = tcp.PauliOperator.from_code_arrays(codes, coef)
        v = _state(n, dtype, seed)
        expected = torch.empty_like(v)
        op.native_mvp_plan().apply_into(v, expected)
        got = torch.full_like(v, float("nan"))
        assert _static_apply(op, v, got) == 0, _lib.last_error()
        torch.cuda.synchronize()
        assert torch.equal(got, expected), (precision, n)


def test_static_entry_empty_operator_and_argument_errors():
    import torch

    op = tcp.PauliOperator.empty(12)
    v = _state(12, torch.complex128, 4)
    y = torch.full_like(v, 1.0)
    assert _static_apply(op, v, y) == 0
    torch.cuda.synchronize()
    assert torch.count_nonzero(y).item() == 0
    L = _lib.lib()
    bad = L.tcp_mvp_apply_static(op._handle.ptr, v.data_ptr(), y.data_ptr(), v.numel() - 1, y.numel(),
                                 _lib.TCP_C128, None)
    assert bad == _lib.TCP_ERR_INVALID_STRUCTURE_LENGTH
    assert L.tcp_mvp_apply_static(op._handle.ptr, v.data_ptr(), v.data_ptr(), v.numel(), v.numel(),
                                  _lib.TCP_C128, None) == _lib.TCP_ERR_INVALID_ARGUMENT


def test_static_entry_refuses_tables_beyond_its_workspace():
    import torch

    rng_codes, rng_coef = lattice.random_pauli_sum(16, 300000, seed=35, p_identity=0.2)
    op = tcp.PauliOperator.from_code_arrays(rng_codes, rng_coef)
    assert op.term_count > 270000
    v = _state(16, torch.complex128, 5)
    y = torch.zeros_like(v)
    status = _static_apply(op, v, y)
    assert status == _lib.TCP_ERR_UNSUPPORTED
    assert "static workspace" in _lib.last_error()
    # the plan path serves the same operator
    expected = torch.empty_like(v)
    op.native_mvp_plan().apply_into(v, expected)
    torch.cuda.synchronize()
    assert np.isfinite(expected.abs().max().item())
