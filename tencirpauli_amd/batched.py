"""Batched MVP: many Pauli operators of uneven size, each on a block of right-hand
sides, in one persistent GPU launch (``tcp_batch_plan_*`` in the C ABI).

No reference counterpart: the reference applies one plan per call
(core/hamiltonian.rs:204-270).  The batch buffer holds, for operator p, its
``nrhs`` vectors back to back starting at ``offsets[p]``.  ``autotune`` times
every tile geometry (threads per workgroup x amplitudes per thread) on the
caller's buffers and keeps the fastest, then the classes per work item on
it (``splits``: an operator with more classes is applied as several items
per output tile added with atomic adds, whose summation order varies from
run to run within rounding); ``capture`` records one apply into a
hipGraph that ``replay`` launches.  A captured graph is bound to the input
and output tensors it was captured with: the plan keeps references to them
(so their memory is not handed to other tensors while the graph can still
run) until the next ``capture`` or ``release_graph_buffers``.
"""

from __future__ import annotations

import ctypes
from typing import Any, Dict, List, Sequence

from . import _lib
from ._plan import current_stream_handle


def _dtype_code(tensor: Any) -> int:
    import torch

    if tensor.dtype == torch.complex128:
        return _lib.TCP_C128
    if tensor.dtype == torch.complex64:
        return _lib.TCP_C64
    raise TypeError("batch buffers must be complex128 or complex64 device tensors")


class BatchedMVP:
    def __init__(self, operators: Sequence[Any], nrhs: int) -> None:
        _lib.require_device()
        self.operators = list(operators)
        self.nrhs = int(nrhs)
        handles = (ctypes.c_void_p * max(1, len(self.operators)))(*[op._handle.ptr for op in self.operators])
        out = ctypes.c_void_p()
        _lib.check(_lib.lib().tcp_batch_plan_create(handles, len(self.operators), self.nrhs, ctypes.byref(out)))
        self._ptr = out
        L = _lib.lib()
        self.offsets: List[int] = [int(L.tcp_batch_plan_offset(out, p)) for p in range(len(self.operators))]
        self.total_amplitudes = int(L.tcp_batch_plan_total_amplitudes(out))
        self.work = float(L.tcp_batch_plan_work(out))  # term * amp * rhs per apply
        self._graph_buffers: Any = None

    def __del__(self) -> None:
        if getattr(self, "_ptr", None) and getattr(_lib, "_lib", None) is not None:
            _lib.lib().tcp_batch_plan_destroy(self._ptr)
            self._ptr = None

    @staticmethod
    def configs() -> List[Dict[str, int]]:
        out = []
        for i in range(int(_lib.lib().tcp_batch_config_count())):
            t, a = ctypes.c_uint32(), ctypes.c_uint32()
            _lib.check(_lib.lib().tcp_batch_config(i, ctypes.byref(t), ctypes.byref(a)))
            out.append({"threads": t.value, "amplitudes_per_thread": a.value})
        return out

    @staticmethod
    def splits() -> List[int]:
        """Classes per work item the autotuner tries (0: whole operators)."""
        out = []
        for i in range(int(_lib.lib().tcp_batch_split_count())):
            c = ctypes.c_uint32()
            _lib.check(_lib.lib().tcp_batch_split(i, ctypes.byref(c)))
            out.append(int(c.value))
        return out

    @property
    def selected(self) -> int:
        return int(_lib.lib().tcp_batch_plan_selected(self._ptr))

    @property
    def selected_split(self) -> int:
        return int(_lib.lib().tcp_batch_plan_selected_split(self._ptr))

    def select_split(self, index: int) -> None:
        _lib.check(_lib.lib().tcp_batch_plan_select_split(self._ptr, index))

    def select(self, config: int) -> None:
        _lib.check(_lib.lib().tcp_batch_plan_select(self._ptr, config))

    def block(self, buffer: Any, index: int) -> Any:
        """View of operator `index`'s right-hand sides, shape (nrhs, 2^n)."""
        dim = 1 << self.operators[index].nqubits
        start = self.offsets[index]
        return buffer[start:start + self.nrhs * dim].view(self.nrhs, dim)

    def _check(self, inp: Any, out: Any) -> int:
        if inp.numel() != self.total_amplitudes or out.numel() != self.total_amplitudes:
            raise ValueError(f"batch buffers must hold {self.total_amplitudes} amplitudes")
        if inp.dtype != out.dtype:
            raise TypeError("input and output must have the same dtype")
        if not (inp.is_contiguous() and out.is_contiguous()):
            raise ValueError("batch buffers must be contiguous")
        return _dtype_code(inp)

    def apply(self, inp: Any, out: Any) -> None:
        code = self._check(inp, out)
        _lib.check(_lib.lib().tcp_batch_plan_apply(self._ptr, inp.data_ptr(), out.data_ptr(), code,
                                                   current_stream_handle(inp)))

    def autotune(self, inp: Any, out: Any, repeats: int = 3) -> Dict[str, Any]:
        import numpy as np

        code = self._check(inp, out)
        count = int(_lib.lib().tcp_batch_config_count())
        times = np.zeros(count, dtype=np.float32)
        chosen = ctypes.c_uint32()
        _lib.check(_lib.lib().tcp_batch_plan_autotune(self._ptr, inp.data_ptr(), out.data_ptr(), code, repeats,
                                                      times.ctypes.data, ctypes.byref(chosen)))
        split_times = np.zeros(int(_lib.lib().tcp_batch_split_count()), dtype=np.float32)
        _lib.check(_lib.lib().tcp_batch_plan_split_times(self._ptr, split_times.ctypes.data))
        return {"times_ms": times.tolist(), "chosen": int(chosen.value), "configs": self.configs(),
                "split_times_ms": split_times.tolist(), "chosen_split": self.selected_split,
                "classes_per_item": self.splits()}

    def capture(self, inp: Any, out: Any, stream: Any = None) -> None:
        """Record one apply of `inp` -> `out` into a hipGraph; `replay` reruns it on these tensors."""
        code = self._check(inp, out)
        handle = int(stream.cuda_stream) if stream is not None else 0
        _lib.check(_lib.lib().tcp_batch_plan_capture(self._ptr, inp.data_ptr(), out.data_ptr(), code,
                                                     handle or None))
        self._graph_buffers = (inp, out)

    def replay(self, stream: Any = None) -> None:
        """Launch the captured graph on `stream` (reads and writes the captured tensors)."""
        if self._graph_buffers is None:
            raise ValueError("no captured graph")
        handle = int(stream.cuda_stream) if stream is not None else 0
        _lib.check(_lib.lib().tcp_batch_plan_replay(self._ptr, handle or None))
        if stream is not None:
            for tensor in self._graph_buffers:  # the allocator waits for this stream before reusing them
                tensor.record_stream(stream)

    def release_graph_buffers(self) -> None:
        """Drop the references to the captured tensors; replay is refused until the next capture."""
        self._graph_buffers = None
