"""ctypes binding of ``libtencirpauli_amd.so`` (the C ABI in include/tencirpauli_amd.h).

The shared library is required: there is no CPU fallback.  Importing a
function that needs it raises ``NativeLibraryError`` when the library is
missing, and every status code is turned into the exception class the
reference's binding uses for the same ``PauliError`` variant
(native/convert.rs:35-42: MemoryLimit -> MemoryError, Overflow ->
OverflowError, everything else -> ValueError).
"""

from __future__ import annotations

import ctypes
import os
import threading
from typing import Optional

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtencirpauli_amd.so")

TCP_OK = 0
TCP_ERR_INVALID_STRUCTURE_LENGTH = 1
TCP_ERR_INVALID_CODE = 2
TCP_ERR_NON_FINITE_COEFFICIENT = 3
TCP_ERR_MEMORY_LIMIT = 4
TCP_ERR_OVERFLOW = 5
TCP_ERR_INVALID_SECTOR = 6
TCP_ERR_SECTOR_LEAKAGE = 7
TCP_ERR_INVALID_INDEX = 8
TCP_ERR_INVALID_ARGUMENT = 9
TCP_ERR_INCOMPATIBLE_QUBIT_COUNTS = 10
TCP_ERR_INVALID_WORD_LENGTH = 11
TCP_ERR_HIP = 12
TCP_ERR_UNSUPPORTED = 13

TCP_C128 = 0
TCP_C64 = 1
TCP_STORAGE_LAZY = 0
TCP_STORAGE_EAGER = 1
TCP_TILE_AUTO = 0
TCP_TILE_512x8 = 1
TCP_TILE_1024x8 = 2
TCP_TILE_512x16 = 3
TCP_TILE_512x4 = 4
TCP_TILE_1024x4 = 5
TCP_TILE_1024x8x2 = 6
TCP_TILE_512x16x2 = 7
TCP_TILE_XCD_ORDER = 0x100
TCP_TILE_LINEAR_ORDER = 0x200
TCP_TILE_PREFETCH_SHIFT = 10
TCP_TILE_PREFETCH_MASK = 0xC00
TCP_TILE_NT_LOADS = 0x1000

# Every exported entry point: (name, restype, argtypes).
_c = ctypes
_vp = _c.c_void_p
_u8p = _c.POINTER(_c.c_uint8)
_u32p = _c.POINTER(_c.c_uint32)
_u64p = _c.POINTER(_c.c_uint64)
_dp = _c.POINTER(_c.c_double)
_ip = _c.POINTER(_c.c_int)

SIGNATURES = {
    "tcp_version": (_c.c_char_p, []),
    "tcp_last_error": (_c.c_size_t, [_c.c_char_p, _c.c_size_t]),
    "tcp_device_count": (_c.c_int, [_ip]),
    "tcp_operator_from_codes": (_c.c_int, [_c.c_uint32, _c.c_uint64, _u8p, _dp, _c.POINTER(_vp)]),
    "tcp_fermion_from_ladder": (_c.c_int, [_c.c_uint32, _c.c_uint64, _u64p, _c.POINTER(_c.c_uint32),
                                            _c.POINTER(_c.c_uint8), _dp, _c.c_uint64, _c.POINTER(_vp)]),
    "tcp_fermion_n_modes": (_c.c_uint32, [_vp]),
    "tcp_fermion_term_count": (_c.c_uint64, [_vp]),
    "tcp_fermion_factor_count": (_c.c_uint64, [_vp]),
    "tcp_fermion_export": (_c.c_int, [_vp, _u64p, _c.POINTER(_c.c_uint32), _c.POINTER(_c.c_uint8), _dp]),
    "tcp_fermion_jordan_wigner": (_c.c_int, [_vp, _c.c_uint64, _c.POINTER(_vp)]),
    "tcp_fermion_destroy": (None, [_vp]),
    "tcp_operator_add": (_c.c_int, [_vp, _vp, _c.POINTER(_vp)]),
    "tcp_operator_scale": (_c.c_int, [_vp, _c.c_double, _c.c_double, _c.POINTER(_vp)]),
    "tcp_operator_multiply": (_c.c_int, [_vp, _vp, _c.c_uint64, _c.POINTER(_vp)]),
    "tcp_operator_adjoint": (_c.c_int, [_vp, _c.POINTER(_vp)]),
    "tcp_operator_from_words": (_c.c_int, [_c.c_uint32, _c.c_uint64, _u64p, _u64p, _dp, _c.POINTER(_vp)]),
    "tcp_operator_nqubits": (_c.c_uint32, [_vp]),
    "tcp_operator_term_count": (_c.c_uint64, [_vp]),
    "tcp_operator_distinct_x_mask_count": (_c.c_uint64, [_vp]),
    "tcp_operator_export_codes": (_c.c_int, [_vp, _u8p, _dp]),
    "tcp_operator_export_words": (_c.c_int, [_vp, _u64p, _u64p, _dp]),
    "tcp_operator_analyze_charge": (_c.c_int, [_vp, _dp, _ip, _u64p]),
    "tcp_operator_destroy": (None, [_vp]),
    "tcp_mvp_plan_create": (_c.c_int, [_vp, _c.c_int, _c.c_uint64, _c.POINTER(_vp)]),
    "tcp_mvp_plan_create_from_words": (
        _c.c_int, [_c.c_uint32, _c.c_uint64, _u64p, _u64p, _dp, _c.c_int, _c.c_uint64, _c.POINTER(_vp)]),
    "tcp_mvp_plan_nqubits": (_c.c_uint32, [_vp]),
    "tcp_mvp_plan_term_count": (_c.c_uint64, [_vp]),
    "tcp_mvp_plan_group_count": (_c.c_uint64, [_vp]),
    "tcp_mvp_plan_class_count": (_c.c_uint64, [_vp]),
    "tcp_mvp_plan_pass_count": (_c.c_uint64, [_vp]),
    "tcp_mvp_plan_class_count_for": (_c.c_uint64, [_vp, _c.c_int]),
    "tcp_mvp_plan_create_tiled": (_c.c_int, [_vp, _c.c_int, _c.c_uint64, _c.c_uint32, _c.POINTER(_vp)]),
    "tcp_mvp_plan_tiling": (_c.c_uint32, [_vp, _c.c_int]),
    "tcp_mvp_plan_tile_bits": (_c.c_uint32, [_vp, _c.c_int]),
    "tcp_mvp_plan_dimension": (_c.c_uint64, [_vp]),
    "tcp_mvp_plan_storage": (_c.c_int, [_vp]),
    "tcp_mvp_plan_transition_count": (_c.c_uint64, [_vp]),
    "tcp_mvp_plan_estimated_bytes": (_c.c_uint64, [_vp]),
    "tcp_mvp_plan_bytes_alg": (_c.c_uint64, [_vp, _c.c_int]),
    "tcp_mvp_plan_apply_into": (_c.c_int, [_vp, _vp, _vp, _c.c_uint64, _c.c_uint64, _c.c_int, _vp]),
    "tcp_mvp_plan_apply_batch": (_c.c_int, [_vp, _vp, _vp, _c.c_uint64, _c.c_int, _vp]),
    "tcp_mvp_plan_apply_host": (_c.c_int, [_vp, _dp, _dp, _c.c_uint64, _c.c_uint64]),
    "tcp_mvp_plan_destroy": (None, [_vp]),
    "tcp_mvp": (_c.c_int, [_vp, _dp, _dp, _c.c_uint64, _c.c_uint64]),
    "tcp_mvp_apply_static": (_c.c_int, [_vp, _vp, _vp, _c.c_uint64, _c.c_uint64, _c.c_int, _vp]),
    "tcp_u1_dimension": (_c.c_int, [_c.c_uint32, _c.c_uint32, _u64p]),
    "tcp_u1_rank": (_c.c_int, [_c.c_uint32, _c.c_uint32, _c.c_uint64, _u64p]),
    "tcp_u1_unrank": (_c.c_int, [_c.c_uint32, _c.c_uint32, _c.c_uint64, _u64p]),
    "tcp_u1_rank_words": (_c.c_int, [_c.c_uint32, _c.c_uint32, _u64p, _u64p]),
    "tcp_u1_unrank_words": (_c.c_int, [_c.c_uint32, _c.c_uint32, _c.c_uint64, _u64p]),
    "tcp_u1_basis_words": (_c.c_int, [_c.c_uint32, _c.c_uint32, _u64p, _c.c_uint64]),
    "tcp_blocks_dimension": (_c.c_int, [_c.c_uint32, _u32p, _u32p, _u64p]),
    "tcp_blocks_rank": (_c.c_int, [_c.c_uint32, _u32p, _u32p, _c.c_uint64, _u64p]),
    "tcp_blocks_unrank": (_c.c_int, [_c.c_uint32, _u32p, _u32p, _c.c_uint64, _u64p]),
    "tcp_sector_plan_create_u1": (_c.c_int, [_vp, _c.c_uint32, _c.c_uint64, _c.POINTER(_vp)]),
    "tcp_sector_plan_create_blocks": (
        _c.c_int, [_vp, _c.c_uint32, _u32p, _u32p, _c.c_uint64, _c.POINTER(_vp)]),
    "tcp_sector_plan_dimension": (_c.c_uint64, [_vp]),
    "tcp_sector_plan_nqubits": (_c.c_uint32, [_vp]),
    "tcp_sector_plan_term_count": (_c.c_uint64, [_vp]),
    "tcp_sector_plan_group_count": (_c.c_uint64, [_vp]),
    "tcp_sector_plan_estimated_bytes": (_c.c_uint64, [_vp]),
    "tcp_sector_plan_apply_into": (_c.c_int, [_vp, _vp, _vp, _c.c_uint64, _c.c_uint64, _c.c_int, _vp]),
    "tcp_sector_plan_apply_host": (_c.c_int, [_vp, _dp, _dp, _c.c_uint64, _c.c_uint64]),
    "tcp_sector_plan_rank_device": (_c.c_int, [_vp, _vp, _vp, _c.c_uint64, _vp]),
    "tcp_sector_plan_unrank_device": (_c.c_int, [_vp, _vp, _vp, _c.c_uint64, _vp]),
    "tcp_sector_plan_walk_device": (_c.c_int, [_vp, _c.c_uint64, _vp, _c.c_uint64, _c.c_uint32, _vp]),
    "tcp_sector_plan_destroy": (None, [_vp]),
    "tcp_csr_plan_create_u1": (_c.c_int, [_vp, _c.c_uint32, _c.c_uint64, _c.POINTER(_vp)]),
    "tcp_csr_plan_create_blocks": (_c.c_int, [_vp, _c.c_uint32, _c.POINTER(_c.c_uint32), _c.POINTER(_c.c_uint32),
                                               _c.c_uint64, _c.POINTER(_vp)]),
    "tcp_csr_plan_create_full": (_c.c_int, [_vp, _c.c_uint64, _c.POINTER(_vp)]),
    "tcp_csr_plan_create_fermion": (_c.c_int, [_vp, _c.c_uint32, _c.POINTER(_c.c_uint32), _c.POINTER(_c.c_uint32),
                                                _c.c_uint64, _c.POINTER(_vp)]),
    "tcp_sector_plan_create_fermion": (_c.c_int, [_vp, _c.c_uint32, _c.POINTER(_c.c_uint32),
                                                   _c.POINTER(_c.c_uint32), _c.c_uint64, _c.POINTER(_vp)]),
    "tcp_charge_sector_create": (_c.c_int, [_c.c_uint32, _c.c_uint32, _c.POINTER(_c.c_int64),
                                            _c.POINTER(_c.c_int64), _c.c_uint64, _c.POINTER(_vp)]),
    "tcp_charge_sector_dimension": (_c.c_uint64, [_vp]),
    "tcp_charge_sector_rank": (_c.c_int, [_vp, _c.c_uint64, _u64p]),
    "tcp_charge_sector_unrank": (_c.c_int, [_vp, _c.c_uint64, _u64p]),
    "tcp_charge_sector_basis_words": (_c.c_int, [_vp, _c.c_uint64, _c.c_uint64, _u64p]),
    "tcp_charge_sector_destroy": (None, [_vp]),
    "tcp_sector_plan_create_charge": (_c.c_int, [_vp, _vp, _c.c_uint64, _c.POINTER(_vp)]),
    "tcp_sector_plan_create_fermion_charge": (_c.c_int, [_vp, _vp, _c.c_uint64, _c.POINTER(_vp)]),
    "tcp_csr_plan_create_charge": (_c.c_int, [_vp, _vp, _c.c_uint64, _c.POINTER(_vp)]),
    "tcp_csr_plan_create_fermion_charge": (_c.c_int, [_vp, _vp, _c.c_uint64, _c.POINTER(_vp)]),
    "tcp_csr_plan_dimension": (_c.c_uint64, [_vp]),
    "tcp_csr_plan_transition_count": (_c.c_uint64, [_vp]),
    "tcp_csr_plan_estimated_bytes": (_c.c_uint64, [_vp]),
    "tcp_csr_plan_is_compact": (_c.c_int, [_vp]),
    "tcp_csr_plan_storage_form": (_c.c_int, [_vp]),
    "tcp_csr_plan_slab_groups": (_c.c_uint32, [_vp]),
    "tcp_csr_plan_apply_into": (_c.c_int, [_vp, _vp, _vp, _c.c_uint64, _c.c_uint64, _c.c_int, _vp]),
    "tcp_csr_plan_apply_host": (_c.c_int, [_vp, _dp, _dp, _c.c_uint64, _c.c_uint64]),
    "tcp_csr_plan_export": (_c.c_int, [_vp, _c.POINTER(_c.c_uint64), _c.POINTER(_c.c_uint64), _dp]),
    "tcp_csr_plan_destroy": (None, [_vp]),
    "tcp_batch_plan_create": (_c.c_int, [_vp, _c.c_uint32, _c.c_uint32, _c.POINTER(_vp)]),
    "tcp_batch_plan_count": (_c.c_uint32, [_vp]),
    "tcp_batch_plan_offset": (_c.c_uint64, [_vp, _c.c_uint32]),
    "tcp_batch_plan_total_amplitudes": (_c.c_uint64, [_vp]),
    "tcp_batch_plan_work": (_c.c_double, [_vp]),
    "tcp_batch_config_count": (_c.c_uint32, []),
    "tcp_batch_config": (_c.c_int, [_c.c_uint32, _u32p, _u32p]),
    "tcp_batch_plan_selected": (_c.c_uint32, [_vp]),
    "tcp_batch_plan_select": (_c.c_int, [_vp, _c.c_uint32]),
    "tcp_batch_plan_autotune": (_c.c_int, [_vp, _vp, _vp, _c.c_int, _c.c_uint32, _vp, _u32p]),
    "tcp_batch_split_count": (_c.c_uint32, []),
    "tcp_batch_split": (_c.c_int, [_c.c_uint32, _u32p]),
    "tcp_batch_plan_selected_split": (_c.c_uint32, [_vp]),
    "tcp_batch_plan_select_split": (_c.c_int, [_vp, _c.c_uint32]),
    "tcp_batch_plan_split_times": (_c.c_int, [_vp, _vp]),
    "tcp_batch_plan_apply": (_c.c_int, [_vp, _vp, _vp, _c.c_int, _vp]),
    "tcp_batch_plan_capture": (_c.c_int, [_vp, _vp, _vp, _c.c_int, _vp]),
    "tcp_batch_plan_replay": (_c.c_int, [_vp, _vp]),
    "tcp_batch_plan_destroy": (None, [_vp]),
    "tcp_davidson_default_options": (None, [_vp]),
    "tcp_davidson_sector": (_c.c_int, [_vp, _vp, _vp, _vp]),
    "tcp_davidson_full": (_c.c_int, [_vp, _vp, _vp, _vp]),
    "tcp_davidson_csr": (_c.c_int, [_vp, _vp, _vp, _vp]),
    "tcp_lanczos_default_options": (None, [_vp]),
    "tcp_lanczos_full": (_c.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "tcp_lanczos_sector": (_c.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "tcp_lanczos_csr": (_c.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "tcp_sector_plan_diagonal": (_c.c_int, [_vp, _vp, _vp]),
    "tcp_mvp_plan_diagonal": (_c.c_int, [_vp, _vp, _vp]),
    "tcp_csr_plan_diagonal": (_c.c_int, [_vp, _vp, _vp]),
    "tcp_mlp_parameter_count": (_c.c_uint32, []),
    "tcp_mlp_create": (_c.c_int, [_c.c_uint64, _c.POINTER(_vp)]),
    "tcp_mlp_get_parameters": (_c.c_int, [_vp, _vp]),
    "tcp_mlp_set_parameters": (_c.c_int, [_vp, _vp]),
    "tcp_mlp_forward": (_c.c_int, [_vp, _vp, _c.c_uint64, _vp, _vp]),
    "tcp_mlp_gradient": (_c.c_int, [_vp, _vp, _c.c_uint64, _vp, _vp, _vp]),
    "tcp_mlp_adam_step": (_c.c_int, [_vp, _c.c_float, _c.c_float, _vp]),
    "tcp_mlp_destroy": (None, [_vp]),
    "tcp_learned_precond_create": (_c.c_int, [_vp, _vp, _c.POINTER(_vp)]),
    "tcp_learned_precond_log_multiplier": (_vp, [_vp, _c.c_double, _vp]),
    "tcp_learned_precond_set_refresh": (_c.c_int, [_vp, _c.c_double]),
    "tcp_learned_precond_ritz_loss": (_c.c_int, [_vp, _c.c_double, _vp, _vp, _vp, _vp, _dp, _dp, _vp]),
    "tcp_learned_precond_prepare": (_c.c_int, [_vp, _c.c_double, _vp]),
    "tcp_learned_precond_buffers": (_c.c_int, [_vp, _c.POINTER(_vp), _c.POINTER(_vp), _dp]),
    "tcp_learned_precond_destroy": (None, [_vp]),
}


class NativeLibraryError(RuntimeError):
    """The compiled gfx950 library is missing or cannot be loaded."""


class UnsupportedOnDeviceError(NotImplementedError):
    """Valid input outside what the GPU kernels implement."""


class DeviceError(RuntimeError):
    """A HIP runtime call failed."""


_lib: Optional[ctypes.CDLL] = None
_lock = threading.Lock()


def lib() -> ctypes.CDLL:
    """Load the shared library once; raise loudly when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            # torch ships its own libamdhip64.so.7 (same soname as the system
            # ROCm one).  Only one runtime can be resident in a process: load
            # torch's first so device tensors and this library share it.
            try:
                import torch  # noqa: F401
            except ImportError:  # pragma: no cover - torch-less C-ABI users
                pass
            if not os.path.exists(LIB_PATH):
                raise NativeLibraryError(
                    f"{LIB_PATH} is not built; run tencirpauli_amd.build.build() "
                    "(or __graft_entry__.build()) first"
                )
            try:
                handle = ctypes.CDLL(LIB_PATH)
            except OSError as exc:  # pragma: no cover - depends on the host
                raise NativeLibraryError(f"cannot load {LIB_PATH}: {exc}") from exc
            for name, (restype, argtypes) in SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = restype
                fn.argtypes = argtypes
            _lib = handle
    return _lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().tcp_last_error(buf, len(buf))
    return buf.value.decode("utf-8", "replace")


def check(status: int) -> None:
    """Raise the reference's exception class for a nonzero status."""
    if status == TCP_OK:
        return
    message = last_error()
    if status == TCP_ERR_MEMORY_LIMIT:
        raise MemoryError(message)
    if status == TCP_ERR_OVERFLOW:
        raise OverflowError(message)
    if status == TCP_ERR_HIP:
        raise DeviceError(message)
    if status == TCP_ERR_UNSUPPORTED:
        raise UnsupportedOnDeviceError(message)
    raise ValueError(message)


def device_count() -> int:
    count = ctypes.c_int(0)
    check(lib().tcp_device_count(ctypes.byref(count)))
    return count.value


def require_device() -> None:
    """Fail loudly when no HIP device is visible (no CPU fallback exists)."""
    if device_count() < 1:
        raise DeviceError("no HIP device is visible; the MVP runs only on the GPU")
