"""ctypes declarations of libspined (include/lattice_symmetries.h, include/spined.h).

This plays the role of the `foreign import ccall` block of
src/SpinED/Internal.hs:39-384: one Python declaration per C symbol, with the
same argument meaning.  The library is built in-tree (spin-ed_amd/libspined.so)
by ``__graft_entry__.build()``; importing this module never builds it, and a
missing library is an error (there is no fallback path).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = Path(os.environ.get("SPINED_LIBRARY", PKG_ROOT / "libspined.so"))

LS_SYMBOLS = (
    "ls_error_to_string", "ls_destroy_string", "ls_enable_logging", "ls_disable_logging",
    "ls_create_symmetry", "ls_destroy_symmetry", "ls_get_sector", "ls_get_phase",
    "ls_get_periodicity", "ls_create_group", "ls_destroy_group", "ls_get_group_size",
    "ls_create_spin_basis", "ls_destroy_spin_basis", "ls_build", "ls_build_unsafe",
    "ls_get_number_states", "ls_get_states", "ls_states_get_data", "ls_states_get_size",
    "ls_destroy_states", "ls_create_interaction1", "ls_create_interaction2",
    "ls_create_interaction3", "ls_create_interaction4", "ls_interaction_is_real",
    "ls_destroy_interaction", "ls_create_operator", "ls_destroy_operator",
    "ls_operator_matmat", "ls_operator_expectation", "ls_operator_is_real",
)

SPINED_SYMBOLS = (
    "spined_eigsh_options_init", "spined_eigsh", "spined_get_last_stats", "spined_stream",
    "spined_synchronize", "spined_basis_device_states", "spined_basis_get_stabilizers",
    "spined_operator_matmat_device", "spined_operator_matmat_rows_device",
    "spined_operator_expectation_device", "spined_operator_expectation_device_noalloc",
    "spined_operator_count_offdiag", "spined_operator_orbit_method",
    "spined_operator_set_orbit_method", "spined_operator_set_schedule",
    "spined_operator_set_rows_per_wave", "spined_operator_get_rows_per_wave",
    "spined_operator_get_schedule", "spined_basis_set_bucket_bits",
    "spined_basis_get_bucket_bits", "spined_basis_set_index_kind", "spined_basis_get_index_kind",
    "spined_basis_is_complete", "spined_operator_set_gather_cache",
    "spined_operator_get_gather_cache",
    "spined_basis_sector_dimension", "spined_debug_orbit",
    "spined_debug_ring_available", "spined_debug_torus_shape", "spined_debug_orbit_hop", "spined_device_synchronize",
    "spined_device_malloc", "spined_device_release_pool",
    "spined_device_free", "spined_memcpy", "spined_memset", "spined_cnn_num_params",
    "spined_cnn_create", "spined_cnn_destroy", "spined_cnn_get_params", "spined_cnn_set_params",
    "spined_cnn_forward_device", "spined_cnn_backward_device", "spined_cnn_adam_step",
    "spined_cnn_start_vector", "spined_debug_mfma_probe", "spined_debug_copy_device",
    "spined_debug_gather_device", "spined_debug_gather_out_words",
    "spined_batch_eigsh", "spined_operator_autotune",
)


class SpinedAutotuneReport(C.Structure):
    _fields_ = [
        ("index_kind", C.c_int32),
        ("bucket_bits", C.c_int32),
        ("schedule", C.c_int32),
        ("candidates", C.c_int32),
        ("ms_per_matvec", C.c_double),
        ("ms_per_matvec_first", C.c_double),
        ("ms_per_vector_block", C.c_double * 4),
        ("best_block", C.c_int32),
        ("rows_per_wave", C.c_int32),
        ("ms_per_rows_per_wave", C.c_double * 3),
    ]


class SpinedMonitorInfo(C.Structure):
    _fields_ = [
        ("matvecs", C.c_int64),
        ("restarts", C.c_int32),
        ("phase", C.c_int32),
        ("nconverged", C.c_int32),
        ("k", C.c_int32),
        ("ritz_values", C.POINTER(C.c_double)),
        ("residuals", C.POINTER(C.c_double)),
    ]


# spined_checkpoint_fn(ritz_device, dtype, n, info, ctx)
CHECKPOINT_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_uint64, C.POINTER(SpinedMonitorInfo),
                            C.c_void_p)


class SpinedEigshOptions(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("start", C.c_int32),
        ("seed", C.c_uint64),
        ("max_basis_size", C.c_int32),
        ("min_restart_size", C.c_int32),
        ("max_block_size", C.c_int32),
        ("verbose", C.c_int32),
        ("max_matvecs", C.c_int64),
        ("start_vector", C.c_void_p),
        ("model", C.c_void_p),
        ("start_sign_mask", C.c_uint64),
        ("checkpoint", CHECKPOINT_FN),
        ("checkpoint_ctx", C.c_void_p),
        ("checkpoint_every", C.c_int32),
    ]


class SpinedStats(C.Structure):
    _fields_ = [
        ("matvecs", C.c_int64),
        ("matvecs_mixed", C.c_int64),
        ("restarts", C.c_int32),
        ("seconds_total", C.c_double),
        ("seconds_matvec", C.c_double),
        ("seconds_start", C.c_double),
        ("seconds_matvec_mixed", C.c_double),
    ]


MONITOR_FN = C.CFUNCTYPE(C.c_int, C.POINTER(SpinedMonitorInfo), C.c_void_p)

_lib = None


def load():
    """Load libspined; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise OSError(f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C spin-ed_amd/csrc)")
    L = C.CDLL(str(LIB_PATH))
    P, I, U, U64, D = C.c_void_p, C.c_int, C.c_uint, C.c_uint64, C.c_double
    PP = C.POINTER(C.c_void_p)
    sig = {
        "ls_error_to_string": (C.c_void_p, [I]),
        "ls_destroy_string": (None, [C.c_void_p]),
        "ls_enable_logging": (None, []),
        "ls_disable_logging": (None, []),
        "ls_create_symmetry": (I, [PP, U, P, U]),
        "ls_destroy_symmetry": (None, [P]),
        "ls_get_sector": (U, [P]),
        "ls_get_phase": (D, [P]),
        "ls_get_periodicity": (U, [P]),
        "ls_create_group": (I, [PP, U, P]),
        "ls_destroy_group": (None, [P]),
        "ls_get_group_size": (U, [P]),
        "ls_create_spin_basis": (I, [PP, P, U, I, I]),
        "ls_destroy_spin_basis": (None, [P]),
        "ls_build": (I, [P]),
        "ls_build_unsafe": (I, [P, U64, P]),
        "ls_get_number_states": (I, [P, C.POINTER(U64)]),
        "ls_get_states": (I, [PP, P]),
        "ls_states_get_data": (C.POINTER(U64), [P]),
        "ls_states_get_size": (U64, [P]),
        "ls_destroy_states": (None, [P]),
        "ls_create_interaction1": (I, [PP, P, U, P]),
        "ls_create_interaction2": (I, [PP, P, U, P]),
        "ls_create_interaction3": (I, [PP, P, U, P]),
        "ls_create_interaction4": (I, [PP, P, U, P]),
        "ls_interaction_is_real": (C.c_bool, [P]),
        "ls_destroy_interaction": (None, [P]),
        "ls_create_operator": (I, [PP, P, U, P]),
        "ls_destroy_operator": (None, [P]),
        "ls_operator_matmat": (I, [P, I, U64, U64, P, U64, P, U64]),
        "ls_operator_expectation": (I, [P, I, U64, U64, P, U64, P]),
        "ls_operator_is_real": (C.c_bool, [P]),
        "spined_eigsh_options_init": (None, [C.POINTER(SpinedEigshOptions)]),
        "spined_eigsh": (I, [P, I, U64, D, C.POINTER(SpinedEigshOptions), P, P, U64, P,
                             MONITOR_FN, P]),
        "spined_get_last_stats": (I, [P, C.POINTER(SpinedStats)]),
        "spined_stream": (P, []),
        "spined_synchronize": (I, []),
        "spined_basis_device_states": (I, [P, PP, PP, C.POINTER(U64)]),
        "spined_basis_get_stabilizers": (I, [P, P]),
        "spined_operator_matmat_device": (I, [P, I, U64, U64, P, U64, P, U64, I]),
        "spined_operator_matmat_rows_device": (I, [P, I, U64, U64, P, P, I]),
        "spined_operator_expectation_device": (I, [P, I, U64, U64, P, U64, P, P]),
        "spined_operator_expectation_device_noalloc": (I, [P, I, U64, U64, P, U64, P, P]),
        "spined_operator_count_offdiag": (I, [P, C.POINTER(U64)]),
        "spined_operator_orbit_method": (I, [P]),
        "spined_operator_set_orbit_method": (I, [P, I]),
        "spined_operator_set_schedule": (I, [P, I]),
        "spined_operator_autotune": (I, [P, C.c_int32, C.POINTER(SpinedAutotuneReport)]),
        "spined_operator_get_schedule": (I, [P]),
        "spined_operator_set_rows_per_wave": (I, [P, I]),
        "spined_operator_get_rows_per_wave": (I, [P]),
        "spined_operator_set_gather_cache": (I, [P, I]),
        "spined_operator_get_gather_cache": (I, [P]),
        "spined_basis_set_bucket_bits": (I, [P, I]),
        "spined_basis_get_bucket_bits": (I, [P]),
        "spined_basis_set_index_kind": (I, [P, I]),
        "spined_basis_get_index_kind": (I, [P]),
        "spined_basis_is_complete": (I, [P]),
        "spined_basis_sector_dimension": (I, [P, C.POINTER(D)]),
        "spined_debug_orbit": (I, [P, I, U64, P, P, P, P, P]),
        "spined_debug_ring_available": (I, [P]),
        "spined_debug_torus_shape": (I, [P, P, P]),
        "spined_debug_orbit_hop": (I, [P, U64, P, P, P, P, P, P]),
        "spined_device_synchronize": (I, []),
        "spined_device_malloc": (I, [PP, U64]),
        "spined_device_free": (I, [P]),
        "spined_device_release_pool": (I, []),
        "spined_memcpy": (I, [P, P, U64, I]),
        "spined_memset": (I, [P, I, U64]),
        "spined_cnn_num_params": (I, []),
        "spined_cnn_create": (I, [PP, P]),
        "spined_cnn_destroy": (None, [P]),
        "spined_cnn_get_params": (I, [P, P]),
        "spined_cnn_set_params": (I, [P, P]),
        "spined_cnn_forward_device": (I, [P, I, U64, P, P, P, P]),
        "spined_cnn_backward_device": (I, [P, I, U64, P, P, P, P, P]),
        "spined_cnn_adam_step": (I, [P, C.c_float]),
        "spined_cnn_start_vector": (I, [P, P, U64, P]),
        "spined_debug_mfma_probe": (I, [P]),
        "spined_debug_copy_device": (I, [P, P, U64]),
        "spined_debug_gather_device": (I, [P, U64, P, U64, C.c_int32]),
        "spined_debug_gather_out_words": (U64, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def error_message(code: int) -> str:
    """getErrorMessage (Internal.hs:48-54): owned string, freed after reading."""
    L = load()
    p = L.ls_error_to_string(code)
    try:
        return C.cast(p, C.c_char_p).value.decode()
    finally:
        L.ls_destroy_string(p)
