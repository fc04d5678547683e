"""ctypes declarations of the C ABI in include/nttanafis.h.

The shared library is built in-tree (nttanafis_amd/libnttanafis_amd.so) by
nttanafis_amd.build. Loading fails loudly when it is missing: there is no
CPU fallback behind this package.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnttanafis_amd.so")

u64 = ctypes.c_uint64
u32 = ctypes.c_uint32
size_t = ctypes.c_size_t
P64 = ctypes.POINTER(ctypes.c_uint64)


class MulPlanC(ctypes.Structure):
    _fields_ = [
        ("len_a", u64),
        ("len_b", u64),
        ("primes", u32),
        ("chunk_bits", u32),
        ("log_len", u32),
        ("valid", u32),
        ("row_tail", u32),
        ("schoolbook", u32),
    ]


class ConvPlanC(ctypes.Structure):
    _fields_ = [
        ("len_a", u64),
        ("len_b", u64),
        ("max_a", u64),
        ("max_b", u64),
        ("stride", u64),
        ("output_len", u64),
        ("is_signed", u32),
        ("schoolbook", u32),
        ("primes", u32),
        ("log_len", u32),
        ("bound_bits", u32),
        ("value_bits_a", u32),
        ("value_bits_b", u32),
        ("valid", u32),
        ("row_tail", u32),
        ("split_blocks", u32),
        ("block_len", u64),
    ]


class RequestC(ctypes.Structure):
    _fields_ = [
        ("kind", u32),
        ("flags", u32),
        ("a", ctypes.c_void_p),
        ("len_a", u64),
        ("b", ctypes.c_void_p),
        ("len_b", u64),
        ("max_a", u64),
        ("max_b", u64),
        ("dst", ctypes.c_void_p),
        ("dst_len", u64),
    ]


class ServerStatsC(ctypes.Structure):
    _fields_ = [
        ("submitted", u64),
        ("completed", u64),
        ("failed", u64),
        ("groups", u64),
        ("grouped", u64),
        ("budget_bytes", u64),
        ("peak_bytes", u64),
        ("waits_for_memory", u64),
    ]


REQ_MUL, REQ_SQR, REQ_CONV = 0, 1, 2
REQ_SIGNED, REQ_DEVICE = 1, 2

# name -> (restype, argtypes)
class PrimeFieldC(ctypes.Structure):
    _fields_ = [("modulus", u64), ("reciprocal", ctypes.c_double)]


class CrtPairC(ctypes.Structure):
    _fields_ = [("first", PrimeFieldC), ("second", PrimeFieldC), ("inverse", u64)]


_PF = ctypes.POINTER(PrimeFieldC)
_PC = ctypes.POINTER(CrtPairC)

SIGNATURES = {
    "ntt_field_new": (ctypes.c_bool, [u64, _PF]),
    "ntt_field_below": (ctypes.c_bool, [u64, _PF]),
    "ntt_field_catalog": (size_t, [_PF, size_t]),
    "ntt_field_residue": (ctypes.c_bool, [_PF, u64, P64]),
    "ntt_field_reduce": (u64, [_PF, ctypes.c_int64]),
    "ntt_field_reduce_u64": (u64, [_PF, u64]),
    "ntt_field_signed": (ctypes.c_int64, [_PF, u64]),
    "ntt_field_add": (u64, [_PF, u64, u64]),
    "ntt_field_sub": (u64, [_PF, u64, u64]),
    "ntt_field_neg": (u64, [_PF, u64]),
    "ntt_field_mul": (u64, [_PF, u64, u64]),
    "ntt_field_pow": (u64, [_PF, u64, u64]),
    "ntt_field_inv": (ctypes.c_bool, [_PF, u64, P64]),
    "ntt_field_mul_sub": (None, [_PF, P64, P64, size_t, u64]),
    "ntt_field_scale": (None, [_PF, P64, size_t, u64]),
    "ntt_field_evaluate": (u64, [_PF, P64, size_t, u64]),
    "ntt_crt_pair_new": (ctypes.c_bool, [_PF, _PF, _PC]),
    "ntt_crt_pair_modulus": (None, [_PC, P64, P64]),
    "ntt_crt_pair_combine": (None, [_PC, u64, u64, P64, ctypes.POINTER(ctypes.c_int64)]),
    "ntt_mul_plan_new": (ctypes.c_bool, [size_t, size_t, ctypes.POINTER(MulPlanC)]),
    "ntt_mul_plan_scratch_len": (size_t, [ctypes.POINTER(MulPlanC)]),
    "ntt_mul_plan_sqr_scratch_len": (size_t, [ctypes.POINTER(MulPlanC)]),
    "ntt_mul_plan_mul": (
        ctypes.c_bool,
        [ctypes.POINTER(MulPlanC), P64, size_t, P64, size_t, P64, size_t, P64, size_t],
    ),
    "ntt_mul_plan_sqr": (ctypes.c_bool, [ctypes.POINTER(MulPlanC), P64, size_t, P64, size_t, P64, size_t]),
    "ntt_mul_workspace_new": (ctypes.c_void_p, []),
    "ntt_mul_workspace_free": (None, [ctypes.c_void_p]),
    "ntt_mul_workspace_prepare": (None, [ctypes.c_void_p, ctypes.POINTER(MulPlanC)]),
    "ntt_mul_workspace_mul": (ctypes.c_bool, [ctypes.c_void_p, P64, size_t, P64, size_t, P64, size_t]),
    "ntt_mul_workspace_sqr": (ctypes.c_bool, [ctypes.c_void_p, P64, size_t, P64, size_t]),
    "ntt_conv_plan_new": (ctypes.c_bool, [size_t, size_t, u64, u64, ctypes.c_bool, ctypes.POINTER(ConvPlanC)]),
    "ntt_conv_plan_stride": (size_t, [ctypes.POINTER(ConvPlanC)]),
    "ntt_conv_plan_output_len": (size_t, [ctypes.POINTER(ConvPlanC)]),
    "ntt_conv_plan_scratch_len": (size_t, [ctypes.POINTER(ConvPlanC)]),
    "ntt_conv_plan_convolve": (
        ctypes.c_bool,
        [ctypes.POINTER(ConvPlanC), P64, size_t, P64, size_t, P64, size_t, P64, size_t],
    ),
    "ntt_conv_workspace_new": (ctypes.c_void_p, []),
    "ntt_conv_workspace_free": (None, [ctypes.c_void_p]),
    "ntt_conv_workspace_prepare": (None, [ctypes.c_void_p, ctypes.POINTER(ConvPlanC)]),
    "ntt_conv_workspace_convolve": (ctypes.c_bool, [ctypes.c_void_p, P64, size_t, P64, size_t, P64, size_t]),
    "ntt_mul_workspace_mul_device": (
        ctypes.c_bool,
        [ctypes.c_void_p, ctypes.c_void_p, size_t, ctypes.c_void_p, size_t, ctypes.c_void_p, size_t, ctypes.c_void_p],
    ),
    "ntt_mul_workspace_sqr_device": (
        ctypes.c_bool,
        [ctypes.c_void_p, ctypes.c_void_p, size_t, ctypes.c_void_p, size_t, ctypes.c_void_p],
    ),
    "ntt_conv_workspace_convolve_device": (
        ctypes.c_bool,
        [ctypes.c_void_p, ctypes.c_void_p, size_t, ctypes.c_void_p, size_t, ctypes.c_void_p, size_t, ctypes.c_void_p],
    ),
    "ntt_mul_device_noalloc": (
        ctypes.c_bool,
        [ctypes.c_void_p, size_t, ctypes.c_void_p, size_t, ctypes.c_void_p, size_t, ctypes.c_void_p],
    ),
    "ntt_batch_new": (ctypes.c_void_p, []),
    "ntt_batch_free": (None, [ctypes.c_void_p]),
    "ntt_batch_plan": (
        ctypes.c_bool,
        [ctypes.c_void_p, size_t, ctypes.POINTER(size_t), ctypes.POINTER(size_t), ctypes.POINTER(ctypes.c_uint8)],
    ),
    "ntt_batch_count": (size_t, [ctypes.c_void_p]),
    "ntt_batch_output_len": (size_t, [ctypes.c_void_p, size_t]),
    "ntt_batch_bind": (
        ctypes.c_bool,
        [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
         ctypes.POINTER(ctypes.c_void_p)],
    ),
    "ntt_batch_run": (ctypes.c_bool, [ctypes.c_void_p, ctypes.c_void_p]),
    "ntt_batch_run_plain": (ctypes.c_bool, [ctypes.c_void_p, ctypes.c_void_p]),
    "ntt_batch_launches": (ctypes.c_int, [ctypes.c_void_p]),
    "ntt_batch_mul_host": (
        ctypes.c_bool,
        [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
         ctypes.POINTER(ctypes.c_void_p)],
    ),
    "ntt_server_new": (ctypes.c_void_p, [u64, u32]),
    "ntt_server_free": (None, [ctypes.c_void_p]),
    "ntt_server_submit": (ctypes.c_int64, [ctypes.c_void_p, ctypes.POINTER(RequestC)]),
    "ntt_server_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "ntt_server_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "ntt_server_drain": (ctypes.c_bool, [ctypes.c_void_p]),
    "ntt_server_times": (
        ctypes.c_bool,
        [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
         ctypes.POINTER(ctypes.c_double)],
    ),
    "ntt_server_release": (None, [ctypes.c_void_p, ctypes.c_int64]),
    "ntt_server_get_stats": (None, [ctypes.c_void_p, ctypes.POINTER(ServerStatsC)]),
    "ntt_clock_s": (ctypes.c_double, []),
    "ntt_host_alloc": (ctypes.c_void_p, [size_t]),
    "ntt_host_free": (None, [ctypes.c_void_p]),
    "ntt_device_available": (ctypes.c_int, []),
    "ntt_device_alloc": (ctypes.c_void_p, [size_t]),
    "ntt_device_free": (None, [ctypes.c_void_p]),
    "ntt_copy_to_device": (ctypes.c_bool, [ctypes.c_void_p, ctypes.c_void_p, size_t]),
    "ntt_copy_to_host": (ctypes.c_bool, [ctypes.c_void_p, ctypes.c_void_p, size_t]),
    "ntt_device_synchronize": (ctypes.c_bool, []),
    "ntt_last_error": (ctypes.c_char_p, []),
    "ntt_profile_enable": (None, [ctypes.c_int]),
    "ntt_profile_read": (
        ctypes.c_int,
        [ctypes.c_char_p, size_t, ctypes.POINTER(ctypes.c_double), P64, ctypes.c_int],
    ),
    "ntt_profile_reset": (None, []),
    "ntt_profile_replay": (ctypes.c_bool, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "ntt_transform_split": (ctypes.c_int, [u32, ctypes.c_int, ctypes.POINTER(u32), ctypes.POINTER(u32)]),
    "ntt_row_shape": (ctypes.c_bool, [u64, u64, P64, P64, P64]),
    "ntt_mul_split": (ctypes.c_bool, [size_t, size_t, ctypes.POINTER(u32), P64]),
    "ntt_mul_schoolbook": (ctypes.c_int, [size_t, size_t]),
    "ntt_mul_geometry": (
        ctypes.c_bool,
        [size_t, size_t, ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32)],
    ),
    "ntt_selftest_host": (ctypes.c_int, [u64]),
    "ntt_debug_copy": (ctypes.c_bool, [ctypes.c_void_p, ctypes.c_void_p, u64]),
    "ntt_debug_tables": (ctypes.c_bool, [u64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "ntt_debug_residues": (
        ctypes.c_bool,
        [P64, size_t, P64, size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_double), size_t, ctypes.POINTER(u32)],
    ),
    "ntt_debug_crt": (ctypes.c_bool, [size_t, size_t, ctypes.POINTER(ctypes.c_double), P64]),
}

_lib = None


def load(path: str | None = None) -> ctypes.CDLL:
    """Loads the engine library and declares every entry point."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise ImportError(
            f"nttanafis_amd: {path} is missing; build it with nttanafis_amd.build.build()"
        )
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib
