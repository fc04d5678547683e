"""ctypes wrapper of the CPU restatement in ntt_oracle.c.

TEST INFRASTRUCTURE: the checker for the HIP engine. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module. The engine never calls it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

u64 = ctypes.c_uint64
u32 = ctypes.c_uint32
P64 = ctypes.POINTER(ctypes.c_uint64)
PD = ctypes.POINTER(ctypes.c_double)


class Geometry(ctypes.Structure):
    _fields_ = [
        ("primes", u32), ("chunk_bits", u32), ("digits_a", u64), ("digits_b", u64),
        ("value_bits_a", u32), ("value_bits_b", u32), ("conv_len", u64), ("transform_len", u64),
        ("disjoint", u32), ("sign_bit", ctypes.c_int32), ("output_stride", u64), ("capacity", u32),
    ]


class MulPlan(ctypes.Structure):
    _fields_ = [("len_a", u64), ("len_b", u64), ("geometry", Geometry)]


class ConvPlan(ctypes.Structure):
    _fields_ = [
        ("len_a", u64), ("len_b", u64), ("max_a", u64), ("max_b", u64), ("is_signed", u32),
        ("schoolbook", u32), ("stride", u64), ("output_len", u64), ("geometry", Geometry),
    ]


_lib = None
_lib_path = LIB_PATH


def use_native() -> bool:
    """Switches this module to liboracle_native.so (the same source built
    with -march=native on the calling host, as the CPU baseline wants);
    False, keeping the portable build, when it cannot be built here. Call it
    before the first use of the module."""
    global _lib_path
    native = os.path.join(HERE, "liboracle_native.so")
    proc = subprocess.run(["make", "-s", "-C", HERE, "liboracle_native.so"], capture_output=True, text=True)
    if proc.returncode != 0 or not os.path.exists(native):
        return False
    _lib_path = native
    return True


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if _lib_path == LIB_PATH and not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = ctypes.CDLL(_lib_path)
        L.oracle_prime.restype = u64
        L.oracle_generator.restype = u64
        L.oracle_mulmod.restype = ctypes.c_double
        L.oracle_mulmod.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.oracle_reduce.restype = ctypes.c_double
        L.oracle_reduce.argtypes = [ctypes.c_double, ctypes.c_int]
        L.oracle_capacity.restype = u32
        L.oracle_root_table.argtypes = [ctypes.c_int, u64, ctypes.c_int, PD]
        L.oracle_forward_row.argtypes = [PD, u64, ctypes.c_int]
        L.oracle_inverse_row.argtypes = [PD, u64, ctypes.c_int]
        L.oracle_mul_plan_new.argtypes = [u64, u64, ctypes.POINTER(MulPlan)]
        L.oracle_mul.argtypes = [ctypes.POINTER(MulPlan), P64, u64, P64, u64, P64, u64, ctypes.c_int]
        L.oracle_sqr.argtypes = [ctypes.POINTER(MulPlan), P64, u64, P64, u64, ctypes.c_int]
        L.oracle_workspace_mul.argtypes = [P64, u64, P64, u64, P64, u64, ctypes.c_int, ctypes.c_int]
        L.oracle_conv_plan_new.argtypes = [u64, u64, u64, u64, ctypes.c_int, ctypes.POINTER(ConvPlan)]
        L.oracle_convolve.argtypes = [ctypes.POINTER(ConvPlan), P64, u64, P64, u64, P64, u64, ctypes.c_int]
        L.oracle_schoolbook.argtypes = [P64, P64, u64, P64, u64]
        L.oracle_operand.argtypes = [P64, u64, u64]
        L.oracle_mod_limbs.restype = u64
        L.oracle_mod_limbs.argtypes = [P64, u64, u64]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(P64)


def prime(i: int) -> int:
    return int(lib().oracle_prime(i))


def generator(i: int) -> int:
    return int(lib().oracle_generator(i))


def capacity(primes: int) -> int:
    return int(lib().oracle_capacity(primes))


def mulmod(a: float, b: float, i: int) -> float:
    return lib().oracle_mulmod(a, b, i)


def reduce(a: float, i: int) -> float:
    return lib().oracle_reduce(a, i)


def root_table(primes: int, table_len: int, inverse: bool = False) -> np.ndarray:
    out = np.zeros(primes * table_len, dtype=np.float64)
    lib().oracle_root_table(primes, table_len, 1 if inverse else 0, out.ctypes.data_as(PD))
    return out.reshape(primes, table_len)


def forward_row(row: np.ndarray, i: int) -> np.ndarray:
    r = np.array(row, dtype=np.float64)
    lib().oracle_forward_row(r.ctypes.data_as(PD), len(r), i)
    return r


def inverse_row(row: np.ndarray, i: int) -> np.ndarray:
    r = np.array(row, dtype=np.float64)
    lib().oracle_inverse_row(r.ctypes.data_as(PD), len(r), i)
    return r


def operand(length: int, seed: int) -> np.ndarray:
    """SplitMix64 limbs of bench/src/lib.rs:13-29."""
    out = np.zeros(length, dtype=np.uint64)
    lib().oracle_operand(_p(out), length, seed)
    return out


def mod_limbs(limbs: np.ndarray, q: int) -> int:
    a = np.ascontiguousarray(limbs, dtype=np.uint64)
    return int(lib().oracle_mod_limbs(_p(a), len(a), q))


def mul_plan(len_a: int, len_b: int):
    plan = MulPlan()
    if not lib().oracle_mul_plan_new(len_a, len_b, ctypes.byref(plan)):
        return None
    return plan


def mul(a: np.ndarray, b: np.ndarray, threads: int = 1) -> np.ndarray:
    plan = mul_plan(len(a), len(b))
    assert plan is not None
    dst = np.zeros(len(a) + len(b), dtype=np.uint64)
    bb = a if b is a else b
    assert lib().oracle_mul(ctypes.byref(plan), _p(dst), len(dst), _p(a), len(a), _p(bb), len(bb), threads)
    return dst


def sqr(a: np.ndarray, threads: int = 1) -> np.ndarray:
    plan = mul_plan(len(a), len(a))
    dst = np.zeros(2 * len(a), dtype=np.uint64)
    assert lib().oracle_sqr(ctypes.byref(plan), _p(dst), len(dst), _p(a), len(a), threads)
    return dst


def workspace_mul(dst: np.ndarray, a: np.ndarray, b: np.ndarray, square: bool = False, threads: int = 1) -> bool:
    """MulWorkspace::run (src/plan.rs:182-205): an empty operand writes zeros."""
    aa = np.ascontiguousarray(a, dtype=np.uint64)
    bb = aa if square else np.ascontiguousarray(b, dtype=np.uint64)
    ok = lib().oracle_workspace_mul(_p(dst), len(dst), _p(aa), len(aa), _p(bb), len(bb), 1 if square else 0,
                                    threads)
    return bool(ok)


def conv_plan(len_a: int, len_b: int, max_a: int, max_b: int, signed: bool):
    plan = ConvPlan()
    if not lib().oracle_conv_plan_new(len_a, len_b, max_a, max_b, 1 if signed else 0, ctypes.byref(plan)):
        return None
    return plan


def convolve(a: np.ndarray, b: np.ndarray, max_a: int, max_b: int, signed: bool, threads: int = 1):
    plan = conv_plan(len(a), len(b), max_a, max_b, signed)
    assert plan is not None
    dst = np.zeros(plan.output_len, dtype=np.uint64)
    bb = a if b is a else b
    ok = lib().oracle_convolve(ctypes.byref(plan), _p(dst), len(dst), _p(a), len(a), _p(bb), len(bb), threads)
    return dst if ok else None


def schoolbook(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    dst = np.zeros(len(a) + len(b), dtype=np.uint64)
    lib().oracle_schoolbook(_p(dst), _p(a), len(a), _p(b), len(b))
    return dst
