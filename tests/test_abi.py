"""The C ABI library loads, exports every entry point of include/nttanafis.h,
and its host-only planning matches the reference's admission rules. No test
here launches device work."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

import nttanafis_amd as ntt
from nttanafis_amd import _ffi

HEADER = os.path.join(ROOT, "include", "nttanafis.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ntt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_ffi.LIB_PATH), "build the library first (nttanafis_amd.build)"
    out = subprocess.run(["nm", "-D", "--defined-only", _ffi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\b(ntt_[a-z0-9_]+)\b", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    assert set(declared_symbols()) == set(_ffi.SIGNATURES)


def test_library_loads_and_binds():
    lib = ntt.load()
    for name in _ffi.SIGNATURES:
        assert getattr(lib, name)


def test_mul_plan_admission():
    assert ntt.MulPlan.new(0, 5) is None
    assert ntt.MulPlan.new(5, 0) is None
    plan = ntt.MulPlan.new(1024, 1024)
    assert plan.len_a == 1024 and plan.len_b == 1024
    assert plan.scratch_len() == 0 and plan.sqr_scratch_len() == 0
    assert ntt.MulPlan.new(1, 1) is not None
    assert ntt.MulPlan.new(1 << 24, 1 << 24) is not None


def test_conv_plan_admission(kat):
    for terms, ma, mb, stride in kat["stride_table"]:
        plan = ntt.ConvPlan.new(terms, terms, ma, mb, False)
        assert plan.stride() == stride
        assert plan.output_len() == (2 * terms - 1) * stride
    assert ntt.ConvPlan.new(3, 3, 1 << 63, 1, True) is None
    assert ntt.ConvPlan.new(0, 3, 1, 1, False) is None
    plan = ntt.ConvPlan.new(1 << 22, 1 << 22, 1 << 31, 1 << 31, True)
    assert plan.stride() == 2 and plan.primes == 2 and plan.log_len == 23
    plan = ntt.ConvPlan.new(1 << 22, 1 << 22, 65535, 65535, False)
    assert plan.stride() == 1 and plan.primes == 2 and plan.log_len == 23
    assert ntt.ConvPlan.new(24, 1000, 5, 5, False).schoolbook
    assert not ntt.ConvPlan.new(25, 1000, 5, 5, False).schoolbook
    assert ntt.ConvPlan.new(25, 1000, 5, 5, False).scratch_len() == 0


def test_conv_plan_rejects_transforms_beyond_the_largest():
    """Transforms are at most 2^30 points: a convolution of 2^31 - 1 output
    coefficients has no admitted row and is refused (its plan would index the
    per-length split table past its end), while a composite row of 2^30 cyclic
    points still admits a slightly longer convolution than 2^30."""
    assert ntt.ConvPlan.new(1 << 30, 1 << 30, 1, 1, False) is None
    assert ntt.ConvPlan.new((1 << 40) + 1, 1 << 40, 1, 1, False) is None
    n = (1 << 29) + (1 << 27)  # 2^30 + 2^28 - 1 coefficients: a 5N/4 row, N = 2^30
    plan = ntt.ConvPlan.new(n, n, 1, 1, False)
    assert plan is not None and plan.log_len == 30 and plan.row_tail == 1
    plan = ntt.ConvPlan.new(1 << 29, 1 << 29, 1, 1, False)
    assert plan is not None and plan.log_len == 30 and plan.row_tail == 0


@pytest.mark.parametrize("la,lb", [(1, 1), (3, 700), (192, 65536), (1024, 1024), (1 << 20, 1 << 20),
                                   (1 << 24, 1 << 24), (4194304, 256), (65535, 65537)])
def test_device_geometry_fits_capacity(kat, la, lb):
    """Every coefficient of the engine's geometry stays below the CRT capacity:
    2w + ceil(log2(terms)) <= capacity (src/profile/planner.rs:558-563)."""
    primes, w, log_len, tail = ntt.mul_geometry(la, lb)
    da, db = -(-64 * la // w), -(-64 * lb // w)
    terms = min(da, db)
    assert 2 * w + (terms - 1).bit_length() <= kat["capacities"][primes - 1]
    blocks, d = ntt.mul_split(la, lb)
    if blocks:
        # A split runs blocks of d coefficients of the longer operand (d at
        # least the shorter operand's count, a multiple of 64) on rows that
        # hold a block's convolution with the shorter operand.
        dl, ds = max(da, db), min(da, db)
        assert blocks >= 2 and d % 64 == 0 and d >= ds and (blocks - 1) * d < dl <= blocks * d
        assert tail == 0 and (1 << log_len) >= d + ds - 1
        return
    assert ntt.row_len(log_len, tail) >= da + db - 1
    if tail:
        # A composite row keeps both operands in its cyclic part.
        assert max(da, db) <= 1 << log_len


def test_schoolbook_admission(monkeypatch):
    """The schoolbook takes shorter operands of at most 512 limbs: forced on,
    every such shape runs it and no other; forced off, none; by price, tiny
    products do and large balanced ones do not."""
    monkeypatch.setenv("NTTAMD_SCHOOL", "1")
    for la, lb in [(1, 1), (512, 512), (512, 1 << 20), (1 << 20, 7)]:
        assert ntt.mul_schoolbook(la, lb) and ntt.MulPlan.new(la, lb).schoolbook, (la, lb)
    for la, lb in [(513, 513), (1000, 600), (1 << 20, 1 << 20)]:
        assert not ntt.mul_schoolbook(la, lb), (la, lb)
    monkeypatch.setenv("NTTAMD_SCHOOL", "0")
    assert not any(ntt.mul_schoolbook(la, lb) for la, lb in [(1, 1), (64, 64), (512, 512)])
    monkeypatch.delenv("NTTAMD_SCHOOL")
    assert ntt.mul_schoolbook(1, 1) and ntt.mul_schoolbook(8, 8)
    assert not ntt.mul_schoolbook(1 << 16, 1 << 16) and not ntt.mul_schoolbook(0, 5)


def test_headline_geometry():
    # 2^20 x 2^20 limbs: the engine's geometry holds the whole product in a
    # power-of-two transform of at most 2^21 points.
    primes, w, log_len, tail = ntt.mul_geometry(1 << 20, 1 << 20)
    assert primes * ntt.row_len(log_len, tail) <= 6 * (1 << 20)


# The reference's own row-shape table (src/tests/plans.rs:46-81): operand
# lengths of byte-valued residues -> (row length, cyclic length, table length).
ROW_SHAPES = [((1, 1), (32, 32, 16)), ((30, 30), (64, 64, 32)), ((40, 40), (96, 64, 64)),
              ((49, 48), (96, 64, 64)), ((49, 49), (128, 128, 64)), ((64, 33), (96, 64, 64)),
              ((65, 32), (128, 128, 64)), ((70, 20), (128, 128, 64)), ((80, 80), (160, 128, 128)),
              ((100, 100), (224, 128, 128)), ((2100, 2100), (5120, 4096, 4096)),
              ((3000, 2500), (6144, 4096, 4096)), ((3600, 3400), (7168, 4096, 4096)),
              ((4097, 1000), (8192, 8192, 4096))]


@pytest.mark.parametrize("lengths,expected", ROW_SHAPES)
def test_row_shape_matches_the_reference_table(lengths, expected):
    left, right = lengths
    assert ntt.row_shape(left + right - 1, max(left, right)) == expected


def test_row_shape_is_the_shortest_admitted_row():
    """Brute force over small shapes: the row is the shortest of 2^k, 5N/4,
    3N/2, 7N/4 that holds the convolution with both operands inside N and
    every tail of at least 32 points."""
    for conv in range(1, 3000, 7):
        for longest in {1, conv // 3 + 1, (conv + 1) // 2, conv}:
            power = max(32, 1 << (conv - 1).bit_length())
            n = power // 2
            rows = [(power, power, power // 2)]
            if longest <= n:
                for tail, part in ((1, n // 4), (2, n // 2), (3, n // 4)):
                    if part >= 32 and conv <= n + tail * n // 4:
                        rows.append((n + tail * n // 4, n, n))
            assert ntt.row_shape(conv, longest) == min(rows), (conv, longest)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_host_selftest_of_shared_arithmetic(seed):
    """Residue products, packing, Garner and carry functions, compiled for
    the host from the same source as the kernels, against exact integers."""
    assert ntt.load().ntt_selftest_host(seed) == 0


def test_pinned_array_owns_its_memory(monkeypatch):
    """pinned_array views stay valid without other references and free the
    allocation exactly once, after the last view is gone (host logic; libc
    malloc stands in for hipHostMalloc)."""
    import ctypes
    import gc

    import nttanafis_amd.api as api

    libc = ctypes.CDLL(None)
    libc.malloc.restype = ctypes.c_void_p
    libc.malloc.argtypes = [ctypes.c_size_t]
    libc.free.argtypes = [ctypes.c_void_p]
    freed = []

    class FakeLib:
        @staticmethod
        def ntt_host_alloc(n):
            return libc.malloc(n)

        @staticmethod
        def ntt_host_free(p):
            freed.append(p)
            libc.free(p)

    monkeypatch.setattr(api, "_lib", lambda: FakeLib)
    arr = api.pinned_array(1000)
    view = arr[10:20]
    arr[:] = np.arange(1000, dtype=np.uint64)
    del arr
    gc.collect()
    assert not freed and int(view[0]) == 10
    del view
    gc.collect()
    assert len(freed) == 1
    empty = api.pinned_array(0)
    assert len(empty) == 0
