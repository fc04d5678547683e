import hashlib
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

MASK = (1 << 64) - 1


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def digest(limbs) -> str:
    return hashlib.sha256(np.ascontiguousarray(limbs, dtype="<u8").tobytes()).hexdigest()


def to_int(limbs) -> int:
    return int.from_bytes(np.asarray(limbs, dtype="<u8").tobytes(), "little")


def from_int(value: int, length: int) -> np.ndarray:
    return np.frombuffer(value.to_bytes(8 * length, "little"), dtype="<u8").astype(np.uint64)


def all_ones_product(n: int, m: int) -> np.ndarray:
    """(B^n - 1)(B^m - 1) in limbs, B = 2^64 (src/tests/scheduling.rs:35-45)."""
    if m > n:
        n, m = m, n
    out = np.empty(n + m, dtype=np.uint64)
    out[0] = 1
    out[1:m] = 0
    out[m:n] = np.uint64(MASK)
    out[n] = np.uint64(MASK - 1)
    out[n + 1:] = np.uint64(MASK)
    return out


@pytest.fixture(scope="session")
def golden_config1():
    with np.load(os.path.join(GOLDEN, "config1.npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.fixture(scope="session")
def kat():
    with open(os.path.join(GOLDEN, "kat.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def digests():
    with open(os.path.join(GOLDEN, "digests.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def oracle():
    import oracle as o

    o.lib()
    return o


def check_record(limbs, rec):
    assert len(limbs) == rec["len"]
    for idx, val in rec["samples"].items():
        assert int(limbs[int(idx)]) == int(val, 16), f"limb {idx}"
    assert digest(limbs) == rec["sha256"]


def conv16_operands(n):
    import oracle as o

    return o.operand(n, 1) & np.uint64(0xFFFF), o.operand(n, 2) & np.uint64(0xFFFF)


def conv32s_operands(n):
    import oracle as o

    def signed(seed):
        v = (o.operand(n, seed) & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.int32)
        return v.astype(np.int64).astype(np.uint64)

    return signed(1), signed(2)


def signed_conv_int(a, b, stride):
    """Exact signed convolution by Python ints, packed at `stride` limbs."""
    sa = [int(x) - (1 << 64) if int(x) >> 63 else int(x) for x in a]
    sb = [int(x) - (1 << 64) if int(x) >> 63 else int(x) for x in b]
    out = [0] * (len(a) + len(b) - 1)
    for i, x in enumerate(sa):
        for j, y in enumerate(sb):
            out[i + j] += x * y
    words = []
    for c in out:
        c %= 1 << (64 * stride)
        words.extend((c >> (64 * w)) & MASK for w in range(stride))
    return np.array(words, dtype=np.uint64)


def unsigned_conv_int(a, b, stride):
    out = [0] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] += int(x) * int(y)
    words = []
    for c in out:
        words.extend((c >> (64 * w)) & MASK for w in range(stride))
    return np.array(words, dtype=np.uint64)
