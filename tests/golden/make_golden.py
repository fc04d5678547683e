"""Writes the golden fixtures under tests/golden/ from independent arithmetic.

Products come from GMP (mpn_mul / mpn_sqr through ctypes on libgmp.so.10)
and Python integers; convolutions from GMP by Kronecker substitution; the
known-answer vectors from the reference crate's documentation and tests are
restated as data. Operands follow the reference benchmarks' SplitMix64
generator (bench/src/lib.rs:13-29) with seeds 1 and 2.

    python tests/golden/make_golden.py

Full-size fixtures keep a SHA-256 digest of the little-endian result limbs
plus sampled limbs; small ones keep every limb.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MASK = (1 << 64) - 1


def splitmix_operand(length: int, seed: int) -> np.ndarray:
    """SplitMix64 limbs, vectorized: state_i = seed + (i+1)*gamma."""
    gamma = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(1, length + 1, dtype=np.uint64) * gamma)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z.astype(np.uint64)


def gmp():
    return ctypes.CDLL("libgmp.so.10")


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def gmp_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    g = gmp()
    if len(a) < len(b):
        a, b = b, a
    out = np.zeros(len(a) + len(b), dtype=np.uint64)
    g.__gmpn_mul(_p(out), _p(a), ctypes.c_long(len(a)), _p(b), ctypes.c_long(len(b)))
    return out


def gmp_sqr(a: np.ndarray) -> np.ndarray:
    out = np.zeros(2 * len(a), dtype=np.uint64)
    gmp().__gmpn_sqr(_p(out), _p(a), ctypes.c_long(len(a)))
    return out


def to_int(limbs) -> int:
    return int.from_bytes(np.asarray(limbs, dtype="<u8").tobytes(), "little")


def from_int(value: int, length: int) -> np.ndarray:
    return np.frombuffer(value.to_bytes(8 * length, "little"), dtype="<u8").astype(np.uint64)


def digest(limbs: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(limbs, dtype="<u8").tobytes()).hexdigest()


def samples(limbs: np.ndarray, count: int = 32, seed: int = 7) -> dict:
    n = len(limbs)
    rng = np.random.default_rng(seed)
    idx = sorted(set(list(range(min(16, n))) + list(range(max(0, n - 16), n))
                     + [int(i) for i in rng.integers(0, n, count)]))
    return {str(i): hex(int(limbs[i])) for i in idx}


def record(limbs: np.ndarray, **extra) -> dict:
    out = {"len": int(len(limbs)), "sha256": digest(limbs), "samples": samples(limbs)}
    out.update(extra)
    return out


def conv16(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Unsigned 16-bit convolution: coefficients below 2^54 fit one limb
    each, so the integer product of the limb arrays is the convolution."""
    return gmp_mul(a, b)[: len(a) + len(b) - 1]


def conv_signed(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Signed 32-bit convolution at a stride of two limbs: positive and
    negative parts in 128-bit slots, four GMP products, slot differences."""
    sa, sb = a.astype(np.int64), b.astype(np.int64)

    def slots(v):
        out = np.zeros(2 * len(v), dtype=np.uint64)
        out[0::2] = v.astype(np.uint64)
        return out

    ap, an = slots(np.maximum(sa, 0)), slots(np.maximum(-sa, 0))
    bp, bn = slots(np.maximum(sb, 0)), slots(np.maximum(-sb, 0))
    n = len(a) + len(b) - 1

    def prod(x, y):
        r = gmp_mul(x, y)[: 2 * n]
        return r[0::2].copy(), r[1::2].copy()

    def add(x, y):
        lo = x[0] + y[0]
        carry = (lo < x[0]).astype(np.uint64)
        return lo, x[1] + y[1] + carry

    pos = add(prod(ap, bp), prod(an, bn))
    neg = add(prod(ap, bn), prod(an, bp))
    lo = pos[0] - neg[0]
    borrow = (pos[0] < neg[0]).astype(np.uint64)
    hi = pos[1] - neg[1] - borrow
    out = np.zeros(2 * n, dtype=np.uint64)
    out[0::2], out[1::2] = lo, hi
    return out


def main() -> None:
    np.seterr(over="ignore")
    # Config 1: 1,024 x 1,024 limbs, every limb kept.
    a = splitmix_operand(1024, 1)
    b = splitmix_operand(1024, 2)
    prod = gmp_mul(a, b)
    assert to_int(prod) == to_int(a) * to_int(b)
    sq = gmp_sqr(a)
    assert to_int(sq) == to_int(a) ** 2
    np.savez(os.path.join(HERE, "config1.npz"), a=a, b=b, mul=prod, sqr=sq)

    kat = {}
    # src/lib.rs:43-120 and src/convolve.rs:43-49 examples.
    kat["conv_123_45"] = {"a": [1, 2, 3], "b": [4, 5], "max_a": 3, "max_b": 5, "out": [4, 13, 22, 15]}
    ex_a, ex_b = [0xFFFFFFFF] * 64, [0x12345678] * 48
    kat["doc_mul"] = {"a": ex_a, "b": ex_b,
                      "out": [int(x) for x in from_int(to_int(ex_a) * to_int(ex_b), 112)]}
    kat["doc_workspace"] = {"a": [42] * 128, "b": [7] * 128,
                            "out": [int(x) for x in from_int(to_int([42] * 128) * to_int([7] * 128), 256)]}
    kat["doc_sqr"] = {"a": [3] * 96, "out": [int(x) for x in from_int(to_int([3] * 96) ** 2, 192)]}
    # src/tests/convolution.rs:284-298: stride of (terms, max_a, max_b).
    half = (1 << 32) - 1
    kat["stride_table"] = [[1, 0, 0, 1], [1, 0, MASK, 1], [1, MASK, 1, 1], [1, half, half, 1],
                           [2, half, half, 2], [1, MASK, MASK, 2], [2, MASK, MASK, 3]]
    # Capacities floor(log2(2P/3)) of the catalog prefixes, from Python ints.
    primes = [1108307720798209, 1086317488242689, 910395627798529, 659706976665601,
              699289395265537, 1022545813831681, 1013749720809473, 868614185943041]
    caps, P = [], 1
    for p in primes:
        P *= p
        caps.append((2 * P // 3).bit_length() - 1)
    kat["capacities"] = caps
    with open(os.path.join(HERE, "kat.json"), "w") as f:
        json.dump(kat, f, indent=1)

    dig = {}
    # Config 2: 2^20 x 2^20, and the square of a 2^24-limb operand.
    a20, b20 = splitmix_operand(1 << 20, 1), splitmix_operand(1 << 20, 2)
    dig["mul_2p20"] = record(gmp_mul(a20, b20), len_a=1 << 20, len_b=1 << 20, seeds=[1, 2])
    a24 = splitmix_operand(1 << 24, 1)
    dig["sqr_2p24"] = record(gmp_sqr(a24), len_a=1 << 24, seeds=[1])
    del a24
    # Config 3: length-2^22 convolutions.
    n = 1 << 22
    c16a = splitmix_operand(n, 1) & np.uint64(0xFFFF)
    c16b = splitmix_operand(n, 2) & np.uint64(0xFFFF)
    dig["conv16_2p22"] = record(conv16(c16a, c16b), len_a=n, len_b=n, max=65535, signed=False,
                                stride=1, seeds=[1, 2], mask="0xffff")
    s32a = (splitmix_operand(n, 1) & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.int32).astype(np.int64)
    s32b = (splitmix_operand(n, 2) & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.int32).astype(np.int64)
    dig["conv32s_2p22"] = record(conv_signed(s32a.astype(np.uint64), s32b.astype(np.uint64)), len_a=n,
                                 len_b=n, max=1 << 31, signed=True, stride=2, seeds=[1, 2],
                                 mask="low 32 bits, sign-extended")
    with open(os.path.join(HERE, "digests.json"), "w") as f:
        json.dump(dig, f, indent=1)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
