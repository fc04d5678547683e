"""Synthetic operands and workload shapes shared by the tests and the benchmark.

Operands follow the reference benchmark crate's generator: limb i (1-based)
of ``operand(len, seed)`` is SplitMix64 of ``seed + i * 0x9e3779b97f4a7c15``
(bench/src/lib.rs:13-29, used as operand(len, 1) and operand(len, 2) in
bench/benches/multiply.rs:30-38).
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def operand(length: int, seed: int) -> np.ndarray:
    """``length`` SplitMix64 limbs of stream ``seed``, least significant first."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + np.arange(1, length + 1, dtype=np.uint64) * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def log_uniform(rng: np.random.Generator, lo: int, hi: int, size: int) -> np.ndarray:
    """Integers drawn log-uniformly in [lo, hi]."""
    x = np.exp(rng.uniform(np.log(lo), np.log(hi + 1), size=size))
    return np.clip(np.floor(x).astype(np.int64), lo, hi)


def batch_shapes(count: int = 4096, lo: int = 192, hi: int = 65536, square_frac: float = 0.2,
                 seed: int = 4096):
    """Shapes of the uneven batch: len_a and len_b drawn independently and
    log-uniformly in [lo, hi]; a fraction ``square_frac`` of the entries are
    squares (len_b = len_a). Returns (len_a, len_b, square) lists."""
    rng = np.random.default_rng(seed)
    la = log_uniform(rng, lo, hi, count)
    lb = log_uniform(rng, lo, hi, count)
    sq = rng.uniform(size=count) < square_frac
    lb = np.where(sq, la, lb)
    return [int(x) for x in la], [int(x) for x in lb], [bool(s) for s in sq]


def batch_operands(len_a, len_b, square, seed: int = 7):
    """Operands of a batch: entry i uses SplitMix64 streams seed+2i and
    seed+2i+1. Returns (a, b) lists; b[i] is a[i] for a square."""
    a, b = [], []
    for i, (x, y, s) in enumerate(zip(len_a, len_b, square)):
        ai = operand(x, seed + 2 * i)
        a.append(ai)
        b.append(ai if s else operand(y, seed + 2 * i + 1))
    return a, b
