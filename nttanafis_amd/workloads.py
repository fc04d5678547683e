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


# Request kinds of a served stream.
MUL, SQR, CONV = "mul", "sqr", "conv"


def served_stream(count: int, rate_per_s: float, lo: int = 256, hi: int = 4194304, special_frac: float = 0.1,
                  seed: int = 5):
    """A served request stream: Poisson arrivals at ``rate_per_s``, kinds
    mul / sqr / convolve in equal shares, operand lengths log-uniform in
    [lo, hi] (limbs, or 16-bit coefficients for a convolution), and a fraction
    ``special_frac`` of operands sparse or all-ones (half each).

    Returns a list of dicts: kind, arrival (s), len_a, len_b, fill_a, fill_b
    (fill "random", "sparse" or "ones"), off_a, off_b (offsets into operand
    pools of ``hi`` entries)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, size=count) if rate_per_s > 0 else np.zeros(count)
    arrivals = np.cumsum(gaps) - gaps[0]
    kinds = rng.choice([MUL, SQR, CONV], size=count)
    la = log_uniform(rng, lo, hi, count)
    lb = log_uniform(rng, lo, hi, count)

    def fills():
        u = rng.uniform(size=count)
        return np.where(u < special_frac / 2, "sparse", np.where(u < special_frac, "ones", "random"))

    fa, fb = fills(), fills()
    out = []
    for i in range(count):
        kind = str(kinds[i])
        x = int(la[i])
        y = x if kind == SQR else int(lb[i])
        out.append({
            "kind": kind,
            "arrival": float(arrivals[i]),
            "len_a": x,
            "len_b": y,
            "fill_a": str(fa[i]),
            "fill_b": str(fa[i]) if kind == SQR else str(fb[i]),
            "off_a": int(rng.integers(0, hi - x + 1)),
            "off_b": int(rng.integers(0, hi - y + 1)),
        })
    return out


def operand_pools(size: int, alloc=None):
    """Operand pools of ``size`` entries that served requests slice:
    limbs ("random", "sparse", "ones") and 16-bit coefficients (same fills).
    ``alloc(n)`` returns a writable uint64 array of n entries (pinned memory,
    for instance); numpy memory by default."""
    alloc = alloc or (lambda n: np.empty(n, dtype=np.uint64))
    pools = {}
    rand = operand(size, 1)
    rand2 = operand(size, 2)
    sparse = np.zeros(size, dtype=np.uint64)
    idx = np.arange(0, size, 977)
    sparse[idx] = rand2[idx]
    for key, arr in [("limbs", {"random": rand, "random_b": rand2, "sparse": sparse,
                                "ones": np.full(size, np.uint64((1 << 64) - 1), dtype=np.uint64)}),
                     ("coeffs", {"random": rand & np.uint64(0xFFFF), "random_b": rand2 & np.uint64(0xFFFF),
                                 "sparse": sparse & np.uint64(0xFFFF),
                                 "ones": np.full(size, np.uint64(0xFFFF), dtype=np.uint64)})]:
        pools[key] = {}
        for name, src in arr.items():
            dst = alloc(size)
            dst[:] = src
            pools[key][name] = dst
    return pools


def request_operands(req, pools):
    """(a, b) array views of a served request into the pools (b is a for a
    square). The second operand of two random ones reads the second stream."""
    space = pools["coeffs" if req["kind"] == CONV else "limbs"]
    a = space[req["fill_a"]][req["off_a"]:req["off_a"] + req["len_a"]]
    if req["kind"] == SQR:
        return a, a
    name_b = "random_b" if req["fill_b"] == "random" else req["fill_b"]
    b = space[name_b][req["off_b"]:req["off_b"] + req["len_b"]]
    return a, b


def output_len(req) -> int:
    """Limbs a served request writes (16-bit convolutions of at most 2^22
    terms have one limb per coefficient: their bound is below 2^64)."""
    if req["kind"] == CONV:
        return req["len_a"] + req["len_b"] - 1
    return req["len_a"] + req["len_b"]
