"""Prime chunks give the same limbs as one launch over every prime.

Multi-pass transforms run their passes per chunk of primes (sized so that a
chunk's rows fit the 256 MiB memory-side cache); each chunk's launches shift
the rows, root tables and per-prime constants (field, T^-1, packing powers,
fold multipliers of composite parts) to its first prime. NTTAMD_CHUNK_PRIMES
forces a chunk size: chunks of 1..4 primes, including uneven last chunks
(5 primes in chunks of 2: 2, 2, 1), must reproduce the single-chunk result
bit for bit for products, squares, composite rows and convolutions; the
single-chunk result is checked against the oracle (MulPlan::mul,
src/plan.rs:112-150; ConvPlan::convolve, src/convolve.rs:148-162).
"""
import os
import sys

import numpy as np
import pytest

import nttanafis_amd as ntt
from nttanafis_amd.workloads import operand

pytestmark = pytest.mark.gpu

ENV = ("NTTAMD_CHUNK_PRIMES", "NTTAMD_PRIMES", "NTTAMD_COMPOSITE", "NTTAMD_SCHOOL", "NTTAMD_SPLIT")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def clean_env():
    for k in ENV:
        os.environ.pop(k, None)
    yield
    for k in ENV:
        os.environ.pop(k, None)


@pytest.fixture(scope="module", autouse=True)
def device():
    ntt.load()
    assert ntt.device_available(), "no gfx950 device visible"
    yield


def product(a, b):
    plan = ntt.MulPlan.new(len(a), len(b))
    dst = np.zeros(len(a) + len(b), dtype=np.uint64)
    assert plan.mul(dst, a, b), ntt.last_error()
    return dst


def square(a):
    plan = ntt.MulPlan.new(len(a), len(a))
    dst = np.zeros(2 * len(a), dtype=np.uint64)
    assert plan.sqr(dst, a), ntt.last_error()
    return dst


@pytest.mark.parametrize("primes,la,lb,composite", [(5, 60000, 50000, "0"), (6, 120000, 120000, "0"),
                                                    (3, 70000, 70000, "1"), (8, 30000, 25000, "0")])
def test_chunks_match_one_launch(oracle, primes, la, lb, composite):
    os.environ["NTTAMD_PRIMES"] = str(primes)
    os.environ["NTTAMD_COMPOSITE"] = composite
    os.environ["NTTAMD_SCHOOL"] = "0"
    os.environ["NTTAMD_SPLIT"] = "0"
    g = ntt.mul_geometry(la, lb)
    say(f"{la}x{lb}: geometry {g}, composite {composite}")
    assert g[0] == primes and g[2] >= 13, g  # a multi-pass transform of the forced prime count
    a, b = operand(la, 41), operand(lb, 42)
    os.environ["NTTAMD_CHUNK_PRIMES"] = "8"
    ref, ref_sq = product(a, b), square(a)
    assert (ref == oracle.mul(a, b, threads=8)).all(), "single chunk differs from the oracle"
    assert (ref_sq == oracle.sqr(a, threads=8)).all(), "single-chunk square differs from the oracle"
    for chunk in (1, 2, 3, 4):
        if chunk >= primes:
            continue
        os.environ["NTTAMD_CHUNK_PRIMES"] = str(chunk)
        say(f"  chunks of {chunk}")
        assert (product(a, b) == ref).all(), f"{primes} primes in chunks of {chunk}"
        assert (square(a) == ref_sq).all(), f"square, {primes} primes in chunks of {chunk}"


@pytest.mark.parametrize("signed", [False, True])
def test_chunked_convolutions(oracle, signed):
    n = 300000
    rng = np.random.default_rng(7 + signed)
    if signed:
        a = rng.integers(-(1 << 31), 1 << 31, n, dtype=np.int64).astype(np.uint64)
        b = rng.integers(-(1 << 31), 1 << 31, n, dtype=np.int64).astype(np.uint64)
        mx = 1 << 31
    else:
        a = rng.integers(0, 1 << 48, n, dtype=np.uint64)
        b = rng.integers(0, 1 << 48, n, dtype=np.uint64)
        mx = (1 << 48) - 1
    plan = ntt.ConvPlan.new(n, n, mx, mx, signed)
    out = []
    for chunk in ("8", "1"):
        os.environ["NTTAMD_CHUNK_PRIMES"] = chunk
        dst = np.zeros(plan.output_len(), dtype=np.uint64)
        assert plan.convolve(dst, a, b), ntt.last_error()
        out.append(dst)
    say(f"conv signed={signed}: {plan.primes} primes")
    assert (out[0] == out[1]).all(), "chunked convolution differs from one launch"
    import oracle as o
    assert (out[0] == o.convolve(a, b, mx, mx, signed, threads=8)).all()
