"""Composite transform rows: 5N/4, 3N/2 and 7N/4 points.

A composite row evaluates the product modulo x^N - 1 (the cyclic part) and
modulo the factors of x^N + 1 at root-table nodes 4, 2, or 2 and 6 (the
tails), each by its own transform; the inverse recombines them into the
row's coefficients (src/transform/forward.rs:95-197, src/transform/stage.rs:
175-264, src/transform/finish.rs:106-158). Forced with NTTAMD_COMPOSITE=1
and a fixed prime count (NTTAMD_PRIMES), every shape below runs on the row
form it names; each result must equal the oracle's limbs and the same
product on a power-of-two row (NTTAMD_COMPOSITE=0), bit for bit. Parts take
one pass (2^12 and below), two or three, in every mix the shapes allow.
"""
import contextlib
import os
import sys

import numpy as np
import pytest

from conftest import all_ones_product

import nttanafis_amd as ntt
from nttanafis_amd.workloads import operand

pytestmark = pytest.mark.gpu

Q61 = (1 << 61) - 1


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    ntt.load()
    assert ntt.device_available(), "no gfx950 device visible"
    yield
    os.environ.pop("NTTAMD_COMPOSITE", None)
    os.environ.pop("NTTAMD_PRIMES", None)
    os.environ.pop("NTTAMD_SPLIT", None)
    os.environ.pop("NTTAMD_SCHOOL", None)


@contextlib.contextmanager
def rows(composite, primes=None):
    saved = {k: os.environ.get(k) for k in ("NTTAMD_COMPOSITE", "NTTAMD_PRIMES", "NTTAMD_SPLIT", "NTTAMD_SCHOOL")}
    os.environ["NTTAMD_COMPOSITE"] = "1" if composite else "0"
    os.environ["NTTAMD_SPLIT"] = "0"  # whole products; splits have their own tests
    os.environ["NTTAMD_SCHOOL"] = "0"  # transforms, also for the shortest operands
    if primes:
        os.environ["NTTAMD_PRIMES"] = str(primes)
    else:
        os.environ.pop("NTTAMD_PRIMES", None)
    try:
        yield
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def product(a, b, square=False):
    plan = ntt.MulPlan.new(len(a), len(b))
    dst = np.zeros(len(a) + len(b), dtype=np.uint64)
    ok = plan.sqr(dst, a) if square else plan.mul(dst, a, b)
    assert ok, ntt.last_error()
    return dst


def mod_q(limbs):
    return int.from_bytes(np.ascontiguousarray(limbs, dtype="<u8").tobytes(), "little") % Q61


# (len_a, len_b, row tail) with three primes: the convolution of the
# operands' coefficients sits in (N, 5N/4], (5N/4, 3N/2] or (3N/2, 7N/4].
PRODUCTS = [
    (2500, 2500, 1), (3000, 2900, 2), (3500, 3400, 3),          # N = 2^12: one-pass parts
    (4000, 1500, 1),                                            # unequal operands
    (9000, 9000, 1), (11000, 11000, 2), (13000, 13000, 3),      # N = 2^14: two passes and one
    (600000, 600000, 1), (700000, 700000, 2), (850000, 800000, 3),  # N = 2^20
]


@pytest.mark.parametrize("la,lb,tail", PRODUCTS)
def test_composite_products(oracle, la, lb, tail):
    a, b = operand(la, 11), operand(lb, 12)
    with rows(True, primes=3):
        primes, width, log_n, t = ntt.mul_geometry(la, lb)
        say(f"{la}x{lb}: {primes} primes, {width}-bit coefficients, row {ntt.row_len(log_n, t)} (N = 2^{log_n})")
        assert (primes, t) == (3, tail)
        got = product(a, b)
    assert (got == oracle.mul(a, b, threads=8)).all(), "composite row differs from the oracle"
    with rows(False, primes=3):
        assert ntt.mul_geometry(la, lb)[3] == 0
        assert (product(a, b) == got).all(), "composite and power-of-two rows differ"


@pytest.mark.parametrize("n,tail", [(2600, 1), (3000, 2), (3500, 3), (11000, 2), (13000, 3)])
def test_composite_squares(oracle, n, tail):
    a = operand(n, 13)
    with rows(True, primes=3):
        assert ntt.mul_geometry(n, n)[3] == tail
        got = product(a, a, square=True)
    assert (got == oracle.sqr(a, threads=8)).all()


@pytest.mark.parametrize("la,lb", [(3500, 3400), (13000, 13000)])
def test_composite_all_ones(la, lb):
    """The longest carry chains, against the closed form (B^n-1)(B^m-1)
    (src/tests/scheduling.rs:35-45)."""
    a = np.full(la, np.uint64((1 << 64) - 1), dtype=np.uint64)
    b = np.full(lb, np.uint64((1 << 64) - 1), dtype=np.uint64)
    with rows(True, primes=3):
        assert ntt.mul_geometry(la, lb)[3] == 3
        assert (product(a, b) == all_ones_product(la, lb)).all()


@pytest.mark.parametrize("n,tail", [(5000, 1), (6000, 2), (7000, 3), (40000, 1)])
@pytest.mark.parametrize("bits,signed", [(16, False), (32, True)])
def test_composite_convolutions(oracle, n, tail, bits, signed):
    rng = np.random.default_rng(n + bits)
    if signed:
        lim = 1 << (bits - 1)
        a = rng.integers(-lim + 1, lim, n, dtype=np.int64).astype(np.uint64)
        b = rng.integers(-lim + 1, lim, n - 1, dtype=np.int64).astype(np.uint64)
        mx = lim - 1
    else:
        mx = (1 << bits) - 1
        a = rng.integers(0, mx, n, dtype=np.uint64, endpoint=True)
        b = rng.integers(0, mx, n - 1, dtype=np.uint64, endpoint=True)
    with rows(True):
        plan = ntt.ConvPlan.new(n, n - 1, mx, mx, signed)
        assert plan.row_tail == tail, plan.row_tail
        dst = np.zeros(plan.output_len(), dtype=np.uint64)
        assert plan.convolve(dst, a, b), ntt.last_error()
    assert (dst == oracle.convolve(a, b, mx, mx, signed, threads=8)).all()


@pytest.mark.parametrize("la,lb,tail", [(12000000, 12000000, 2), (6500000, 6500000, 3)])
def test_composite_large(la, lb, tail):
    """2^24- and 2^23-point cyclic parts (three and two passes) with their
    tails: equal to the power-of-two row and to the residue modulo 2^61 - 1."""
    a, b = operand(la, 1), operand(lb, 2)
    with rows(True, primes=3):
        primes, width, log_n, t = ntt.mul_geometry(la, lb)
        say(f"{la}x{lb}: {primes} primes, {width}-bit coefficients, row {ntt.row_len(log_n, t)} (N = 2^{log_n})")
        assert t == tail
        got = product(a, b)
    assert mod_q(got) == mod_q(a) * mod_q(b) % Q61
    with rows(False, primes=3):
        assert (product(a, b) == got).all()


def test_composite_workspace_and_plan_agree(oracle):
    """A workspace prepared for a composite shape (its autotuner runs the
    composite row) gives the plan's limbs."""
    a, b = operand(11000, 21), operand(11000, 22)
    with rows(True, primes=3):
        plan = ntt.MulPlan.new(len(a), len(b))
        ws = ntt.MulWorkspace()
        ws.prepare(plan)
        dst = np.zeros(len(a) + len(b), dtype=np.uint64)
        assert ws.mul(dst, a, b), ntt.last_error()
    assert (dst == oracle.mul(a, b, threads=8)).all()


# Row-shape boundaries through 16-bit convolutions (one coefficient per
# element, so the convolution length is la + lb - 1): just above N, exactly
# 5N/4, 3N/2 and 7N/4, and an operand of exactly N coefficients (N = 8192).
BOUNDARIES = [(4097, 4097, 1), (5121, 5120, 1), (5121, 5121, 2), (6145, 6144, 2), (6145, 6145, 3),
              (7169, 7168, 3), (8192, 1000, 1), (8192, 3073, 2)]


@pytest.mark.parametrize("la,lb,tail", BOUNDARIES)
def test_composite_row_boundaries(oracle, la, lb, tail):
    rng = np.random.default_rng(la * 7 + lb)
    a = rng.integers(0, 65535, la, dtype=np.uint64, endpoint=True)
    b = rng.integers(0, 65535, lb, dtype=np.uint64, endpoint=True)
    with rows(True):
        plan = ntt.ConvPlan.new(la, lb, 65535, 65535, False)
        assert (plan.log_len, plan.row_tail) == (13, tail), (plan.log_len, plan.row_tail)
        dst = np.zeros(plan.output_len(), dtype=np.uint64)
        assert plan.convolve(dst, a, b), ntt.last_error()
    assert (dst == oracle.convolve(a, b, 65535, 65535, False, threads=8)).all()
