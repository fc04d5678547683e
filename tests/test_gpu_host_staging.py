"""Host-buffer calls through the pinned staging ring.

Pageable host operands and results of calls moving more than 4 MiB go
through four pinned 4 MiB chunk buffers, smaller calls run their kernels on
mapped pinned memory, and pinned arrays are copied directly. Every route
must give the oracle's product,
including sizes that end inside a chunk, a destination longer than the
product (its suffix untouched), a convolution and a square.
"""
import sys

import numpy as np
import pytest

import nttanafis_amd as ntt

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    ntt.load()
    assert ntt.device_available(), "no gfx950 device visible"
    yield


def pinned_copy(x):
    y = ntt.pinned_array(len(x))
    y[:] = x
    return y


# 524,288 limbs fill one 4 MiB chunk exactly; the others end inside one.
@pytest.mark.parametrize("la,lb", [(524288, 524288), (600001, 333333), (1048577, 65536), (300000, 250001)])
def test_pageable_and_pinned_host_buffers_match_oracle(oracle, la, lb):
    say(f"host buffers {la} x {lb}")
    a, b = oracle.operand(la, 11), oracle.operand(lb, 12)
    want = oracle.mul(a, b, threads=4)
    plan = ntt.MulPlan.new(la, lb)
    dst = np.full(la + lb + 5, 7, dtype=np.uint64)
    assert plan.mul(dst, a, b), ntt.last_error()
    assert (dst[:la + lb] == want).all()
    assert (dst[la + lb:] == 7).all()
    pa, pb = pinned_copy(a), pinned_copy(b)
    pdst = ntt.pinned_array(la + lb)
    assert plan.mul(pdst, pa, pb), ntt.last_error()
    assert (pdst == want).all()
    # Mixed: pinned operands, pageable result and the reverse.
    dst[:] = 0
    assert plan.mul(dst[:la + lb], pa, pb), ntt.last_error()
    assert (dst[:la + lb] == want).all()
    pdst[:] = 0
    assert plan.mul(pdst, a, b), ntt.last_error()
    assert (pdst == want).all()


def test_pageable_square_and_convolution(oracle):
    say("host buffers: square 700,000 limbs, 16-bit convolution of 2^20 terms")
    a = oracle.operand(700000, 13)
    plan = ntt.MulPlan.new(len(a), len(a))
    dst = np.zeros(2 * len(a), dtype=np.uint64)
    assert plan.sqr(dst, a), ntt.last_error()
    assert (dst == oracle.mul(a, a, threads=4)).all()
    x = oracle.operand(1 << 20, 14) & np.uint64(0xFFFF)
    y = oracle.operand(1 << 20, 15) & np.uint64(0xFFFF)
    cplan = ntt.ConvPlan.new(len(x), len(y), 65535, 65535, False)
    out = np.zeros(cplan.output_len(), dtype=np.uint64)
    assert cplan.convolve(out, x, y), ntt.last_error()
    assert (out == oracle.convolve(x, y, 65535, 65535, False)).all()
    # The same convolution from pinned buffers (the direct route).
    px, py, pout = pinned_copy(x), pinned_copy(y), ntt.pinned_array(cplan.output_len())
    assert cplan.convolve(pout, px, py), ntt.last_error()
    assert (out == pout).all()


def test_repeated_calls_reuse_the_ring(oracle):
    """Calls back to back (the ring's buffers are refilled only after their
    previous DMA finished) and a short call between long ones."""
    say("host buffers: repeated calls")
    plan = ntt.MulPlan.new(800000, 800000)
    small = ntt.MulPlan.new(1000, 999)
    for seed in range(3):
        a, b = oracle.operand(800000, 20 + seed), oracle.operand(800000, 30 + seed)
        dst = np.zeros(1600000, dtype=np.uint64)
        assert plan.mul(dst, a, b), ntt.last_error()
        assert (dst == oracle.mul(a, b, threads=4)).all(), seed
        c, d = oracle.operand(1000, seed), oracle.operand(999, seed + 1)
        sdst = np.zeros(1999, dtype=np.uint64)
        assert small.mul(sdst, c, d), ntt.last_error()
        assert (sdst == oracle.mul(c, d)).all(), seed


# Calls moving at most 4 MiB (operands and product) run their kernels on
# mapped pinned host memory; (131072, 131072) is the largest square shape at
# that limit, (131072, 131073) just past it.
@pytest.mark.parametrize("la,lb", [(1, 1), (64, 64), (1024, 1024), (1024, 192), (5000, 17), (8000, 8400),
                                   (131072, 131072), (131072, 131073)])
def test_small_host_calls_match_oracle(oracle, la, lb):
    say(f"small host call {la} x {lb}")
    a, b = oracle.operand(la, 40), oracle.operand(lb, 41)
    plan = ntt.MulPlan.new(la, lb)
    dst = np.full(la + lb + 3, 9, dtype=np.uint64)
    assert plan.mul(dst, a, b), ntt.last_error()
    assert (dst[:la + lb] == oracle.mul(a, b, threads=8)).all()
    assert (dst[la + lb:] == 9).all()
    sq = ntt.MulPlan.new(la, la)
    sdst = np.zeros(2 * la, dtype=np.uint64)
    assert sq.sqr(sdst, a), ntt.last_error()
    assert (sdst == oracle.mul(a, a, threads=8)).all()
    # The same product from pinned arrays and through a workspace.
    pdst = ntt.pinned_array(la + lb)
    assert plan.mul(pdst, pinned_copy(a), pinned_copy(b)), ntt.last_error()
    assert (pdst == dst[:la + lb]).all()
    ws = ntt.MulWorkspace()
    wdst = np.zeros(la + lb, dtype=np.uint64)
    assert ws.mul(wdst, a, b), ntt.last_error()
    assert (wdst == dst[:la + lb]).all()


@pytest.mark.parametrize("n", [24, 25, 1024, 5000])
def test_small_host_convolutions_match_oracle(oracle, n):
    say(f"small host convolution {n}")
    x = oracle.operand(n, 42) & np.uint64(0xFFFF)
    y = oracle.operand(n, 43) & np.uint64(0xFFFF)
    plan = ntt.ConvPlan.new(n, n, 65535, 65535, False)
    out = np.zeros(plan.output_len(), dtype=np.uint64)
    assert plan.convolve(out, x, y), ntt.last_error()
    assert (out == oracle.convolve(x, y, 65535, 65535, False)).all()


def test_mapped_calls_in_sequence(oracle):
    """Hundreds of mapped calls of changing shapes back to back through one
    thread's executor (the served stream's per-call pattern): each call
    rewrites the mapped buffer that the previous call's kernels read and
    wrote, and must see only its own operands. Products, squares and 16-bit
    convolutions from pageable and pinned arrays, against the oracle."""
    say("mapped calls in sequence")
    rng = np.random.default_rng(1131)
    for i in range(240):
        la = int(np.exp(rng.uniform(0, np.log(60000))))
        lb = int(np.exp(rng.uniform(0, np.log(60000))))
        kind = i % 3
        a = oracle.operand(la, 3000 + i)
        b = oracle.operand(lb, 4000 + i)
        if i % 4 == 1:
            a = pinned_copy(a)
            b = pinned_copy(b)
        if kind == 0:
            dst = np.zeros(la + lb, dtype=np.uint64)
            assert ntt.MulPlan.new(la, lb).mul(dst, a, b), ntt.last_error()
            assert (dst == oracle.mul(a, b, threads=8)).all(), (i, la, lb)
        elif kind == 1:
            dst = np.zeros(2 * la, dtype=np.uint64)
            assert ntt.MulPlan.new(la, la).sqr(dst, a), ntt.last_error()
            assert (dst == oracle.mul(a, a, threads=8)).all(), (i, la)
        else:
            x, y = a & np.uint64(0xFFFF), b & np.uint64(0xFFFF)
            plan = ntt.ConvPlan.new(la, lb, 65535, 65535, False)
            out = np.zeros(plan.output_len(), dtype=np.uint64)
            assert plan.convolve(out, x, y), ntt.last_error()
            assert (out == oracle.convolve(x, y, 65535, 65535, False, threads=8)).all(), (i, la, lb)
