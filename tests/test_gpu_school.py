"""Schoolbook products below the transform crossover.

Products whose shorter operand has at most 512 limbs may run as direct sums
of the 128-bit limb products (one launch of 1024-limb output blocks and the
carry pass) instead of transforms: the basecase the reference leaves to its
callers below the Karatsuba crossover (README.md:6-8). Every schoolbook
product must equal the oracle, the closed forms and the transform product of
the same operands (NTTAMD_SCHOOL=0) bit for bit, alone, through workspaces and
device buffers, and inside batches.
"""
import contextlib
import os
import sys

import numpy as np
import pytest

from conftest import MASK, all_ones_product

import nttanafis_amd as ntt
from nttanafis_amd.workloads import operand

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    ntt.load()
    assert ntt.device_available(), "no gfx950 device visible"
    yield
    os.environ.pop("NTTAMD_SCHOOL", None)


@contextlib.contextmanager
def school(mode):
    saved = os.environ.get("NTTAMD_SCHOOL")
    os.environ["NTTAMD_SCHOOL"] = str(mode)
    try:
        yield
    finally:
        if saved is None:
            os.environ.pop("NTTAMD_SCHOOL", None)
        else:
            os.environ["NTTAMD_SCHOOL"] = saved


def product(a, b, square=False):
    plan = ntt.MulPlan.new(len(a), len(b))
    dst = np.zeros(len(a) + len(b), dtype=np.uint64)
    ok = plan.sqr(dst, a) if square else plan.mul(dst, a, b)
    assert ok, ntt.last_error()
    return dst


# One block and many, either operand shorter, the longest admitted shorter
# operand, single limbs, lengths around the 1024-limb blocks and the
# 16-term chunks.
SHAPES = [(1, 1), (1, 7), (7, 1), (5, 9), (16, 16), (17, 33), (192, 192), (511, 513), (512, 512),
          (1023, 1), (1024, 1), (1022, 3), (300, 2049), (2049, 300), (100000, 3), (512, 70000),
          (4096, 512)]


@pytest.mark.parametrize("la,lb", SHAPES)
def test_school_products(oracle, la, lb):
    a, b = operand(la, 61), operand(lb, 62)
    with school(1):
        assert ntt.mul_schoolbook(la, lb)
        assert ntt.MulPlan.new(la, lb).schoolbook
        got = product(a, b)
    assert (got == oracle.mul(a, b, threads=8)).all(), "schoolbook product differs from the oracle"
    with school(0):
        assert not ntt.mul_schoolbook(la, lb)
        assert (product(a, b) == got).all(), "schoolbook and transform products differ"


@pytest.mark.parametrize("n", [1, 2, 31, 192, 512])
def test_school_squares(oracle, n):
    a = operand(n, 63)
    with school(1):
        got = product(a, a, square=True)
        alias = product(a, a)
    want = oracle.sqr(a, threads=8)
    assert (got == want).all() and (alias == want).all()


@pytest.mark.parametrize("la,lb", [(512, 512), (1, 5000), (511, 3000), (3000, 511), (1024, 1024 - 512),
                                   (200, 2048)])
def test_school_all_ones(la, lb):
    """(B^n-1)(B^m-1): every column sum and carry at its largest
    (src/tests/scheduling.rs:35-45), carries through every block boundary."""
    a = np.full(la, MASK, dtype=np.uint64)
    b = np.full(lb, MASK, dtype=np.uint64)
    with school(1):
        assert (product(a, b) == all_ones_product(la, lb)).all()
        if la == lb:
            assert (product(a, a, square=True) == all_ones_product(la, la)).all()


def test_school_carry_ripples(oracle):
    """A carry out of the low limbs running through long stretches of
    all-ones limbs across block boundaries, and sparse operands."""
    la, lb = 5000, 200
    a = np.full(la, MASK, dtype=np.uint64)
    a[0] = np.uint64(3)
    b = np.zeros(lb, dtype=np.uint64)
    b[0] = np.uint64(MASK)
    b[lb - 1] = np.uint64(1)
    with school(1):
        got = product(a, b)
    assert (got == oracle.mul(a, b, threads=8)).all()
    rng = np.random.default_rng(5)
    a = np.zeros(3000, dtype=np.uint64)
    a[rng.integers(0, 3000, 20)] = rng.integers(0, 1 << 63, 20, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    b = operand(77, 9)
    with school(1):
        got = product(a, b)
    assert (got == oracle.mul(a, b, threads=8)).all()


def test_school_workspace_and_device_path(oracle):
    la, lb = 20000, 300
    a, b = operand(la, 71), operand(lb, 72)
    want = oracle.mul(a, b, threads=8)
    with school(1):
        plan = ntt.MulPlan.new(la, lb)
        assert plan.schoolbook
        ws = ntt.MulWorkspace()
        ws.prepare(plan)
        dst = np.zeros(la + lb, dtype=np.uint64)
        assert ws.mul(dst, a, b), ntt.last_error()
        da, db = ntt.DeviceBuffer.from_host(a), ntt.DeviceBuffer.from_host(b)
        dout = ntt.DeviceBuffer(8 * (la + lb))
        assert ws.mul_device(dout.ptr, la + lb, da.ptr, la, db.ptr, lb), ntt.last_error()
        assert ntt.synchronize()
        dev = dout.download(la + lb)
        # A second, smaller product through the same workspace.
        c = operand(40, 73)
        small = np.zeros(la + 40, dtype=np.uint64)
        assert ws.mul(small, a, c), ntt.last_error()
    assert (dst == want).all() and (dev == want).all()
    assert (small == oracle.mul(a, c, threads=8)).all()


def test_school_default_choice(oracle):
    """Shapes the engine's own pricing gives to the schoolbook and to the
    transforms; both equal the oracle."""
    assert not ntt.mul_schoolbook(1 << 20, 1 << 20)
    assert not ntt.mul_schoolbook(600, 600)
    for la, lb in [(8, 8), (1, 1000), (64, 64), (192, 192), (512, 512), (40, 100000)]:
        a, b = operand(la, 81), operand(lb, 82)
        say(f"{la}x{lb}: schoolbook={ntt.mul_schoolbook(la, lb)}")
        assert (product(a, b) == oracle.mul(a, b, threads=8)).all()


def test_school_in_batches(oracle):
    """Schoolbook products inside a batch (forced): one launch over their
    output blocks beside the transform buckets of the other products, graph
    replays included; equal to the oracle and to the batch without them."""
    la = [1, 192, 512, 3000, 20000, 65536, 777, 100, 4096, 5]
    lb = [9, 192, 512, 300, 64, 65536, 1000, 100, 4096, 70000]
    sq = [False, False, True, False, False, False, False, True, False, False]
    lb = [x if s else y for x, y, s in zip(la, lb, sq)]
    rng = np.random.default_rng(11)
    a = [rng.integers(0, 1 << 64, n, dtype=np.uint64, endpoint=False) for n in la]
    b = [a[i] if sq[i] else rng.integers(0, 1 << 64, lb[i], dtype=np.uint64, endpoint=False)
         for i in range(len(la))]
    a[3][:] = np.uint64(MASK)
    b[3][:] = np.uint64(MASK)
    want = [oracle.sqr(a[i], threads=8) if sq[i] else oracle.mul(a[i], b[i], threads=8) for i in range(len(la))]
    with school(1):
        assert sum(ntt.mul_schoolbook(x, y) for x, y in zip(la, lb)) == 7
        batch = ntt.Batch()
        assert batch.plan(la, lb, sq), ntt.last_error()
        out = [np.zeros(batch.output_len(i), dtype=np.uint64) for i in range(len(la))]
        for rep in range(3):
            for o in out:
                o[:] = 0
            assert batch.mul_host(a, b, out), ntt.last_error()
            for i in range(len(la)):
                assert (out[i] == want[i]).all(), f"replay {rep}, product {i}: {la[i]}x{lb[i]}"
    with school(0):
        whole = ntt.Batch()
        assert whole.plan(la, lb, sq), ntt.last_error()
        out2 = [np.zeros(whole.output_len(i), dtype=np.uint64) for i in range(len(la))]
        assert whole.mul_host(a, b, out2), ntt.last_error()
    for i in range(len(la)):
        assert (out2[i] == want[i]).all(), f"product {i} without the schoolbook"


@pytest.mark.parametrize("mode", [1, 0])
def test_carries_across_few_and_many_blocks(mode):
    """All-ones products (a carry into every block, rippling through it) with
    outputs of 2 to 40 carry blocks, of the schoolbook (64 limbs) and of the
    reconstruction (1024 limbs), each run twice in a row and in both orders,
    equal the closed form."""
    bl = 64 if mode else 1024
    shapes = []
    for blocks in (2, 3, 31, 32, 33, 40):
        out = blocks * bl - (bl // 2 if blocks > 2 else 0)
        la = min(out // 2, 512) if mode else out // 2
        shapes.append((la, out - la))
    with school(mode):
        for la, lb in shapes + shapes[::-1]:
            assert ntt.mul_schoolbook(la, lb) == bool(mode), (la, lb)
            a = np.full(la, MASK, dtype=np.uint64)
            b = np.full(lb, MASK, dtype=np.uint64)
            want = all_ones_product(la, lb)
            for _ in range(2):
                assert (product(a, b) == want).all(), (la, lb)
