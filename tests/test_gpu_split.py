"""Unbalanced products and convolutions split into blocks of the longer
operand.

The longer operand runs in blocks of D coefficients (D at least the shorter
operand's count, a multiple of 64), every block multiplied on its own row by
the shorter operand's spectrum, which is transformed once; the
reconstruction adds the residues of the (at most two) blocks that cover each
coefficient before the CRT, which is exact because their sum is the
product's coefficient (src/profile/planner.rs:283-393,
src/multiply/execute.rs:131-411). Every split product must equal the oracle
and the same product without a split (NTTAMD_SPLIT=0), bit for bit.
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
    os.environ.pop("NTTAMD_SPLIT", None)


@contextlib.contextmanager
def split(mode):
    """Forces the split (or not) of transform products: the schoolbook, which
    takes some of these shapes by price, stays off."""
    saved = {k: os.environ.get(k) for k in ("NTTAMD_SPLIT", "NTTAMD_SCHOOL")}
    if mode is None:
        os.environ.pop("NTTAMD_SPLIT", None)
    else:
        os.environ["NTTAMD_SPLIT"] = str(mode)
    os.environ["NTTAMD_SCHOOL"] = "0"
    try:
        yield
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def product(a, b):
    plan = ntt.MulPlan.new(len(a), len(b))
    dst = np.zeros(len(a) + len(b), dtype=np.uint64)
    assert plan.mul(dst, a, b), ntt.last_error()
    return dst


def mod_q(limbs):
    return int.from_bytes(np.ascontiguousarray(limbs, dtype="<u8").tobytes(), "little") % Q61


# One-pass block rows (up to 2^12 points), two-pass rows, the shorter operand
# first, a longer operand of exactly twice the shorter one, ragged lengths.
SHAPES = [(20000, 1000), (1000, 20000), (65536, 192), (65536, 2000), (300000, 1000), (5000, 2500),
          (1048583, 3001), (4000000, 256), (2000000, 30000)]


@pytest.mark.parametrize("la,lb", SHAPES)
def test_split_products(oracle, la, lb):
    a, b = operand(la, 31), operand(lb, 32)
    with split(1):
        blocks, digits = ntt.mul_split(la, lb)
        primes, width, log_n, _ = ntt.mul_geometry(la, lb)
        say(f"{la}x{lb}: {blocks} blocks of {digits} coefficients, {primes} primes x {width} bits, rows 2^{log_n}")
        assert blocks >= 2
        got = product(a, b)
    assert (got == oracle.mul(a, b, threads=8)).all(), "split product differs from the oracle"
    with split(0):
        assert ntt.mul_split(la, lb)[0] == 0
        assert (product(a, b) == got).all(), "split and whole products differ"


def test_split_all_ones():
    """Carry chains through every block boundary: (B^n-1)(B^m-1)
    (src/tests/scheduling.rs:35-45)."""
    la, lb = 65536, 2000
    a = np.full(la, np.uint64((1 << 64) - 1), dtype=np.uint64)
    b = np.full(lb, np.uint64((1 << 64) - 1), dtype=np.uint64)
    with split(1):
        assert ntt.mul_split(la, lb)[0] >= 2
        assert (product(a, b) == all_ones_product(la, lb)).all()
        assert (product(b, a) == all_ones_product(la, lb)).all()


def test_split_default_choice(oracle):
    """Shapes that split under the engine's own pricing."""
    for la, lb in [(65536, 192), (300000, 1000)]:
        assert ntt.mul_split(la, lb)[0] >= 2, (la, lb)
        a, b = operand(la, 41), operand(lb, 42)
        with split(None):
            assert (product(a, b) == oracle.mul(a, b, threads=8)).all()


def test_split_workspace_and_device_path(oracle):
    la, lb = 400000, 3000
    a, b = operand(la, 51), operand(lb, 52)
    with split(1):
        assert ntt.mul_split(la, lb)[0] >= 2
        plan = ntt.MulPlan.new(la, lb)
        ws = ntt.MulWorkspace()
        ws.prepare(plan)
        dst = np.zeros(la + lb, dtype=np.uint64)
        assert ws.mul(dst, a, b), ntt.last_error()
        da, db = ntt.DeviceBuffer.from_host(a), ntt.DeviceBuffer.from_host(b)
        dout = ntt.DeviceBuffer(8 * (la + lb))
        assert ws.mul_device(dout.ptr, la + lb, da.ptr, la, db.ptr, lb), ntt.last_error()
        assert ntt.synchronize()
        dev = dout.download(la + lb)
    want = oracle.mul(a, b, threads=8)
    assert (dst == want).all() and (dev == want).all()


def test_split_large_rows():
    """Blocks on 2^24-point rows (three passes) against the whole product and
    the residue modulo 2^61 - 1."""
    la, lb = 60000000, 12000000
    a, b = operand(la, 1), operand(lb, 2)
    with split(1):
        blocks, digits = ntt.mul_split(la, lb)
        primes, width, log_n, _ = ntt.mul_geometry(la, lb)
        say(f"{la}x{lb}: {blocks} blocks of {digits} coefficients on rows of 2^{log_n} points")
        assert blocks >= 2 and log_n >= 24
        got = product(a, b)
    assert mod_q(got) == mod_q(a) * mod_q(b) % Q61
    with split(0):
        assert (product(a, b) == got).all()


@pytest.mark.parametrize("la,lb,bits,signed", [(200000, 3000, 16, False), (3000, 200000, 16, False),
                                               (100000, 2000, 32, True), (65536, 100, 64, False),
                                               (50000, 777, 50, True)])
def test_split_convolutions(oracle, la, lb, bits, signed):
    """ConvPlan::convolve with the longer operand split (the reference adds
    disjoint blocks modulo their width, src/multiply/execute.rs:370-411):
    signed and unsigned coefficients up to 64 bits, either operand longer."""
    rng = np.random.default_rng(la + lb + bits)
    if signed:
        lim = 1 << (bits - 1)
        a = rng.integers(-lim + 1, lim, la, dtype=np.int64).astype(np.uint64)
        b = rng.integers(-lim + 1, lim, lb, dtype=np.int64).astype(np.uint64)
        mx = lim - 1
    else:
        mx = (1 << bits) - 1
        a = rng.integers(0, mx, la, dtype=np.uint64, endpoint=True)
        b = rng.integers(0, mx, lb, dtype=np.uint64, endpoint=True)
    with split(1):
        plan = ntt.ConvPlan.new(la, lb, mx, mx, signed)
        say(f"conv {la}x{lb} {bits}-bit signed={signed}: {plan.split_blocks} blocks, rows 2^{plan.log_len}")
        assert plan.split_blocks >= 2
        dst = np.zeros(plan.output_len(), dtype=np.uint64)
        assert plan.convolve(dst, a, b), ntt.last_error()
    assert (dst == oracle.convolve(a, b, mx, mx, signed, threads=8)).all()
    with split(0):
        whole = ntt.ConvPlan.new(la, lb, mx, mx, signed)
        assert whole.split_blocks == 0
        dst2 = np.zeros(whole.output_len(), dtype=np.uint64)
        assert whole.convolve(dst2, a, b), ntt.last_error()
    assert (dst2 == dst).all()


# Split boundaries through 16-bit convolutions: a last block of one
# coefficient, blocks exactly as long as the shorter operand, exactly two
# blocks.
@pytest.mark.parametrize("la,lb", [(3073, 1000), (5000, 1024), (2048, 1024), (4097, 64), (1000, 3073)])
def test_split_boundaries(oracle, la, lb):
    rng = np.random.default_rng(la * 3 + lb)
    a = rng.integers(0, 65535, la, dtype=np.uint64, endpoint=True)
    b = rng.integers(0, 65535, lb, dtype=np.uint64, endpoint=True)
    with split(1):
        plan = ntt.ConvPlan.new(la, lb, 65535, 65535, False)
        blocks, d = plan.split_blocks, plan._c.block_len
        dl, ds = max(la, lb), min(la, lb)
        say(f"conv {la}x{lb}: {blocks} blocks of {d}, last block {dl - (blocks - 1) * d}")
        assert blocks >= 2 and d >= ds and (blocks - 1) * d < dl <= blocks * d
        dst = np.zeros(plan.output_len(), dtype=np.uint64)
        assert plan.convolve(dst, a, b), ntt.last_error()
    assert (dst == oracle.convolve(a, b, 65535, 65535, False, threads=8)).all()
