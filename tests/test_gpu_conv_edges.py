"""Convolutions at the edges of their bounds, as the reference's own
convolution tests hold them (src/tests/convolution.rs):

- every coefficient at its bound, across the shapes around the 24-term
  column path, the transforms and the split blocks
  (saturated_operands_reach_the_bound_at_every_block_edge, :184-214);
- signed coefficients at minus the bound and alternating +-bound
  (signed_extremes_reach_the_bound_in_columns_transforms_and_blocks,
  :379-408);
- unbalanced saturated and mixed operands in either order
  (unbalanced_convolutions_match_the_oracle_across_block_windows, :217-239);
- one array used twice, refused calls and signed bounds (:160-180,
  :311-342, :413-424).

Each convolution runs through ConvPlan and through a prepared ConvWorkspace
and is compared with exact Python integers at the plan's stride.
"""
import sys

import numpy as np
import pytest

import nttanafis_amd as ntt

pytestmark = pytest.mark.gpu

MASK = (1 << 64) - 1


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    ntt.load()
    assert ntt.device_available(), "no gfx950 device visible"
    yield


def bound_of(bits: int) -> int:
    """The largest value of `bits` bits (0 for none), as the reference's
    tests build bounds."""
    return (1 << bits) - 1


def exact(a, b, signed: bool, stride: int) -> np.ndarray:
    """The convolution in Python integers (two's complement words when
    signed), `stride` 64-bit words per coefficient."""
    def value(x):
        x = int(x)
        return x - (1 << 64) if signed and x >> 63 else x

    va = np.array([value(x) for x in a], dtype=object)
    vb = np.array([value(x) for x in b], dtype=object)
    out = np.convolve(va, vb)
    mod = 1 << (64 * stride)
    words = []
    for c in out:
        c = int(c) % mod
        words.extend((c >> (64 * w)) & MASK for w in range(stride))
    return np.array(words, dtype=np.uint64)


def check(a, b, max_a: int, max_b: int, signed: bool):
    a = np.asarray(a, dtype=np.uint64)
    b = np.asarray(b, dtype=np.uint64)
    plan = ntt.ConvPlan.new(len(a), len(b), max_a, max_b, signed)
    assert plan is not None, (len(a), len(b), max_a, max_b, signed)
    want = exact(a, b, signed, plan.stride())
    dst = np.full(plan.output_len(), 7, dtype=np.uint64)
    assert plan.convolve(dst, a, b), ntt.last_error()
    assert (dst == want).all(), (len(a), len(b), max_a, max_b, signed, "plan")
    ws = ntt.ConvWorkspace()
    ws.prepare(plan)
    dst[:] = 7
    assert ws.convolve(dst, a, b), ntt.last_error()
    assert (dst == want).all(), (len(a), len(b), max_a, max_b, signed, "workspace")


SATURATED_SHAPES = [(1, 1), (1, 40), (8, 8), (9, 9), (24, 24), (24, 100), (25, 25), (25, 100), (31, 2), (32, 1),
                    (33, 33)]


@pytest.mark.parametrize("left,right", SATURATED_SHAPES)
def test_saturated_operands_reach_the_bound(left, right):
    say(f"saturated {left}x{right}")
    for bits_a, bits_b in [(0, 0), (1, 1), (32, 32), (32, 64), (64, 1), (64, 64)]:
        max_a, max_b = bound_of(bits_a), bound_of(bits_b)
        check([max_a] * left, [max_b] * right, max_a, max_b, False)


@pytest.mark.parametrize("left,right", [(1, 1), (8, 8), (24, 100), (25, 25), (33, 33), (1000, 25), (2048, 64)])
def test_signed_extremes_reach_the_bound(left, right):
    say(f"signed extremes {left}x{right}")
    for bits_a, bits_b in [(0, 0), (1, 1), (63, 1), (63, 63)]:
        max_a, max_b = bound_of(bits_a), bound_of(bits_b)
        neg_a = [(-max_a) & MASK] * left
        neg_b = [(-max_b) & MASK] * right
        check(neg_a, neg_b, max_a, max_b, True)
        alternating = [max_a if i % 2 == 0 else (-max_a) & MASK for i in range(left)]
        check(alternating, neg_b, max_a, max_b, True)
        check(neg_b, alternating, max_b, max_a, True)


@pytest.mark.parametrize("long,short", [(1000, 25), (1001, 30), (2048, 64), (2500, 100), (3000, 200)])
def test_unbalanced_saturated_and_mixed(long, short):
    say(f"unbalanced {long}x{short}")
    for bits_long, bits_short in [(1, 1), (8, 64), (64, 8), (64, 64)]:
        max_long, max_short = bound_of(bits_long), bound_of(bits_short)
        sat_long, sat_short = [max_long] * long, [max_short] * short
        check(sat_long, sat_short, max_long, max_short, False)
        check(sat_short, sat_long, max_short, max_long, False)
        mixed = [(i * 0x9E37) & max_long for i in range(long)]
        check(mixed, sat_short, max_long, max_short, False)


def test_one_operand_used_twice_matches_two_copies():
    """The same array twice (the square path, src/convolve.rs:156) equals two
    distinct copies (one_operand_slice_used_twice_matches_two_distinct_copies,
    src/tests/convolution.rs:160-180), for the column path, transforms and
    every coefficient width."""
    rng = np.random.default_rng(160)
    for n in (1, 7, 24, 25, 70, 3000):
        for bits in (0, 1, 16, 33, 63, 64):
            mx = bound_of(bits)
            a = rng.integers(0, 2**64, n, dtype=np.uint64, endpoint=False) & np.uint64(mx)
            plan = ntt.ConvPlan.new(n, n, mx, mx, False)
            distinct = np.zeros(plan.output_len(), dtype=np.uint64)
            aliased = np.full(plan.output_len(), MASK, dtype=np.uint64)
            assert plan.convolve(distinct, a, a.copy()), ntt.last_error()
            assert plan.convolve(aliased, a, a), ntt.last_error()
            assert (aliased == distinct).all(), (n, bits)
            ws = ntt.ConvWorkspace()
            ws.prepare(plan)
            aliased[:] = MASK
            assert ws.convolve(aliased, a, a), ntt.last_error()
            assert (aliased == distinct).all(), (n, bits, "workspace")


def test_rejected_calls_preserve_the_destination():
    """Wrong operand lengths, a coefficient above its bound, a short
    destination and an unprepared workspace are refused and leave the
    destination untouched (rejected_calls_preserve_the_destination,
    src/tests/convolution.rs:311-342), on the column path and the transforms."""
    for left, right in [(3, 2), (30, 30)]:
        plan = ntt.ConvPlan.new(left, right, 7, 7, False)
        a = np.full(left, 7, dtype=np.uint64)
        b = np.full(right, 7, dtype=np.uint64)
        above = a.copy()
        above[-1] = 8
        untouched = np.full(plan.output_len() + 1, MASK, dtype=np.uint64)
        ws = ntt.ConvWorkspace()
        assert not ws.convolve(untouched, a, b)
        ws.prepare(plan)
        for first, second in [(a[1:], b), (a, b[1:]), (above, b)]:
            assert not plan.convolve(untouched, first, second)
            assert not ws.convolve(untouched, first, second)
        short = untouched[:plan.output_len() - 1]
        assert not plan.convolve(short, a, b)
        assert not ws.convolve(short, a, b)
        assert (untouched == np.uint64(MASK)).all()


def test_signed_plans_reject_wide_bounds_and_values_beyond_them():
    """A signed bound needs a sign bit in one limb, and a value beyond the
    bound is refused without writing (src/tests/convolution.rs:413-424)."""
    assert ntt.ConvPlan.new(4, 4, (1 << 63), 1, True) is None
    plan = ntt.ConvPlan.new(4, 4, 3, 3, True)
    dst = np.full(plan.output_len(), 7, dtype=np.uint64)
    beyond = np.array([1, MASK - 3, 0, 2], dtype=np.uint64)
    assert not plan.convolve(dst, beyond, np.array([1, 2, 3, 0], dtype=np.uint64))
    assert (dst == 7).all()
