"""Product edge cases as the reference's own product tests hold them
(src/tests/products.rs):

- saturated long x short operands whose carries cross block windows
  (saturated_block_windows_carry_into_the_next_window, :18-30), against the
  closed form (B^n - 1)(B^m - 1);
- products and squares of 1- to 40-limb operands on every prime count 1..8
  (multiplication_and_squaring_match_schoolbook_for_every_prime_count,
  :131-150; NTTAMD_PRIMES forces the count, NTTAMD_SCHOOL=0 keeps them on
  the transforms), against Python integers;
- unequal operands in either order, through a plan and a workspace
  (unbalanced_products_match_schoolbook_in_either_operand_order, :178-198).
"""
import os
import sys

import numpy as np
import pytest

from conftest import all_ones_product, from_int, to_int

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
    for k in ("NTTAMD_PRIMES", "NTTAMD_SCHOOL"):
        os.environ.pop(k, None)


def mul(a, b):
    plan = ntt.MulPlan.new(len(a), len(b))
    dst = np.full(len(a) + len(b), 7, dtype=np.uint64)
    assert plan.mul(dst, a, b), ntt.last_error()
    return dst


@pytest.mark.parametrize("long,short", [(4096, 64), (4097, 64), (5000, 300), (8192, 25), (3000, 1)])
def test_saturated_block_windows(long, short):
    say(f"saturated {long}x{short}")
    a = np.full(long, MASK, dtype=np.uint64)
    b = np.full(short, MASK, dtype=np.uint64)
    want = all_ones_product(long, short)
    assert (mul(a, b) == want).all()
    assert (mul(b, a) == want).all()


def test_every_prime_count(monkeypatch):
    monkeypatch.setenv("NTTAMD_SCHOOL", "0")
    rng = np.random.default_rng(131)
    cases = [(int(rng.integers(1, 41)), int(rng.integers(1, 41))) for _ in range(24)] + [(1, 1), (40, 40), (1, 40)]
    for primes in range(1, 9):
        monkeypatch.setenv("NTTAMD_PRIMES", str(primes))
        say(f"prime count {primes}")
        for la, lb in cases:
            a = rng.integers(0, 2**64, la, dtype=np.uint64, endpoint=False)
            b = rng.integers(0, 2**64, lb, dtype=np.uint64, endpoint=False)
            if rng.random() < 0.25:
                a[:] = np.uint64(MASK)
            assert ntt.mul_geometry(la, lb)[0] == primes
            got = mul(a, b)
            assert to_int(got) == to_int(a) * to_int(b), (primes, la, lb)
            plan = ntt.MulPlan.new(la, la)
            sq = np.zeros(2 * la, dtype=np.uint64)
            assert plan.sqr(sq, a), ntt.last_error()
            assert to_int(sq) == to_int(a) ** 2, (primes, la, "square")


@pytest.mark.parametrize("la,lb", [(1, 300), (17, 5000), (100, 70000), (3000, 200000)])
def test_unbalanced_either_order(oracle, la, lb):
    say(f"unbalanced {la}x{lb}")
    a, b = oracle.operand(la, 51), oracle.operand(lb, 52)
    want = from_int(to_int(a) * to_int(b), la + lb)
    assert (mul(a, b) == want).all()
    assert (mul(b, a) == want).all()
    ws = ntt.MulWorkspace()
    dst = np.zeros(la + lb, dtype=np.uint64)
    assert ws.mul(dst, b, a), ntt.last_error()
    assert (dst == want).all()
