"""Every kernel form the autotuner may choose gives the same limbs.

Beside the pass split, the autotuner picks six kernel forms per transform
length and use: the alternative strided tiles (other line counts per tile),
radix-16 instead of radix-8 inverse rows in the product row kernel, radix-8
row rounds with twice the threads, first-pass packing per (tile, prime)
instead of one packing workgroup shared by the primes of a tile, radix-8
strided passes with twice the threads, paired product rows (one thread
per product group, both operands' groups, the pointwise product between the
last forward and first inverse round in registers), and 2^5- and 2^6-point
strided passes in wave-owned lines. NTTAMD_ALT_TILES, NTTAMD_INV16,
NTTAMD_ROWS8, NTTAMD_PER_PRIME_PACK, NTTAMD_STRIDED8, NTTAMD_PAIRED and
NTTAMD_WAVES force them. Each forced form must reproduce the
default form's product bit for bit under every admissible split, for single
products, squares and batches; the default itself is checked against the
oracle (MulPlan::mul, src/plan.rs:112-150) or, for the largest shapes,
against the product's residue modulo 2^61 - 1.
"""
import os
import sys

import numpy as np
import pytest

from test_gpu_batch import expected, mixed_shapes, plan_batch
from test_gpu_splits import mod_q, product, split_candidates, Q61

import nttanafis_amd as ntt
from nttanafis_amd.workloads import batch_operands, operand

pytestmark = pytest.mark.gpu

FORM_ENV = ("NTTAMD_ALT_TILES", "NTTAMD_INV16", "NTTAMD_PER_PRIME_PACK", "NTTAMD_ROWS8", "NTTAMD_ROW_LOG",
            "NTTAMD_PASSES", "NTTAMD_STRIDED8", "NTTAMD_PAIRED", "NTTAMD_WAVES")

SHAPES = [(4000, 4000, 3), (30000, 30000, 0), (800000, 800000, 0), (2500000, 2500000, 0),
          (8000000, 8000000, 0)]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    ntt.load()
    assert ntt.device_available(), "no gfx950 device visible"
    yield
    for k in FORM_ENV:
        os.environ.pop(k, None)


@pytest.fixture(autouse=True)
def clean_env():
    for k in FORM_ENV:
        os.environ.pop(k, None)
    yield
    for k in FORM_ENV:
        os.environ.pop(k, None)


def alt_lines(logl):
    """Lines per alternative strided tile (device.h alt_strided_lines)."""
    return 32 if logl <= 7 else 16 if logl <= 9 else 4 if logl == 10 else 2


def alt_admissible(K, passes, r):
    logs = [K - r] if passes == 2 else [(K - r + 1) // 2, K - r - (K - r + 1) // 2]
    level = 0
    for lg in logs:
        if (1 << (K - level - lg)) < alt_lines(lg):
            return False
        level += lg
    return True


def square(a):
    plan = ntt.MulPlan.new(len(a), len(a))
    dst = np.zeros(2 * len(a), dtype=np.uint64)
    assert plan.sqr(dst, a), ntt.last_error()
    return dst


def force(**forms):
    for k, v in forms.items():
        os.environ[k] = str(v)


@pytest.mark.parametrize("la,lb,force_primes", SHAPES)
def test_forms_match_the_default(oracle, monkeypatch, la, lb, force_primes):
    monkeypatch.setenv("NTTAMD_COMPOSITE", "0")
    monkeypatch.setenv("NTTAMD_SCHOOL", "0")
    if force_primes:
        monkeypatch.setenv("NTTAMD_PRIMES", str(force_primes))
    primes, width, K, _ = ntt.mul_geometry(la, lb)
    say(f"{la}x{lb}: {primes} primes, 2^{K} points")
    a, b = operand(la, 5), operand(lb, 6)
    force(NTTAMD_ALT_TILES=0, NTTAMD_INV16=0, NTTAMD_PER_PRIME_PACK=0, NTTAMD_ROWS8=0, NTTAMD_STRIDED8=0,
          NTTAMD_PAIRED=0, NTTAMD_WAVES=0)
    ref = product(a, b)
    ref_sq = square(a)
    if K <= 21:
        assert (ref == oracle.mul(a, b, threads=8)).all(), "default form differs from the oracle"
    else:
        assert mod_q(ref) == mod_q(a) * mod_q(b) % Q61, "default form fails its residue check"
        assert mod_q(ref_sq) == mod_q(a) * mod_q(a) % Q61, "default square fails its residue check"
    for passes, r in split_candidates(K):
        force(NTTAMD_PASSES=passes, NTTAMD_ROW_LOG=r)
        if alt_admissible(K, passes, r):
            force(NTTAMD_ALT_TILES=1, NTTAMD_INV16=0, NTTAMD_PER_PRIME_PACK=0)
            form = ntt.transform_form(K)
            assert form["passes"][-1] == r and "alt_tiles" in form["forms"], form
            say(f"  {passes} passes, rows 2^{r}: alternative tiles")
            assert (product(a, b) == ref).all(), f"alternative tiles, {passes} passes, rows 2^{r}"
        force(NTTAMD_ALT_TILES=0, NTTAMD_INV16=1, NTTAMD_PER_PRIME_PACK=1, NTTAMD_ROWS8=0)
        say(f"  {passes} passes, rows 2^{r}: radix-16 inverse rows, per-prime packing")
        assert (product(a, b) == ref).all(), f"inv16 + per-prime packing, {passes} passes, rows 2^{r}"
        force(NTTAMD_INV16=0, NTTAMD_PER_PRIME_PACK=0, NTTAMD_ROWS8=1)
        say(f"  {passes} passes, rows 2^{r}: radix-8 rows")
        assert (product(a, b) == ref).all(), f"radix-8 rows, {passes} passes, rows 2^{r}"
        assert (square(a) == ref_sq).all(), f"square, radix-8 rows, {passes} passes, rows 2^{r}"
        force(NTTAMD_ROWS8=0, NTTAMD_PAIRED=1, NTTAMD_STRIDED8=1)
        say(f"  {passes} passes, rows 2^{r}: paired rows, radix-8 strided passes")
        assert (product(a, b) == ref).all(), f"paired rows + radix-8 strided, {passes} passes, rows 2^{r}"
        assert (square(a) == ref_sq).all(), f"square, radix-8 strided, {passes} passes, rows 2^{r}"
        force(NTTAMD_PAIRED=0, NTTAMD_STRIDED8=0, NTTAMD_WAVES=1)
        say(f"  {passes} passes, rows 2^{r}: strided passes in wave-owned lines")
        assert (product(a, b) == ref).all(), f"wave-owned strided lines, {passes} passes, rows 2^{r}"
        assert (square(a) == ref_sq).all(), f"square, wave-owned strided lines, {passes} passes, rows 2^{r}"
        force(NTTAMD_WAVES=0)
    os.environ.pop("NTTAMD_PASSES")
    os.environ.pop("NTTAMD_ROW_LOG")
    force(NTTAMD_ALT_TILES=1, NTTAMD_INV16=1, NTTAMD_PER_PRIME_PACK=1, NTTAMD_ROWS8=0)
    say("  every form at the default split; square")
    assert (product(a, b) == ref).all(), "all forms at the default split"
    assert (square(a) == ref_sq).all(), "square with alternative tiles and per-prime packing"
    force(NTTAMD_ROWS8=1)
    assert (product(a, b) == ref).all(), "all forms and radix-8 rows at the default split"
    assert (square(a) == ref_sq).all(), "square with every form and radix-8 rows"
    force(NTTAMD_ROWS8=0, NTTAMD_PAIRED=1, NTTAMD_STRIDED8=1)
    assert (product(a, b) == ref).all(), "all forms with paired rows and radix-8 strided passes"
    assert (square(a) == ref_sq).all(), "square with every form and radix-8 strided passes"


@pytest.mark.parametrize("la,lb", [(1024, 1024), (100, 3000), (700, 701), (2000, 2000)])
def test_single_pass_forms(oracle, monkeypatch, la, lb):
    """Products of one fused pass (up to 2^12 points) in each row form."""
    monkeypatch.setenv("NTTAMD_SCHOOL", "0")
    a, b = operand(la, 7), operand(lb, 8)
    want = oracle.mul(a, b)
    for inv, r8, pr in ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)):
        force(NTTAMD_INV16=inv, NTTAMD_ROWS8=r8, NTTAMD_PAIRED=pr)
        forms = ntt.transform_form(ntt.mul_geometry(la, lb)[2])["forms"]
        assert forms == (["inv16"] if inv else []) + (["rows8"] if r8 else []) + (["paired"] if pr else []), forms
        assert (product(a, b) == want).all(), f"{la}x{lb} inv16={inv} rows8={r8}"
        if la == lb:
            assert (square(a) == oracle.sqr(a)).all(), f"{la} square inv16={inv} rows8={r8}"


@pytest.mark.parametrize("split", [False, True])
def test_batch_forms(oracle, monkeypatch, split):
    """A mixed batch (every transform length up to two passes, squares,
    ragged shapes; with split, unbalanced products split into blocks) under
    the alternative tiles, radix-16 inverse rows and per-prime packing (the
    default packs every prime of a tile in one workgroup)."""
    if split:
        monkeypatch.setenv("NTTAMD_SPLIT", "1")
        monkeypatch.setenv("NTTAMD_SCHOOL", "0")
        la = [20000, 1000, 65536, 192, 65536, 5000, 300000, 12345, 4096, 3000, 7000]
        lb = [1000, 20000, 192, 65536, 2000, 2500, 1000, 777, 4096, 3000, 64]
        sq = [False] * 8 + [True, False, False]
    else:
        la, lb, sq = mixed_shapes()
    a, b = batch_operands(la, lb, sq, seed=31)
    want = [expected(oracle, a[i], b[i], sq[i]) for i in range(len(la))]
    for alt, inv, ppp, r8, pr in ((0, 0, 0, 0, 0), (1, 0, 0, 0, 0), (0, 1, 0, 0, 0), (1, 1, 0, 0, 0), (0, 0, 1, 0, 0),
                                  (1, 1, 1, 0, 0), (0, 0, 0, 1, 0), (1, 0, 1, 1, 0), (0, 0, 0, 0, 1), (1, 0, 1, 0, 1)):
        force(NTTAMD_ALT_TILES=alt, NTTAMD_INV16=inv, NTTAMD_PER_PRIME_PACK=ppp, NTTAMD_ROWS8=r8, NTTAMD_PAIRED=pr)
        say(f"batch alt_tiles={alt} inv16={inv} per_prime_pack={ppp} rows8={r8}")
        batch = plan_batch(la, lb, sq)
        out = [np.zeros(batch.output_len(i), dtype=np.uint64) for i in range(len(la))]
        assert batch.mul_host(a, b, out), ntt.last_error()
        for i in range(len(la)):
            assert (out[i] == want[i]).all(), (f"product {i}: {la[i]}x{lb[i]} square={sq[i]} alt={alt} "
                                               f"inv16={inv} per_prime_pack={ppp}")


@pytest.mark.parametrize("n", [1024, 300000])
def test_tuned_forms_are_reported(oracle, n):
    """After a workspace's first prepare has tuned its length (single-pass
    lengths try the inverse radix only), the split export marks it tuned, the
    tuned forms are admissible ones and the tuned product is exact."""
    plan = ntt.MulPlan.new(n, n)
    ws = ntt.MulWorkspace()
    ws.prepare(plan)
    a, b = operand(n, 9), operand(n, 10)
    dst = np.zeros(2 * n, dtype=np.uint64)
    assert ws.mul(dst, a, b), ntt.last_error()
    K = ntt.mul_geometry(n, n)[2]
    form = ntt.transform_form(K, "product")
    say(f"2^{K} product: {form}")
    assert form["tuned"]
    assert set(form["forms"]) <= {"alt_tiles", "inv16", "per_prime_pack", "rows8", "strided8", "paired", "waves"}
    assert sum(form["passes"]) == K
    sq = ntt.transform_form(K, "square")
    assert sq["tuned"] and "inv16" not in sq["forms"], sq
    if n <= 4096:
        assert (dst == oracle.mul(a, b)).all()
    else:
        assert mod_q(dst) == mod_q(a) * mod_q(b) % Q61
