"""Prime fields and two-prime CRT (src/field/prime_field.rs:91-442), host only.

Restates the reference's own tests (src/field/tests/prime_field.rs): field
arithmetic against wide integers for every catalog prime and the largest
admitted one, admission against trial division, the rejected moduli, the
search below a bound, and the documentation examples (100 * 25 = 2500, the
CrtPair round trip of -12345).
"""
import numpy as np
import pytest

import nttanafis_amd as ntt

MIN, MAX = 1 << 49, 1 << 50


@pytest.fixture(scope="module")
def small_primes():
    """Primes up to 2^25 (sqrt of 2^50) by a sieve, for trial division."""
    n = (1 << 25) + 1
    sieve = np.ones(n, dtype=bool)
    sieve[:2] = False
    for i in range(2, int(n ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i::i] = False
    return np.nonzero(sieve)[0].astype(np.uint64)


def is_prime_trial(n, primes):
    if n < 2:
        return False
    cand = primes[primes * primes <= n]
    return not np.any(np.uint64(n) % cand == 0) or n in cand


def fields():
    return list(ntt.PrimeField.catalog()) + [ntt.PrimeField.below(MAX)]


def test_documentation_examples():
    f = next(ntt.PrimeField.catalog())
    a, b = f.reduce(100), f.reduce(25)
    assert f.add(a, b) == f.reduce(125)
    assert f.mul(a, b) == f.reduce(2500)
    cat = ntt.PrimeField.catalog()
    f1, f2 = next(cat), next(cat)
    pair = ntt.CrtPair.new(f1, f2)
    assert pair.combine(f1.reduce(-12345), f2.reduce(-12345)) == -12345
    assert pair.combine(f1.reduce(12345), f2.reduce(12345)) == 12345


def test_catalog_is_the_transform_catalog():
    assert [f.modulus for f in ntt.PrimeField.catalog()] == [
        1108307720798209, 1086317488242689, 910395627798529, 659706976665601,
        699289395265537, 1022545813831681, 1013749720809473, 868614185943041]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_field_arithmetic_matches_wide_integers(seed):
    rng = np.random.default_rng(seed)
    for f in fields():
        p = f.modulus
        for _ in range(64):
            left, right = int(rng.integers(0, 1 << 63)) * 2 + int(rng.integers(0, 2)), int(rng.integers(0, 1 << 63))
            signed = int(rng.integers(-(1 << 63), (1 << 63) - 1))
            exponent = int(rng.integers(0, 1 << 63))
            assert f.reduce_u64(left) == left % p
            assert f.reduce(signed) == signed % p
            for x, y in [(left, right), (0, right), (1, right), (p - 1, right), (p - 1, p - 1), (p >> 1, p >> 1)]:
                a, b = x % p, y % p
                assert f.add(a, b) == (a + b) % p
                assert f.sub(a, b) == (a - b) % p
                assert f.add(f.neg(a), a) == 0
                assert f.mul(a, b) == a * b % p
                c = f.signed(a)
                assert abs(c) <= (p - 1) >> 1 and c % p == a
                assert f.reduce(c) == a
                inv = f.inv(a)
                if a == 0:
                    assert inv is None
                else:
                    assert f.mul(a, inv) == 1
                assert f.pow(a, exponent) == pow(a, exponent, p)
                assert f.residue(a) == a and f.residue(p) is None


def test_vector_operations_match_wide_integers():
    rng = np.random.default_rng(9)
    for f in fields():
        p = f.modulus
        src = np.array([int(v) % p for v in rng.integers(0, 1 << 62, 200)], dtype=np.uint64)
        tgt = np.array([int(v) % p for v in rng.integers(0, 1 << 62, 200)], dtype=np.uint64)
        factor = int(rng.integers(0, 1 << 62)) % p
        want = [(int(t) - int(s) * factor) % p for t, s in zip(tgt, src)]
        got = tgt.copy()
        f.mul_sub(got, src, factor)
        assert [int(v) for v in got] == want
        got = tgt.copy()
        f.scale(got, factor)
        assert [int(v) for v in got] == [int(t) * factor % p for t in tgt]
        point = int(rng.integers(0, 1 << 62)) % p
        assert f.evaluate(src, point) == sum(int(c) * pow(point, i, p) for i, c in enumerate(src)) % p
        assert f.evaluate([], point) == 0


def test_admission_and_search_respect_the_prime_range(small_primes):
    for f in ntt.PrimeField.catalog():
        assert ntt.PrimeField.new(f.modulus) == f
    for rejected in [0, 1, 2, 1 << 49, (1 << 50) + 1, 28_000_001 * 28_000_003]:
        assert ntt.PrimeField.new(rejected) is None, rejected
    largest = ntt.PrimeField.below(MAX)
    assert is_prime_trial(largest.modulus, small_primes)
    assert all(not is_prime_trial(n, small_primes) for n in range(largest.modulus + 1, MAX))
    assert ntt.PrimeField.below(largest.modulus).modulus < largest.modulus
    assert ntt.PrimeField.below(MIN + 2) is None
    assert ntt.PrimeField.below(0) is None
    assert ntt.CrtPair.new(largest, largest) is None


def test_admission_matches_trial_division(small_primes):
    """Odd candidates just above 2^49, around the middle and just below 2^50."""
    for start in (MIN + 1, MIN + (MIN >> 1) + 1, MAX - 301):
        for n in range(start, start + 300, 2):
            assert (ntt.PrimeField.new(n) is not None) == is_prime_trial(n, small_primes), n


def test_crt_pair_matches_wide_integers():
    rng = np.random.default_rng(4)
    fs = fields()
    for i in range(len(fs)):
        for j in range(len(fs)):
            if i == j:
                continue
            f1, f2 = fs[i], fs[j]
            pair = ntt.CrtPair.new(f1, f2)
            m = f1.modulus * f2.modulus
            assert pair.modulus() == m
            for _ in range(16):
                x = int(rng.integers(-(1 << 62), 1 << 62)) * int(rng.integers(1, 1 << 36))
                x = (x + (m >> 1)) % m - (m >> 1)  # inside -(M-1)/2..=(M-1)/2
                assert pair.combine(x % f1.modulus, x % f2.modulus) == x
            half = (m - 1) // 2
            assert pair.combine(half % f1.modulus, half % f2.modulus) == half
            assert pair.combine(-half % f1.modulus, -half % f2.modulus) == -half
