"""Pins the CPU oracle (oracle/ntt_oracle.c) before it is trusted as the
checker of the GPU engine: against the GMP/closed-form golden fixtures,
Python integer arithmetic, and the mathematics of each kernel."""
import numpy as np
import pytest

from conftest import (MASK, all_ones_product, check_record, conv16_operands, conv32s_operands,
                      from_int, signed_conv_int, to_int, unsigned_conv_int)

PRIMES = [1108307720798209, 1086317488242689, 910395627798529, 659706976665601,
          699289395265537, 1022545813831681, 1013749720809473, 868614185943041]


def is_prime(n):
    if n < 2:
        return False
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37]:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def test_catalog_is_the_reference_catalog(oracle):
    for i, p in enumerate(PRIMES):
        assert oracle.prime(i) == p
        assert 2 ** 49 < p < 2 ** 50 and is_prime(p)
        g = oracle.generator(i)
        assert pow(g, (p - 1) // 2, p) == p - 1  # quadratic nonresidue
        assert ((p - 1) & -(p - 1)).bit_length() - 1 >= 41


def test_capacities_match_python_integers(oracle, kat):
    assert [oracle.capacity(k) for k in range(1, 9)] == kat["capacities"]
    assert kat["capacities"] == [49, 99, 149, 198, 247, 297, 347, 396]


@pytest.mark.parametrize("prime", range(8))
def test_mulmod_is_exact_and_bounded(oracle, prime):
    p = PRIMES[prime]
    rng = np.random.default_rng(prime)
    edges = [0, 1, -1, p - 1, -(p - 1), (p - 1) // 2, -(p - 1) // 2, 2 * p - 1, -(2 * p - 1)]
    vals = edges + [int(x) for x in rng.integers(-(2 * p - 1), 2 * p, 200)]
    for a in vals:
        for b in vals[:12] + [int(x) for x in rng.integers(-(p // 2), p // 2 + 1, 4)]:
            if abs(a * b) >= 4 * p * p:
                continue
            r = oracle.mulmod(float(a), float(b), prime)
            assert r == int(r)
            assert (int(r) - a * b) % p == 0
            bound = 15 * p / 16 if abs(a * b) < 2 * p * p else 11 * p / 8
            assert abs(r) < bound


@pytest.mark.parametrize("prime", [0, 3, 7])
def test_reduce_centers(oracle, prime):
    p = PRIMES[prime]
    rng = np.random.default_rng(9)
    for x in [0, 1, p - 1, p, 8 * p - 1, -(8 * p - 1)] + [int(v) for v in rng.integers(-8 * p + 1, 8 * p, 300)]:
        r = oracle.reduce(float(x), prime)
        assert (int(r) - x) % p == 0 and abs(r) < 9 * p / 16
        if abs(x) < p:
            assert abs(r) <= p / 2


def rev(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


@pytest.mark.parametrize("prime", [0, 5])
def test_root_tables_are_bit_reversed_powers(oracle, prime):
    p, g = PRIMES[prime], oracle.generator(prime)
    n = 64
    fwd = oracle.root_table(prime + 1, n, False)[prime]
    inv = oracle.root_table(prime + 1, n, True)[prime]
    omega = pow(g, (p - 1) // (2 * n), p)
    for j in range(n):
        w = pow(omega, rev(j, 6), p)
        assert int(fwd[j]) % p == w and abs(fwd[j]) <= p / 2
        assert int(inv[j]) * w % p == 1 and abs(inv[j]) <= p / 2


@pytest.mark.parametrize("logn", [5, 6, 7, 8])
def test_forward_is_bit_reversed_evaluation(oracle, logn):
    p = PRIMES[2]
    n = 1 << logn
    rng = np.random.default_rng(logn)
    x = [int(v) for v in rng.integers(-(p // 2), p // 2, n)]
    y = oracle.forward_row(np.array(x, dtype=np.float64), 2)
    omega = pow(oracle.generator(2), (p - 1) // n, p)
    for i in range(n):
        point = pow(omega, rev(i, logn), p)
        expect = sum(c * pow(point, k, p) for k, c in enumerate(x)) % p
        assert int(y[i]) % p == expect
        assert abs(y[i]) < 3 * p / 2


@pytest.mark.parametrize("logn", [5, 6, 9, 12])
def test_inverse_undoes_forward(oracle, logn):
    p = PRIMES[1]
    n = 1 << logn
    rng = np.random.default_rng(logn + 100)
    x = rng.integers(0, p, n).astype(np.float64)
    y = oracle.inverse_row(oracle.forward_row(x, 1), 1)
    assert all((int(a) - n * int(b)) % p == 0 for a, b in zip(y, x))


def test_config1_golden(oracle, golden_config1):
    a, b = golden_config1["a"], golden_config1["b"]
    assert (oracle.mul(a, b) == golden_config1["mul"]).all()
    assert (oracle.sqr(a) == golden_config1["sqr"]).all()
    assert (oracle.mul(a, a) == golden_config1["sqr"]).all()  # same slice squares


def test_documentation_examples(oracle, kat):
    for key in ["doc_mul", "doc_workspace"]:
        a = np.array(kat[key]["a"], dtype=np.uint64)
        b = np.array(kat[key]["b"], dtype=np.uint64)
        assert [int(x) for x in oracle.mul(a, b)] == kat[key]["out"]
    a = np.array(kat["doc_sqr"]["a"], dtype=np.uint64)
    assert [int(x) for x in oracle.sqr(a)] == kat["doc_sqr"]["out"]
    c = kat["conv_123_45"]
    out = oracle.convolve(np.array(c["a"], dtype=np.uint64), np.array(c["b"], dtype=np.uint64),
                          c["max_a"], c["max_b"], False)
    assert [int(x) for x in out] == c["out"]


def test_stride_table(oracle, kat):
    for terms, ma, mb, stride in kat["stride_table"]:
        assert oracle.conv_plan(terms, terms, ma, mb, False).stride == stride


@pytest.mark.parametrize("seed", range(6))
def test_random_products_match_integers(oracle, seed):
    rng = np.random.default_rng(seed)
    for _ in range(20):
        la, lb = (int(v) for v in rng.integers(1, 700, 2))
        a = rng.integers(0, 2 ** 64, la, dtype=np.uint64)
        b = rng.integers(0, 2 ** 64, lb, dtype=np.uint64)
        if seed % 3 == 1:
            a[rng.random(la) < 0.7] = 0
        if seed % 3 == 2:
            b[:] = np.uint64(MASK)
        assert to_int(oracle.mul(a, b, threads=2)) == to_int(a) * to_int(b)


def test_products_against_schoolbook(oracle):
    rng = np.random.default_rng(3)
    for la, lb in [(1, 1), (2, 3), (24, 25), (100, 3), (257, 511)]:
        a = rng.integers(0, 2 ** 64, la, dtype=np.uint64)
        b = rng.integers(0, 2 ** 64, lb, dtype=np.uint64)
        assert (oracle.mul(a, b) == oracle.schoolbook(a, b)).all()


@pytest.mark.parametrize("n,m", [(1, 1), (5, 3), (1000, 1000), (4097, 100), (1 << 16, 1 << 16)])
def test_all_ones_closed_form(oracle, n, m):
    a = np.full(n, MASK, dtype=np.uint64)
    b = np.full(m, MASK, dtype=np.uint64)
    assert (oracle.mul(a, b, threads=4) == all_ones_product(n, m)).all()


def test_mul_2p20_digest(oracle, digests):
    a, b = oracle.operand(1 << 20, 1), oracle.operand(1 << 20, 2)
    check_record(oracle.mul(a, b, threads=8), digests["mul_2p20"])


def test_all_ones_2p20(oracle):
    a = np.full(1 << 20, MASK, dtype=np.uint64)
    assert (oracle.mul(a, a.copy(), threads=8) == all_ones_product(1 << 20, 1 << 20)).all()


def test_sqr_2p24_digest(oracle, digests):
    a = oracle.operand(1 << 24, 1)
    check_record(oracle.sqr(a, threads=8), digests["sqr_2p24"])


def test_conv16_2p22_digest(oracle, digests):
    a, b = conv16_operands(1 << 22)
    check_record(oracle.convolve(a, b, 65535, 65535, False, threads=8), digests["conv16_2p22"])


def test_conv32s_2p22_digest(oracle, digests):
    a, b = conv32s_operands(1 << 22)
    plan = oracle.conv_plan(len(a), len(b), 1 << 31, 1 << 31, True)
    assert plan.stride == 2 and plan.geometry.primes == 2
    check_record(oracle.convolve(a, b, 1 << 31, 1 << 31, True, threads=8), digests["conv32s_2p22"])


@pytest.mark.parametrize("bits,signed", [(1, False), (16, False), (32, True), (49, False), (50, True),
                                         (63, True), (64, False)])
def test_convolutions_match_integers(oracle, bits, signed):
    rng = np.random.default_rng(bits)
    for la, lb in [(3, 2), (24, 30), (25, 25), (200, 77)]:
        if signed:
            lim = 1 << (bits - 1)
            a = rng.integers(-lim + 1, lim, la, dtype=np.int64).astype(np.uint64)
            b = rng.integers(-lim + 1, lim, lb, dtype=np.int64).astype(np.uint64)
            mx = lim - 1
        else:
            mx = (1 << bits) - 1
            a = rng.integers(0, mx, la, dtype=np.uint64, endpoint=True)
            b = rng.integers(0, mx, lb, dtype=np.uint64, endpoint=True)
        plan = oracle.conv_plan(la, lb, mx, mx, signed)
        out = oracle.convolve(a, b, mx, mx, signed, threads=2)
        expect = (signed_conv_int if signed else unsigned_conv_int)(a, b, plan.stride)
        assert (out == expect).all()


def test_rejections(oracle):
    assert oracle.mul_plan(0, 5) is None and oracle.mul_plan(5, 0) is None
    assert oracle.conv_plan(3, 3, 1 << 63, 1, True) is None
    assert oracle.conv_plan(0, 3, 1, 1, False) is None
    # A coefficient above its admitted bound fails (src/convolve.rs:140-150).
    a = np.array([1, 2, 9], dtype=np.uint64)
    assert oracle.convolve(a, np.array([1, 1], dtype=np.uint64), 8, 1, False) is None


def test_digest_check_catches_a_planted_error(oracle, digests):
    """The digest and sample checks fail on a product with one wrong limb."""
    a, b = oracle.operand(1 << 20, 1), oracle.operand(1 << 20, 2)
    out = oracle.mul(a, b, threads=8)
    out[123457] ^= np.uint64(1)
    with pytest.raises(AssertionError):
        check_record(out, digests["mul_2p20"])


def test_mod_limbs_matches_integers(oracle):
    """oracle_mod_limbs (Horner over limbs, most significant first) against
    Python's exact remainder, for moduli of 2..64 bits and edge operands."""
    rng = np.random.default_rng(61)
    moduli = [2, 3, 65537, (1 << 61) - 1, (1 << 63) + 29, (1 << 64) - 59] + [oracle.prime(i) for i in range(8)]
    operands = [np.zeros(1, dtype=np.uint64), np.full(7, MASK, dtype=np.uint64),
                np.array([MASK, 0, 1], dtype=np.uint64)]
    operands += [rng.integers(0, MASK, n, dtype=np.uint64, endpoint=True) for n in (1, 2, 5, 64, 1000)]
    for q in moduli:
        for a in operands:
            assert oracle.mod_limbs(a, q) == to_int(a) % q, (q, len(a))
    # A plausible slip (limbs taken least significant first) would differ.
    a = np.array([1, 2], dtype=np.uint64)
    assert oracle.mod_limbs(a, (1 << 61) - 1) == (1 + (2 << 64)) % ((1 << 61) - 1)


def test_workspace_mul_empty_operand_rule(oracle):
    """MulWorkspace::run (src/plan.rs:182-205): an empty operand gives a zero
    result of a_len + b_len limbs and returns true; nothing beyond it is
    written; non-empty operands give the product, the square flag squares a."""
    empty = np.zeros(0, dtype=np.uint64)
    a = oracle.operand(37, 5)
    b = oracle.operand(12, 6)
    for x, y in [(empty, b), (a, empty), (empty, empty)]:
        n = len(x) + len(y)
        dst = np.full(n + 3, 0xDEAD, dtype=np.uint64)
        assert oracle.workspace_mul(dst, x, y)
        assert (dst[:n] == 0).all() and (dst[n:] == 0xDEAD).all()
    # Too short a destination for the zero result: false, dst untouched.
    dst = np.full(4, 7, dtype=np.uint64)
    assert not oracle.workspace_mul(dst, empty, b)
    assert (dst == 7).all()
    dst = np.zeros(len(a) + len(b), dtype=np.uint64)
    assert oracle.workspace_mul(dst, a, b)
    assert to_int(dst) == to_int(a) * to_int(b)
    dst = np.zeros(2 * len(a), dtype=np.uint64)
    assert oracle.workspace_mul(dst, a, a, square=True)
    assert to_int(dst) == to_int(a) ** 2
