"""Parity of the gfx950 engine with the oracle and the golden fixtures.

Every test calls the engine through the C ABI (nttanafis_amd over ctypes)
and needs an MI355X. Results are compared bit for bit.
"""
import sys

import numpy as np
import pytest

from conftest import (MASK, all_ones_product, check_record, conv16_operands, conv32s_operands, from_int,
                      signed_conv_int, to_int, unsigned_conv_int)

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


def engine_mul(a, b):
    plan = ntt.MulPlan.new(len(a), len(b))
    dst = np.zeros(len(a) + len(b), dtype=np.uint64)
    assert plan.mul(dst, a, b), ntt.last_error()
    return dst


def engine_sqr(a):
    plan = ntt.MulPlan.new(len(a), len(a))
    dst = np.zeros(2 * len(a), dtype=np.uint64)
    assert plan.sqr(dst, a), ntt.last_error()
    return dst


def test_config1_golden(golden_config1):
    say("config1 1024x1024")
    a, b = golden_config1["a"], golden_config1["b"]
    assert (engine_mul(a, b) == golden_config1["mul"]).all()
    assert (engine_sqr(a) == golden_config1["sqr"]).all()
    ws = ntt.MulWorkspace()
    ws.prepare(ntt.MulPlan.new(1024, 1024))
    dst = np.zeros(2048, dtype=np.uint64)
    assert ws.mul(dst, a, b) and (dst == golden_config1["mul"]).all()
    assert ws.sqr(dst, a) and (dst == golden_config1["sqr"]).all()
    assert ws.mul(dst, a, a) and (dst == golden_config1["sqr"]).all()


def test_documentation_examples(kat):
    for key in ["doc_mul", "doc_workspace"]:
        a = np.array(kat[key]["a"], dtype=np.uint64)
        b = np.array(kat[key]["b"], dtype=np.uint64)
        assert [int(x) for x in engine_mul(a, b)] == kat[key]["out"]
    a = np.array(kat["doc_sqr"]["a"], dtype=np.uint64)
    assert [int(x) for x in engine_sqr(a)] == kat["doc_sqr"]["out"]
    c = kat["conv_123_45"]
    plan = ntt.ConvPlan.new(3, 2, 3, 5, False)
    dst = np.zeros(plan.output_len(), dtype=np.uint64)
    assert plan.convolve(dst, np.array(c["a"], dtype=np.uint64), np.array(c["b"], dtype=np.uint64))
    assert [int(x) for x in dst] == c["out"]


@pytest.mark.parametrize("seed", range(4))
def test_random_shapes_match_oracle(oracle, seed):
    rng = np.random.default_rng(seed)
    shapes = [(1, 1), (1, 2), (2, 2), (3, 7), (31, 33), (64, 48), (100, 1000), (192, 192),
              (500, 9000), (4096, 4096), (4097, 4095), (12345, 678), (40000, 40000)]
    for la, lb in shapes:
        say(f"seed {seed} shape {la}x{lb}")
        a = rng.integers(0, 2 ** 64, la, dtype=np.uint64)
        b = rng.integers(0, 2 ** 64, lb, dtype=np.uint64)
        if seed == 1:
            a[rng.random(la) < 0.9] = 0
        if seed == 2:
            a[:] = np.uint64(MASK)
        if seed == 3:
            b[rng.random(lb) < 0.5] = np.uint64(MASK)
        assert (engine_mul(a, b) == oracle.mul(a, b, threads=4)).all(), (la, lb)


def test_default_forms_over_random_shapes(oracle):
    """Shapes log-uniform in [1, 200,000] limbs each way run in whatever form
    the engine's pricing picks (schoolbook, unbalanced split, whole
    power-of-two row); every form occurs and every product equals the
    oracle."""
    rng = np.random.default_rng(77)
    seen = {}
    for k in range(48):
        la = int(np.exp(rng.uniform(0, np.log(200000))))
        lb = int(np.exp(rng.uniform(0, np.log(200000))))
        form = ("schoolbook" if ntt.mul_schoolbook(la, lb) else
                "split" if ntt.mul_split(la, lb)[0] else "row")
        seen[form] = seen.get(form, 0) + 1
        say(f"shape {la}x{lb}: {form}")
        a = rng.integers(0, 2 ** 64, la, dtype=np.uint64)
        b = rng.integers(0, 2 ** 64, lb, dtype=np.uint64)
        if k % 4 == 1:
            a[:] = np.uint64(MASK)
        assert (engine_mul(a, b) == oracle.mul(a, b, threads=8)).all(), (la, lb, form)
    assert set(seen) == {"schoolbook", "split", "row"}, seen


@pytest.mark.parametrize("n", [1, 5, 1000, 1024, 4096, 65536, 1 << 18])
def test_squares_match_oracle(oracle, n):
    say(f"square {n}")
    a = oracle.operand(n, 11)
    expect = oracle.sqr(a, threads=8)
    assert (engine_sqr(a) == expect).all()
    assert (engine_mul(a, a) == expect).all()


# The last four shapes sit at the edge of the carry pass's direct form
# (2,048 output blocks of 1,024 limbs, or of 64 limbs for a schoolbook
# product) and one block past it, where the block functions are scanned first.
@pytest.mark.parametrize("n,m", [(1, 1), (7, 3), (700, 701), (1024, 1024), (1500, 548), (5000, 17), (65536, 65536),
                                 (300000, 200000), (1 << 20, 1 << 20), ((1 << 20) + 1024, 1 << 20), (8, 131064),
                                 (8, 131128)])
def test_all_ones_carry_chains(n, m):
    say(f"all ones {n}x{m}")
    a = np.full(n, MASK, dtype=np.uint64)
    b = np.full(m, MASK, dtype=np.uint64)
    assert (engine_mul(a, b) == all_ones_product(n, m)).all()


def test_carry_ripples_across_blocks():
    """(B^n - 1) * (B^k + 1): long runs of all-ones limbs straddle the
    1024-limb carry blocks."""
    for n, k in [(5000, 1), (5000, 1023), (5000, 1024), (70000, 3000)]:
        a = np.full(n, MASK, dtype=np.uint64)
        b = np.zeros(k + 1, dtype=np.uint64)
        b[0] = 1
        b[k] = 1
        got = engine_mul(a, b)
        assert to_int(got) == to_int(a) * to_int(b), (n, k)


def test_mul_2p20_golden(oracle, digests):
    say("mul 2^20")
    a, b = oracle.operand(1 << 20, 1), oracle.operand(1 << 20, 2)
    check_record(engine_mul(a, b), digests["mul_2p20"])


def test_all_ones_2p20():
    say("all ones 2^20")
    a = np.full(1 << 20, MASK, dtype=np.uint64)
    assert (engine_mul(a, a.copy()) == all_ones_product(1 << 20, 1 << 20)).all()


def test_sparse_operands_2p20(oracle):
    """2^20-limb operands that are mostly zero (the served stream's sparse
    operands, and every zero group the packing skips): a few nonzero limbs
    against a random operand, and two operands whose only nonzero limb is
    the top one; checked against Python integers."""
    n = 1 << 20
    say("sparse 2^20")
    rng = np.random.default_rng(20)
    b = oracle.operand(n, 2)
    a = np.zeros(n, dtype=np.uint64)
    spots = [0, 1, 12345, 524287, 524288, n - 2, n - 1]
    a[spots] = rng.integers(1, 2**64, len(spots), dtype=np.uint64, endpoint=False)
    bi = to_int(b)
    want = sum(int(a[i]) * bi << (64 * i) for i in spots)
    assert to_int(engine_mul(a, b)) == want
    c = np.zeros(n, dtype=np.uint64)
    d = np.zeros(n, dtype=np.uint64)
    c[-1] = np.uint64(0xFEDCBA9876543210)
    d[-1] = np.uint64(0x0123456789ABCDEF)
    got = engine_mul(c, d)
    assert (got == from_int(int(c[-1]) * int(d[-1]) << (64 * (2 * n - 2)), 2 * n)).all()
    assert (engine_sqr(c) == from_int(int(c[-1]) ** 2 << (64 * (2 * n - 2)), 2 * n)).all()


def test_largest_tested_product():
    """2^26 x 2^26 limbs (512 MiB operands, six primes on 2^26-point rows in
    three passes, 1 GiB result) against its residue modulo 2^61 - 1, and
    2^26 all-ones limbs squared against the closed form."""
    n = 1 << 26
    a = oracle_operand(n, 5)
    b = oracle_operand(n, 6)
    primes, width, log_n, tail = ntt.mul_geometry(n, n)
    say(f"2^26 x 2^26: {primes} primes x {width} bits, row 2^{log_n} (+{tail}/4)")
    got = engine_mul(a, b)
    q = (1 << 61) - 1
    assert residue(got, q) == residue(a, q) * residue(b, q) % q
    del got, b
    a[:] = np.uint64(MASK)
    assert (engine_sqr(a) == all_ones_product(n, n)).all()


def oracle_operand(n, seed):
    import oracle
    return oracle.operand(n, seed)


def residue(limbs, q):
    return int.from_bytes(np.ascontiguousarray(limbs, dtype="<u8").tobytes(), "little") % q


def test_sqr_2p24_golden(oracle, digests):
    say("sqr 2^24")
    a = oracle.operand(1 << 24, 1)
    check_record(engine_sqr(a), digests["sqr_2p24"])


def test_conv16_2p22_golden(digests):
    say("conv16 2^22")
    a, b = conv16_operands(1 << 22)
    plan = ntt.ConvPlan.new(len(a), len(b), 65535, 65535, False)
    dst = np.zeros(plan.output_len(), dtype=np.uint64)
    assert plan.convolve(dst, a, b), ntt.last_error()
    check_record(dst, digests["conv16_2p22"])


def test_conv32s_2p22_golden(digests):
    say("conv32s 2^22")
    a, b = conv32s_operands(1 << 22)
    plan = ntt.ConvPlan.new(len(a), len(b), 1 << 31, 1 << 31, True)
    ws = ntt.ConvWorkspace()
    ws.prepare(plan)
    dst = np.zeros(plan.output_len(), dtype=np.uint64)
    assert ws.convolve(dst, a, b), ntt.last_error()
    check_record(dst, digests["conv32s_2p22"])


def conv_counts(n):
    """Terms of coefficient k of a convolution of two length-n vectors:
    min(k + 1, 2n - 1 - k)."""
    k = np.arange(2 * n - 1, dtype=np.uint64)
    return np.minimum(k + np.uint64(1), np.uint64(2 * n - 1) - k)


def test_conv16_2p22_all_max():
    """Config-3 shape with every coefficient at its bound (65,535): output
    coefficient k is 65535^2 times its term count, the largest values the
    plan's bound admits (closed form)."""
    n = 1 << 22
    say("conv16 2^22 all 65535")
    a = np.full(n, 65535, dtype=np.uint64)
    plan = ntt.ConvPlan.new(n, n, 65535, 65535, False)
    assert plan.stride() == 1
    dst = np.zeros(plan.output_len(), dtype=np.uint64)
    assert plan.convolve(dst, a, a.copy()), ntt.last_error()
    assert (dst == conv_counts(n) * np.uint64(65535 * 65535)).all()


def test_conv32s_2p22_all_min():
    """Signed 32-bit config-3 shape with every coefficient -2^31: output
    coefficient k is 2^62 times its term count (up to 2^84, two limbs in
    two's complement), the largest magnitude the plan's bound admits."""
    n = 1 << 22
    say("conv32s 2^22 all -2^31")
    a = np.full(n, np.uint64((1 << 64) - (1 << 31)), dtype=np.uint64)
    plan = ntt.ConvPlan.new(n, n, 1 << 31, 1 << 31, True)
    assert plan.stride() == 2
    dst = np.zeros(plan.output_len(), dtype=np.uint64)
    assert plan.convolve(dst, a, a.copy()), ntt.last_error()
    m = conv_counts(n)
    assert (dst[0::2] == (m & np.uint64(3)) << np.uint64(62)).all()
    assert (dst[1::2] == m >> np.uint64(2)).all()


@pytest.mark.parametrize("bits,signed", [(1, False), (16, False), (32, True), (49, False), (50, True),
                                         (51, False), (52, False), (52, True), (53, False), (63, True),
                                         (64, False)])
def test_convolutions_match_integers(bits, signed):
    """Convolutions of 1- to 64-bit coefficients (ConvPlan::convolve,
    src/convolve.rs:75-203) against exact integers or the oracle; 50 to 52
    bits put one to three bits in the top 49-bit part, the SHORT_HIGH
    packing form (src/arch/schedule/expression/packing.rs:12, 99-101)."""
    rng = np.random.default_rng(bits)
    for la, lb in [(3, 2), (24, 30), (25, 25), (200, 77), (3000, 2900)]:
        say(f"conv bits {bits} signed {signed} {la}x{lb}")
        if signed:
            lim = 1 << (bits - 1)
            a = rng.integers(-lim + 1, lim, la, dtype=np.int64).astype(np.uint64)
            b = rng.integers(-lim + 1, lim, lb, dtype=np.int64).astype(np.uint64)
            mx = lim - 1
        else:
            mx = (1 << bits) - 1
            a = rng.integers(0, mx, la, dtype=np.uint64, endpoint=True)
            b = rng.integers(0, mx, lb, dtype=np.uint64, endpoint=True)
        plan = ntt.ConvPlan.new(la, lb, mx, mx, signed)
        dst = np.zeros(plan.output_len(), dtype=np.uint64)
        assert plan.convolve(dst, a, b), ntt.last_error()
        if la * lb <= 100000:
            expect = (signed_conv_int if signed else unsigned_conv_int)(a, b, plan.stride())
            assert (dst == expect).all()
        else:
            import oracle as o
            assert (dst == o.convolve(a, b, mx, mx, signed, threads=4)).all()


def test_failures_leave_dst_unchanged():
    a = np.arange(1, 101, dtype=np.uint64)
    b = np.arange(1, 51, dtype=np.uint64)
    plan = ntt.MulPlan.new(100, 50)
    short = np.full(149, 7, dtype=np.uint64)
    assert not plan.mul(short, a, b)
    assert (short == 7).all()
    dst = np.full(150, 7, dtype=np.uint64)
    assert not plan.mul(dst, a[:99], b)  # length differs from the plan
    assert (dst == 7).all()
    assert not plan.sqr(dst, a)  # plan is not square
    cplan = ntt.ConvPlan.new(3, 2, 8, 1, False)
    cdst = np.full(4, 9, dtype=np.uint64)
    assert not cplan.convolve(cdst, np.array([1, 2, 9], dtype=np.uint64), np.array([1, 1], dtype=np.uint64))
    assert (cdst == 9).all()
    assert not ntt.ConvWorkspace().convolve(cdst, np.array([1, 2, 3], dtype=np.uint64),
                                            np.array([1, 1], dtype=np.uint64))


def test_success_preserves_suffix():
    a = np.arange(1, 101, dtype=np.uint64)
    b = np.arange(1, 51, dtype=np.uint64)
    dst = np.full(160, 7, dtype=np.uint64)
    assert ntt.MulPlan.new(100, 50).mul(dst, a, b)
    assert to_int(dst[:150]) == to_int(a) * to_int(b)
    assert (dst[150:] == 7).all()
    plan = ntt.ConvPlan.new(100, 50, 100, 50, False)
    cdst = np.full(plan.output_len() + 3, 7, dtype=np.uint64)
    assert plan.convolve(cdst, a, b)
    assert (cdst[plan.output_len():] == 7).all()


def test_workspace_empty_operands_write_zero():
    ws = ntt.MulWorkspace()
    dst = np.full(5, 7, dtype=np.uint64)
    assert ws.mul(dst, np.zeros(0, dtype=np.uint64), np.arange(3, dtype=np.uint64))
    assert (dst[:3] == 0).all() and (dst[3:] == 7).all()
    assert ws.sqr(dst, np.zeros(0, dtype=np.uint64))


def test_workspace_serves_plans_in_any_order(oracle):
    ws = ntt.MulWorkspace()
    rng = np.random.default_rng(5)
    for la, lb in [(5000, 5000), (10, 20), (70000, 3), (5000, 5000)]:
        a = rng.integers(0, 2 ** 64, la, dtype=np.uint64)
        b = rng.integers(0, 2 ** 64, lb, dtype=np.uint64)
        dst = np.zeros(la + lb, dtype=np.uint64)
        assert ws.mul(dst, a, b)
        assert (dst == oracle.mul(a, b)).all()


def test_device_resident_path_matches_host_path(oracle):
    a, b = oracle.operand(100000, 1), oracle.operand(77777, 2)
    da, db = ntt.DeviceBuffer.from_host(a), ntt.DeviceBuffer.from_host(b)
    dout = ntt.DeviceBuffer((len(a) + len(b)) * 8)
    ws = ntt.MulWorkspace()
    assert ws.mul_device(dout.ptr, len(a) + len(b), da.ptr, len(a), db.ptr, len(b))
    assert ntt.synchronize()
    assert (dout.download(len(a) + len(b)) == oracle.mul(a, b, threads=4)).all()
