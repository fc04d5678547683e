"""Concurrent plan calls. MulPlan::mul / sqr and ConvPlan::convolve take
&self and run concurrently in the reference (src/plan.rs:94-114,
src/convolve.rs:148-162); here each calling thread has its own executor
(stream and device arenas), so four host threads calling at once must each
get the oracle's result, and together move more products per second than
one thread calling alone."""
import sys
import threading
import time

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


def run_threads(work, threads):
    errors = []

    def body(items):
        try:
            for fn in items:
                fn()
        except Exception as e:  # reported by the main thread
            errors.append(repr(e))

    ts = [threading.Thread(target=body, args=(work[i::threads],)) for i in range(threads)]
    t0 = time.perf_counter()
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    return time.perf_counter() - t0


def test_concurrent_calls_match_the_oracle(oracle):
    shapes = [(3000, 70000), (20000, 30000), (100000, 100000), (1024, 1024), (500, 500), (65536, 1 << 17)]
    cases = []
    for i, (la, lb) in enumerate(shapes * 2):
        a, b = oracle.operand(la, 10 + i), oracle.operand(lb, 40 + i)
        cases.append(("mul", a, b, oracle.mul(a, b, threads=8)))
        cases.append(("sqr", a, a, oracle.sqr(a, threads=8)))
    x = oracle.operand(50000, 7) & np.uint64(0xFFFF)
    y = oracle.operand(40000, 8) & np.uint64(0xFFFF)
    cases.append(("conv", x, y, oracle.convolve(x, y, 65535, 65535, False, threads=8)))
    outs = [np.zeros(len(w), dtype=np.uint64) for _, _, _, w in cases]

    def job(k):
        kind, a, b, _ = cases[k]

        def run():
            if kind == "conv":
                assert ntt.ConvPlan.new(len(a), len(b), 65535, 65535, False).convolve(outs[k], a, b), ntt.last_error()
            elif kind == "sqr":
                assert ntt.MulPlan.new(len(a), len(a)).sqr(outs[k], a), ntt.last_error()
            else:
                assert ntt.MulPlan.new(len(a), len(b)).mul(outs[k], a, b), ntt.last_error()
        return run

    say(f"{len(cases)} calls on 4 threads")
    run_threads([job(k) for k in range(len(cases))], 4)
    for k, (kind, a, b, want) in enumerate(cases):
        assert (outs[k] == want).all(), f"call {k} ({kind} {len(a)}x{len(b)})"


def test_threads_beat_one_thread():
    n = 1 << 17
    a = ntt.pinned_array(n)
    b = ntt.pinned_array(n)
    a[:] = np.arange(n, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    b[:] = a[::-1]
    plan = ntt.MulPlan.new(n, n)
    dsts = [ntt.pinned_array(2 * n) for _ in range(4)]
    count = 32

    def work(threads):
        jobs = []
        for i in range(count):
            d = dsts[i % threads]
            jobs.append(lambda d=d: plan.mul(d, a, b) or (_ for _ in ()).throw(AssertionError(ntt.last_error())))
        return run_threads(jobs, threads)

    work(1)
    work(4)  # every thread's executor warm
    one = min(work(1) for _ in range(3))
    four = min(work(4) for _ in range(3))
    say(f"{count} products of {n} x {n} limbs: 1 thread {one * 1e3:.1f} ms, 4 threads {four * 1e3:.1f} ms")
    ref = dsts[0].copy()
    for d in dsts[1:]:
        assert (d == ref).all()
    assert four < one, "four threads calling at once are no faster than one"
