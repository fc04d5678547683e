"""The served form: continuous batching under a device-memory budget.

Every request served in a mix (products, squares, convolutions; random,
sparse and all-ones operands; small products grouped, large ones admitted as
memory frees up) must give the same limbs as the same request run alone
through the plain per-call path (MulPlan::mul / sqr, ConvPlan::convolve,
src/plan.rs:94-114, src/convolve.rs:148-162), and a sample is checked
against the oracle.
"""
import sys

import numpy as np
import pytest

import nttanafis_amd as ntt
from nttanafis_amd.workloads import CONV, MUL, SQR, operand_pools, output_len, request_operands, served_stream

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    ntt.load()
    assert ntt.device_available(), "no gfx950 device visible"
    yield


@pytest.fixture(scope="module")
def pools():
    return operand_pools(1 << 18)


def alone(req, a, b):
    """The request through the plain per-call path."""
    dst = np.zeros(output_len(req), dtype=np.uint64)
    if req["kind"] == CONV:
        plan = ntt.ConvPlan.new(len(a), len(b), 0xFFFF, 0xFFFF, False)
        assert plan.convolve(dst, a, b), ntt.last_error()
    elif req["kind"] == SQR:
        assert ntt.MulPlan.new(len(a), len(a)).sqr(dst, a), ntt.last_error()
    else:
        assert ntt.MulPlan.new(len(a), len(b)).mul(dst, a, b), ntt.last_error()
    return dst


def submit(server, req, a, b, dst):
    if req["kind"] == CONV:
        return server.submit_conv(a, b, 0xFFFF, 0xFFFF, False, dst)
    if req["kind"] == SQR:
        return server.submit_sqr(a, dst)
    return server.submit_mul(a, b, dst)


def test_served_mix_equals_alone(pools, oracle, monkeypatch):
    # Transforms for every request, so that the large working sets (spectra
    # included) exceed the budget together; the schoolbook runs in the other
    # served tests.
    monkeypatch.setenv("NTTAMD_SCHOOL", "0")
    # Two of the largest products first: the second cannot be admitted while
    # the first holds its working set (each needs about 33 MiB of the 48).
    big = [{"kind": MUL, "arrival": 0.0, "len_a": 1 << 18, "len_b": 1 << 18, "fill_a": "random", "fill_b": "random",
            "off_a": 0, "off_b": 0} for _ in range(2)]
    reqs = big + served_stream(96, 0.0, lo=256, hi=1 << 18, seed=21)
    say(f"served mix: {len(reqs)} requests")
    # A budget of 48 MiB holds only a few large working sets at once, so
    # admissions wait for memory while small products run in groups.
    server = ntt.Server(budget_bytes=48 << 20, slots=4)
    outs, tickets = [], []
    for req in reqs:
        a, b = request_operands(req, pools)
        dst = np.zeros(output_len(req), dtype=np.uint64)
        t = submit(server, req, a, b, dst)
        assert t >= 0, ntt.last_error()
        outs.append(dst)
        tickets.append(t)
    assert server.drain()
    st = server.stats()
    say(f"stats {st}")
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["groups"] >= 1 and st["grouped"] >= 2
    assert st["waits_for_memory"] >= 1
    assert st["peak_bytes"] <= st["budget_bytes"]
    for i, req in enumerate(reqs):
        assert server.wait(tickets[i]) == 1
        a, b = request_operands(req, pools)
        assert (outs[i] == alone(req, a, b)).all(), f"request {i} ({req['kind']} {req['len_a']}x{req['len_b']})"
    rng = np.random.default_rng(1)
    for i in rng.choice(len(reqs), size=12, replace=False).tolist():
        req = reqs[i]
        a, b = request_operands(req, pools)
        if req["kind"] == CONV:
            want = oracle.convolve(a, b, 0xFFFF, 0xFFFF, False, threads=8)
        elif req["kind"] == SQR:
            want = oracle.sqr(a)
        else:
            want = oracle.mul(a, b)
        assert (outs[i] == want).all(), f"request {i} vs oracle"
    server.close()


def test_served_paced_arrivals(pools):
    """Requests submitted while earlier ones are still on the device."""
    reqs = served_stream(40, 2000.0, lo=256, hi=1 << 18, seed=33)
    server = ntt.Server(budget_bytes=256 << 20, slots=4)
    t0 = ntt.clock()
    outs, tickets = [], []
    for req in reqs:
        while ntt.clock() - t0 < req["arrival"]:
            pass
        a, b = request_operands(req, pools)
        dst = np.zeros(output_len(req), dtype=np.uint64)
        tickets.append(submit(server, req, a, b, dst))
        outs.append(dst)
    for i, req in enumerate(reqs):
        assert server.wait(tickets[i]) == 1
        s, b0, d = server.times(tickets[i])
        assert s <= b0 <= d
        a, b = request_operands(req, pools)
        assert (outs[i] == alone(req, a, b)).all(), f"request {i}"
        server.release(tickets[i])
        assert server.poll(tickets[i]) == -1
    server.close()


def test_served_device_requests(pools):
    """Device-resident operands and results (no copies through the host)."""
    server = ntt.Server(budget_bytes=64 << 20, slots=2)
    a = pools["limbs"]["random"][:3000].copy()
    b = pools["limbs"]["random_b"][:70000].copy()
    da, db = ntt.DeviceBuffer.from_host(a), ntt.DeviceBuffer.from_host(b)
    d1, d2, d3 = ntt.DeviceBuffer(8 * 73000), ntt.DeviceBuffer(8 * 6000), ntt.DeviceBuffer(8 * 5999)
    t1 = server.submit_device(0, da.ptr, 3000, db.ptr, 70000, d1.ptr, 73000)
    t2 = server.submit_device(1, da.ptr, 3000, 0, 0, d2.ptr, 6000)
    t3 = server.submit_device(2, da.ptr, 3000, da.ptr, 3000, d3.ptr, 5999, max_a=(1 << 64) - 1,
                              max_b=(1 << 64) - 1)
    assert min(t1, t2) >= 0
    assert server.wait(t1) == 1 and server.wait(t2) == 1
    mul = alone({"kind": MUL, "len_a": 3000, "len_b": 70000}, a, b)
    sqr = alone({"kind": SQR, "len_a": 3000, "len_b": 3000}, a, a)
    assert (d1.download(73000) == mul).all()
    assert (d2.download(6000) == sqr).all()
    # A full-width convolution of 3000 terms needs more than one limb per
    # coefficient: 5999 limbs are too few, so the request is refused.
    assert t3 == -1
    server.close()


def test_served_failures(pools):
    server = ntt.Server(budget_bytes=16 << 20, slots=1)
    a = pools["limbs"]["random"][:1000].copy()
    short = np.zeros(1999, dtype=np.uint64)
    assert server.submit_mul(a, a, short) == -1  # dst one limb short
    # Larger than the whole budget: refused at once.
    big = pools["limbs"]["random"][:1 << 18].copy()
    dst = np.zeros(1 << 19, dtype=np.uint64)
    assert server.submit_sqr(big, dst) == -1
    # A coefficient above its bound fails the request and leaves dst alone.
    c = np.full(500, 70000, dtype=np.uint64)
    out = np.full(999, 7, dtype=np.uint64)
    t = server.submit_conv(c, c, 0xFFFF, 0xFFFF, False, out)
    assert t >= 0
    assert server.wait(t) == -1
    assert (out == 7).all()
    # The server keeps serving after a failed request.
    ok = np.zeros(2000, dtype=np.uint64)
    t = server.submit_mul(a, a, ok)
    assert server.wait(t) == 1
    assert (ok == alone({"kind": MUL, "len_a": 1000, "len_b": 1000}, a, a)).all()
    assert not server.drain()  # one request failed
    server.close()


def test_served_full_size_range(oracle):
    """Config 5's whole size range: 48 requests with operands log-uniform in
    [256, 4,194,304] limbs (16-bit coefficients for convolutions), 10 % sparse
    or all-ones, all submitted at once under a 2 GiB budget. Every result
    equals the same request alone through the per-call path; the largest
    requests of each kind and a random sample also match the oracle."""
    hi = 1 << 22
    pools = operand_pools(hi)
    reqs = served_stream(48, 0.0, lo=256, hi=hi, seed=55)
    sizes = [max(r["len_a"], r["len_b"]) for r in reqs]
    say(f"served full range: {len(reqs)} requests, largest operand {max(sizes)}, "
        f"{sum(s > (1 << 18) for s in sizes)} above 2^18 limbs")
    assert max(sizes) > (1 << 21), "the draw must reach the top of the range"
    server = ntt.Server(budget_bytes=2 << 30, slots=4)
    outs, tickets = [], []
    for req in reqs:
        a, b = request_operands(req, pools)
        dst = np.zeros(output_len(req), dtype=np.uint64)
        t = submit(server, req, a, b, dst)
        assert t >= 0, ntt.last_error()
        outs.append(dst)
        tickets.append(t)
    assert server.drain()
    st = server.stats()
    say(f"stats {st}")
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["peak_bytes"] <= st["budget_bytes"]
    for i, req in enumerate(reqs):
        assert server.wait(tickets[i]) == 1
        a, b = request_operands(req, pools)
        assert (outs[i] == alone(req, a, b)).all(), f"request {i} ({req['kind']} {req['len_a']}x{req['len_b']})"
    say("served full range: every request equals the per-call path")
    largest = {}
    for i, req in enumerate(reqs):
        k = req["kind"]
        if k not in largest or sizes[i] > sizes[largest[k]]:
            largest[k] = i
    rng = np.random.default_rng(2)
    sample = sorted(set(largest.values()) | set(rng.choice(len(reqs), size=4, replace=False).tolist()))
    for i in sample:
        req = reqs[i]
        a, b = request_operands(req, pools)
        if req["kind"] == CONV:
            want = oracle.convolve(a, b, 0xFFFF, 0xFFFF, False, threads=16)
        elif req["kind"] == SQR:
            want = oracle.sqr(a, threads=16)
        else:
            want = oracle.mul(a, b, threads=16)
        assert (outs[i] == want).all(), f"request {i} ({req['kind']} {req['len_a']}x{req['len_b']}) vs oracle"
        say(f"oracle ok: request {i} {req['kind']} {req['len_a']}x{req['len_b']}")
    server.close()


def test_large_request_not_starved_by_small_traffic(pools):
    """Arrival order: with one slot, a large request queued behind a run of
    small products and ahead of another run starts before every small
    product submitted after it (small products are grouped only from those
    queued ahead of it)."""
    server = ntt.Server(budget_bytes=256 << 20, slots=1)
    small = [{"kind": MUL, "arrival": 0.0, "len_a": 700 + k, "len_b": 900, "fill_a": "random", "fill_b": "random",
              "off_a": k, "off_b": 2 * k} for k in range(120)]
    large = {"kind": MUL, "arrival": 0.0, "len_a": 1 << 17, "len_b": 1 << 17, "fill_a": "random",
             "fill_b": "random", "off_a": 0, "off_b": 0}
    reqs = small[:60] + [large] + small[60:]
    outs, tickets = [], []
    for req in reqs:
        a, b = request_operands(req, pools)
        dst = np.zeros(output_len(req), dtype=np.uint64)
        t = submit(server, req, a, b, dst)
        assert t >= 0, ntt.last_error()
        outs.append(dst)
        tickets.append(t)
    assert server.drain()
    t_large = server.times(tickets[60])[1]
    later = [server.times(t)[1] for t in tickets[61:]]
    say(f"large request started at {t_large:.6f}; first later small start {min(later):.6f}")
    assert all(t_large <= t for t in later), "a small product submitted later started before the large request"
    for i in (0, 60, 61, len(reqs) - 1):
        a, b = request_operands(reqs[i], pools)
        assert (outs[i] == alone(reqs[i], a, b)).all(), f"request {i}"
    server.close()


def test_served_stream_window_with_long_by_short_split_product(oracle):
    """Requests 1100-1159 of the 1,200-request served stream (sizes up to
    4,194,304 limbs, operands and results in pinned memory, as the served
    measurement runs them), served twice into adjacent slices of one output
    pool: request 1131, a 3,245,183 x 387-limb split product, against the
    oracle; every other request against the plain path."""
    base = served_stream(1200, 1.0, hi=4194304)
    big_pools = operand_pools(4194304, alloc=ntt.pinned_array)
    window = list(range(1100, 1160))
    ops = {i: request_operands(base[i], big_pools) for i in window}
    a, b = ops[1131]
    assert (len(a), len(b)) == (3245183, 387)
    want = {i: alone(base[i], *ops[i]) for i in window}
    assert (want[1131] == oracle.mul(np.ascontiguousarray(a), np.ascontiguousarray(b), threads=16)).all()
    lens = [output_len(base[i]) for i in window]
    pool = ntt.pinned_array(int(sum(lens)))
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    dsts = {i: pool[int(o):int(o) + n] for i, o, n in zip(window, offs, lens)}
    server = ntt.Server(budget_bytes=0, slots=4)
    for rep in range(2):
        say(f"stream window, pass {rep}")
        pool[:] = 0
        tickets = []
        for i in window:
            t = submit(server, base[i], *ops[i], dsts[i])
            assert t >= 0, ntt.last_error()
            tickets.append(t)
        assert server.drain()
        for t in tickets:
            server.release(t)
        bad = [i for i in window if not (dsts[i] == want[i]).all()]
        assert not bad, (rep, bad)
    server.close()
