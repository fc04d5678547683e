#!/usr/bin/env python3
"""Benchmark of the gfx950 multi-prime NTT engine.

Default workload: one exact product of two 2^20-limb operands (SplitMix64
limbs, seeds 1 and 2, as the reference benchmarks use), operands and result
resident in HBM. A step is one MulWorkspace product on the device; the metric
is operand limbs per second (Mlimbs/s). Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--workload mul_2p20|sqr_2p24|conv16_2p22|conv16_5x2p20|conv32s_2p22|batch4096|served]

The line also carries the batched form ("batched": 4,096 uneven products in
one graph launch against the plain one-call-per-product loop) unless
--no-batch is given.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import nttanafis_amd as ntt  # noqa: E402
from nttanafis_amd.workloads import (CONV, SQR, batch_operands, batch_shapes, operand, operand_pools,  # noqa: E402
                                     output_len, request_operands, served_stream)

HBM_PEAK_GBS = 8000.0
VALU_PEAK_G = 1024 * 2.4e9 / 4 / 1e9  # wave64 FP64 instructions per second, G


def splitmix(length: int, seed: int) -> np.ndarray:
    gamma = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(1, length + 1, dtype=np.uint64) * gamma)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def workload(name: str):
    """Returns (operands, step builder description)."""
    if name == "mul_2p20":
        return dict(kind="mul", a=splitmix(1 << 20, 1), b=splitmix(1 << 20, 2))
    if name == "sqr_2p24":
        return dict(kind="sqr", a=splitmix(1 << 24, 1))
    if name == "conv16_2p22":
        n = 1 << 22
        return dict(kind="conv", a=splitmix(n, 1) & np.uint64(0xFFFF), b=splitmix(n, 2) & np.uint64(0xFFFF),
                    max=65535, signed=False)
    if name == "conv16_5x2p20":
        # 5*2^20 coefficients: a convolution of 5*2^21 - 1 terms, just above
        # 2^23, on a composite row of 5*2^21 points.
        n = 5 << 20
        return dict(kind="conv", a=splitmix(n, 1) & np.uint64(0xFFFF), b=splitmix(n, 2) & np.uint64(0xFFFF),
                    max=65535, signed=False)
    if name == "conv32s_2p22":
        n = 1 << 22

        def s(seed):
            return (splitmix(n, seed) & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.int32).astype(
                np.int64).astype(np.uint64)

        return dict(kind="conv", a=s(1), b=s(2), max=1 << 31, signed=True)
    raise SystemExit(f"unknown workload {name}")


def row_parts(log_len: int, tail: int):
    """Transform lengths of the parts of a row: 2^log_len, plus the tails of a
    composite row (N/4 for tail 1, N/2 for tail 2, both for tail 3)."""
    n = 1 << log_len
    return [n] + {0: [], 1: [n // 4], 2: [n // 2], 3: [n // 2, n // 4]}[tail]


def passes_of(n: int) -> int:
    k = n.bit_length() - 1
    return 1 if k <= 12 else (2 if k <= 23 else 3)


def step_bytes(name: str, primes: int, log_len: int, tail: int, la: int, lb: int, out_len: int,
               square: bool) -> float:
    """Algorithmic HBM bytes that the launches of one kernel kind move in one
    step: every residue of a pass read and written once, packing reads the
    limbs once per transform it feeds, and the recombination of a composite
    row reads the values it combines and writes the ones it changes."""
    ops = 1 if square else 2
    limbs = 8.0 * (la if square else la + lb)
    n = 1 << log_len
    T = n + tail * (n // 4)
    total = 0.0
    for m in row_parts(log_len, tail):
        passes = passes_of(m)
        if name == "fused_single" and passes == 1:
            total += limbs + 8.0 * primes * m
        elif name == "fwd_pack_strided" and passes > 1:
            total += limbs + 8.0 * ops * primes * m
        elif name == "fwd_strided" and passes == 3:
            total += 16.0 * ops * primes * m
        elif name == "inv_strided" and passes == 3:
            total += 16.0 * primes * m
        elif name == "fused_rows" and passes > 1:
            total += (16.0 if square else 24.0) * primes * m
        elif name == "inv_strided_canon" and passes > 1:
            total += 16.0 * primes * m
    if name == "recombine" and tail:
        total = 8.0 * primes * {1: 7 * n / 4, 2: 5 * n / 2, 3: 13 * n / 4}[tail]
    if name == "crt_product":
        total = 8.0 * primes * min(T, out_len * 64) + 8.0 * out_len
    if name == "crt_disjoint":
        total = 8.0 * primes * (la + lb - 1) + 8.0 * out_len
    return total


def measured_valu(workload: str) -> dict:
    """VALU wave-instructions per launch of each kernel kind (SQ_INSTS_VALU
    from the committed rocprofv3 counter summary of this workload), lines of
    the form `VALU <kind> <instructions>`; empty when none are committed."""
    path = os.path.join(ROOT, "profiles", f"{workload}_counters.txt")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            parts = line.split()
            if len(parts) == 3 and parts[0] == "VALU":
                out[parts[1]] = int(float(parts[2]))
    return out


def measured_traffic(workload: str) -> dict:
    """HBM bytes per launch of each kernel kind, from the committed rocprofv3
    counter summary of this workload (FETCH_SIZE and WRITE_SIZE, each scaled
    by what the counter reads for the plain copy kernel). Lines of the form
    `TRAFFIC <kind> <bytes>`; empty when no summary is committed."""
    path = os.path.join(ROOT, "profiles", f"{workload}_counters.txt")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            parts = line.split()
            if len(parts) == 3 and parts[0] == "TRAFFIC":
                out[parts[1]] = int(float(parts[2]))
    return out


def cpu_baseline(wl, sample_seconds: float, threads: int = 1):
    """Times the oracle (the C restatement of the reference CPU path) on a
    bounded sample of the same workload, with `threads` host threads (the
    oracle parallelises over primes and tiles, as the reference's rayon
    schedule does over primes)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    a = wl["a"]
    t0 = time.perf_counter()
    runs = 0
    tag = f"{threads} thread" + ("s" if threads > 1 else "")
    if wl["kind"] == "mul":
        b = wl["b"]
        while True:
            oracle.mul(a, b, threads=threads)
            runs += 1
            if time.perf_counter() - t0 > sample_seconds:
                break
        per = (time.perf_counter() - t0) / runs
        return {"value": len(a) / per / 1e6, "unit": "Mlimbs/s", "cores": threads, "kind": "port",
                "sample": f"{runs} products of 2^20 x 2^20 limbs, oracle/ntt_oracle.c, {tag}"}
    if wl["kind"] == "sqr":
        n = 1 << 20
        sub = a[:n]
        while True:
            oracle.sqr(sub, threads=threads)
            runs += 1
            if time.perf_counter() - t0 > sample_seconds:
                break
        per = (time.perf_counter() - t0) / runs
        return {"value": n / per / 1e6, "unit": "Mlimbs/s", "cores": threads, "kind": "port",
                "sample": f"{runs} squares of 2^20 limbs (sub-sample of the 2^24 square), {tag}"}
    b = wl["b"]
    while True:
        oracle.convolve(a, b, wl["max"], wl["max"], wl["signed"], threads=threads)
        runs += 1
        if time.perf_counter() - t0 > sample_seconds:
            break
    per = (time.perf_counter() - t0) / runs
    return {"value": len(a) / per / 1e6, "unit": "Mcoeff/s", "cores": threads, "kind": "port",
            "sample": f"{runs} convolutions of {len(a)} coefficients, {tag}"}


def mod_q61(limbs) -> int:
    q = (1 << 61) - 1
    return int.from_bytes(np.ascontiguousarray(limbs, dtype="<u8").tobytes(), "little") % q


# The reference's published single-core figures for the same workloads
# (README.md:28 and :36 of the reference: one Zen 5 core at 5.09 GHz, AVX-512
# IFMA, MulWorkspace / ConvWorkspace with retained tables), as operand length
# over time.
PUBLISHED = {
    "mul_2p20": {"value": 24.2, "unit": "Mlimbs/s", "time_ms": 43.4,
                 "source": "reference README.md:28, one Zen 5 core (AVX-512 IFMA)"},
    "conv16_2p22": {"value": 31.3, "unit": "Mcoeff/s", "time_ms": 134.0,
                    "source": "reference README.md:36, one Zen 5 core (AVX-512 IFMA)"},
}

SMALL_SHAPES = [(64, 64), (192, 192), (384, 384), (512, 512), (8, 200000), (192, 65536)]


def bench_small(reps: int = 200) -> dict:
    """Per-call latency of single small and very unequal products through the
    device path, each shape run as the engine chooses, as a schoolbook
    (NTTAMD_SCHOOL=1) and through transforms (NTTAMD_SCHOOL=0): reps calls
    queued back to back, then reps calls each waited for. Both forms must
    give the same limbs."""
    rows = []
    saved = os.environ.get("NTTAMD_SCHOOL")
    try:
        for la, lb in SMALL_SHAPES:
            a, b = operand(la, 1), operand(lb, 2)
            da, db = ntt.DeviceBuffer.from_host(a), ntt.DeviceBuffer.from_host(b)
            dout = ntt.DeviceBuffer(8 * (la + lb))
            os.environ.pop("NTTAMD_SCHOOL", None)
            row = {"shape": [la, lb], "default": "schoolbook" if ntt.mul_schoolbook(la, lb) else "transforms"}
            results = []
            for mode, name in (("1", "schoolbook"), ("0", "transforms")):
                os.environ["NTTAMD_SCHOOL"] = mode
                ws = ntt.MulWorkspace()
                ws.prepare(ntt.MulPlan.new(la, lb))
                for _ in range(5):
                    assert ws.mul_device(dout.ptr, la + lb, da.ptr, la, db.ptr, lb), ntt.last_error()
                assert ntt.synchronize()
                t0 = time.perf_counter()
                for _ in range(reps):
                    ws.mul_device(dout.ptr, la + lb, da.ptr, la, db.ptr, lb)
                assert ntt.synchronize()
                queued = (time.perf_counter() - t0) / reps
                t0 = time.perf_counter()
                for _ in range(reps):
                    ws.mul_device(dout.ptr, la + lb, da.ptr, la, db.ptr, lb)
                    ntt.synchronize()
                waited = (time.perf_counter() - t0) / reps
                row[name] = {"us_per_call": round(queued * 1e6, 1), "us_waited": round(waited * 1e6, 1)}
                results.append(dout.download(la + lb))
            assert (results[0] == results[1]).all(), f"schoolbook and transform products differ at {la}x{lb}"
            rows.append(row)
    finally:
        if saved is None:
            os.environ.pop("NTTAMD_SCHOOL", None)
        else:
            os.environ["NTTAMD_SCHOOL"] = saved
    return {"reps": reps, "shapes": rows}


def bench_batch(steps: int, warmup: int, plain_steps: int) -> dict:
    """4,096 products, len_a and len_b log-uniform in [192, 65,536], 20%
    squares, operands resident in HBM. A step is one graph launch computing
    every product; the comparison is the plain per-call loop (one
    MulWorkspace device call per product) over the same operands."""
    la, lb, sq = batch_shapes()
    a, b = batch_operands(la, lb, sq)
    n = len(la)
    batch = ntt.Batch()
    assert batch.plan(la, lb, sq), ntt.last_error()
    parts, pa, pb, ofs = [], [], [], 0
    for i in range(n):
        parts.append(a[i])
        pa.append(ofs)
        ofs += la[i]
        if sq[i]:
            pb.append(pa[-1])
        else:
            parts.append(b[i])
            pb.append(ofs)
            ofs += lb[i]
    din = ntt.DeviceBuffer.from_host(np.concatenate(parts))
    out_sizes = [batch.output_len(i) for i in range(n)]
    out_ofs = np.concatenate([[0], np.cumsum(out_sizes)[:-1]]).astype(np.int64)
    total_out = int(sum(out_sizes))
    dout = ntt.DeviceBuffer(8 * total_out)
    assert batch.bind([din.ptr + 8 * x for x in pa], [din.ptr + 8 * x for x in pb],
                      [dout.ptr + 8 * int(x) for x in out_ofs]), ntt.last_error()
    t_tune = time.perf_counter()
    assert batch.run(), ntt.last_error()  # autotunes and captures the graph
    assert ntt.synchronize()
    tune_s = time.perf_counter() - t_tune
    for _ in range(warmup):
        assert batch.run()
    assert ntt.synchronize()
    # Per-kernel-kind times (the batch runs its launches eagerly while
    # profiling, with events around each launch).
    ntt.Profile.reset()
    ntt.Profile.enable(True)
    assert batch.run()
    assert ntt.synchronize()
    ntt.Profile.enable(False)
    prof = ntt.Profile.read()
    t0 = time.perf_counter()
    for _ in range(steps):
        if not batch.run():
            raise SystemExit("batch run failed: " + ntt.last_error())
    assert ntt.synchronize()
    batched_s = (time.perf_counter() - t0) / steps
    got = dout.download(total_out)
    # Plain per-call loop over the same operands and outputs.
    assert batch.run_plain(), ntt.last_error()
    assert ntt.synchronize()
    t1 = time.perf_counter()
    for _ in range(plain_steps):
        if not batch.run_plain():
            raise SystemExit("plain loop failed: " + ntt.last_error())
    assert ntt.synchronize()
    plain_s = (time.perf_counter() - t1) / plain_steps
    plain = dout.download(total_out)
    assert (got == plain).all(), "batched results differ from the plain per-call path"
    for i in range(0, n, 97):
        ra = mod_q61(a[i])
        rb = ra if sq[i] else mod_q61(b[i])
        seg = got[int(out_ofs[i]):int(out_ofs[i]) + out_sizes[i]]
        assert mod_q61(seg) == ra * rb % ((1 << 61) - 1), f"batch product {i} failed its modular check"
    limbs = int(sum(la) + sum(y for y, s in zip(lb, sq) if not s))
    return {
        "products": n,
        "products_per_s": round(n / batched_s, 1),
        "ms_per_batch": round(1000 * batched_s, 3),
        "plain_products_per_s": round(n / plain_s, 1),
        "plain_ms_per_batch": round(1000 * plain_s, 3),
        "speedup_vs_plain": round(plain_s / batched_s, 2),
        "operand_limbs": limbs,
        "Mlimbs_per_s": round(limbs / batched_s / 1e6, 1),
        "graph_launches": batch.launches(),
        "first_run_s": round(tune_s, 3),
        "kernels_ms": {k: round(v[0], 3) for k, v in prof.items()},
        "squares": int(sum(sq)),
    }


def bench_host_path(a, b, steps: int) -> dict:
    """The same product through the host-buffer API (MulPlan::mul with host
    slices): operands cross PCIe in, the product back out. Pageable numpy
    arrays and pinned (hipHostMalloc) arrays."""
    plan = ntt.MulPlan.new(len(a), len(b))
    res = {}
    for name, alloc in (("pageable", lambda n: np.empty(n, dtype=np.uint64)), ("pinned", ntt.pinned_array)):
        ha, hb, dst = alloc(len(a)), alloc(len(b)), alloc(len(a) + len(b))
        ha[:] = a
        hb[:] = b
        assert plan.mul(dst, ha, hb), ntt.last_error()
        t0 = time.perf_counter()
        for _ in range(steps):
            assert plan.mul(dst, ha, hb)
        per = (time.perf_counter() - t0) / steps
        res[name] = {"ms_per_call": round(1000 * per, 3), "Mlimbs_per_s": round(len(a) / per / 1e6, 1)}
    return res


def _pct(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if xs else None


def bench_served(count: int, load: float, hi: int = 4194304) -> dict:
    """Served request stream: Poisson arrivals of mul / sqr / 16-bit convolve
    requests (equal shares), sizes log-uniform in [256, hi], 10% of operands
    sparse or all-ones, host operands in pinned memory. The plain per-call
    path (one MulPlan / ConvPlan call at a time, in arrival order) and the
    server (continuous batching, admission by device-memory budget) serve the
    same stream at the same offered load, `load` times the plain path's
    capacity. Latency runs from a request's arrival to its result in host
    memory."""
    base = served_stream(count, 1.0, hi=hi)
    pools = operand_pools(hi, alloc=ntt.pinned_array)
    outs_len = [output_len(r) for r in base]
    out_pool = ntt.pinned_array(int(sum(outs_len)))
    ofs = np.concatenate([[0], np.cumsum(outs_len)[:-1]]).astype(np.int64)
    ops = [request_operands(r, pools) for r in base]
    dsts = [out_pool[int(o):int(o) + n] for o, n in zip(ofs, outs_len)]
    plans = []
    for r, (a, b) in zip(base, ops):
        if r["kind"] == CONV:
            plans.append(ntt.ConvPlan.new(len(a), len(b), 0xFFFF, 0xFFFF, False))
        else:
            plans.append(ntt.MulPlan.new(len(a), len(b)))

    def plain_call(i):
        r, (a, b), dst, plan = base[i], ops[i], dsts[i], plans[i]
        if r["kind"] == CONV:
            ok = plan.convolve(dst, a, b)
        elif r["kind"] == SQR:
            ok = plan.sqr(dst, a)
        else:
            ok = plan.mul(dst, a, b)
        if not ok:
            raise SystemExit(f"plain request {i} failed: {ntt.last_error()}")

    # Warm both paths (tables, arenas, autotuning) on the whole stream once.
    for i in range(count):
        plain_call(i)
    # Plain capacity: the stream back to back.
    t0 = ntt.clock()
    for i in range(count):
        plain_call(i)
    plain_busy = ntt.clock() - t0
    capacity = count / plain_busy
    digests = [hash(d.tobytes()) for d in dsts]
    rate = load * capacity
    arrivals = [r["arrival"] / rate for r in base]

    # Plain per-call path at the offered load.
    lat_plain = []
    t0 = ntt.clock()
    for i in range(count):
        while ntt.clock() - t0 < arrivals[i]:
            pass
        plain_call(i)
        lat_plain.append(ntt.clock() - t0 - arrivals[i])
    plain_span = ntt.clock() - t0

    server = ntt.Server(budget_bytes=0, slots=4)

    def serve(arr):
        out_pool[:] = 0
        t0 = ntt.clock()
        tickets = []
        for i in range(count):
            while ntt.clock() - t0 < arr[i]:
                pass
            r, (a, b), dst = base[i], ops[i], dsts[i]
            if r["kind"] == CONV:
                t = server.submit_conv(a, b, 0xFFFF, 0xFFFF, False, dst)
            elif r["kind"] == SQR:
                t = server.submit_sqr(a, dst)
            else:
                t = server.submit_mul(a, b, dst)
            if t < 0:
                raise SystemExit(f"served request {i} refused: {ntt.last_error()}")
            tickets.append(t)
        if not server.drain():
            raise SystemExit("a served request failed")
        lat, last = [], 0.0
        for i, t in enumerate(tickets):
            _, _, done = server.times(t)
            lat.append(done - t0 - arr[i])
            last = max(last, done - t0)
            server.release(t)
        bad = [i for i in range(count) if hash(dsts[i].tobytes()) != digests[i]]
        if bad:
            raise SystemExit(f"served results differ from the plain path for requests {bad[:5]}")
        return lat, last

    serve([0.0] * count)  # warm the server's slots and groups
    lat_served, served_span = serve(arrivals)
    _, burst_span = serve([0.0] * count)
    stats = server.stats()
    server.close()
    ms = lambda x: round(1000 * x, 3)  # noqa: E731
    return {
        "requests": count,
        "offered_load_per_s": round(rate, 1),
        "load_vs_plain_capacity": load,
        "plain_capacity_per_s": round(capacity, 1),
        "served_capacity_per_s": round(count / burst_span, 1),
        "capacity_ratio": round(plain_busy / burst_span, 2),
        "served": {"products_per_s": round(count / served_span, 1), "p50_ms": ms(_pct(lat_served, 50)),
                   "p99_ms": ms(_pct(lat_served, 99)), "mean_ms": ms(float(np.mean(lat_served)))},
        "plain": {"products_per_s": round(count / plain_span, 1), "p50_ms": ms(_pct(lat_plain, 50)),
                  "p99_ms": ms(_pct(lat_plain, 99)), "mean_ms": ms(float(np.mean(lat_plain)))},
        "server_stats": stats,
        "mix": {k: sum(1 for r in base if r["kind"] == k) for k in ("mul", "sqr", "conv")},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="mul_2p20")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-small", action="store_true", help="skip the small-product latency table")
    ap.add_argument("--no-host", action="store_true", help="skip the host-buffer API timing")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=240)
    ap.add_argument("--serve-load", type=float, default=0.7)
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="after the timed region, run the plain copy kernel over this many bytes "
                         "(the calibration reference for hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        return 2
    ntt.load()
    if not ntt.device_available():
        print("bench.py: no gfx950 device", file=sys.stderr)
        return 1

    if args.workload == "served":
        res = bench_served(args.serve_requests, args.serve_load)
        print(json.dumps({"metric": "served products/s and p99 latency", "value": res["served"]["products_per_s"],
                          "unit": "requests/s", "n_gpus": 1, "steps": args.serve_requests, "warmup": 1,
                          "ms_per_step": res["served"]["p99_ms"], "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "dtype": "f64", "data": "synthetic (SplitMix64 limbs)",
                          "config": {"workload": "served"}, "served": res}), flush=True)
        return 0
    if args.workload == "small":
        res = bench_small()
        rows = res["shapes"]
        print(json.dumps({"metric": "per-call latency of small products", "value": rows[1]["schoolbook"]["us_per_call"],
                          "unit": "us per call (192x192)", "n_gpus": 1, "steps": res["reps"], "warmup": 5,
                          "ms_per_step": rows[1]["schoolbook"]["us_per_call"] / 1000, "higher_is_better": False,
                          "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                          "data": "synthetic (SplitMix64 limbs)", "config": {"workload": "small"}, "small": res}),
              flush=True)
        return 0
    if args.workload == "batch4096":
        res = bench_batch(args.steps, args.warmup, max(1, args.steps // 5))
        print(json.dumps({"metric": "batched products/s (4096 uneven products)", "value": res["products_per_s"],
                          "unit": "products/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": res["ms_per_batch"], "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "dtype": "f64", "data": "synthetic (SplitMix64 limbs)",
                          "config": {"workload": "batch4096"}, "batched": res}), flush=True)
        return 0

    wl = workload(args.workload)
    a = np.ascontiguousarray(wl["a"])
    kind = wl["kind"]
    b = np.ascontiguousarray(wl["b"]) if kind != "sqr" else a
    la, lb = len(a), len(b)
    da = ntt.DeviceBuffer.from_host(a)
    db = ntt.DeviceBuffer.from_host(b) if kind != "sqr" else da
    if kind == "conv":
        plan = ntt.ConvPlan.new(la, lb, wl["max"], wl["max"], wl["signed"])
        out_len = plan.output_len()
        ws = ntt.ConvWorkspace()
        ws.prepare(plan)
        primes, log_len, tail = plan.primes, plan.log_len, plan.row_tail
        chunk = 64

        def step():
            return ws.convolve_device(dout.ptr, out_len, da.ptr, la, db.ptr, lb)
    else:
        plan = ntt.MulPlan.new(la, lb)
        out_len = la + lb
        ws = ntt.MulWorkspace()
        ws.prepare(plan)
        primes, log_len, chunk = plan.primes, plan.log_len, plan.chunk_bits
        tail = plan.row_tail

        if kind == "sqr":
            def step():
                return ws.sqr_device(dout.ptr, out_len, da.ptr, la)
        else:
            def step():
                return ws.mul_device(dout.ptr, out_len, da.ptr, la, db.ptr, lb)
    dout = ntt.DeviceBuffer(out_len * 8)
    T = ntt.row_len(log_len, tail)

    for _ in range(args.warmup):
        assert step(), ntt.last_error()
    assert ntt.synchronize()

    # Timed region: exactly K steps, synchronized on both sides. Every kernel
    # launch inside is bracketed by HIP events on its own stream.
    ntt.Profile.reset()
    ntt.Profile.enable(True)
    assert ntt.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        if not step():
            raise SystemExit("step failed: " + ntt.last_error())
    assert ntt.synchronize()
    elapsed = time.perf_counter() - t0
    ntt.Profile.enable(False)
    prof = ntt.Profile.read()

    # The same K steps without events, for the event overhead.
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step()
    assert ntt.synchronize()
    clean = time.perf_counter() - t1

    if args.copy_bytes:
        n = (args.copy_bytes // 16) * 16
        src, dst = ntt.DeviceBuffer(n), ntt.DeviceBuffer(n)
        assert ntt.load().ntt_debug_copy(dst.ptr, src.ptr, n)
        src.free()
        dst.free()

    # Result check on the last step: the product modulo a 61-bit prime.
    got = dout.download(out_len)
    if kind != "conv":
        q = (1 << 61) - 1
        ia = int.from_bytes(a.astype("<u8").tobytes(), "little") % q
        ib = ia if kind == "sqr" else int.from_bytes(b.astype("<u8").tobytes(), "little") % q
        ig = int.from_bytes(got.astype("<u8").tobytes(), "little") % q
        assert ig == ia * ib % q, "device product failed its modular check"

    ms_step = 1000.0 * elapsed / args.steps
    ms_clean = 1000.0 * clean / args.steps
    per_kernel = {}
    dominant, dom_total = None, -1.0
    for name, (ms, count) in prof.items():
        avg_ms = ms / max(count, 1)
        # Bytes of one step's launches of this kind over their time in a step.
        nbytes = step_bytes(name, primes, log_len, tail, la, lb, out_len, kind == "sqr")
        step_ms = ms / args.steps
        per_kernel[name] = {"avg_us": round(avg_ms * 1000, 2), "launches": int(count),
                            "GBps": round(nbytes / (step_ms * 1e-3) / 1e9, 1) if nbytes else None}
        if ms > dom_total and nbytes:
            dominant, dom_total = name, ms
    traffic = measured_traffic(args.workload)
    roofline = None
    result_valu = None
    if dominant:
        ms, count = prof[dominant]
        per_step = count // args.steps
        nbytes = step_bytes(dominant, primes, log_len, tail, la, lb, out_len, kind == "sqr")
        achieved = nbytes / (ms / args.steps * 1e-3) / 1e9
        # A kind launched once per step: bytes and traffic per launch. A
        # composite row launches some kinds once per part; then the figures
        # cover the step's launches of that kind together.
        roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic.get(dominant) if per_step == 1 else None,
                    "algorithmic_bytes": nbytes, "launches_per_step": per_step}
        # The FP64 vector issue rate of the same kernel: its VALU instructions
        # per launch (committed counters) over the live launch duration,
        # against 1024 SIMDs issuing one wave64 FP64 instruction per 4 cycles
        # at 2.4 GHz (78.6 TFLOP/s of FMA).
        insts = measured_valu(args.workload).get(dominant)
        if insts and per_step == 1:
            rate = insts / (prof[dominant][0] / prof[dominant][1] * 1e-3) / 1e9
            result_valu = {"kernel": dominant, "insts_per_launch": insts, "achieved": round(rate, 1),
                           "peak": VALU_PEAK_G, "unit": "G wave-instr/s", "frac": round(rate / VALU_PEAK_G, 4)}
        else:
            result_valu = None

    unit = "Mcoeff/s" if kind == "conv" else "Mlimbs/s"
    value = la / (ms_clean * 1e-3) / 1e6
    published = PUBLISHED.get(args.workload)
    result = {
        "metric": "Mlimbs/s at 2^20-limb mul" if args.workload == "mul_2p20" else f"{unit} ({args.workload})",
        "value": round(value, 2),
        "unit": unit,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_clean, 4),
        "ms_per_step_with_events": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / published["value"], 1) if published else None,
        "baseline": published,
        "dtype": "f64",
        "data": "synthetic (SplitMix64 limbs, seeds 1 and 2), operands resident in HBM",
        "config": {"workload": args.workload, "len_a": la, "len_b": lb, "primes": primes,
                   "chunk_bits": chunk, "transform_len": T, "cyclic_len": 1 << log_len},
        "kernels": per_kernel,
        "roofline": roofline,
        "valu": result_valu,
    }
    if kind == "mul" and not args.no_host:
        result["host_buffers"] = bench_host_path(a, b, max(3, args.steps // 5))
    if not args.no_small:
        result["small"] = bench_small()
    if not args.no_batch:
        result["batched"] = bench_batch(args.batch_steps, 2, 2)
    if not args.no_serve:
        result["served"] = bench_served(args.serve_requests, args.serve_load)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(wl, args.cpu_seconds)
        # The same oracle with the host threads one GPU's share of the box
        # offers (16), as a second reported baseline.
        result["cpu_baseline_threads"] = cpu_baseline(wl, args.cpu_seconds / 2, min(16, os.cpu_count() or 1))
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
