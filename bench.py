#!/usr/bin/env python3
"""Benchmark of the gfx950 multi-prime NTT engine.

Default workload: one exact product of two 2^20-limb operands (SplitMix64
limbs, seeds 1 and 2, as the reference benchmarks use), operands and result
resident in HBM. A step is one MulWorkspace product on the device; the metric
is operand limbs per second (Mlimbs/s). Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--workload mul_2p20|sqr_2p24|ones_mul_2p20|ones_sqr_2p24|
                                                          conv16_2p22|conv16_5x2p20|conv32s_2p22|split_mul|
                                                          mul_1024|sqr_1024|batch4096|served|small]

Besides `value` (K steps between device synchronizations), the line carries:
per kernel kind, the launch duration bracketed by HIP events and the same
launch replayed back to back (`kernels`); the roofline of the dominant kernel
from its replayed duration; the NTT passes' algorithmic bytes over their
summed durations (`ntt_passes`); the other configurations of the workload
set (`configs`: 1,024-limb products per call, the 2^24-limb square, all-ones
operands, both 2^22-coefficient convolutions); the host-buffer path; small
products; the 4,096-product batch against the per-call loop; the served
stream; and CPU baselines (the C restatement of the reference single-threaded
and on every usable host core, and GMP's mpn_mul). --no-* flags drop legs.
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
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector peak (an FMA counted as two flops)


def splitmix(length: int, seed: int) -> np.ndarray:
    gamma = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(1, length + 1, dtype=np.uint64) * gamma)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def workload(name: str):
    """Operands of a named workload: kind (mul, sqr, conv), a, b, and the
    coefficient bounds of a convolution."""
    ones = lambda n: np.full(n, np.uint64((1 << 64) - 1), dtype=np.uint64)  # noqa: E731
    if name == "mul_2p20":
        return dict(kind="mul", a=splitmix(1 << 20, 1), b=splitmix(1 << 20, 2))
    if name == "ones_mul_2p20":
        return dict(kind="mul", a=ones(1 << 20), b=ones(1 << 20))
    if name == "split_mul":
        # 65,536 x 4,194,304 limbs: an unbalanced product the engine runs as
        # 9 blocks of the longer operand against one spectrum of the shorter.
        return dict(kind="mul", a=splitmix(1 << 22, 1), b=splitmix(1 << 16, 2))
    if name == "mul_1024":
        return dict(kind="mul", a=splitmix(1024, 1), b=splitmix(1024, 2))
    if name == "sqr_1024":
        return dict(kind="sqr", a=splitmix(1024, 1))
    if name == "sqr_2p24":
        return dict(kind="sqr", a=splitmix(1 << 24, 1))
    if name == "ones_sqr_2p24":
        return dict(kind="sqr", a=ones(1 << 24))
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


def step_bytes(name: str, primes: int, log_len: int, tail: int, la: int, lb: int, out_len: int,
               square: bool, blocks: int = 0) -> float:
    """Algorithmic HBM bytes that the launches of one kernel kind move in one
    step: every residue of a pass read and written once, packing reads the
    limbs once per transform it feeds, and the recombination of a composite
    row reads the values it combines and writes the ones it changes. A split
    product (blocks > 0) transforms each block of the longer operand and the
    shorter operand once (blocks + 1 rows forward), multiplies every block by
    the shorter operand's spectrum and inverts every block; its
    reconstruction reads every block's rows."""
    ops = 1 if square else (blocks + 1 if blocks else 2)
    rows = blocks if blocks else 1  # product rows of the row and inverse passes
    limbs = 8.0 * (la if square else la + lb)
    n = 1 << log_len
    T = n + tail * (n // 4)
    total = 0.0
    for m in row_parts(log_len, tail):
        # The passes the engine runs for this length (its tuned split).
        passes = len(ntt.transform_form(m.bit_length() - 1, "square" if square else "product")["passes"])
        if name == "fused_single" and passes == 1:
            total += limbs + 8.0 * primes * m
        elif name == "fwd_pack_strided" and passes > 1:
            total += limbs + 8.0 * ops * primes * m
        elif name == "fwd_strided" and passes == 3:
            total += 16.0 * ops * primes * m
        elif name == "inv_strided" and passes == 3:
            total += 16.0 * rows * primes * m
        elif name == "fused_rows" and passes > 1:
            total += (16.0 if square else 24.0) * rows * primes * m
        elif name == "inv_strided_canon" and passes > 1:
            total += 16.0 * rows * primes * m
    if name == "recombine" and tail:
        total = 8.0 * primes * {1: 7 * n / 4, 2: 5 * n / 2, 3: 13 * n / 4}[tail]
    if name == "crt_product":
        total = 8.0 * primes * (rows * T if blocks else min(T, out_len * 64)) + 8.0 * out_len
    if name == "crt_disjoint":
        total = 8.0 * primes * (la + lb - 1) + 8.0 * out_len
    return total


FORM_ENV = {"alt_tiles": "NTTAMD_ALT_TILES", "inv16": "NTTAMD_INV16", "per_prime_pack": "NTTAMD_PER_PRIME_PACK",
            "rows8": "NTTAMD_ROWS8", "strided8": "NTTAMD_STRIDED8", "paired": "NTTAMD_PAIRED", "waves": "NTTAMD_WAVES"}


def pin_env(split) -> dict:
    """Environment that forces a split [log_len, pass levels, forms]: the
    pass count and row size, and every kernel-form flag on or off."""
    _, logs, forms = split
    env = {}
    if len(logs) > 1:
        env["NTTAMD_PASSES"] = str(len(logs))
        env["NTTAMD_ROW_LOG"] = str(logs[-1])
    for f, e in FORM_ENV.items():
        env[e] = "1" if f in forms else "0"
    return env


# Workloads that run the transforms of another workload's profile: the same
# lengths, prime counts and kinds of operation.
PROFILE_OF = {"ones_mul_2p20": "mul_2p20", "ones_sqr_2p24": "sqr_2p24", "conv32s_2p22": "conv16_2p22"}


def pinned_workload(name: str, steps: int, warmup: int, pin: bool, **kw) -> dict:
    """gpu_workload with the split and kernel forms of the workload's
    committed profile forced for its duration (when pin and a profile
    exists), so that the line and the profile describe the same kernels;
    `split_source` says which."""
    committed = committed_profile(PROFILE_OF.get(name, name))
    env = pin_env(committed["split"]) if pin and committed["split"] else {}
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        r = gpu_workload(name, steps, warmup, **kw)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    r["split_source"] = "committed profile" if env else "autotuned"
    return r


def committed_profile(workload: str) -> dict:
    """The committed rocprofv3 summary of a workload
    (profiles/<workload>_kernels.txt, written by --profile-summary): the
    `SPLIT <log_len> <pass levels> <forms>` the profiled runs used, and per
    kernel kind `KERNEL <kind> <launches per step> <average us> <algorithmic
    bytes per launch>` (kernel trace), `TRAFFIC <kind> <bytes>` (FETCH_SIZE
    and WRITE_SIZE per dispatch, each scaled by what it reads for the plain
    copy kernel of the same run), `VALU <kind> <wave instructions>` and
    `MIX <kind> valu <n> fma64 <share> mul64 <share> ...` (the instruction
    mix of the kind's VALU instructions)."""
    path = os.path.join(ROOT, "profiles", f"{workload}_kernels.txt")
    out = {"valu": {}, "traffic": {}, "kernels": {}, "split": None, "mix": {}}
    if os.path.exists(path):
        for line in open(path):
            parts = line.split()
            if len(parts) > 3 and parts[0] == "MIX":
                # MIX <kind> valu <n> <name> <share> ...: shares of the kind's VALU instructions
                vals = parts[2:]
                out["mix"][parts[1]] = {vals[i]: float(vals[i + 1]) for i in range(0, len(vals) - 1, 2)}
            elif len(parts) == 3 and parts[0] == "VALU":
                out["valu"][parts[1]] = int(float(parts[2]))
            elif len(parts) == 3 and parts[0] == "TRAFFIC":
                out["traffic"][parts[1]] = int(float(parts[2]))
            elif len(parts) == 5 and parts[0] == "KERNEL":
                out["kernels"][parts[1]] = {"launches_per_step": int(parts[2]), "us": float(parts[3]),
                                            "bytes": int(float(parts[4]))}
            elif len(parts) == 4 and parts[0] == "SPLIT":
                forms = [] if parts[3] == "-" else parts[3].split(",")
                out["split"] = [int(parts[1]), [int(x) for x in parts[2].split(",")], forms]
    return out


# Kernel kinds of the engine's launches, from the kernel names rocprofv3
# records (template arguments: k_fwd_strided<LOGL, LINES, NT, PACK, ...>,
# k_inv_strided<LOGL, LINES, NT, CANON, ...>, k_fused<LOGL, LINES, NT, PACK, ...>).
def kernel_kind(name: str):
    import re
    m = re.search(r"nttamd::(k_[a-z_]+)(?:<([^>]*)>)?", name)
    if not m:
        return None
    k, args = m.group(1), [x.strip() for x in (m.group(2) or "").split(",")]
    if k in ("k_pack_primes", "k_pack_primes_batch"):
        return "fwd_pack_strided"
    if k == "k_fwd_strided":
        return "fwd_pack_strided" if args[3] == "true" else "fwd_strided"
    if k == "k_inv_strided":
        return "inv_strided_canon" if args[3] == "true" else "inv_strided"
    if k == "k_fused":
        return "fused_single" if args[3] == "true" else "fused_rows"
    return {"k_crt_product": "crt_product", "k_carry_apply": "carry_apply", "k_carry_scan": "carry_scan",
            "k_crt_disjoint": "crt_disjoint", "k_recombine": "recombine", "k_school": "school",
            "k_copy": "copy"}.get(k)


def stats_summary(d: str) -> str:
    """Per-kernel statistics of a profiled run without a single timed
    workload (batch, served stream, small products): rocprofv3's
    <d>/trace/run_kernel_stats.csv with the kernel kind of each name."""
    import csv
    rows = list(csv.DictReader(open(os.path.join(d, "trace", "run_kernel_stats.csv"))))
    lines = [f"{'kernel':64s} {'kind':18s} {'calls':>7s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s} {'pct':>6s}"]
    for r in rows:
        name = r["Name"].replace("void ", "").replace("nttamd::", "").replace("(PassArgs)", "").replace("(CrtArgs)", "")
        kind = kernel_kind(r["Name"]) or "-"
        lines.append(f"{name[:64]:64s} {kind:18s} {int(r['Calls']):7d} {float(r['AverageNs']) / 1000:9.2f} "
                     f"{float(r['MinNs']) / 1000:9.2f} {float(r['MaxNs']) / 1000:9.2f} {float(r['Percentage']):6.2f}")
    return "\n".join(lines) + "\n"


def profile_summary(d: str, workload_name: str, steps: int) -> str:
    """Text summary of a profiled run directory: <d>/trace.json (the bench
    line of the kernel-trace run), <d>/trace/run_kernel_trace.csv, and
    <d>/pmc_<counter>/run_counter_collection.csv of the counter runs (each
    its own process with the same split forced). Per kernel kind it averages
    the timed steps' dispatches: the last (5 + 1) * steps * launches-per-step
    of each kernel (five timed runs and the events run of gpu_workload; the
    profiled runs replay nothing), and scales FETCH_SIZE and WRITE_SIZE by
    what they read for the plain copy kernel the counter runs end with."""
    import csv
    import glob
    if not os.path.exists(os.path.join(d, "trace.json")):
        return stats_summary(d)
    bench_line = json.loads([x for x in open(os.path.join(d, "trace.json")) if x.startswith("{")][-1])
    lps = {k: r["launches_per_step"] for k, r in bench_line["kernels"].items()}
    cfg = bench_line["config"]

    def timed(rows):
        # The timed steps are the last dispatches of the run (after the
        # prepare-time autotuner's trials; the copy kernel of a counter run
        # excluded): (5 + 1) * steps * the launches of one step.
        rows = [r for r in rows if kernel_kind(r["Kernel_Name"]) in lps]
        tail = rows[-6 * steps * sum(lps.values()):]
        out = {}
        for r in tail:
            out.setdefault(kernel_kind(r["Kernel_Name"]), []).append(r)
        return out

    trace = list(csv.DictReader(open(os.path.join(d, "trace", "run_kernel_trace.csv"))))
    tk = timed(trace)
    avg = {k: sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in v) / len(v) / 1000.0
           for k, v in tk.items()}
    counters = {}
    copy_scale = {}
    for path in sorted(glob.glob(os.path.join(d, "pmc_*", "run_counter_collection.csv"))):
        rows = list(csv.DictReader(open(path)))
        per = {}
        for r in rows:
            per.setdefault((r["Kernel_Name"], int(r["Dispatch_Id"])), {}).setdefault(r["Counter_Name"], 0.0)
            per[(r["Kernel_Name"], int(r["Dispatch_Id"]))][r["Counter_Name"]] += float(r["Counter_Value"])
        flat = [dict(Kernel_Name=name, Dispatch_Id=did, **vals) for (name, did), vals in sorted(per.items(),
                key=lambda kv: kv[0][1])]
        for kind, lst in timed(flat).items():
            for cname in lst[0]:
                if cname in ("Kernel_Name", "Dispatch_Id"):
                    continue
                counters.setdefault(kind, {})[cname] = sum(x[cname] for x in lst) / len(lst)
        for x in flat:
            if kernel_kind(x["Kernel_Name"]) == "copy":
                for cname in ("FETCH_SIZE", "WRITE_SIZE"):
                    if cname in x:
                        copy_scale[cname] = x[cname]
    bytes_of = {k: r.get("bytes_per_launch") for k, r in bench_line["kernels"].items()}
    copy_bytes = 268435456
    fscale = copy_bytes / (copy_scale["FETCH_SIZE"] * 1024) if copy_scale.get("FETCH_SIZE") else None
    wscale = copy_bytes / (copy_scale["WRITE_SIZE"] * 1024) if copy_scale.get("WRITE_SIZE") else None
    forms = ",".join(bench_line.get("config", {}).get("transform_forms") or []) or "-"
    split = cfg.get("transform_split")
    lines = [f"# {workload_name}: {cfg}",
             f"# ms per step (this run, median of five timed runs): {bench_line['ms_per_step']}",
             "# kind, dispatches averaged, launches per step, trace us, algorithmic bytes per launch, GB/s, "
             "fraction of 8 TB/s; traffic (scaled counters) against the algorithmic bytes; VALU wave "
             "instructions per dispatch, issue fraction at 2.4 GHz (1024 SIMDs, 4 cycles each), clock",
             f"#   counter scaling from the plain copy of {copy_bytes} bytes: FETCH_SIZE x{fscale:.3f}, "
             f"WRITE_SIZE x{wscale:.3f}" if fscale and wscale else "#   no copy-kernel calibration in this run"]
    if split:
        lines.append(f"SPLIT {bench_line['config']['cyclic_len'].bit_length() - 1} {','.join(map(str, split))} "
                     f"{forms}")
    pass_bytes = pass_us = 0.0
    for kind in sorted(avg, key=lambda k: -avg[k] * lps[k]):
        us, b = avg[kind], bytes_of.get(kind)
        c = counters.get(kind, {})
        row = f"{kind:18s} n {len(tk[kind]):5d} lps {lps[kind]} us {us:9.2f}"
        if b:
            row += f" bytes {b:11d} GBps {b / (us * 1e-6) / 1e9:7.1f} frac {b / (us * 1e-6) / 8e12:.3f}"
            if kind in NTT_KINDS:
                pass_bytes += b * lps[kind]
                pass_us += us * lps[kind]
        traffic = None
        if fscale and wscale and "FETCH_SIZE" in c and "WRITE_SIZE" in c:
            traffic = c["FETCH_SIZE"] * 1024 * fscale + c["WRITE_SIZE"] * 1024 * wscale
            row += f" traffic_MB {traffic / 1e6:8.1f}" + (f" x{traffic / b:.2f}" if b else "")
        if "SQ_INSTS_VALU" in c:
            row += f" valu {c['SQ_INSTS_VALU']:.0f} issue {c['SQ_INSTS_VALU'] * 4 / (1024 * us * 1e-6 * 2.4e9):.3f}"
            if "SQ_INSTS_VALU_FMA_F64" in c:
                row += f" fma_share {c['SQ_INSTS_VALU_FMA_F64'] / max(c['SQ_INSTS_VALU'], 1):.2f}"
        if "GRBM_GUI_ACTIVE" in c:
            row += f" clock_GHz {c['GRBM_GUI_ACTIVE'] / 8 / (us * 1e3):.2f}"
        lines.append(row)
        if "SQ_INSTS_VALU_MUL_F64" in c and c.get("SQ_INSTS_VALU"):
            # Instruction mix: FP64 FMA, MUL, ADD and transcendental-unit
            # (rint) instructions, 32- and 64-bit integer ones, conversions,
            # and the rest (moves, selects, bit operations).
            v = c["SQ_INSTS_VALU"]
            mix = {n: c.get("SQ_INSTS_VALU_" + k, 0.0) for n, k in
                   (("fma64", "FMA_F64"), ("mul64", "MUL_F64"), ("add64", "ADD_F64"), ("trans64", "TRANS_F64"),
                    ("int32", "INT32"), ("int64", "INT64"), ("cvt", "CVT"))}
            mix["other"] = max(v - sum(mix.values()), 0.0)
            lines.append(f"MIX {kind} valu {v:.0f} " + " ".join(f"{n} {x / v:.3f}" for n, x in mix.items()))
        if "SQ_WAVE_CYCLES" in c and "SQ_WAIT_ANY" in c and c["SQ_WAVE_CYCLES"]:
            # Where the waves' time goes (shares of SQ_WAVE_CYCLES): waiting
            # on a counter (s_waitcnt), waiting to issue, issuing any
            # instruction, issuing VALU, issuing LDS.
            w = c["SQ_WAVE_CYCLES"]
            lines.append(f"STALL {kind} wait_any {c['SQ_WAIT_ANY'] / w:.3f} wait_inst_any "
                         f"{c.get('SQ_WAIT_INST_ANY', 0.0) / w:.3f} active_any {c.get('SQ_ACTIVE_INST_ANY', 0.0) / w:.3f} "
                         f"active_valu {c.get('SQ_ACTIVE_INST_VALU', 0.0) / w:.3f} active_lds "
                         f"{c.get('SQ_ACTIVE_INST_LDS', 0.0) / w:.3f} wait_inst_lds {c.get('SQ_WAIT_INST_LDS', 0.0) / w:.3f}")
        if b:
            lines.append(f"KERNEL {kind} {lps[kind]} {us:.2f} {b}")
        if traffic:
            lines.append(f"TRAFFIC {kind} {traffic:.0f}")
        if "SQ_INSTS_VALU" in c:
            lines.append(f"VALU {kind} {c['SQ_INSTS_VALU']:.0f}")
    if pass_us:
        lines.append(f"NTT_PASSES bytes {pass_bytes:.0f} us {pass_us:.1f} TBps {pass_bytes / (pass_us * 1e-6) / 1e12:.3f} "
                     f"frac {pass_bytes / (pass_us * 1e-6) / 8e12:.4f}")
    return "\n".join(lines) + "\n"


def _median_time(fn, runs: int, budget_s: float) -> tuple:
    """Median wall time of fn over `runs` runs after one warm-up (fewer runs
    when the budget is spent, at least three)."""
    fn()
    times = []
    t_start = time.perf_counter()
    while len(times) < runs:
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
        if len(times) >= 3 and time.perf_counter() - t_start > budget_s:
            break
    return float(np.median(times)), len(times)


def cpu_baselines(a, b, budget_s: float) -> dict:
    """The reference CPU path restated in C (oracle/ntt_oracle.c, built with
    -march=native on this host when gcc is here), timed on the same 2^20 x
    2^20-limb operands single-threaded and on every host core this process
    may run on (the oracle parallelises over primes and tiles, as the
    reference's rayon schedule does); and GMP's mpn_mul (libgmp.so.10, one
    thread) when the library loads. Median of at least five runs each after
    one warm-up, within `budget_s` seconds per leg."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    native = oracle.use_native()
    build = "-march=native" if native else "-march=x86-64-v3"
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    n = len(a)
    t1, r1 = _median_time(lambda: oracle.mul(a, b, threads=1), 5, budget_s)
    tn, rn = _median_time(lambda: oracle.mul(a, b, threads=cores), 7, budget_s)
    single = {"value": n / t1 / 1e6, "unit": "Mlimbs/s", "cores": 1, "kind": "port",
              "sample": f"median of {r1} products of 2^20 x 2^20 limbs, oracle/ntt_oracle.c ({build}), 1 thread",
              "ms_per_product": round(1000 * t1, 2),
              "vs_reference_published": round(n / t1 / 1e6 / PUBLISHED["mul_2p20"]["value"], 3)}
    threads = {"value": n / tn / 1e6, "unit": "Mlimbs/s", "cores": cores, "kind": "port",
               "sample": f"median of {rn} products of 2^20 x 2^20 limbs, oracle/ntt_oracle.c ({build}), "
                         f"{cores} threads", "ms_per_product": round(1000 * tn, 2),
               "host_cpu_count": os.cpu_count(), "host_cores_usable": cores}
    out = {"cpu_baseline": single, "cpu_baseline_threads": threads}
    try:
        import ctypes

        gmp = ctypes.CDLL("libgmp.so.10")
        dst = np.zeros(2 * n, dtype=np.uint64)
        P = ctypes.POINTER(ctypes.c_uint64)
        pa, pb, pd = a.ctypes.data_as(P), b.ctypes.data_as(P), dst.ctypes.data_as(P)
        tg, rg = _median_time(lambda: gmp.__gmpn_mul(pd, pa, ctypes.c_long(n), pb, ctypes.c_long(n)), 5, budget_s)
        out["cpu_gmp"] = {"value": n / tg / 1e6, "unit": "Mlimbs/s", "cores": 1, "kind": "gmp mpn_mul",
                          "sample": f"median of {rg} products of 2^20 x 2^20 limbs, libgmp.so.10, 1 thread",
                          "ms_per_product": round(1000 * tg, 2)}
        want = oracle.mul(a, b, threads=cores)
        out["cpu_gmp"]["matches_oracle"] = bool((dst == want).all())
    except OSError as e:
        out["cpu_gmp"] = {"value": None, "unavailable": str(e)}
    return out


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
        "tuned_buckets": tuned_forms("batch"),
    }


def tuned_forms(use: str) -> dict:
    """Split and kernel forms the autotuner chose, per tuned transform length
    ("2^K": {"passes": [...], "forms": [...]})."""
    out = {}
    for k in range(5, 31):
        f = ntt.transform_form(k, use)
        if f["tuned"]:
            out[f"2^{k}"] = {"passes": f["passes"], "forms": f["forms"]}
    return out


def bench_host_path(a, b, steps: int) -> dict:
    """The same product through the host-buffer API (MulPlan::mul with host
    slices): operands cross PCIe in, the product back out. Pageable numpy
    arrays (staged through the engine's pinned chunk buffers) and pinned
    (hipHostMalloc) arrays (copied directly); three untimed calls first,
    then `steps` timed calls."""
    plan = ntt.MulPlan.new(len(a), len(b))
    res = {}
    for name, alloc in (("pageable", lambda n: np.empty(n, dtype=np.uint64)), ("pinned", ntt.pinned_array)):
        ha, hb, dst = alloc(len(a)), alloc(len(b)), alloc(len(a) + len(b))
        ha[:] = a
        hb[:] = b
        for _ in range(3):
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
        for i in bad[:5]:
            # Which side differs: the request again through the plain path,
            # against the served result and the plain digest.
            served_copy = dsts[i].copy()
            plain_call(i)
            again = hash(dsts[i].tobytes())
            mismatches.append({"request": i, "kind": base[i]["kind"], "len_a": base[i]["len_a"],
                               "len_b": base[i]["len_b"], "plain_repeat_matches_digest": again == digests[i],
                               "served_wrong_limbs": int((served_copy != dsts[i]).sum())})
            print(f"served request {i} differs from the plain path: {mismatches[-1]}", file=sys.stderr, flush=True)
        return lat, last

    mismatches = []

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
        "mismatches": mismatches,
        "mix": {k: sum(1 for r in base if r["kind"] == k) for k in ("mul", "sqr", "conv")},
    }


NTT_KINDS = ("fused_single", "fwd_pack_strided", "fwd_strided", "fused_rows", "inv_strided", "inv_strided_canon")


def load_digests() -> dict:
    path = os.path.join(ROOT, "tests", "golden", "digests.json")
    return json.load(open(path)) if os.path.exists(path) else {}


def check_result(name: str, wl: dict, got: np.ndarray) -> str:
    """Checks a device result: against the committed golden digest of the
    workload (GMP products and Kronecker convolutions), the all-ones closed
    form (B^n - 1)(B^m - 1), or the product modulo 2^61 - 1. Returns the
    check that passed; raises on a mismatch."""
    import hashlib

    rec = load_digests().get(name)
    if rec is not None:
        h = hashlib.sha256(np.ascontiguousarray(got, dtype="<u8").tobytes()).hexdigest()
        if h != rec["sha256"]:
            raise SystemExit(f"{name}: result differs from the golden digest")
        return "golden digest"
    a = wl["a"]
    b = a if wl["kind"] == "sqr" else wl["b"]
    if name.startswith("ones_"):
        n, m = max(len(a), len(b)), min(len(a), len(b))
        mx = np.uint64((1 << 64) - 1)
        ok = (got[0] == 1 and not got[1:m].any() and (got[m:n] == mx).all() and got[n] == mx - np.uint64(1)
              and (got[n + 1:] == mx).all())
        if not ok:
            raise SystemExit(f"{name}: result differs from the all-ones closed form")
        return "all-ones closed form"
    if wl["kind"] != "conv":
        q = (1 << 61) - 1
        if mod_q61(got) != mod_q61(a) * mod_q61(b) % q:
            raise SystemExit(f"{name}: result fails its residue modulo 2^61 - 1")
        return "residue modulo 2^61 - 1"
    return "none"


def gpu_workload(name: str, steps: int, warmup: int, replay_reps: int = 20, runs: int = 5) -> dict:
    """One workload on the device path with operands resident in HBM.

    - timed region: exactly `steps` steps between device synchronizations
      (ms_per_step), repeated `runs` times for a median;
    - the same steps again with every launch bracketed by HIP events (per
      kernel kind: event_us, the launch time including the event packets);
    - per kernel kind, the last launch re-issued `replay_reps` times back to
      back between two events on its stream (replay_us: the kernel's own
      duration, what rocprofv3 reports);
    - the result checked (golden digest, closed form or residue)."""
    wl = workload(name)
    a = np.ascontiguousarray(wl["a"])
    kind = wl["kind"]
    b = a if kind == "sqr" else np.ascontiguousarray(wl["b"])
    la, lb = len(a), len(b)
    da = ntt.DeviceBuffer.from_host(a)
    db = da if kind == "sqr" else ntt.DeviceBuffer.from_host(b)
    if kind == "conv":
        plan = ntt.ConvPlan.new(la, lb, wl["max"], wl["max"], wl["signed"])
        out_len = plan.output_len()
        ws = ntt.ConvWorkspace()
        ws.prepare(plan)
        primes, log_len, tail, chunk = plan.primes, plan.log_len, plan.row_tail, 64
        blocks = plan.split_blocks
        dout = ntt.DeviceBuffer(out_len * 8)

        def step():
            return ws.convolve_device(dout.ptr, out_len, da.ptr, la, db.ptr, lb)
    else:
        plan = ntt.MulPlan.new(la, lb)
        out_len = la + lb
        ws = ntt.MulWorkspace()
        ws.prepare(plan)
        primes, log_len, tail, chunk = plan.primes, plan.log_len, plan.row_tail, plan.chunk_bits
        blocks = 0 if kind == "sqr" else (ntt.mul_split(la, lb) or (0, 0))[0]
        dout = ntt.DeviceBuffer(out_len * 8)
        if kind == "sqr":
            def step():
                return ws.sqr_device(dout.ptr, out_len, da.ptr, la)
        else:
            def step():
                return ws.mul_device(dout.ptr, out_len, da.ptr, la, db.ptr, lb)
    for _ in range(warmup):
        if not step():
            raise SystemExit(f"{name}: step failed: {ntt.last_error()}")
    assert ntt.synchronize()
    region_ms = []
    for _ in range(runs):
        assert ntt.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            if not step():
                raise SystemExit(f"{name}: step failed: {ntt.last_error()}")
        assert ntt.synchronize()
        region_ms.append(1000.0 * (time.perf_counter() - t0) / steps)
    ntt.Profile.reset()
    ntt.Profile.enable(True)
    assert ntt.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    assert ntt.synchronize()
    events_ms = 1000.0 * (time.perf_counter() - t0) / steps
    ntt.Profile.enable(False)
    prof = ntt.Profile.read()
    checked = check_result(name, wl, dout.download(out_len))
    kernels = {}
    for k, (ms, count) in prof.items():
        per_step = max(1, int(round(count / steps)))
        rep = ntt.Profile.replay(k, replay_reps)
        nbytes = step_bytes(k, primes, log_len, tail, la, lb, out_len, kind == "sqr", blocks) / per_step
        ev_us = 1000.0 * ms / max(count, 1)
        row = {"event_us": round(ev_us, 2), "launches_per_step": per_step}
        if rep is not None:
            row["replay_us"] = round(1000.0 * rep, 2)
        if nbytes:
            row["bytes_per_launch"] = int(nbytes)
            if rep:
                row["GBps"] = round(nbytes / (rep * 1e-3) / 1e9, 1)
        kernels[k] = row
    # The NTT passes' algorithmic bytes over their summed launch durations in
    # the timed steps (events), and over their replayed durations.
    passes = [k for k in kernels if k in NTT_KINDS and kernels[k].get("bytes_per_launch")]
    pass_bytes = sum(kernels[k]["bytes_per_launch"] * kernels[k]["launches_per_step"] for k in passes)
    pass_us = sum(kernels[k]["event_us"] * kernels[k]["launches_per_step"] for k in passes)
    pass_us_rep = sum(kernels[k].get("replay_us", 0.0) * kernels[k]["launches_per_step"] for k in passes)
    form = ntt.transform_form(log_len, "square" if kind == "sqr" else "product") if tail == 0 else None
    split = form["passes"] if form else None
    units = la if kind != "sqr" else la
    return {
        "workload": name,
        "kind": kind,
        "len_a": la,
        "len_b": lb,
        "primes": primes,
        "chunk_bits": chunk,
        "log_len": log_len,
        "row_tail": tail,
        "transform_split": split,
        "transform_forms": form["forms"] if form else None,
        "steps": steps,
        "ms_per_step": round(float(np.median(region_ms)), 4),
        "ms_per_step_runs": [round(x, 4) for x in region_ms],
        "ms_per_step_median": round(float(np.median(region_ms)), 4),
        "ms_per_step_with_events": round(events_ms, 4),
        "rate": round(units / (float(np.median(region_ms)) * 1e-3) / 1e6, 2),
        "rate_unit": "Mcoeff/s" if kind == "conv" else "Mlimbs/s",
        "kernels": kernels,
        "ntt_passes": {"kinds": passes, "bytes_per_step": int(pass_bytes), "us_per_step": round(pass_us, 1),
                       "TBps": round(pass_bytes / (pass_us * 1e-6) / 1e12, 3) if pass_us else None,
                       "frac_of_8TBps": round(pass_bytes / (pass_us * 1e-6) / 8e12, 4) if pass_us else None,
                       "us_per_step_replayed": round(pass_us_rep, 1),
                       "frac_of_8TBps_replayed": round(pass_bytes / (pass_us_rep * 1e-6) / 8e12, 4) if pass_us_rep else None},
        "check": checked,
    }


def bench_config1(reps: int = 200) -> dict:
    """Config 1: MulPlan::mul of two 1,024-limb operands and MulPlan::sqr of
    the first, per call. Device path (operands in HBM): `reps` calls queued
    back to back, and `reps` calls each waited for; host path (MulPlan with
    host arrays, PCIe included): calls each waited for. Results are checked
    against the committed GMP fixture."""
    gold = np.load(os.path.join(ROOT, "tests", "golden", "config1.npz"), allow_pickle=False)
    a, b = splitmix(1024, 1), splitmix(1024, 2)
    out = {}
    for name, square in (("mul", False), ("sqr", True)):
        bb = a if square else b
        want = gold["sqr" if square else "mul"]
        da, db = ntt.DeviceBuffer.from_host(a), ntt.DeviceBuffer.from_host(bb)
        dout = ntt.DeviceBuffer(8 * 2048)
        ws = ntt.MulWorkspace()
        ws.prepare(ntt.MulPlan.new(1024, 1024))
        if square:
            call = lambda: ws.sqr_device(dout.ptr, 2048, da.ptr, 1024)  # noqa: E731
        else:
            call = lambda: ws.mul_device(dout.ptr, 2048, da.ptr, 1024, db.ptr, 1024)  # noqa: E731
        for _ in range(10):
            assert call(), ntt.last_error()
        assert ntt.synchronize()
        if not (dout.download(2048) == want).all():
            raise SystemExit(f"config 1 {name}: device result differs from the GMP fixture")
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
        assert ntt.synchronize()
        queued = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
            ntt.synchronize()
        waited = (time.perf_counter() - t0) / reps
        plan = ntt.MulPlan.new(1024, 1024)
        dst = np.zeros(2048, dtype=np.uint64)
        host = (lambda: plan.sqr(dst, a)) if square else (lambda: plan.mul(dst, a, b))
        for _ in range(10):
            assert host(), ntt.last_error()
        if not (dst == want).all():
            raise SystemExit(f"config 1 {name}: host-path result differs from the GMP fixture")
        t0 = time.perf_counter()
        for _ in range(reps):
            host()
        host_s = (time.perf_counter() - t0) / reps
        out[name] = {"device_us_per_call": round(queued * 1e6, 2), "device_us_waited": round(waited * 1e6, 2),
                     "host_us_per_call": round(host_s * 1e6, 2),
                     "schoolbook": bool(ntt.mul_schoolbook(1024, 1024)),
                     "geometry": list(ntt.mul_geometry(1024, 1024)),
                     "transform_form": ntt.transform_form(ntt.mul_geometry(1024, 1024)[2],
                                                          "square" if square else "product")}
    out["reference_us"] = {"value": 23.1, "source": "reference README.md:25, mul 1,024 x 1,024, one Zen 5 core"}
    out["reps"] = reps
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="mul_2p20")
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="time budget of each CPU-baseline leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-small", action="store_true", help="skip the small-product latency table")
    ap.add_argument("--no-host", action="store_true", help="skip the host-buffer API timing")
    ap.add_argument("--no-configs", action="store_true", help="skip the other configurations' legs")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=1200)
    ap.add_argument("--serve-load", type=float, default=0.7)
    ap.add_argument("--replay-reps", type=int, default=20,
                    help="replays of each kernel kind's last launch after the timed steps (0 in profiled runs)")
    ap.add_argument("--profile-summary", default=None, metavar="DIR",
                    help="print the text summary of a profiled run directory (kernel trace and counter runs "
                         "of this workload with these --steps) instead of running anything")
    ap.add_argument("--no-pin", action="store_true",
                    help="autotune the timed workload's split instead of running the split and kernel forms of "
                         "its committed profile (profiles/<workload>_kernels.txt)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="after the timed region, run the plain copy kernel over this many bytes "
                         "(the calibration reference for hardware byte counters)")
    args = ap.parse_args()
    if args.profile_summary:
        sys.stdout.write(profile_summary(args.profile_summary, args.workload, args.steps))
        return 0
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

    # The timed workload (default: the headline, 2^20 x 2^20 limbs).
    # The timed workload runs the split and kernel forms of its committed
    # profile (when there is one), so that the roofline below and the
    # profile describe the same kernels: the autotuner's choice between
    # near-equal candidates differs from box to box.
    committed = committed_profile(args.workload)
    head = pinned_workload(args.workload, args.steps, args.warmup, not args.no_pin, replay_reps=args.replay_reps)
    pinned = head["split_source"] == "committed profile"
    if args.copy_bytes:
        n = (args.copy_bytes // 16) * 16
        src, dst = ntt.DeviceBuffer(n), ntt.DeviceBuffer(n)
        assert ntt.load().ntt_debug_copy(dst.ptr, src.ptr, n)
        src.free()
        dst.free()
    kernels = head["kernels"]
    # Roofline of the dominant kernel kind: its algorithmic bytes per launch
    # over its launch duration inside the timed steps (HIP events bracketing
    # each launch on its own stream); beside it, the same launch replayed
    # back to back between two events.
    dominant, dom_us = None, -1.0
    for k, row in kernels.items():
        us = row["event_us"] * row["launches_per_step"]
        if row.get("bytes_per_launch") and us > dom_us:
            dominant, dom_us = k, us
    roofline, valu, fp64 = None, None, None
    split_matches = committed["split"] == [head["log_len"], head["transform_split"], head["transform_forms"] or []]
    if dominant:
        row = kernels[dominant]
        achieved = row["bytes_per_launch"] / (row["event_us"] * 1e-6) / 1e9
        replay = row.get("replay_us")
        prof = committed["kernels"].get(dominant) if split_matches else None
        roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": committed["traffic"].get(dominant) if split_matches else None,
                    "algorithmic_bytes": row["bytes_per_launch"], "launch_us": row["event_us"],
                    "duration_source": "HIP events around each launch on its stream, timed steps",
                    "launch_us_replayed": replay,
                    "achieved_replayed": round(row["bytes_per_launch"] / (replay * 1e-6) / 1e9, 1) if replay else None,
                    "profile": {"file": f"profiles/{args.workload}_kernels.txt", "split": committed["split"],
                                "split_matches": split_matches,
                                "launch_us": prof["us"] if prof else None,
                                "frac": round(prof["bytes"] / (prof["us"] * 1e-6) / 8e12, 4) if prof else None,
                                "event_vs_profile": round(row["event_us"] / prof["us"], 3) if prof else None}}
        insts = committed["valu"].get(dominant) if split_matches else None
        if insts:
            rate = insts / (row["event_us"] * 1e-6) / 1e9
            valu = {"kernel": dominant, "insts_per_launch": insts, "achieved": round(rate, 1), "peak": VALU_PEAK_G,
                    "unit": "G wave-instr/s", "frac": round(rate / VALU_PEAK_G, 4)}
            mix = committed["mix"].get(dominant)
            if mix and "fma64" in mix:
                # FP64 work of the dominant kernel: its instruction mix (the
                # committed profile's counters) over its live launch time; an
                # FMA is two flops, a multiply or an add one, 64 lanes each.
                t = row["event_us"] * 1e-6
                fma = insts * mix["fma64"]
                flops = 64 * insts * (2 * mix["fma64"] + mix.get("mul64", 0.0) + mix.get("add64", 0.0))
                fp64 = {"kernel": dominant, "fma_wave_insts_per_launch": round(fma),
                        "fma_tflops": round(128 * fma / t / 1e12, 2),
                        "tflops": round(flops / t / 1e12, 2), "peak_tflops": FP64_PEAK_TFLOPS,
                        "frac": round(flops / t / 1e12 / FP64_PEAK_TFLOPS, 4),
                        "source": "instruction mix of the committed profile over the live launch time"}

    kind = head["kind"]
    unit = head["rate_unit"]
    # The workload's K steps are timed five times (each between device
    # synchronizations); the line reports the median run.
    ms_step = head["ms_per_step_median"]
    value = head["len_a"] / (ms_step * 1e-3) / 1e6
    published = PUBLISHED.get(args.workload)
    result = {
        "metric": "Mlimbs/s at 2^20-limb mul" if args.workload == "mul_2p20" else f"{unit} ({args.workload})",
        "value": round(value, 2),
        "unit": unit,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "ms_per_step_runs": head["ms_per_step_runs"],
        "ms_per_step_with_events": head["ms_per_step_with_events"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / published["value"], 1) if published else None,
        "baseline": published,
        "dtype": "f64",
        "data": "synthetic (SplitMix64 limbs, seeds 1 and 2), operands resident in HBM",
        "config": {"workload": args.workload, "len_a": head["len_a"], "len_b": head["len_b"],
                   "primes": head["primes"], "chunk_bits": head["chunk_bits"],
                   "transform_len": ntt.row_len(head["log_len"], head["row_tail"]),
                   "cyclic_len": 1 << head["log_len"], "transform_split": head["transform_split"],
                   "transform_forms": head["transform_forms"],
                   "split_source": head["split_source"]},
        "kernels": kernels,
        "ntt_passes": head["ntt_passes"],
        "check": head["check"],
        "roofline": roofline,
        "valu": valu,
        "fp64": fp64,
    }
    if kind == "mul" and not args.no_host:
        wl = workload(args.workload)
        result["host_buffers"] = bench_host_path(wl["a"], wl["b"], max(10, args.steps // 2))
    if not args.no_configs and args.workload == "mul_2p20":
        configs = {"config1_1024": bench_config1()}
        for name, steps in (("sqr_2p24", 10), ("ones_mul_2p20", 20), ("ones_sqr_2p24", 10), ("conv16_2p22", 10),
                            ("conv32s_2p22", 10)):
            r = pinned_workload(name, steps, 2, not args.no_pin)
            pub = PUBLISHED.get(name)
            if pub:
                r["vs_reference_published"] = round(r["rate"] / pub["value"], 1)
            configs[name] = r
        configs["ones_vs_random"] = {
            "mul_2p20": round(configs["ones_mul_2p20"]["ms_per_step_median"] / head["ms_per_step_median"], 3),
            "sqr_2p24": round(configs["ones_sqr_2p24"]["ms_per_step_median"] / configs["sqr_2p24"]["ms_per_step_median"], 3),
        }
        result["configs"] = configs
    if not args.no_small:
        result["small"] = bench_small()
    if not args.no_batch:
        result["batched"] = bench_batch(args.batch_steps, 2, 2)
    if not args.no_serve:
        result["served"] = bench_served(args.serve_requests, args.serve_load)
    if not args.no_cpu and args.workload == "mul_2p20":
        wl = workload(args.workload)
        result.update(cpu_baselines(np.ascontiguousarray(wl["a"]), np.ascontiguousarray(wl["b"]), args.cpu_seconds))
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
