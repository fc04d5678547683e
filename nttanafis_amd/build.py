"""In-tree build of the gfx950 engine library (and, for tests, the oracle).

    python -m nttanafis_amd.build

compiles nttanafis_amd/csrc/*.hip with hipcc for gfx950 into
nttanafis_amd/libnttanafis_amd.so. hipcc cross-compiles, so no GPU is needed.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnttanafis_amd.so")
SOURCES = ["kernels_fused.hip", "kernels_fused_sq.hip", "kernels_strided.hip", "kernels.hip", "engine.hip", "server.hip", "field.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # Exactness of the residue arithmetic depends on every product and sum
    # rounding separately; only explicit fma() calls may fuse.
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


def build_library(verbose: bool = False) -> str:
    """Compiles every source in parallel, then links the shared library."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    with tempfile.TemporaryDirectory(prefix="nttamd_build_") as tmpdir:
        objs, procs = [], []
        for src in srcs:
            obj = os.path.join(tmpdir, os.path.basename(src) + ".o")
            cmd = [HIPCC, *FLAGS, "-c", "-o", obj, src]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
            objs.append(obj)
        failed = False
        for src, proc in procs:
            out, _ = proc.communicate()
            if proc.returncode != 0:
                sys.stderr.write(out)
                failed = True
        if failed:
            raise RuntimeError("hipcc failed to build libnttanafis_amd.so")
        tmp = LIB + ".tmp"
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
        if verbose:
            print(" ".join(cmd), flush=True)
        proc = subprocess.run(cmd, capture_output=True, text=True)
        if proc.returncode != 0:
            sys.stderr.write(proc.stdout + proc.stderr)
            raise RuntimeError("hipcc failed to link libnttanafis_amd.so")
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = False) -> str:
    """Builds the CPU restatement used by tests and the CPU baseline."""
    oracle = os.path.join(ROOT, "oracle")
    proc = subprocess.run(["make", "-s", "-B", "-C", oracle], capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("oracle build failed")
    return os.path.join(oracle, "liboracle.so")


def build(verbose: bool = False) -> None:
    build_library(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
