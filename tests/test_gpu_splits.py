"""Every admissible pass split gives the same product.

The autotuner may pick any admissible row size for a transform length, and
NTTAMD_ROW_LOG forces one, so each split (one strided pass and a row pass of
2^5..2^12 points, or two strided passes and a row pass) must reproduce the
default split's limbs bit for bit; the default split itself is checked against
the oracle (up to 2^21-point transforms) or, for the largest shapes, against
the product's residue modulo 2^61 - 1.
"""
import os
import sys

import numpy as np
import pytest

import nttanafis_amd as ntt
from nttanafis_amd.workloads import operand

pytestmark = pytest.mark.gpu

Q61 = (1 << 61) - 1

# (len_a, len_b, prime count or 0 for the engine's choice) per transform
# length, found with ntt.mul_geometry (three primes put 4000 x 4000 limbs on
# 2^13 points, where the engine's own choice is a one-pass 2^12).
SHAPES = [(4000, 4000, 3), (8000, 8000, 0), (30000, 30000, 0), (200000, 200000, 0), (800000, 800000, 0),
          (1 << 20, 1 << 20, 0), (2500000, 2500000, 0), (4000000, 4000000, 0), (8000000, 8000000, 0)]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    ntt.load()
    assert ntt.device_available(), "no gfx950 device visible"
    yield
    os.environ.pop("NTTAMD_ROW_LOG", None)


def row_candidates(K):
    """Admissible row sizes, largest first (at most eight), as the engine's
    autotuner enumerates them."""
    out = []
    for r in range(12, 4, -1):
        rest = K - r
        if K <= 23:
            ok = 5 <= rest <= 11
        else:
            k1 = (rest + 1) // 2
            ok = k1 <= 11 and rest - k1 >= 5
        if ok:
            out.append(r)
        if len(out) == 8:
            break
    return out


def mod_q(limbs):
    return int.from_bytes(np.ascontiguousarray(limbs, dtype="<u8").tobytes(), "little") % Q61


def product(a, b):
    plan = ntt.MulPlan.new(len(a), len(b))
    dst = np.zeros(len(a) + len(b), dtype=np.uint64)
    assert plan.mul(dst, a, b), ntt.last_error()
    return dst


@pytest.mark.parametrize("la,lb,force_primes", SHAPES)
def test_every_split_matches_the_default(oracle, monkeypatch, la, lb, force_primes):
    monkeypatch.setenv("NTTAMD_COMPOSITE", "0")  # power-of-two rows; composite rows have their own tests
    monkeypatch.setenv("NTTAMD_SCHOOL", "0")
    if force_primes:
        monkeypatch.setenv("NTTAMD_PRIMES", str(force_primes))
    primes, width, K, _ = ntt.mul_geometry(la, lb)
    say(f"{la}x{lb}: {primes} primes, {width}-bit coefficients, 2^{K} points")
    a, b = operand(la, 3), operand(lb, 4)
    os.environ.pop("NTTAMD_ROW_LOG", None)
    ref = product(a, b)
    if K <= 21:
        assert (ref == oracle.mul(a, b, threads=8)).all(), "default split differs from the oracle"
    else:
        assert mod_q(ref) == mod_q(a) * mod_q(b) % Q61, "default split fails its residue check"
    cands = row_candidates(K)
    assert cands, f"no admissible split for 2^{K}"
    for r in cands:
        os.environ["NTTAMD_ROW_LOG"] = str(r)
        say(f"  rows 2^{r}")
        got = product(a, b)
        assert (got == ref).all(), f"rows 2^{r} differ from the default split"
    os.environ.pop("NTTAMD_ROW_LOG", None)
