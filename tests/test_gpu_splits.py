"""Every admissible pass split gives the same product.

The autotuner may pick any admissible split for a transform length, and
NTTAMD_PASSES / NTTAMD_ROW_LOG force one, so each split (one strided pass and
a row pass of 2^5..2^12 points, or two strided passes and a row pass) must
reproduce the default split's limbs bit for bit; the default split itself is checked against
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
    os.environ.pop("NTTAMD_PASSES", None)


def strided_lines(logl):
    """Lines per strided tile (device.h strided_lines)."""
    return 64 if logl <= 7 else 32 if logl == 8 else 8 if logl <= 10 else 4


def split_candidates(K):
    """Admissible (passes, row size) splits as the engine's autotuner
    enumerates them: two passes up to 2^23 points, three from 2^18 on, row
    pass 2^5..2^12, strided passes of 5..11 levels whose line stride holds
    their tile's lines."""
    out = []
    for passes in (2, 3):
        if (passes == 2 and K > 23) or (passes == 3 and K < 18):
            continue
        for r in range(12, 4, -1):
            if passes == 2:
                logs = [K - r]
            else:
                k1 = (K - r + 1) // 2
                logs = [k1, K - r - k1]
            level, ok = 0, True
            for lg in logs:
                stride = K - level - lg
                ok = ok and 5 <= lg <= 11 and stride >= 0 and (1 << stride) >= strided_lines(lg)
                level += lg
            if ok:
                out.append((passes, r))
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
    cands = split_candidates(K)
    assert cands, f"no admissible split for 2^{K}"
    for passes, r in cands:
        os.environ["NTTAMD_PASSES"] = str(passes)
        os.environ["NTTAMD_ROW_LOG"] = str(r)
        assert ntt.transform_split(K)[-1] == r and len(ntt.transform_split(K)) == passes
        say(f"  {passes} passes, rows 2^{r}")
        got = product(a, b)
        assert (got == ref).all(), f"{passes} passes with rows 2^{r} differ from the default split"
    os.environ.pop("NTTAMD_ROW_LOG", None)
    os.environ.pop("NTTAMD_PASSES", None)
