"""The allocation-free product entry point (ntt_mul_device_noalloc) against
the workspace device path it mirrors (ntt_mul_workspace_mul_device), bit for
bit, on calls in a row that shrink, a square, a schoolbook-sized product and
a split product; a product beyond the fixed scratch is refused with dst
unchanged. Everything goes through the C ABI."""
import sys

import numpy as np
import pytest

import nttanafis_amd as ntt

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lib():
    lib = ntt.load()
    assert ntt.device_available(), "no gfx950 device visible"
    return lib


def reference(a, b, square):
    """The existing entry point on the same device operands."""
    ws = ntt.MulWorkspace()
    da = ntt.DeviceBuffer.from_host(a)
    db = da if square else ntt.DeviceBuffer.from_host(b)
    n = len(a) + len(b)
    dd = ntt.DeviceBuffer(8 * n)
    assert ws.mul_device(dd.ptr, n, da.ptr, len(a), db.ptr, len(b)), ntt.last_error()
    assert ntt.synchronize()
    return dd.download(n)


def noalloc(lib, a, b, square, stream=None):
    da = ntt.DeviceBuffer.from_host(a)
    db = da if square else ntt.DeviceBuffer.from_host(b)
    n = len(a) + len(b)
    dd = ntt.DeviceBuffer(8 * (n + 2))
    dd.upload(np.full(n + 2, 0x5A5A, dtype=np.uint64))
    ok = lib.ntt_mul_device_noalloc(dd.ptr, n, da.ptr, len(a), db.ptr, len(b), stream)
    assert ntt.synchronize()
    out = dd.download(n + 2)
    assert (out[n:] == 0x5A5A).all(), "the call wrote past its result"
    return bool(ok), out[:n]


@pytest.mark.parametrize("shapes", [[(40000, 30000), (1024, 1024)], [(65536, 65536), (777, 5000)]])
def test_noalloc_equals_workspace_path_on_shrinking_calls(lib, oracle, shapes):
    for k, (la, lb) in enumerate(shapes):
        say(f"noalloc call {k}: {la} x {lb}")
        a, b = oracle.operand(la, 11 + k), oracle.operand(lb, 12 + k)
        ok, got = noalloc(lib, a, b, False)
        assert ok, ntt.last_error()
        assert (got == reference(a, b, False)).all(), f"call {k} ({la} x {lb}) differs from the workspace path"
        if la * lb <= 40000 * 30000:
            assert (got == oracle.mul(a, b, threads=8)).all(), f"call {k} ({la} x {lb}) differs from the oracle"


def test_noalloc_square_schoolbook_split_and_all_ones(lib, oracle):
    cases = [("square", 30000, 30000, True), ("schoolbook", 100, 300, False), ("split", 300, 200000, False),
             ("ones", 20000, 9000, False)]
    for name, la, lb, square in cases:
        say(f"noalloc {name}: {la} x {lb}")
        if name == "ones":
            a = np.full(la, (1 << 64) - 1, dtype=np.uint64)
            b = np.full(lb, (1 << 64) - 1, dtype=np.uint64)
        else:
            a = oracle.operand(la, 21)
            b = a if square else oracle.operand(lb, 22)
        ok, got = noalloc(lib, a, b, square)
        assert ok, ntt.last_error()
        assert (got == reference(a, b, square)).all(), name
        assert (got == (oracle.sqr(a, threads=8) if square else oracle.mul(a, b, threads=8))).all(), name


def test_noalloc_refuses_products_beyond_its_scratch(lib, oracle):
    a, b = oracle.operand(1 << 18, 1), oracle.operand(1 << 18, 2)
    ok, got = noalloc(lib, a, b, False)
    assert not ok and "scratch" in ntt.last_error()
    assert (got == 0x5A5A).all(), "a refused call left dst unchanged"
    # The next call that fits still works.
    a, b = oracle.operand(3000, 3), oracle.operand(2000, 4)
    ok, got = noalloc(lib, a, b, False)
    assert ok and (got == oracle.mul(a, b)).all()
