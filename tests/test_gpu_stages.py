"""Stage-level parity: the transform pipeline and the reconstruction, each
checked on its own against exact integers.

- ntt_debug_residues runs packing, forward passes, pointwise product and
  inverse passes and returns the canonical residue rows; row y must be the
  linear convolution of the operands' coefficients (width-bit digits, as
  src/limb/packing.rs:113-185 cuts them) reduced mod prime y, zero past the
  convolution length.
- ntt_debug_crt runs the reconstruction and carry kernels on given residue
  rows; fed the exact rows it must produce the product limbs
  (src/limb/reconstruction.rs:77-165).
"""
import ctypes

import numpy as np
import pytest

from conftest import to_int

import nttanafis_amd as ntt
from nttanafis_amd import _ffi
from nttanafis_amd.workloads import operand

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    ntt.load()
    assert ntt.device_available(), "no gfx950 device visible"
    yield


def digits(limbs, width, count):
    x = to_int(limbs)
    mask = (1 << width) - 1
    return [(x >> (n * width)) & mask for n in range(count)]


def conv_mod(da, db, p):
    """Linear convolution of two digit lists mod p, exact: residues are cut
    into 25-bit halves so every int64 convolution sum stays below 2^63."""
    ra = np.array([d % p for d in da], dtype=object)
    rb = np.array([d % p for d in db], dtype=object)
    mask = (1 << 25) - 1
    ah = np.array([int(v) >> 25 for v in ra], dtype=np.int64)
    al = np.array([int(v) & mask for v in ra], dtype=np.int64)
    bh = np.array([int(v) >> 25 for v in rb], dtype=np.int64)
    bl = np.array([int(v) & mask for v in rb], dtype=np.int64)
    hh, hl, lh, ll = (np.convolve(x, y) for x, y in ((ah, bh), (ah, bl), (al, bh), (al, bl)))
    return [((int(hh[k]) << 50) + ((int(hl[k]) + int(lh[k])) << 25) + int(ll[k])) % p for k in range(len(hh))]


def geometry(la, lb):
    """(primes, coefficient bits, row length)."""
    geo = (ctypes.c_uint32 * 4)()
    lib = ntt.load()
    assert lib.ntt_debug_residues(None, la, None, lb, 0, None, 0, geo)
    return int(geo[0]), int(geo[1]), ntt.row_len(int(geo[2]), int(geo[3]))


def residue_rows(a, b):
    np_, width, T = geometry(len(a), len(b))
    out = np.zeros(np_ * T, dtype=np.float64)
    geo = (ctypes.c_uint32 * 4)()
    lib = ntt.load()
    ok = lib.ntt_debug_residues(a.ctypes.data_as(_ffi.P64), len(a), b.ctypes.data_as(_ffi.P64), len(b),
                                1 if a is b else 0, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                out.size, geo)
    assert ok, ntt.last_error()
    return np_, width, T, out.reshape(np_, T)


def expected_rows(oracle, a, b, np_, width, T):
    na = (len(a) * 64 + width - 1) // width
    nb = (len(b) * 64 + width - 1) // width
    da, db = digits(a, width, na), digits(b, width, nb)
    rows = np.zeros((np_, T), dtype=np.float64)
    for y in range(np_):
        c = conv_mod(da, db, oracle.prime(y))
        rows[y, :len(c)] = np.array(c, dtype=np.float64)
    return rows


# Single-pass transforms, two-pass transforms, squares, a ragged operand.
SHAPES = [(40, 30, False), (300, 300, False), (1000, 777, False), (2000, 2000, True), (4000, 3000, False)]


@pytest.mark.parametrize("la,lb,square", SHAPES)
def test_residue_rows_are_the_convolution(oracle, la, lb, square):
    a = operand(la, 5)
    b = a if square else operand(lb, 6)
    np_, width, T, rows = residue_rows(a, b)
    want = expected_rows(oracle, a, b, np_, width, T)
    for y in range(np_):
        bad = np.nonzero(rows[y] != want[y])[0]
        assert bad.size == 0, f"prime {y}: {bad.size} residues differ (first at {bad[0]})"


# Composite rows (forced): after the recombination (src/transform/finish.rs:
# 106-158) the row holds the convolution's T coefficients mod each prime.
COMPOSITE = [(2500, 2500, False, 1), (3000, 2900, False, 2), (3500, 3400, True, 3), (9000, 9000, False, 1),
             (13000, 13000, False, 3)]


@pytest.mark.parametrize("la,lb,square,tail", COMPOSITE)
def test_composite_residue_rows_are_the_convolution(oracle, monkeypatch, la, lb, square, tail):
    monkeypatch.setenv("NTTAMD_COMPOSITE", "1")
    monkeypatch.setenv("NTTAMD_PRIMES", "3")
    assert ntt.mul_geometry(la, lb)[3] == tail
    a = operand(la, 5)
    b = a if square else operand(lb, 6)
    np_, width, T, rows = residue_rows(a, b)
    want = expected_rows(oracle, a, b, np_, width, T)
    for y in range(np_):
        bad = np.nonzero(rows[y] != want[y])[0]
        assert bad.size == 0, f"prime {y}: {bad.size} residues differ (first at {bad[0]})"


@pytest.mark.parametrize("la,lb", [(40, 30), (1000, 777), (4000, 3000)])
def test_reconstruction_of_exact_rows(oracle, la, lb):
    a, b = operand(la, 7), operand(lb, 8)
    np_, width, T = geometry(la, lb)
    rows = expected_rows(oracle, a, b, np_, width, T)
    out = np.zeros(la + lb, dtype=np.uint64)
    lib = ntt.load()
    ok = lib.ntt_debug_crt(la, lb, np.ascontiguousarray(rows).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           out.ctypes.data_as(_ffi.P64))
    assert ok, ntt.last_error()
    assert (out == oracle.mul(a, b)).all()


def test_device_root_tables_are_the_bit_reversed_powers(oracle):
    """ntt_debug_tables: the device root tables (k_tables builds W[j] as a
    product of level roots, the product form of generate_tables,
    src/transform/twiddle.rs:116-156) equal the oracle's bit-reversed powers
    W[j] = w^rev(j) for every catalog prime, centered, and W^-1[j] W[j] = 1;
    the first 64 entries of each are also checked against Python's pow."""
    n = 4096
    lib = ntt.load()
    fwd = np.zeros(8 * n, dtype=np.float64)
    inv = np.zeros(8 * n, dtype=np.float64)
    pd = ctypes.POINTER(ctypes.c_double)
    assert lib.ntt_debug_tables(n, fwd.ctypes.data_as(pd), inv.ctypes.data_as(pd)), ntt.last_error()
    fwd, inv = fwd.reshape(8, n), inv.reshape(8, n)
    want_f = oracle.root_table(8, n, False)
    want_i = oracle.root_table(8, n, True)
    for y in range(8):
        p = oracle.prime(y)
        assert (np.abs(fwd[y]) <= p / 2).all() and (np.abs(inv[y]) <= p / 2).all()
        assert all(int(a) % p == int(b) % p for a, b in zip(fwd[y], want_f[y])), f"prime {y}: forward table"
        assert all(int(a) % p == int(b) % p for a, b in zip(inv[y], want_i[y])), f"prime {y}: inverse table"
        omega = pow(oracle.generator(y), (p - 1) // (2 * n), p)
        for j in range(64):
            w = pow(omega, int(format(j, "012b")[::-1], 2), p)
            assert int(fwd[y][j]) % p == w
            assert int(inv[y][j]) * w % p == 1
