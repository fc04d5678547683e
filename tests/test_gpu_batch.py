"""Parity of the batched form: many uneven products in one graph launch.

Each batch entry must equal what MulWorkspace::mul / sqr computes for the
pair alone (src/plan.rs:164-179): checked bit for bit against the oracle on
small and medium shapes, against the plain per-call engine path on the full
4,096-product batch (log-uniform lengths in [192, 65,536], 20% squares), and
for every entry of that batch through the product's residue modulo the
Mersenne prime 2^61-1, computed from the operands with Python integers.
Graph replays are checked one by one, with other engine work between them.
"""
import sys

import numpy as np
import pytest

from conftest import MASK, all_ones_product

import nttanafis_amd as ntt
from nttanafis_amd.workloads import batch_operands, batch_shapes, operand

pytestmark = pytest.mark.gpu

Q61 = (1 << 61) - 1


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    ntt.load()
    assert ntt.device_available(), "no gfx950 device visible"
    yield


def mod_q(limbs) -> int:
    """Value of a little-endian limb array modulo 2^61-1."""
    return int.from_bytes(np.ascontiguousarray(limbs, dtype="<u8").tobytes(), "little") % Q61


def expected(oracle, a, b, square):
    return oracle.sqr(a) if square else oracle.mul(a, b)


class DeviceBatch:
    """Operands and outputs of a batch in two device buffers."""

    def __init__(self, batch, a, b, square):
        self.n = len(a)
        self.square = square
        in_sizes = [len(a[i]) + (0 if square[i] else len(b[i])) for i in range(self.n)]
        self.out_sizes = [batch.output_len(i) for i in range(self.n)]
        self.din = ntt.DeviceBuffer(8 * max(sum(in_sizes), 1))
        self.dout = ntt.DeviceBuffer(8 * max(sum(self.out_sizes), 1))
        self.pa, self.pb, self.po = [], [], []
        ofs = 0
        for i in range(self.n):
            self.pa.append(self.din.ptr + 8 * ofs)
            ofs += len(a[i])
            if square[i]:
                self.pb.append(self.pa[-1])
            else:
                self.pb.append(self.din.ptr + 8 * ofs)
                ofs += len(b[i])
        ofs = 0
        for i in range(self.n):
            self.po.append(self.dout.ptr + 8 * ofs)
            ofs += self.out_sizes[i]
        self.upload(a, b)

    def upload(self, a, b):
        parts = []
        for i in range(self.n):
            parts.append(a[i])
            if not self.square[i]:
                parts.append(b[i])
        self.din.upload(np.concatenate(parts) if parts else np.zeros(1, dtype=np.uint64))

    def outputs(self):
        flat = self.dout.download(sum(self.out_sizes))
        res, ofs = [], 0
        for s in self.out_sizes:
            res.append(flat[ofs:ofs + s])
            ofs += s
        return res


def plan_batch(la, lb, sq):
    batch = ntt.Batch()
    assert batch.plan(la, lb, sq), ntt.last_error()
    assert batch.count() == len(la)
    return batch


def mixed_shapes():
    """Every transform length from the single-pass sizes to two passes, with
    tiny, ragged, square and very unequal shapes."""
    la = [1, 2, 3, 7, 16, 33, 64, 100, 192, 255, 256, 300, 511, 1000, 1024, 2047, 3000, 4096, 5000,
          8191, 9000, 16384, 20000, 40000, 65536, 1, 5000, 192, 7, 12345]
    lb = [1, 5, 3, 9, 16, 31, 64, 77, 192, 257, 256, 200, 513, 999, 1024, 2049, 2500, 4096, 17,
          8193, 1, 16384, 3, 39999, 65536, 40000, 12, 65536, 7, 12345]
    sq = [False, False, True, False, True, False, False, False, True, False, False, False, False, False, True,
          False, False, True, False, False, False, False, False, False, True, False, False, False, True, True]
    lb = [x if s else y for x, y, s in zip(la, lb, sq)]
    return la, lb, sq


def test_batch_mixed_vs_oracle(oracle):
    la, lb, sq = mixed_shapes()
    say(f"batch mixed {len(la)} products")
    a, b = batch_operands(la, lb, sq, seed=11)
    batch = plan_batch(la, lb, sq)
    out = [np.zeros(batch.output_len(i), dtype=np.uint64) for i in range(len(la))]
    assert batch.mul_host(a, b, out), ntt.last_error()
    for i in range(len(la)):
        want = expected(oracle, a[i], b[i], sq[i])
        assert (out[i] == want).all(), f"product {i}: {la[i]}x{lb[i]} square={sq[i]}"


def test_batch_split_products(oracle, monkeypatch):
    """Unbalanced entries split into blocks of the longer operand inside the
    batch (forced with NTTAMD_SPLIT=1): each block an item of its bucket's
    launches, the shorter operand transformed once, the reconstruction
    summing the blocks' residues; beside whole products and squares, with
    either operand the longer one."""
    monkeypatch.setenv("NTTAMD_SPLIT", "1")
    monkeypatch.setenv("NTTAMD_SCHOOL", "0")
    la = [20000, 1000, 65536, 192, 65536, 5000, 300000, 12345, 4096, 3000, 7000]
    lb = [1000, 20000, 192, 65536, 2000, 2500, 1000, 777, 4096, 3000, 64]
    sq = [False] * 8 + [True, False, False]
    for x, y, s in zip(la, lb, sq):
        if not s and max(x, y) >= 2 * min(x, y):
            assert ntt.mul_split(x, y)[0] >= 2, (x, y)
    a, b = batch_operands(la, lb, sq, seed=23)
    batch = plan_batch(la, lb, sq)
    out = [np.zeros(batch.output_len(i), dtype=np.uint64) for i in range(len(la))]
    assert batch.mul_host(a, b, out), ntt.last_error()
    for i in range(len(la)):
        want = expected(oracle, a[i], b[i], sq[i])
        assert (out[i] == want).all(), f"product {i}: {la[i]}x{lb[i]}"
    # The same entries without splits give the same limbs.
    monkeypatch.setenv("NTTAMD_SPLIT", "0")
    whole = plan_batch(la, lb, sq)
    out2 = [np.zeros(whole.output_len(i), dtype=np.uint64) for i in range(len(la))]
    assert whole.mul_host(a, b, out2), ntt.last_error()
    for i in range(len(la)):
        assert (out2[i] == out[i]).all(), f"product {i} split vs whole"


def test_batch_all_ones_and_carries(oracle):
    """Maximal carry chains in every entry, beside ordinary entries."""
    la = [192, 1000, 4096, 30000, 65536, 777, 2048]
    lb = [192, 3000, 4096, 65536, 65536, 777, 5]
    sq = [False, False, True, False, True, False, False]
    lb = [x if s else y for x, y, s in zip(la, lb, sq)]
    a = [np.full(x, MASK, dtype=np.uint64) for x in la]
    b = [a[i] if sq[i] else np.full(lb[i], MASK, dtype=np.uint64) for i in range(len(la))]
    # Entry 5: a carry ripple through zeros, entry 6: all-ones times random.
    a[5] = np.zeros(777, dtype=np.uint64)
    a[5][0] = np.uint64(MASK)
    a[5][-1] = np.uint64(1)
    b[5] = np.full(777, MASK, dtype=np.uint64)
    b[6] = operand(5, 3)
    say("batch all-ones")
    batch = plan_batch(la, lb, sq)
    out = [np.zeros(batch.output_len(i), dtype=np.uint64) for i in range(len(la))]
    assert batch.mul_host(a, b, out), ntt.last_error()
    for i in range(5):
        assert (out[i] == all_ones_product(la[i], lb[i])).all(), f"all-ones entry {i}"
    for i in (5, 6):
        assert (out[i] == oracle.mul(a[i], b[i])).all(), f"entry {i}"


def test_batch_rebind_and_replays(oracle):
    """One plan, two operand sets, several graph replays with other engine
    work between them; every replay checked."""
    la, lb, sq = batch_shapes(count=48, lo=192, hi=20000, seed=5)
    batch = plan_batch(la, lb, sq)
    a1, b1 = batch_operands(la, lb, sq, seed=100)
    a2, b2 = batch_operands(la, lb, sq, seed=900)
    want1 = [expected(oracle, a1[i], b1[i], sq[i]) for i in range(len(la))]
    want2 = [expected(oracle, a2[i], b2[i], sq[i]) for i in range(len(la))]
    dev = DeviceBatch(batch, a1, b1, sq)
    assert batch.bind(dev.pa, dev.pb, dev.po), ntt.last_error()
    plan = ntt.MulPlan.new(3000, 2000)
    x, y = operand(3000, 1), operand(2000, 2)
    for replay in range(4):
        say(f"replay {replay}")
        if replay == 2:
            dev.upload(a2, b2)
        assert batch.run(), ntt.last_error()
        assert ntt.synchronize()
        got = dev.outputs()
        want = want2 if replay >= 2 else want1
        for i in range(len(la)):
            assert (got[i] == want[i]).all(), f"replay {replay} product {i}"
        # Eager work of the plain path between replays.
        dst = np.zeros(5000, dtype=np.uint64)
        assert plan.mul(dst, x, y)
        assert (dst == oracle.mul(x, y)).all()
    # Re-binding to swapped operand sets without re-planning.
    dev2 = DeviceBatch(batch, a1, b1, sq)
    assert batch.bind(dev2.pa, dev2.pb, dev2.po)
    assert batch.run() and ntt.synchronize()
    got = dev2.outputs()
    for i in range(len(la)):
        assert (got[i] == want1[i]).all(), f"rebound product {i}"


@pytest.mark.parametrize("branches", ["1", "2", "8"])
def test_batch_graph_branches(oracle, monkeypatch, branches):
    """The graph's buckets on one chain or on 2 and 8 forked branches (the
    reconstructions after the join), with schoolbook products as their own
    segment: every replay equals the oracle."""
    monkeypatch.setenv("NTTAMD_BATCH_BRANCHES", branches)
    monkeypatch.setenv("NTTAMD_SCHOOL", "1")
    la, lb, sq = batch_shapes(count=40, lo=20, hi=30000, seed=8)
    batch = plan_batch(la, lb, sq)
    a, b = batch_operands(la, lb, sq, seed=31)
    want = [expected(oracle, a[i], b[i], sq[i]) for i in range(len(la))]
    dev = DeviceBatch(batch, a, b, sq)
    assert batch.bind(dev.pa, dev.pb, dev.po), ntt.last_error()
    for replay in range(3):
        assert batch.run(), ntt.last_error()
        assert ntt.synchronize()
        got = dev.outputs()
        for i in range(len(la)):
            assert (got[i] == want[i]).all(), f"branches {branches} replay {replay} product {i}"


def test_batch_config_4096(oracle):
    """The 4,096-product uneven batch: every entry against the plain per-call
    path bit for bit and against its residue mod 2^61-1; a sample against the
    oracle."""
    la, lb, sq = batch_shapes()
    say(f"batch 4096: {sum(la) + sum(x for x, s in zip(lb, sq) if not s)} operand limbs")
    a, b = batch_operands(la, lb, sq)
    batch = plan_batch(la, lb, sq)
    dev = DeviceBatch(batch, a, b, sq)
    assert batch.bind(dev.pa, dev.pb, dev.po), ntt.last_error()
    assert batch.run(), ntt.last_error()
    assert ntt.synchronize()
    got = dev.outputs()
    say("batch 4096: plain loop")
    assert batch.run_plain(), ntt.last_error()
    assert ntt.synchronize()
    plain = dev.outputs()
    for i in range(len(la)):
        assert (got[i] == plain[i]).all(), f"product {i} differs from the plain path"
    say("batch 4096: modular checks")
    for i in range(len(la)):
        ra = mod_q(a[i])
        rb = ra if sq[i] else mod_q(b[i])
        assert mod_q(got[i]) == ra * rb % Q61, f"product {i} fails its residue check"
    rng = np.random.default_rng(3)
    for i in rng.choice(len(la), size=24, replace=False).tolist():
        assert (got[i] == expected(oracle, a[i], b[i], sq[i])).all(), f"product {i}"


def test_batch_failures():
    batch = ntt.Batch()
    assert not batch.plan([10, 0], [10, 10], [False, False])
    assert "product 1" in ntt.last_error()
    assert not batch.run()  # nothing bound
    assert batch.plan([], [], [])
    assert batch.bind([], [], [])
    assert batch.run()
    assert batch.plan([100], [50], [False])
    out = [np.zeros(149, dtype=np.uint64)]  # one limb short
    assert not batch.mul_host([operand(100, 1)], [operand(50, 2)], out)
