"""Python mirror of the reference crate's public multiplication API.

Names, argument meaning and failure behaviour follow the reference:
``MulPlan.new`` returns ``None`` where ``MulPlan::new`` does (src/plan.rs:53),
calls return ``False`` and leave ``dst`` unchanged on failure, success writes
only the product span, and passing the same array as both operands squares
(src/plan.rs:94-114, src/convolve.rs:148-162). Limbs are numpy ``uint64``
arrays, least significant first. Every call runs on the gfx950 engine
through the C ABI of include/nttanafis.h.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _ffi


def _lib():
    return _ffi.load()


def _ptr(arr: Optional[np.ndarray]):
    if arr is None:
        return ctypes.cast(None, _ffi.P64)
    return arr.ctypes.data_as(_ffi.P64)


def _limbs(arr, writable: bool = False) -> np.ndarray:
    if not isinstance(arr, np.ndarray) or arr.dtype != np.uint64 or arr.ndim != 1:
        raise TypeError("limb arrays are one-dimensional numpy uint64 arrays")
    if not arr.flags["C_CONTIGUOUS"]:
        raise TypeError("limb arrays must be contiguous")
    if writable and not arr.flags["WRITEABLE"]:
        raise TypeError("dst must be writable")
    return arr


def last_error() -> str:
    msg = _lib().ntt_last_error()
    return msg.decode() if msg else ""


class MulPlan:
    """An admitted product of two operand lengths (src/plan.rs:33-120)."""

    __slots__ = ("_c",)

    def __init__(self, c: _ffi.MulPlanC):
        self._c = c

    @classmethod
    def new(cls, len_a: int, len_b: int) -> Optional["MulPlan"]:
        c = _ffi.MulPlanC()
        if not _lib().ntt_mul_plan_new(len_a, len_b, ctypes.byref(c)):
            return None
        return cls(c)

    @property
    def len_a(self) -> int:
        return self._c.len_a

    @property
    def len_b(self) -> int:
        return self._c.len_b

    @property
    def primes(self) -> int:
        return self._c.primes

    @property
    def chunk_bits(self) -> int:
        return self._c.chunk_bits

    @property
    def log_len(self) -> int:
        return self._c.log_len

    @property
    def row_tail(self) -> int:
        """Transform row of N + row_tail*N/4 points, N = 2^log_len."""
        return self._c.row_tail

    @property
    def schoolbook(self) -> bool:
        """The product runs as a schoolbook (mul_schoolbook)."""
        return bool(self._c.schoolbook)

    def scratch_len(self) -> int:
        return _lib().ntt_mul_plan_scratch_len(ctypes.byref(self._c))

    def sqr_scratch_len(self) -> int:
        return _lib().ntt_mul_plan_sqr_scratch_len(ctypes.byref(self._c))

    def mul(self, dst: np.ndarray, a: np.ndarray, b: np.ndarray, scratch: Optional[np.ndarray] = None) -> bool:
        _limbs(dst, True), _limbs(a), _limbs(b)
        if scratch is not None:
            _limbs(scratch, True)
        # The same array twice squares, as core::ptr::eq on slices does.
        bb = a if (b is a or (b.ctypes.data == a.ctypes.data and len(a) == len(b))) else b
        return bool(
            _lib().ntt_mul_plan_mul(
                ctypes.byref(self._c),
                _ptr(dst), len(dst),
                _ptr(a), len(a),
                _ptr(bb), len(bb),
                _ptr(scratch), 0 if scratch is None else len(scratch),
            )
        )

    def sqr(self, dst: np.ndarray, a: np.ndarray, scratch: Optional[np.ndarray] = None) -> bool:
        _limbs(dst, True), _limbs(a)
        return bool(
            _lib().ntt_mul_plan_sqr(
                ctypes.byref(self._c), _ptr(dst), len(dst), _ptr(a), len(a),
                _ptr(scratch), 0 if scratch is None else len(scratch),
            )
        )

    def __repr__(self) -> str:
        return (f"MulPlan(len_a={self.len_a}, len_b={self.len_b}, primes={self.primes}, "
                f"chunk_bits={self.chunk_bits}, log_len={self.log_len})")


class MulWorkspace:
    """Reusable device tables and arenas (src/plan.rs:134-212)."""

    def __init__(self):
        self._h = _lib().ntt_mul_workspace_new()
        if not self._h:
            raise MemoryError("ntt_mul_workspace_new")

    def close(self):
        if getattr(self, "_h", None):
            _lib().ntt_mul_workspace_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare(self, plan: MulPlan) -> None:
        _lib().ntt_mul_workspace_prepare(self._h, ctypes.byref(plan._c))

    def mul(self, dst: np.ndarray, a: np.ndarray, b: np.ndarray) -> bool:
        _limbs(dst, True), _limbs(a), _limbs(b)
        bb = a if (b is a or (b.ctypes.data == a.ctypes.data and len(a) == len(b))) else b
        return bool(_lib().ntt_mul_workspace_mul(self._h, _ptr(dst), len(dst), _ptr(a), len(a), _ptr(bb), len(bb)))

    def sqr(self, dst: np.ndarray, a: np.ndarray) -> bool:
        _limbs(dst, True), _limbs(a)
        return bool(_lib().ntt_mul_workspace_sqr(self._h, _ptr(dst), len(dst), _ptr(a), len(a)))

    # Device-resident operands (addresses from DeviceBuffer.ptr).
    def mul_device(self, d_dst: int, dst_len: int, d_a: int, a_len: int, d_b: int, b_len: int,
                   stream: Optional[int] = None) -> bool:
        return bool(_lib().ntt_mul_workspace_mul_device(self._h, d_dst, dst_len, d_a, a_len, d_b, b_len, stream))

    def sqr_device(self, d_dst: int, dst_len: int, d_a: int, a_len: int, stream: Optional[int] = None) -> bool:
        return bool(_lib().ntt_mul_workspace_sqr_device(self._h, d_dst, dst_len, d_a, a_len, stream))


class ConvPlan:
    """Exact dense convolution plan (src/convolve.rs:51-185)."""

    __slots__ = ("_c",)

    def __init__(self, c: _ffi.ConvPlanC):
        self._c = c

    @classmethod
    def new(cls, len_a: int, len_b: int, max_a: int, max_b: int, signed: bool) -> Optional["ConvPlan"]:
        c = _ffi.ConvPlanC()
        if not _lib().ntt_conv_plan_new(len_a, len_b, max_a, max_b, bool(signed), ctypes.byref(c)):
            return None
        return cls(c)

    def stride(self) -> int:
        return _lib().ntt_conv_plan_stride(ctypes.byref(self._c))

    def output_len(self) -> int:
        return _lib().ntt_conv_plan_output_len(ctypes.byref(self._c))

    def scratch_len(self) -> int:
        return _lib().ntt_conv_plan_scratch_len(ctypes.byref(self._c))

    @property
    def primes(self) -> int:
        return self._c.primes

    @property
    def log_len(self) -> int:
        return self._c.log_len

    @property
    def row_tail(self) -> int:
        """Transform row of N + row_tail*N/4 points, N = 2^log_len."""
        return self._c.row_tail

    @property
    def split_blocks(self) -> int:
        """Blocks of the longer operand in an unbalanced split (0: none)."""
        return self._c.split_blocks

    @property
    def schoolbook(self) -> bool:
        return bool(self._c.schoolbook)

    def convolve(self, dst: np.ndarray, a: np.ndarray, b: np.ndarray, scratch: Optional[np.ndarray] = None) -> bool:
        _limbs(dst, True), _limbs(a), _limbs(b)
        bb = a if (b is a or (b.ctypes.data == a.ctypes.data and len(a) == len(b))) else b
        return bool(
            _lib().ntt_conv_plan_convolve(
                ctypes.byref(self._c), _ptr(dst), len(dst), _ptr(a), len(a), _ptr(bb), len(bb),
                _ptr(scratch), 0 if scratch is None else len(scratch),
            )
        )


class ConvWorkspace:
    """Reusable storage for repeated convolutions (src/convolve.rs:213-281)."""

    def __init__(self):
        self._h = _lib().ntt_conv_workspace_new()
        if not self._h:
            raise MemoryError("ntt_conv_workspace_new")

    def close(self):
        if getattr(self, "_h", None):
            _lib().ntt_conv_workspace_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare(self, plan: ConvPlan) -> None:
        _lib().ntt_conv_workspace_prepare(self._h, ctypes.byref(plan._c))

    def convolve(self, dst: np.ndarray, a: np.ndarray, b: np.ndarray) -> bool:
        _limbs(dst, True), _limbs(a), _limbs(b)
        bb = a if (b is a or (b.ctypes.data == a.ctypes.data and len(a) == len(b))) else b
        return bool(_lib().ntt_conv_workspace_convolve(self._h, _ptr(dst), len(dst), _ptr(a), len(a), _ptr(bb), len(bb)))

    def convolve_device(self, d_dst: int, dst_len: int, d_a: int, a_len: int, d_b: int, b_len: int,
                        stream: Optional[int] = None) -> bool:
        return bool(_lib().ntt_conv_workspace_convolve_device(self._h, d_dst, dst_len, d_a, a_len, d_b, b_len, stream))


def _ptr_array(ptrs):
    arr = (ctypes.c_void_p * max(len(ptrs), 1))()
    for i, v in enumerate(ptrs):
        arr[i] = v
    return arr


class Batch:
    """Many independent products of uneven size computed by one graph launch.

    Each entry is what ``MulWorkspace.mul`` (src/plan.rs:164-170) computes for
    one pair, or ``sqr`` for a square. ``plan`` fixes the shapes; ``bind``
    sets device operands (re-bindable); ``run`` queues one graph replay.
    ``mul_host`` is the synchronous host-array form.
    """

    def __init__(self):
        self._h = _lib().ntt_batch_new()
        if not self._h:
            raise MemoryError("ntt_batch_new failed")
        self._count = 0
        self._square = []

    def close(self):
        if self._h:
            _lib().ntt_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def plan(self, len_a, len_b, square=None) -> bool:
        n = len(len_a)
        la = (_ffi.size_t * max(n, 1))(*[int(x) for x in len_a])
        lb = (_ffi.size_t * max(n, 1))(*[int(x) for x in len_b])
        sq = None
        if square is not None:
            sq = (ctypes.c_uint8 * max(n, 1))(*[1 if s else 0 for s in square])
        ok = bool(_lib().ntt_batch_plan(self._h, n, la, lb, sq))
        if ok:
            self._count = n
            self._square = [bool(s) for s in square] if square is not None else [False] * n
        return ok

    def count(self) -> int:
        return int(_lib().ntt_batch_count(self._h))

    def output_len(self, i: int) -> int:
        return int(_lib().ntt_batch_output_len(self._h, i))

    def launches(self) -> int:
        return int(_lib().ntt_batch_launches(self._h))

    def bind(self, d_a, d_b, d_out) -> bool:
        return bool(_lib().ntt_batch_bind(self._h, _ptr_array(d_a), _ptr_array(d_b), _ptr_array(d_out)))

    def run(self, stream: Optional[int] = None) -> bool:
        return bool(_lib().ntt_batch_run(self._h, stream))

    def run_plain(self, stream: Optional[int] = None) -> bool:
        return bool(_lib().ntt_batch_run_plain(self._h, stream))

    def mul_host(self, a, b, out) -> bool:
        """a, b, out: lists of uint64 arrays (b[i] ignored for squares);
        out[i] must hold output_len(i) limbs."""
        for i in range(self._count):
            _limbs(a[i])
            _limbs(out[i], writable=True)
            if len(out[i]) < self.output_len(i):
                return False
            if not self._square[i]:
                _limbs(b[i])
        pa = _ptr_array([x.ctypes.data for x in a])
        pb = _ptr_array([(x.ctypes.data if x is not None else None) for x in b])
        po = _ptr_array([x.ctypes.data for x in out])
        return bool(_lib().ntt_batch_mul_host(self._h, pa, pb, po))


class _Pinned:
    """Owner of one pinned allocation; freed when the last view is gone."""

    def __init__(self, nbytes: int):
        self.ptr = _lib().ntt_host_alloc(nbytes)
        if not self.ptr:
            raise MemoryError(f"ntt_host_alloc({nbytes})")

    def __del__(self):
        if self.ptr:
            _lib().ntt_host_free(self.ptr)
            self.ptr = None


def pinned_array(count: int) -> np.ndarray:
    """A numpy uint64 array of ``count`` entries in pinned host memory
    (hipHostMalloc): copies between it and the device run at full link speed.
    The memory lives as long as the array or any view of it."""
    count = int(count)
    owner = _Pinned(8 * max(count, 1))
    raw = (ctypes.c_uint64 * max(count, 1)).from_address(owner.ptr)
    raw._owner = owner  # the ctypes array is the numpy base and keeps the owner
    return np.frombuffer(raw, dtype=np.uint64)[:count]


class Server:
    """Continuous-batching request server (include/nttanafis.h, ntt_server_*).

    Each request computes what ``MulWorkspace.mul``/``sqr`` or
    ``ConvPlan.convolve`` computes for it alone (src/plan.rs:164-179,
    src/convolve.rs:148-162). Arrays passed to ``submit_*`` are kept alive by
    the server object until the request is waited for or released.
    """

    def __init__(self, budget_bytes: int = 0, slots: int = 0):
        self._h = _lib().ntt_server_new(int(budget_bytes), int(slots))
        if not self._h:
            raise RuntimeError("ntt_server_new failed: " + last_error())
        self._keep = {}

    def close(self):
        if self._h:
            _lib().ntt_server_free(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _submit(self, kind, a, b, dst, max_a=0, max_b=0, flags=0, keep=None) -> int:
        """Low-level submit: a, b, dst are (address, length) pairs."""
        r = _ffi.RequestC()
        r.kind, r.flags = kind, flags
        r.a, r.len_a = a
        r.b, r.len_b = b if b is not None else (None, 0)
        r.max_a, r.max_b = int(max_a), int(max_b)
        r.dst, r.dst_len = dst
        t = int(_lib().ntt_server_submit(self._h, ctypes.byref(r)))
        if t >= 0 and keep is not None:
            self._keep[t] = keep
        return t

    @staticmethod
    def _host(arr):
        _limbs(arr)
        return arr.ctypes.data, len(arr)

    def submit_mul(self, a: np.ndarray, b: np.ndarray, dst: np.ndarray) -> int:
        _limbs(dst, writable=True)
        return self._submit(_ffi.REQ_MUL, self._host(a), self._host(b), self._host(dst), keep=(a, b, dst))

    def submit_sqr(self, a: np.ndarray, dst: np.ndarray) -> int:
        _limbs(dst, writable=True)
        return self._submit(_ffi.REQ_SQR, self._host(a), None, self._host(dst), keep=(a, dst))

    def submit_conv(self, a: np.ndarray, b: np.ndarray, max_a: int, max_b: int, signed: bool,
                    dst: np.ndarray) -> int:
        _limbs(dst, writable=True)
        flags = _ffi.REQ_SIGNED if signed else 0
        return self._submit(_ffi.REQ_CONV, self._host(a), self._host(b), self._host(dst), max_a, max_b, flags,
                            keep=(a, b, dst))

    def submit_device(self, kind: int, d_a: int, len_a: int, d_b: int, len_b: int, d_dst: int, dst_len: int,
                      max_a: int = 0, max_b: int = 0, signed: bool = False) -> int:
        flags = _ffi.REQ_DEVICE | (_ffi.REQ_SIGNED if signed else 0)
        return self._submit(kind, (d_a, len_a), (d_b, len_b), (d_dst, dst_len), max_a, max_b, flags)

    def poll(self, ticket: int) -> int:
        return int(_lib().ntt_server_poll(self._h, ticket))

    def wait(self, ticket: int) -> int:
        st = int(_lib().ntt_server_wait(self._h, ticket))
        return st

    def drain(self) -> bool:
        return bool(_lib().ntt_server_drain(self._h))

    def times(self, ticket: int):
        s, b, d = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        if not _lib().ntt_server_times(self._h, ticket, ctypes.byref(s), ctypes.byref(b), ctypes.byref(d)):
            return None
        return s.value, b.value, d.value

    def release(self, ticket: int) -> None:
        _lib().ntt_server_release(self._h, ticket)
        self._keep.pop(ticket, None)

    def stats(self) -> dict:
        st = _ffi.ServerStatsC()
        _lib().ntt_server_get_stats(self._h, ctypes.byref(st))
        return {name: int(getattr(st, name)) for name, _ in _ffi.ServerStatsC._fields_}


def clock() -> float:
    """The server's monotonic clock, in seconds."""
    return float(_lib().ntt_clock_s())


class DeviceBuffer:
    """Device memory owned by the engine's allocator (hipMalloc)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = _lib().ntt_device_alloc(self.nbytes)
        if not self.ptr:
            raise MemoryError(f"ntt_device_alloc({nbytes})")

    @classmethod
    def from_host(cls, arr: np.ndarray) -> "DeviceBuffer":
        buf = cls(arr.nbytes)
        buf.upload(arr)
        return buf

    def upload(self, arr: np.ndarray) -> None:
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes or not _lib().ntt_copy_to_device(self.ptr, arr.ctypes.data, arr.nbytes):
            raise RuntimeError("ntt_copy_to_device failed")

    def download(self, count: int, dtype=np.uint64) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes > self.nbytes or not _lib().ntt_copy_to_host(out.ctypes.data, self.ptr, out.nbytes):
            raise RuntimeError("ntt_copy_to_host failed")
        return out

    def free(self) -> None:
        if self.ptr:
            _lib().ntt_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def mul_device_noalloc(d_dst: int, dst_len: int, d_a: int, a_len: int, d_b: int, b_len: int,
                       stream: Optional[int] = None) -> bool:
    """Product of device operands that allocates no device memory (its
    working memory is the library's fixed 32 MiB device array; products of
    transforms up to 2^17 points). The same address and length twice squares."""
    return bool(_lib().ntt_mul_device_noalloc(d_dst, dst_len, d_a, a_len, d_b, b_len, stream))


def device_available() -> bool:
    return bool(_lib().ntt_device_available())


def synchronize() -> bool:
    return bool(_lib().ntt_device_synchronize())


def mul_geometry(len_a: int, len_b: int):
    """(primes, chunk_bits, log_len, row_tail) the engine picks for a product
    shape: the transform row has N + row_tail*N/4 points, N = 2^log_len."""
    p, w, k, t = _ffi.u32(), _ffi.u32(), _ffi.u32(), _ffi.u32()
    if not _lib().ntt_mul_geometry(len_a, len_b, ctypes.byref(p), ctypes.byref(w), ctypes.byref(k),
                                   ctypes.byref(t)):
        return None
    return p.value, w.value, k.value, t.value


def row_shape(conv_len: int, longest: int):
    """(row_len, cyclic_len, table_len) of the reference's row_shape
    (src/profile/planner.rs:565-591) for a convolution of conv_len
    coefficients whose operands have at most `longest` coefficients."""
    r, c, t = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    if not _lib().ntt_row_shape(conv_len, longest, ctypes.byref(r), ctypes.byref(c), ctypes.byref(t)):
        return None
    return r.value, c.value, t.value


def mul_split(len_a: int, len_b: int):
    """(blocks, block_digits) of the unbalanced split the engine picks for a
    product shape; blocks == 0 when the product is not split."""
    n, d = _ffi.u32(), ctypes.c_uint64()
    if not _lib().ntt_mul_split(len_a, len_b, ctypes.byref(n), ctypes.byref(d)):
        return None
    return n.value, d.value


def mul_schoolbook(len_a: int, len_b: int) -> bool:
    """True when the engine runs a product of this shape as a schoolbook
    (direct sums of limb products) instead of transforms."""
    return bool(_lib().ntt_mul_schoolbook(len_a, len_b))


def row_len(log_len: int, row_tail: int) -> int:
    """Points of a transform row: N + row_tail*N/4 with N = 2^log_len."""
    n = 1 << log_len
    return n + row_tail * (n >> 2)


class Profile:
    """Per-kernel event timing of the engine's launches."""

    @staticmethod
    def enable(on: bool = True) -> None:
        _lib().ntt_profile_enable(1 if on else 0)

    @staticmethod
    def reset() -> None:
        _lib().ntt_profile_reset()

    @staticmethod
    def read() -> dict:
        cap = 64
        names = ctypes.create_string_buffer(8192)
        ms = (ctypes.c_double * cap)()
        counts = (ctypes.c_uint64 * cap)()
        n = _lib().ntt_profile_read(names, len(names), ms, counts, cap)
        raw = names.raw.split(b"\0")
        return {raw[i].decode(): (ms[i], counts[i]) for i in range(n)}

    @staticmethod
    def replay(kind: str, reps: int = 20) -> Optional[float]:
        """Average duration (ms) of `reps` back-to-back re-issues of the last
        recorded launch of `kind` on its own stream, or None when no launch of
        that kind was recorded. The replays rewrite that launch's outputs."""
        out = ctypes.c_double(0.0)
        if not _lib().ntt_profile_replay(kind.encode(), int(reps), ctypes.byref(out)):
            return None
        return out.value


SPLIT_USES = {"product": 0, "square": 1, "batch": 2}


FORM_FLAGS = {"alt_tiles": 1, "inv16": 2, "per_prime_pack": 4, "rows8": 8, "strided8": 16, "paired": 32, "waves": 64}
FORM_TUNED = 256


def transform_split(log_len: int, use: str = "product"):
    """Level counts of the passes the engine runs for a transform of
    2^log_len points (strided passes, then the row pass) for a single
    operation's product or square, or a batch bucket, as tuned."""
    return transform_form(log_len, use)["passes"]


def transform_form(log_len: int, use: str = "product") -> dict:
    """The split of transform_split with the kernel forms chosen beside it:
    {"passes": [...], "forms": [names of FORM_FLAGS set], "tuned": bool}
    ("tuned": the autotuner has timed this length and use)."""
    logs = (ctypes.c_uint32 * 3)()
    flags = ctypes.c_uint32(0)
    n = _lib().ntt_transform_split(log_len, SPLIT_USES[use], logs, ctypes.byref(flags))
    return {
        "passes": [int(logs[i]) for i in range(n)],
        "forms": [k for k, bit in FORM_FLAGS.items() if flags.value & bit],
        "tuned": bool(flags.value & FORM_TUNED),
    }
