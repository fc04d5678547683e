"""Prime fields 2^49 < p < 2^50 and two-prime Chinese remaindering.

Mirror of the reference's public scalar API (src/field/prime_field.rs:91-442)
over the library's C ABI (ntt_field_*, ntt_crt_pair_*): a residue is a
canonical int in [0, p); a field's operations expect its own residues.
"""
from __future__ import annotations

import ctypes
from typing import Iterator, Optional, Sequence

import numpy as np

from . import _ffi


def _lib():
    return _ffi.load()


class PrimeField:
    """A prime field Z/pZ with 2^49 < p < 2^50 (PrimeField, prime_field.rs:34-37)."""

    __slots__ = ("_c",)

    def __init__(self, c: _ffi.PrimeFieldC):
        self._c = c

    @classmethod
    def new(cls, modulus: int) -> Optional["PrimeField"]:
        """None unless modulus is a prime with 2^49 < modulus < 2^50."""
        c = _ffi.PrimeFieldC()
        if not 0 <= modulus < 1 << 64 or not _lib().ntt_field_new(modulus, ctypes.byref(c)):
            return None
        return cls(c)

    @classmethod
    def below(cls, bound: int) -> Optional["PrimeField"]:
        """The largest prime below bound and 2^50; None when none exceeds 2^49."""
        c = _ffi.PrimeFieldC()
        if not _lib().ntt_field_below(min(bound, (1 << 64) - 1), ctypes.byref(c)):
            return None
        return cls(c)

    @classmethod
    def catalog(cls) -> Iterator["PrimeField"]:
        """The transform's eight primes in reconstruction order."""
        arr = (_ffi.PrimeFieldC * 8)()
        n = _lib().ntt_field_catalog(arr, 8)
        return iter([cls(arr[i]) for i in range(n)])

    def __eq__(self, other) -> bool:
        return isinstance(other, PrimeField) and self.modulus == other.modulus

    def __hash__(self) -> int:
        return hash(self.modulus)

    def __repr__(self) -> str:
        return f"PrimeField({self.modulus})"

    @property
    def modulus(self) -> int:
        return int(self._c.modulus)

    def _p(self):
        return ctypes.byref(self._c)

    def residue(self, value: int) -> Optional[int]:
        out = ctypes.c_uint64()
        return int(out.value) if _lib().ntt_field_residue(self._p(), value, ctypes.byref(out)) else None

    def reduce(self, value: int) -> int:
        """Canonical residue of a signed 64-bit integer."""
        return int(_lib().ntt_field_reduce(self._p(), value))

    def reduce_u64(self, value: int) -> int:
        return int(_lib().ntt_field_reduce_u64(self._p(), value))

    def signed(self, residue: int) -> int:
        """The representative in -(p-1)/2..=(p-1)/2."""
        return int(_lib().ntt_field_signed(self._p(), residue))

    def add(self, a: int, b: int) -> int:
        return int(_lib().ntt_field_add(self._p(), a, b))

    def sub(self, a: int, b: int) -> int:
        return int(_lib().ntt_field_sub(self._p(), a, b))

    def neg(self, a: int) -> int:
        return int(_lib().ntt_field_neg(self._p(), a))

    def mul(self, a: int, b: int) -> int:
        return int(_lib().ntt_field_mul(self._p(), a, b))

    def pow(self, base: int, exponent: int) -> int:
        return int(_lib().ntt_field_pow(self._p(), base, exponent))

    def inv(self, a: int) -> Optional[int]:
        out = ctypes.c_uint64()
        return int(out.value) if _lib().ntt_field_inv(self._p(), a, ctypes.byref(out)) else None

    def mul_sub(self, target: np.ndarray, source: np.ndarray, factor: int) -> None:
        """target[i] -= source[i] * factor, in place (uint64 arrays)."""
        n = min(len(target), len(source))
        _lib().ntt_field_mul_sub(self._p(), target.ctypes.data_as(_ffi.P64),
                                 np.ascontiguousarray(source, dtype=np.uint64).ctypes.data_as(_ffi.P64), n, factor)

    def scale(self, target: np.ndarray, factor: int) -> None:
        """target[i] *= factor, in place (uint64 array)."""
        _lib().ntt_field_scale(self._p(), target.ctypes.data_as(_ffi.P64), len(target), factor)

    def evaluate(self, coefficients: Sequence[int], point: int) -> int:
        """Sum of coefficients[i] * point^i."""
        c = np.ascontiguousarray(coefficients, dtype=np.uint64)
        return int(_lib().ntt_field_evaluate(self._p(), c.ctypes.data_as(_ffi.P64), len(c), point))


class CrtPair:
    """Chinese remaindering over two distinct prime fields (CrtPair, prime_field.rs:63-68)."""

    __slots__ = ("_c",)

    def __init__(self, c: _ffi.CrtPairC):
        self._c = c

    @classmethod
    def new(cls, first: PrimeField, second: PrimeField) -> Optional["CrtPair"]:
        """None when the two fields coincide."""
        c = _ffi.CrtPairC()
        if not _lib().ntt_crt_pair_new(ctypes.byref(first._c), ctypes.byref(second._c), ctypes.byref(c)):
            return None
        return cls(c)

    def modulus(self) -> int:
        lo, hi = ctypes.c_uint64(), ctypes.c_uint64()
        _lib().ntt_crt_pair_modulus(ctypes.byref(self._c), ctypes.byref(lo), ctypes.byref(hi))
        return int(lo.value) | int(hi.value) << 64

    def combine(self, first: int, second: int) -> int:
        """The integer in -(M-1)/2..=(M-1)/2 congruent to both residues."""
        lo, hi = ctypes.c_uint64(), ctypes.c_int64()
        _lib().ntt_crt_pair_combine(ctypes.byref(self._c), first, second, ctypes.byref(lo), ctypes.byref(hi))
        return int(lo.value) + (int(hi.value) << 64)
