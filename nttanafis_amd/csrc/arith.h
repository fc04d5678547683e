// Arithmetic shared by the gfx950 kernels and the host self-test: limb
// split into residues, Garner recombination and carry functions. Every
// function here compiles for both sides, so the host can check it against
// exact integer arithmetic without a GPU.
#pragma once

#include "device.h"

namespace nttamd {

__host__ __device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) {
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
}

// ---------------------------------------------------------------------------
// Packing (limb split): coefficient n is bits [n*width, n*width + value_bits)
// of the operand, split into 49-bit parts; its residue is
// part_0 + sum_k part_k * (2^(49k) mod p), reduced once. Restates the
// reference's Packing::pack (src/limb/packing.rs:113-292,
// src/arch/schedule/expression/packing.rs:68-109).

// Limb i of an operand. On the device the load goes through the global
// address space: operands reach batched kernels inside structures, where the
// compiler would otherwise load them as generic (flat) accesses, which also
// count against the LDS counter of the packing's LDS waits.
__host__ inline uint64_t limb_at(const uint64_t* limbs, uint64_t i) { return limbs[i]; }
__device__ __forceinline__ uint64_t limb_at(const uint64_t* limbs, uint64_t i) {
    return ((const __attribute__((address_space(1))) uint64_t*)limbs)[i];
}

// 64 bits of the operand from bit `pos`, zero beyond its end. Both loads are
// unconditional (clamped indices), so unrolled callers can issue them early.
__host__ __device__ __forceinline__ uint64_t window64(const uint64_t* limbs, uint64_t len, uint64_t pos) {
    const uint64_t idx = pos >> 6;
    const uint32_t off = (uint32_t)(pos & 63);
    const uint64_t last = len - 1;
    const uint64_t lo_raw = limb_at(limbs, idx < last ? idx : last);
    const uint64_t hi_raw = limb_at(limbs, idx + 1 < last ? idx + 1 : last);
    const uint64_t lo = idx < len ? lo_raw : 0;
    const uint64_t hi = idx + 1 < len ? hi_raw : 0;
    return (lo >> off) | ((hi << 1) << (63 - off));
}

// Parts of a coefficient as doubles: a part v below 2^52 sits in the low
// mantissa bits of 2^52 + v, so that bit pattern minus 2^52 is v exactly (an
// OR and one FP64 subtraction instead of two conversions and an FMA; the
// biased conversion of src/limb/packing.rs:224-230).
__host__ __device__ __forceinline__ double part_to_double(uint64_t v) {
    return __builtin_bit_cast(double, v | 0x4330000000000000ull) - 4503599627370496.0;
}

// Bits of the top part at most three: its product with a centered power is
// exact in binary64 (the SHORT_HIGH form, combine_parts below).
__host__ __device__ __forceinline__ bool short_high_part(int high_bits) { return high_bits <= 3; }

// One limb per coefficient (width 64): at most two parts, both in the limb;
// a signed coefficient is two's complement in its value_bits.
// With CLAMP false the caller guarantees that coefficient n and every limb
// its window reads lie inside the operand (n < digits), so no index is
// clamped and no result masked.
template <bool SIGNED, bool CLAMP = true>
__host__ __device__ __forceinline__ double pack_w64(const Source& s, uint64_t n, const double* pw, double p,
                                                    double pinv) {
    const uint32_t vb = s.value_bits;
    const uint64_t last = s.limb_len - 1;
    const uint64_t limb = CLAMP ? limb_at(s.limbs, n < last ? n : last) : limb_at(s.limbs, n);
    double r;
    if (vb > (uint32_t)kPartBits) {
        const int high_bits = (int)vb - kPartBits;
        const double d0 = part_to_double(limb & ((1ull << kPartBits) - 1));
        const uint64_t masked = (limb >> kPartBits) & ((1ull << high_bits) - 1);
        const double d1 = SIGNED ? (double)((int64_t)(masked ^ (1ull << (high_bits - 1))) -
                                            (int64_t)(1ull << (high_bits - 1)))
                                 : (double)(uint32_t)masked;
        r = reduce(d0 + (short_high_part(high_bits) ? d1 * pw[1] : mulmod(d1, pw[1], p, pinv)), p, pinv);
    } else {
        const uint64_t masked = limb & ((1ull << vb) - 1);
        r = SIGNED ? (double)((int64_t)(masked ^ (1ull << (vb - 1))) - (int64_t)(1ull << (vb - 1)))
                   : part_to_double(masked);
    }
    return !CLAMP || n < s.digits ? r : 0.0;
}

// General width (at most 192 bits, unsigned) with PARTS = ceil(value_bits /
// 49) fixed at compile time. Coefficient n starts at bit off of limb idx; a
// window of W limbs covers it, and part k starts in limb (off + 49k) / 64,
// which is one of two known limbs.
template <int PARTS>
__host__ __device__ constexpr int pack_window() {
    return (63 + PARTS * kPartBits + 63) / 64 < 4 ? (63 + PARTS * kPartBits + 63) / 64 : 4;
}

// Parts of the coefficient that starts at bit `off` of w[0], from a window
// w[0..W] of limbs (w[W] = 0), as doubles; the top part has high_bits bits.
template <int PARTS>
__host__ __device__ __forceinline__ void parts_from_window(const uint64_t* w, uint32_t off, int high_bits,
                                                           double* d) {
    constexpr int W = pack_window<PARTS>();
#pragma unroll
    for (int k = 0; k < PARTS; ++k) {
        const uint32_t bit = off + (uint32_t)(kPartBits * k);
        const int q0 = (kPartBits * k) >> 6;  // limb of the part when off is 0
        const bool next = (bit >> 6) != (uint32_t)q0;
        const uint32_t sh = bit & 63;
        const uint64_t lo = next ? w[q0 + 1 < W ? q0 + 1 : W] : w[q0];
        const uint64_t hi = next ? w[q0 + 2 < W ? q0 + 2 : W] : w[q0 + 1 < W ? q0 + 1 : W];
        const uint64_t word = (lo >> sh) | ((hi << 1) << (63 - sh));
        const uint64_t mask = k + 1 == PARTS ? ((1ull << high_bits) - 1) : ((1ull << kPartBits) - 1);
        d[k] = part_to_double(word & mask);
    }
}

// Residue of a coefficient from its parts: part_0 + sum_k part_k (2^(49k) mod p),
// reduced once. short_high (the SHORT_HIGH form of
// src/arch/schedule/expression/packing.rs:12, 99-101): the top part has at
// most three bits, so its product with the centered power (below 8 * p/2 =
// 4p) is exact in binary64 and needs no residue product; the sum then stays
// below p + 2 * 15p/16 + 4p < 8p for four parts, inside reduce's range.
template <int PARTS>
__host__ __device__ __forceinline__ double combine_parts(const double* d, const double* pw, double p, double pinv,
                                                         bool short_high = false) {
    double sum = d[0];
#pragma unroll
    for (int k = 1; k < PARTS; ++k)
        sum = sum + ((k == PARTS - 1 && short_high) ? d[k] * pw[k] : mulmod(d[k], pw[k], p, pinv));
    return PARTS > 1 ? reduce(sum, p, pinv) : sum;
}

template <int PARTS>
__host__ __device__ __forceinline__ double residue_from_window(const uint64_t* w, uint32_t off, int high_bits,
                                                               const double* pw, double p, double pinv) {
    double d[PARTS];
    parts_from_window<PARTS>(w, off, high_bits, d);
    return combine_parts<PARTS>(d, pw, p, pinv, short_high_part(high_bits));
}

// Parts of coefficient n of a general-width source (zero at or beyond
// digits); CLAMP as for pack_parts.
template <int PARTS, bool CLAMP = true>
__host__ __device__ __forceinline__ void extract_parts(const Source& s, uint64_t n, double* d) {
    constexpr int W = pack_window<PARTS>();
    const uint64_t bitpos = n * (uint64_t)s.width;
    const uint64_t idx = bitpos >> 6;
    const uint32_t off = (uint32_t)bitpos & 63;
    const uint64_t last = s.limb_len - 1;
    uint64_t w[W + 1];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        const uint64_t i = idx + q;
        if (CLAMP) {
            const uint64_t raw = limb_at(s.limbs, i < last ? i : last);
            w[q] = i < s.limb_len && n < s.digits ? raw : 0;
        } else {
            w[q] = limb_at(s.limbs, i);
        }
    }
    w[W] = 0;
    parts_from_window<PARTS>(w, off, (int)s.value_bits - kPartBits * (PARTS - 1), d);
}

template <int PARTS, bool CLAMP = true>
__host__ __device__ __forceinline__ double pack_parts(const Source& s, uint64_t n, const double* pw, double p,
                                                      double pinv) {
    constexpr int W = pack_window<PARTS>();
    const uint64_t bitpos = n * (uint64_t)s.width;
    const uint64_t idx = bitpos >> 6;
    const uint32_t off = (uint32_t)bitpos & 63;
    const uint64_t last = s.limb_len - 1;
    uint64_t w[W + 1];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        const uint64_t i = idx + q;
        if (CLAMP) {
            const uint64_t raw = limb_at(s.limbs, i < last ? i : last);
            w[q] = i < s.limb_len ? raw : 0;
        } else {
            w[q] = limb_at(s.limbs, i);
        }
    }
    w[W] = 0;
    const int high_bits = (int)s.value_bits - kPartBits * (PARTS - 1);
    const double r = residue_from_window<PARTS>(w, off, high_bits, pw, p, pinv);
    return !CLAMP || n < s.digits ? r : 0.0;
}

// Residue of coefficient n of `s` (dispatch over the forms above; callers in
// hot loops dispatch once with pack_kind and call the form directly).
enum PackKind { kPackW64, kPackW64Signed, kPack1, kPack2, kPack3, kPack4 };

__host__ __device__ __forceinline__ int pack_kind(const Source& s) {
    if (s.width == 64) return s.is_signed ? kPackW64Signed : kPackW64;
    const int parts = (int)((s.value_bits + kPartBits - 1) / kPartBits);
    return kPack1 + parts - 1;
}

template <int KIND, bool CLAMP = true>
__host__ __device__ __forceinline__ double pack_as(const Source& s, uint64_t n, const double* pw, double p,
                                                   double pinv) {
    if constexpr (KIND == kPackW64) return pack_w64<false, CLAMP>(s, n, pw, p, pinv);
    if constexpr (KIND == kPackW64Signed) return pack_w64<true, CLAMP>(s, n, pw, p, pinv);
    if constexpr (KIND >= kPack1) return pack_parts<KIND - kPack1 + 1, CLAMP>(s, n, pw, p, pinv);
    return 0.0;
}

// Parts of one coefficient, extracted once and combined per prime (packing
// shared across the primes of a tile, as the reference's digit tiles are,
// src/limb/packing.rs:207-218 and src/arch/schedule/expression/packing.rs:31-60).
// Whole-limb kinds keep two parts: the low 49 bits and the rest (sign-extended
// for a signed coefficient), or the whole value and 0 when it has at most 49
// bits. General widths keep ceil(value_bits / 49) parts.
template <int KIND>
struct PackParts {
    static constexpr int value = (KIND == kPackW64 || KIND == kPackW64Signed) ? 2 : KIND - kPack1 + 1;
};

template <int KIND, bool CLAMP>
__host__ __device__ __forceinline__ void extract_as(const Source& s, uint64_t n, double* d) {
    if constexpr (KIND == kPackW64 || KIND == kPackW64Signed) {
        constexpr bool SIGNED = KIND == kPackW64Signed;
        const uint32_t vb = s.value_bits;
        const uint64_t last = s.limb_len - 1;
        const uint64_t limb = CLAMP ? (n < s.digits ? limb_at(s.limbs, n < last ? n : last) : 0) : limb_at(s.limbs, n);
        if (vb > (uint32_t)kPartBits) {
            const int high_bits = (int)vb - kPartBits;
            d[0] = part_to_double(limb & ((1ull << kPartBits) - 1));
            const uint64_t masked = (limb >> kPartBits) & ((1ull << high_bits) - 1);
            d[1] = SIGNED ? (double)((int64_t)(masked ^ (1ull << (high_bits - 1))) - (int64_t)(1ull << (high_bits - 1)))
                          : part_to_double(masked);
        } else {
            const uint64_t masked = limb & ((1ull << vb) - 1);
            d[0] = SIGNED ? (double)((int64_t)(masked ^ (1ull << (vb - 1))) - (int64_t)(1ull << (vb - 1)))
                          : part_to_double(masked);
            d[1] = 0.0;
        }
    } else {
        extract_parts<KIND - kPack1 + 1, CLAMP>(s, n, d);
    }
}

// Residue of a coefficient from its extracted parts: below 9p/16 in
// magnitude after the one reduction, or the part itself (below 2^49 < p)
// when there is only one.
template <int KIND>
__host__ __device__ __forceinline__ double combine_as(const Source& s, const double* d, const double* pw, double p,
                                                      double pinv) {
    if constexpr (KIND == kPackW64 || KIND == kPackW64Signed) {
        if (s.value_bits <= (uint32_t)kPartBits) return d[0];
        const double high = short_high_part((int)s.value_bits - kPartBits) ? d[1] * pw[1] : mulmod(d[1], pw[1], p, pinv);
        return reduce(d[0] + high, p, pinv);
    } else {
        constexpr int PARTS = KIND - kPack1 + 1;
        return combine_parts<PARTS>(d, pw, p, pinv,
                                    PARTS > 1 && short_high_part((int)s.value_bits - kPartBits * (PARTS - 1)));
    }
}

// Number of leading coefficients whose window lies inside the operand and
// which are below `digits`: pack_as<KIND, false> is exact for every n below
// it.
template <int KIND>
__host__ __device__ __forceinline__ uint64_t pack_safe(const Source& s) {
    if constexpr (KIND == kPackW64 || KIND == kPackW64Signed) {
        return s.digits < s.limb_len ? s.digits : s.limb_len;
    } else {
        constexpr uint64_t W = pack_window<KIND - kPack1 + 1>();
        if (s.limb_len < W) return 0;
        // idx + W - 1 < limb_len with idx = floor(n * width / 64).
        const uint64_t n = ((s.limb_len - W + 1) * 64 + s.width - 1) / s.width;
        return n < s.digits ? n : s.digits;
    }
}

// True when coefficients [0, n_end) may all be packed with
// pack_as<KIND, false>.
template <int KIND>
__host__ __device__ __forceinline__ bool pack_interior(const Source& s, uint64_t n_end) {
    return n_end <= pack_safe<KIND>(s);
}

__host__ __device__ __forceinline__ double pack_coef(const Source& s, uint64_t n, const double* pw, double p,
                                                     double pinv) {
    switch (pack_kind(s)) {
        case kPackW64: return pack_as<kPackW64>(s, n, pw, p, pinv);
        case kPackW64Signed: return pack_as<kPackW64Signed>(s, n, pw, p, pinv);
        case kPack1: return pack_as<kPack1>(s, n, pw, p, pinv);
        case kPack2: return pack_as<kPack2>(s, n, pw, p, pinv);
        case kPack3: return pack_as<kPack3>(s, n, pw, p, pinv);
        default: return pack_as<kPack4>(s, n, pw, p, pinv);
    }
}

template <int NP>
struct CrtWords {
    static constexpr int value = NP == 1 ? 1 : NP == 2 ? 2 : NP == 3 ? 3 : NP <= 5 ? 4 : NP == 6 ? 5 : NP == 7 ? 6 : 7;
};

// Mixed radix to binary for canonical digits y (X = y[0] + p0*(y[1] + p1*(...)))
// in 32-bit words, so every product is one 32x32->64 multiply-add. After the
// step for prime k, X < p_k * ... * p_{NP-1} < 2^(50*(NP-k)), which bounds the
// words each step touches. p = plo + phi*2^32 with phi < 2^18: word i of
// X*p + y takes x_i*plo from the first carry chain and x_{i-1}*phi from the
// second; neither multiply-add can exceed 64 bits (the first chain's carry
// stays below 2^32 + 2^18, the second's below 2^19).
// The integer value of an integer-valued double y in [0, 2^52): y + 2^52 is
// exact and its low 52 mantissa bits are y (two instructions instead of a
// general conversion).
__host__ __device__ __forceinline__ uint64_t exact_u64(double y) {
    return __builtin_bit_cast(uint64_t, y + 4503599627370496.0) - 0x4330000000000000ull;
}

template <int NP>
__host__ __device__ __forceinline__ void horner32(const CrtArgs& a, const double* y, uint64_t* X) {
    constexpr int NW = CrtWords<NP>::value;
    constexpr int N32 = 2 * NW;
    uint32_t x[N32];
    const uint64_t top = exact_u64(y[NP - 1]);
    x[0] = (uint32_t)top;
    x[1] = (uint32_t)(top >> 32);
#pragma unroll
    for (int i = 2; i < N32; ++i) x[i] = 0;
#pragma unroll
    for (int k = NP - 2; k >= 0; --k) {
        const int words = (kPartBits + 1) * (NP - k) / 32 + 1 < N32 ? (kPartBits + 1) * (NP - k) / 32 + 1 : N32;
        const uint64_t p = a.pint[k];
        const uint32_t plo = (uint32_t)p, phi = (uint32_t)(p >> 32);
        const uint64_t yk = exact_u64(y[k]);
        uint64_t c1 = (uint32_t)yk, c2 = 0;
        uint32_t prev = 0;
#pragma unroll
        for (int i = 0; i < N32; ++i) {
            if (i >= words) break;
            const uint32_t xo = x[i];
            const uint64_t t1 = (uint64_t)xo * plo + c1;
            c1 = (t1 >> 32) + (i == 0 ? (yk >> 32) : 0);
            const uint64_t t2 = (uint64_t)prev * phi + ((t1 & 0xFFFFFFFFull) + c2);
            x[i] = (uint32_t)t2;
            c2 = t2 >> 32;
            prev = xo;
        }
    }
#pragma unroll
    for (int w = 0; w < NW; ++w) X[w] = (uint64_t)x[2 * w] | ((uint64_t)x[2 * w + 1] << 32);
}

template <int NP>
__host__ __device__ __forceinline__ void garner_r(const CrtArgs& a, const double* r, uint64_t* X, bool balanced) {
    constexpr int NW = CrtWords<NP>::value;
    double y[NP];
    y[0] = r[0];
#pragma unroll
    for (int k = 1; k < NP; ++k) {
        const double pk = a.field.p[k], pinvk = a.field.pinv[k];
        double u = y[k - 1];
#pragma unroll
        for (int j = k - 2; j >= 0; --j) u = reduce(y[j] + mulmod(u, a.pmod[j][k], pk, pinvk), pk, pinvk);
        y[k] = canonical(mulmod(r[k] - u, a.ginv[k], pk, pinvk), pk);
    }
    if (!balanced) {
        horner32<NP>(a, y, X);
        return;
    }
    int64_t top = (int64_t)y[NP - 1];
    if (2 * (uint64_t)top > a.pint[NP - 1]) top -= (int64_t)a.pint[NP - 1];
    X[0] = (uint64_t)top;
#pragma unroll
    for (int w = 1; w < NW; ++w) X[w] = top < 0 ? ~0ull : 0ull;
#pragma unroll
    for (int k = NP - 2; k >= 0; --k) {
        const uint64_t pk = a.pint[k];
        uint64_t carry = (uint64_t)y[k];
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            uint64_t lo = X[w] * pk;
            uint64_t hi = mulhi64(X[w], pk);
            lo += carry;
            hi += lo < carry;
            X[w] = lo;
            carry = hi;
        }
    }
}

// Residues of coefficient i: one row entry per prime, or for a split
// operation the sum of the entries of the (at most two) blocks covering i:
// block z1 = i / D (the last block for i past it) at i - z1*D, and block
// z1 - 1 at i - z1*D + D while that lies in its row.
template <int NP>
__host__ __device__ __forceinline__ void garner(const CrtArgs& a, uint64_t i, uint64_t* X, bool balanced) {
    double r[NP];
    if (a.split_blocks == 0) {
#pragma unroll
        for (int k = 0; k < NP; ++k) r[k] = a.residues[(uint64_t)k * a.row_stride + i];
    } else {
        const uint64_t D = a.split_digits;
        uint64_t z1 = i / D;
        if (z1 >= a.split_blocks) z1 = a.split_blocks - 1;
        const uint64_t k1 = i - z1 * D;
        const bool two = z1 > 0 && k1 + D < a.row_stride;
        const double* r1 = a.residues + z1 * a.block_stride + k1;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            double v = r1[(uint64_t)k * a.row_stride];
            if (two) {
                v += r1[(uint64_t)k * a.row_stride - a.block_stride + D];
                if (v >= a.field.p[k]) v -= a.field.p[k];
            }
            r[k] = v;
        }
    }
    garner_r<NP>(a, r, X, balanced);
}

// Carry function of a limb range: carry_out(c) = base + (c >= thresh).
struct CarryFn {
    uint32_t base, thresh;
};

// f1 runs first, then f2.
__host__ __device__ __forceinline__ CarryFn compose(CarryFn f1, CarryFn f2) {
    CarryFn r;
    r.base = f2.base + (f1.base >= f2.thresh ? 1u : 0u);
    r.thresh = (f1.base + 1u == f2.thresh) ? f1.thresh : kInf;
    return r;
}

__host__ __device__ __forceinline__ uint32_t limb_thresh(uint64_t v) {
    const uint64_t t = 0 - v;  // 2^64 - v, zero for v == 0
    return (v != 0 && t < (uint64_t)kInf) ? (uint32_t)t : kInf;
}

}  // namespace nttamd
