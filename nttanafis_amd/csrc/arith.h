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

// 64 bits of the operand from bit `pos`, zero beyond its end. Both loads are
// unconditional (clamped indices), so unrolled callers can issue them early.
__host__ __device__ __forceinline__ uint64_t window64(const uint64_t* limbs, uint64_t len, uint64_t pos) {
    const uint64_t idx = pos >> 6;
    const uint32_t off = (uint32_t)(pos & 63);
    const uint64_t last = len - 1;
    const uint64_t lo_raw = limbs[idx < last ? idx : last];
    const uint64_t hi_raw = limbs[idx + 1 < last ? idx + 1 : last];
    const uint64_t lo = idx < len ? lo_raw : 0;
    const uint64_t hi = idx + 1 < len ? hi_raw : 0;
    return (lo >> off) | ((hi << 1) << (63 - off));
}

__host__ __device__ __forceinline__ double pack_coef(const Source& s, uint64_t n, const double* pw, double p,
                                                     double pinv) {
    const uint32_t vb = s.value_bits;
    if (s.width == 64) {
        // One coefficient per limb: at most two parts, both in the limb.
        const uint64_t last = s.limb_len - 1;
        const uint64_t limb = s.limbs[n < last ? n : last];
        const int high_bits = vb > (uint32_t)kPartBits ? (int)vb - kPartBits : (int)vb;
        double r;
        if (vb > (uint32_t)kPartBits) {
            const double d0 = (double)(limb & ((1ull << kPartBits) - 1));
            const uint64_t masked = (limb >> kPartBits) & ((1ull << high_bits) - 1);
            const uint64_t flip = s.is_signed ? 1ull << (high_bits - 1) : 0;
            const double d1 = (double)((int64_t)(masked ^ flip) - (int64_t)flip);
            r = reduce(d0 + mulmod(d1, pw[1], p, pinv), p, pinv);
        } else {
            const uint64_t masked = limb & ((1ull << high_bits) - 1);
            const uint64_t flip = s.is_signed ? 1ull << (high_bits - 1) : 0;
            r = (double)((int64_t)(masked ^ flip) - (int64_t)flip);
        }
        return n < s.digits ? r : 0.0;
    }
    // General width (at most 192 bits): one window of four limbs covers the
    // coefficient at any bit offset; its parts are cut from registers.
    const int parts = (int)((vb + kPartBits - 1) / kPartBits);
    const int high_bits = (int)vb - kPartBits * (parts - 1);
    const uint64_t bitpos = n * (uint64_t)s.width;
    const uint64_t idx = bitpos >> 6;
    const uint32_t off = (uint32_t)(bitpos & 63);
    const uint64_t last = s.limb_len - 1;
    uint64_t w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint64_t i = idx + q;
        const uint64_t raw = s.limbs[i < last ? i : last];
        w[q] = i < s.limb_len ? raw : 0;
    }
    double sum = 0.0;
#pragma unroll
    for (int k = 0; k < kMaxParts; ++k) {
        if (k >= parts) break;
        const uint32_t bit = off + (uint32_t)(kPartBits * k);
        const uint32_t q = bit >> 6, sh = bit & 63;
        const uint64_t lo = q == 0 ? w[0] : q == 1 ? w[1] : q == 2 ? w[2] : w[3];
        const uint64_t hi = q == 0 ? w[1] : q == 1 ? w[2] : q == 2 ? w[3] : 0;
        const uint64_t word = (lo >> sh) | ((hi << 1) << (63 - sh));
        double d;
        if (k + 1 == parts) {
            const uint64_t masked = word & ((1ull << high_bits) - 1);
            if (s.is_signed) {
                const uint64_t flip = 1ull << (high_bits - 1);
                d = (double)((int64_t)(masked ^ flip) - (int64_t)flip);
            } else {
                d = (double)masked;
            }
        } else {
            d = (double)(word & ((1ull << kPartBits) - 1));
        }
        sum = k == 0 ? d : sum + mulmod(d, pw[k], p, pinv);
    }
    const double r = parts > 1 ? reduce(sum, p, pinv) : sum;
    return n < s.digits ? r : 0.0;
}

template <int NP>
struct CrtWords {
    static constexpr int value = NP == 1 ? 1 : NP == 2 ? 2 : NP == 3 ? 3 : NP <= 5 ? 4 : NP == 6 ? 5 : NP == 7 ? 6 : 7;
};

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
    int64_t top = (int64_t)y[NP - 1];
    if (balanced && 2 * (uint64_t)top > a.pint[NP - 1]) top -= (int64_t)a.pint[NP - 1];
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

template <int NP>
__host__ __device__ __forceinline__ void garner(const CrtArgs& a, uint64_t i, uint64_t* X, bool balanced) {
    double r[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) r[k] = a.residues[(uint64_t)k * a.row_stride + i];
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
