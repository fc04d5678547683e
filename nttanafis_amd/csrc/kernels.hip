// gfx950 kernels of the multi-prime NTT engine.
//
// Transform layout: one row of T = 2^K doubles per (operand, prime). The
// forward transform is the reference's remainder tree (node j of every level
// uses the bit-reversed root W[j]; src/transform/twiddle.rs:116-156,
// src/arch/schedule/expression/stages.rs:332-383) split into passes over HBM:
//
//   pass 1  (strided): levels [0, K1) on lines {m * 2^K2 + lo}, 16 lines per
//           workgroup so every HBM access is a 128-byte run;
//   pass 2  (rows):    levels [K1, K) on contiguous rows of 2^K2 points.
//
// Each pass loads its tile into LDS once, runs radix-16 rounds (4 levels per
// LDS round trip, 16 values per thread in registers) and writes the tile back.
// Output stays in tree order, so the inverse (Gentleman-Sande, inverse roots)
// undoes it without any bit reversal. Root of node H*2^l + x of a row whose
// first node is H is W[x] * W[H*2^l]; rows build their roots in LDS that way
// instead of reading a full-length table.
#include "arith.h"

namespace nttamd {

__device__ __forceinline__ int pidx(int e) { return e + (e >> 4); }

template <int LOGL>
struct Lay {
    static constexpr int L = 1 << LOGL;
    static constexpr int STRIDE = L + (L >> 4) + 1;  // padded line in LDS
    static constexpr int ROUNDS = (LOGL + 3) / 4;
    static constexpr int FIRST = LOGL - 4 * (ROUNDS - 1);  // levels of round 0
};

// Roots of the tree are read from the root table of one prime in global
// memory (L1/L2-resident: a row pass reads each root a handful of times).
// Line c of a tile starts at node h0 + c*hstep of the pass's first level, so
// its node x at local level l is table[(h0 + c*hstep) << l | x].
struct Tw {
    const double* table;
    uint64_t h0;
    int hstep;  // 1 when every line is its own node (rows), 0 when lines share one
    __device__ __forceinline__ double at(int c, int lvl, int node) const {
        return table[((h0 + (uint64_t)(c * hstep)) << lvl) + node];
    }
};

// One DIF round: local levels [L0, L0 + RLOG) of every line of the tile.
//
// Reduction schedule, fixed at compile time: the upper input of a butterfly
// is reduced at the odd levels of a round. With |root| <= p/2, a level maps
// inputs below B to outputs below B + 15p/16 (the product stays below 2p^2),
// and a reducing level maps inputs below 27p/8 to outputs below
// 9p/16 + 15p/16 = 3p/2. Every round therefore ends below 3p/2 (its last
// level is odd) or, for a short first round ending on an even level, below
// 39p/16; values never exceed 27p/8, and spectra stay below 3p/2, so a
// pointwise product stays below 9p^2/4.
template <int LOGL, int LINES, int NT, int L0, int RLOG>
__device__ __forceinline__ void fwd_round(double* lds, Tw tw, double p, double pinv) {
    constexpr int L = 1 << LOGL;
    constexpr int RADIX = 1 << RLOG;
    constexpr int D = L >> (L0 + RLOG);
    constexpr int GPL = L / RADIX;
    constexpr int GROUPS = LINES * GPL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    for (int g = threadIdx.x; g < GROUPS; g += NT) {
        const int c = g / GPL;
        const int gl = g - c * GPL;
        const int b = gl / D;
        const int i = gl - b * D;
        double* line = lds + c * LS;
        const int base = b * (L >> L0) + i;
        double x[RADIX];
#pragma unroll
        for (int k = 0; k < RADIX; ++k) x[k] = line[pidx(base + k * D)];
#pragma unroll
        for (int s = 0; s < RLOG; ++s) {
            const int half = RADIX >> (s + 1);
#pragma unroll
            for (int blk = 0; blk < (1 << s); ++blk) {
                const double w = tw.at(c, L0 + s, (b << s) + blk);
#pragma unroll
                for (int j = 0; j < half; ++j) {
                    const int k = blk * 2 * half + j;
                    const double u = (s & 1) ? reduce(x[k], p, pinv) : x[k];
                    const double t = mulmod(x[k + half], w, p, pinv);
                    x[k] = u + t;
                    x[k + half] = u - t;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < RADIX; ++k) line[pidx(base + k * D)] = x[k];
    }
}

// One inverse (DIT) round, levels in decreasing order; the sum is reduced at
// the even levels of a round. Inputs below 11p/8 give differences below
// 11p/4 (products below 2p^2, outputs below 15p/16) and sums below 11p/4 at
// an odd level; an even level then sees inputs below 11p/4, reduces sums
// below 11p/2 to 9p/16 and keeps differences times a root below 4p^2
// (outputs below 11p/8). Every round ends on level 0 below 11p/8.
template <int LOGL, int LINES, int NT, int L0, int RLOG>
__device__ __forceinline__ void inv_round(double* lds, Tw tw, double p, double pinv) {
    constexpr int L = 1 << LOGL;
    constexpr int RADIX = 1 << RLOG;
    constexpr int D = L >> (L0 + RLOG);
    constexpr int GPL = L / RADIX;
    constexpr int GROUPS = LINES * GPL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    for (int g = threadIdx.x; g < GROUPS; g += NT) {
        const int c = g / GPL;
        const int gl = g - c * GPL;
        const int b = gl / D;
        const int i = gl - b * D;
        double* line = lds + c * LS;
        const int base = b * (L >> L0) + i;
        double x[RADIX];
#pragma unroll
        for (int k = 0; k < RADIX; ++k) x[k] = line[pidx(base + k * D)];
#pragma unroll
        for (int s = RLOG - 1; s >= 0; --s) {
            const int half = RADIX >> (s + 1);
#pragma unroll
            for (int blk = 0; blk < (1 << s); ++blk) {
                const double w = tw.at(c, L0 + s, (b << s) + blk);
#pragma unroll
                for (int j = 0; j < half; ++j) {
                    const int k = blk * 2 * half + j;
                    const double u = x[k], v = x[k + half];
                    const double sum = u + v;
                    x[k] = (s & 1) ? sum : reduce(sum, p, pinv);
                    x[k + half] = mulmod(u - v, w, p, pinv);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < RADIX; ++k) line[pidx(base + k * D)] = x[k];
    }
}

template <int LOGL, int LINES, int NT, int R>
__device__ __forceinline__ void fwd_rounds(double* lds, Tw tw, double p, double pinv) {
    if constexpr (R < Lay<LOGL>::ROUNDS) {
        constexpr int L0 = R == 0 ? 0 : Lay<LOGL>::FIRST + 4 * (R - 1);
        constexpr int RL = R == 0 ? Lay<LOGL>::FIRST : 4;
        fwd_round<LOGL, LINES, NT, L0, RL>(lds, tw, p, pinv);
        __syncthreads();
        fwd_rounds<LOGL, LINES, NT, R + 1>(lds, tw, p, pinv);
    }
}

template <int LOGL, int LINES, int NT, int R>
__device__ __forceinline__ void inv_rounds(double* lds, Tw tw, double p, double pinv) {
    if constexpr (R >= 0) {
        constexpr int L0 = R == 0 ? 0 : Lay<LOGL>::FIRST + 4 * (R - 1);
        constexpr int RL = R == 0 ? Lay<LOGL>::FIRST : 4;
        inv_round<LOGL, LINES, NT, L0, RL>(lds, tw, p, pinv);
        __syncthreads();
        inv_rounds<LOGL, LINES, NT, R - 1>(lds, tw, p, pinv);
    }
}

// ---------------------------------------------------------------------------
// Root tables: W[j] = prod over set bits m of j of the level root of order
// 2^(m+2), the product form of generate_tables (src/transform/twiddle.rs:
// 116-156); the inverse table uses the inverse level roots, so
// Winv[j] = W[j]^-1. Values are centered (|w| <= p/2).

__global__ void k_tables(TableArgs a) {
    const int y = blockIdx.y;
    const double p = a.field.p[y], pinv = a.field.pinv[y];
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < a.len;
         j += (uint64_t)gridDim.x * blockDim.x) {
        double w = 1.0, wi = 1.0;
        uint64_t bits = j;
        int m = 0;
        while (bits) {
            if (bits & 1) {
                w = reduce(mulmod(w, a.roots[y][m], p, pinv), p, pinv);
                wi = reduce(mulmod(wi, a.inv_roots[y][m], p, pinv), p, pinv);
            }
            bits >>= 1;
            ++m;
        }
        a.fwd[y * a.table_stride + j] = w;
        a.inv[y * a.table_stride + j] = wi;
    }
}

// ---------------------------------------------------------------------------
// Strided passes: the tile is LINES consecutive lines lo0..lo0+LINES-1 of
// first node H; element m of line lo sits at H*(T >> level_start) + m*S + lo.

template <int LOGL, int LINES, int NT, bool PACK>
__global__ __launch_bounds__(NT) void k_fwd_strided(PassArgs a) {
    constexpr int L = 1 << LOGL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    __shared__ double lds[LINES * LS];
    const int y = blockIdx.y;
    const double p = a.field.p[y], pinv = a.field.pinv[y];
    const uint64_t S = a.line_stride;
    const uint64_t tiles_per_node = S / LINES;
    const uint64_t H = blockIdx.x / tiles_per_node;
    const uint64_t lo0 = (blockIdx.x % tiles_per_node) * LINES;
    const uint64_t span = (uint64_t)1 << (a.log_t - a.level_start);
    const uint64_t base = H * span + lo0;
    double* row = a.data + (uint64_t)y * a.row_stride;
    if constexpr (PACK) {
        const Source& src = a.src_a;
        const double* pw = a.powers[y];
        for (int f = threadIdx.x; f < LINES * L; f += NT) {
            const int c = f % LINES, m = f / LINES;
            lds[c * LS + pidx(m)] = pack_coef(src, base + (uint64_t)m * S + c, pw, p, pinv);
        }
    } else {
        for (int f = threadIdx.x; f < LINES * L; f += NT) {
            const int c = f % LINES, m = f / LINES;
            lds[c * LS + pidx(m)] = row[base + (uint64_t)m * S + c];
        }
    }
    __syncthreads();
    const Tw tw{a.fwd + (uint64_t)y * a.table_stride, H, 0};
    fwd_rounds<LOGL, LINES, NT, 0>(lds, tw, p, pinv);
    for (int f = threadIdx.x; f < LINES * L; f += NT) {
        const int c = f % LINES, m = f / LINES;
        row[base + (uint64_t)m * S + c] = lds[c * LS + pidx(m)];
    }
}

// Inverse strided pass; CANON (level_start == 0) multiplies by T^-1 and stores
// canonical residues in [0, p).
template <int LOGL, int LINES, int NT, bool CANON>
__global__ __launch_bounds__(NT) void k_inv_strided(PassArgs a) {
    constexpr int L = 1 << LOGL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    __shared__ double lds[LINES * LS];
    const int y = blockIdx.y;
    const double p = a.field.p[y], pinv = a.field.pinv[y];
    const uint64_t S = a.line_stride;
    const uint64_t tiles_per_node = S / LINES;
    const uint64_t H = blockIdx.x / tiles_per_node;
    const uint64_t lo0 = (blockIdx.x % tiles_per_node) * LINES;
    const uint64_t span = (uint64_t)1 << (a.log_t - a.level_start);
    const uint64_t base = H * span + lo0;
    double* row = a.data + (uint64_t)y * a.row_stride;
    for (int f = threadIdx.x; f < LINES * L; f += NT) {
        const int c = f % LINES, m = f / LINES;
        lds[c * LS + pidx(m)] = row[base + (uint64_t)m * S + c];
    }
    __syncthreads();
    const Tw tw{a.inv + (uint64_t)y * a.table_stride, H, 0};
    inv_rounds<LOGL, LINES, NT, Lay<LOGL>::ROUNDS - 1>(lds, tw, p, pinv);
    const double tinv = a.tinv[y];
    for (int f = threadIdx.x; f < LINES * L; f += NT) {
        const int c = f % LINES, m = f / LINES;
        double v = lds[c * LS + pidx(m)];
        if constexpr (CANON) v = canonical(mulmod(v, tinv, p, pinv), p);
        row[base + (uint64_t)m * S + c] = v;
    }
}

// ---------------------------------------------------------------------------
// Row passes: rows H0..H0+LINES-1 of 2^LOGL contiguous points, the last
// levels of the transform.

template <int LOGL, int LINES, int NT>
__global__ __launch_bounds__(NT) void k_fwd_rows(PassArgs a) {
    constexpr int L = 1 << LOGL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    __shared__ double lds[LINES * LS];
    const int y = blockIdx.y;
    const double p = a.field.p[y], pinv = a.field.pinv[y];
    const uint64_t H0 = (uint64_t)blockIdx.x * LINES;
    double* row = a.data + (uint64_t)y * a.row_stride + H0 * L;
    for (int f = threadIdx.x; f < LINES * L; f += NT) {
        const int c = f / L, m = f % L;
        lds[c * LS + pidx(m)] = row[f];
    }
    __syncthreads();
    const Tw tw{a.fwd + (uint64_t)y * a.table_stride, H0, 1};
    fwd_rounds<LOGL, LINES, NT, 0>(lds, tw, p, pinv);
    for (int f = threadIdx.x; f < LINES * L; f += NT) {
        const int c = f / L, m = f % L;
        row[f] = lds[c * LS + pidx(m)];
    }
}

// Fused row pass: forward levels of one row, pointwise product with the other
// operand's spectrum (or a square), inverse levels, store. With PACK the row
// is the whole transform (level_start == 0): both operands are packed from
// limbs into LDS and the result is stored as canonical residues times T^-1.
template <int LOGL, int NT, bool PACK, bool SQUARE>
__global__ __launch_bounds__(NT) void k_fused(PassArgs a) {
    constexpr int L = 1 << LOGL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    constexpr int TILES = (PACK && !SQUARE) ? 2 : 1;
    __shared__ double lds[TILES * LS];
    const int y = blockIdx.y;
    const double p = a.field.p[y], pinv = a.field.pinv[y];
    const uint64_t H = blockIdx.x;
    double* row = a.data + (uint64_t)y * a.row_stride + H * L;
    if constexpr (PACK) {
        const double* pw = a.powers[y];
        for (int m = threadIdx.x; m < L; m += NT) {
            lds[pidx(m)] = pack_coef(a.src_a, (uint64_t)m, pw, p, pinv);
            if constexpr (!SQUARE) lds[LS + pidx(m)] = pack_coef(a.src_b, (uint64_t)m, pw, p, pinv);
        }
    } else {
        for (int m = threadIdx.x; m < L; m += NT) lds[pidx(m)] = row[m];
    }
    __syncthreads();
    const Tw twF{a.fwd + (uint64_t)y * a.table_stride, H, 0};
    const Tw twI{a.inv + (uint64_t)y * a.table_stride, H, 0};
    fwd_rounds<LOGL, 1, NT, 0>(lds, twF, p, pinv);
    if constexpr (TILES == 2) fwd_rounds<LOGL, 1, NT, 0>(lds + LS, twF, p, pinv);
    if constexpr (SQUARE) {
        for (int m = threadIdx.x; m < L; m += NT) {
            const double v = lds[pidx(m)];
            lds[pidx(m)] = mulmod(v, v, p, pinv);
        }
    } else if constexpr (PACK) {
        for (int m = threadIdx.x; m < L; m += NT)
            lds[pidx(m)] = mulmod(lds[pidx(m)], lds[LS + pidx(m)], p, pinv);
    } else {
        const double* other = a.other + (uint64_t)y * a.row_stride + H * L;
        for (int m = threadIdx.x; m < L; m += NT) lds[pidx(m)] = mulmod(lds[pidx(m)], other[m], p, pinv);
    }
    __syncthreads();
    inv_rounds<LOGL, 1, NT, Lay<LOGL>::ROUNDS - 1>(lds, twI, p, pinv);
    if constexpr (PACK) {
        const double tinv = a.tinv[y];
        for (int m = threadIdx.x; m < L; m += NT)
            row[m] = canonical(mulmod(lds[pidx(m)], tinv, p, pinv), p);
    } else {
        for (int m = threadIdx.x; m < L; m += NT) row[m] = lds[pidx(m)];
    }
}

// Product reconstruction, one block of kCrtBlockLimbs output limbs: every
// coefficient X_i touching the block is rebuilt once into LDS, every limb
// sums its windows of the shifted coefficients, and carries inside the block
// resolve by a scan of carry functions. Blocks publish their carry function
// for k_carry_scan / k_carry_fix (the block merge of src/limb/
// reconstruction.rs:211-247).
constexpr int kCrtLdsWords = 4096;

template <int NP>
__global__ __launch_bounds__(kCrtThreads) void k_crt_product(CrtArgs a) {
    constexpr int NW = CrtWords<NP>::value;
    __shared__ uint64_t xs[kCrtLdsWords];
    __shared__ CarryFn scan[kCrtThreads];
    __shared__ uint32_t ripple;
    const uint64_t L0 = (uint64_t)blockIdx.x * kCrtBlockLimbs;
    const uint64_t w = a.width;
    const uint64_t xb = a.value_bits;
    const uint64_t lo_bit = L0 * 64;
    const uint64_t i_lo = lo_bit > xb ? (lo_bit - xb) / w : 0;
    uint64_t i_hi = (lo_bit + 64 * (uint64_t)kCrtBlockLimbs - 1) / w + 1;
    if (i_hi > a.coefficients) i_hi = a.coefficients;
    const int count = i_hi > i_lo ? (int)(i_hi - i_lo) : 0;
    for (int idx = threadIdx.x; idx < count; idx += kCrtThreads) {
        uint64_t X[NW];
        garner<NP>(a, i_lo + idx, X, false);
#pragma unroll
        for (int q = 0; q < NW; ++q) xs[idx * NW + q] = X[q];
    }
    if (threadIdx.x == 0) ripple = 0;
    __syncthreads();

    uint64_t v[kCrtLimbsPerThread];
    CarryFn f[kCrtLimbsPerThread];
#pragma unroll
    for (int q = 0; q < kCrtLimbsPerThread; ++q) {
        const uint64_t j = L0 + (uint64_t)threadIdx.x * kCrtLimbsPerThread + q;
        uint64_t sum = 0;
        uint32_t carry = 0;
        if (j < a.out_len && count > 0) {
            const uint64_t jb = 64 * j;
            uint64_t ia = jb + 1 > xb ? (jb + 1 - xb + w - 1) / w : 0;
            if (ia < i_lo) ia = i_lo;
            uint64_t ib = (jb + 63) / w;
            if (ib >= i_hi) ib = i_hi - 1;
            for (uint64_t i = ia; i <= ib && ib < i_hi; ++i) {
                const uint64_t* X = xs + (i - i_lo) * NW;
                const int64_t o = (int64_t)jb - (int64_t)(i * w);
                uint64_t piece;
                if (o >= 0) {
                    const int qw = (int)(o >> 6), sh = (int)(o & 63);
                    const uint64_t lo = qw < NW ? X[qw] : 0;
                    const uint64_t hi = qw + 1 < NW ? X[qw + 1] : 0;
                    piece = sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
                } else {
                    piece = X[0] << (int)(-o);
                }
                sum += piece;
                carry += sum < piece;
            }
        }
        v[q] = sum;
        f[q] = CarryFn{carry, limb_thresh(sum)};
    }
    CarryFn agg = f[0];
#pragma unroll
    for (int q = 1; q < kCrtLimbsPerThread; ++q) agg = compose(agg, f[q]);
    scan[threadIdx.x] = agg;
    __syncthreads();
    for (int d = 1; d < kCrtThreads; d <<= 1) {
        CarryFn mine = scan[threadIdx.x];
        CarryFn prev = threadIdx.x >= (unsigned)d ? scan[threadIdx.x - d] : CarryFn{0, 0};
        __syncthreads();
        if (threadIdx.x >= (unsigned)d) scan[threadIdx.x] = compose(prev, mine);
        __syncthreads();
    }
    // Exclusive prefix: carries into this thread's first limb with zero
    // carry into the block.
    bool have = threadIdx.x > 0;
    CarryFn g = have ? scan[threadIdx.x - 1] : CarryFn{0, kInf};
    const uint32_t first_t = f[0].thresh;
    uint32_t my_ripple = 0;
#pragma unroll
    for (int q = 0; q < kCrtLimbsPerThread; ++q) {
        const uint32_t rel = threadIdx.x * kCrtLimbsPerThread + q;
        const uint64_t j = L0 + rel;
        const uint32_t cin = have ? g.base : 0u;
        if (have && g.thresh != kInf) my_ripple = rel;
        if (j < a.out_len) a.out[j] = v[q] + cin;
        g = have ? compose(g, f[q]) : f[q];
        have = true;
    }
    if (my_ripple) atomicMax(&ripple, my_ripple);
    __syncthreads();
    if (threadIdx.x == kCrtThreads - 1) {
        a.agg_base[blockIdx.x] = scan[kCrtThreads - 1].base;
        a.agg_thresh[blockIdx.x] = scan[kCrtThreads - 1].thresh;
    }
    if (threadIdx.x == 0) {
        a.first_thresh[blockIdx.x] = first_t;
        a.ripple_end[blockIdx.x] = ripple;
    }
}

// Carry into every block: a scan of the blocks' carry functions applied to a
// zero carry into block 0. One workgroup.
constexpr int kScanThreads = 1024;

__global__ __launch_bounds__(kScanThreads) void k_carry_scan(CrtArgs a) {
    __shared__ CarryFn scan[kScanThreads];
    __shared__ uint32_t has[kScanThreads];
    const uint32_t n = a.blocks;
    const uint32_t chunk = (n + kScanThreads - 1) / kScanThreads;
    const uint32_t first = threadIdx.x * chunk;
    const uint32_t end = first + chunk < n ? first + chunk : n;
    CarryFn agg{0, kInf};
    bool any = false;
    for (uint32_t b = first; b < end; ++b) {
        const CarryFn fb{a.agg_base[b], a.agg_thresh[b]};
        agg = any ? compose(agg, fb) : fb;
        any = true;
    }
    scan[threadIdx.x] = agg;
    has[threadIdx.x] = any;
    __syncthreads();
    for (int d = 1; d < kScanThreads; d <<= 1) {
        CarryFn mine = scan[threadIdx.x];
        uint32_t mine_has = has[threadIdx.x];
        CarryFn prev{0, kInf};
        uint32_t prev_has = 0;
        if (threadIdx.x >= (unsigned)d) {
            prev = scan[threadIdx.x - d];
            prev_has = has[threadIdx.x - d];
        }
        __syncthreads();
        if (threadIdx.x >= (unsigned)d && prev_has) {
            scan[threadIdx.x] = mine_has ? compose(prev, mine) : prev;
            has[threadIdx.x] = 1;
        }
        __syncthreads();
    }
    uint32_t c = 0;
    if (threadIdx.x > 0 && has[threadIdx.x - 1]) {
        const CarryFn e = scan[threadIdx.x - 1];
        c = e.base;  // applied to a zero carry into block 0 (thresholds >= 1)
    }
    for (uint32_t b = first; b < end; ++b) {
        a.carry_in[b] = c;
        c = a.agg_base[b] + (c >= a.agg_thresh[b] ? 1u : 0u);
    }
}

// Applies each block's carry: the first limb takes the carry, and limbs up to
// the recorded ripple end take one more when the carry reaches the first
// limb's threshold.
__global__ __launch_bounds__(kCrtThreads) void k_carry_fix(CrtArgs a) {
    const uint32_t cin = a.carry_in[blockIdx.x];
    if (cin == 0) return;
    const uint64_t L0 = (uint64_t)blockIdx.x * kCrtBlockLimbs;
    if (threadIdx.x == 0 && L0 < a.out_len) a.out[L0] += cin;
    if (cin >= a.first_thresh[blockIdx.x]) {
        const uint32_t end = a.ripple_end[blockIdx.x];
        for (uint32_t rel = 1 + threadIdx.x; rel <= end; rel += kCrtThreads) {
            const uint64_t j = L0 + rel;
            if (j < a.out_len) a.out[j] += 1;
        }
    }
}

// Disjoint convolution blocks: coefficient k occupies `stride` limbs, in two's
// complement when signed (src/limb/reconstruction.rs:255-261, 395-425).
template <int NP>
__global__ __launch_bounds__(256) void k_crt_disjoint(CrtArgs a) {
    constexpr int NW = CrtWords<NP>::value;
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.coefficients) return;
    uint64_t X[NW];
    garner<NP>(a, k, X, a.is_signed != 0);
    uint64_t* dst = a.out + k * a.stride;
    for (uint32_t q = 0; q < a.stride; ++q) dst[q] = q < (uint32_t)NW ? X[q] : (a.is_signed && (int64_t)X[NW - 1] < 0 ? ~0ull : 0ull);
}

// ---------------------------------------------------------------------------
// Host-side launchers.

#define NTT_CHECK_LAUNCH() (hipGetLastError() == hipSuccess)

template <int LOGL>
static constexpr int strided_lines() { return LOGL <= 10 ? 16 : (LOGL == 11 ? 8 : 4); }

// Threads of a strided tile: one radix-16 group per thread and round.
template <int LOGL>
static constexpr int strided_threads() {
    return strided_lines<LOGL>() * (1 << LOGL) / 16 >= 512 ? 512 : 256;
}

template <int LOGL>
static bool launch_fwd_strided_t(const PassArgs& a, int primes, bool pack, hipStream_t s) {
    constexpr int LINES = strided_lines<LOGL>();
    const uint64_t nodes = (uint64_t)1 << a.level_start;
    const uint64_t tiles = nodes * (a.line_stride / LINES);
    dim3 grid((unsigned)tiles, (unsigned)primes);
    if (pack)
        hipLaunchKernelGGL((k_fwd_strided<LOGL, LINES, strided_threads<LOGL>(), true>), grid, dim3(strided_threads<LOGL>()), 0, s, a);
    else
        hipLaunchKernelGGL((k_fwd_strided<LOGL, LINES, strided_threads<LOGL>(), false>), grid, dim3(strided_threads<LOGL>()), 0, s, a);
    return NTT_CHECK_LAUNCH();
}

template <int LOGL>
static bool launch_inv_strided_t(const PassArgs& a, int primes, bool canon, hipStream_t s) {
    constexpr int LINES = strided_lines<LOGL>();
    const uint64_t nodes = (uint64_t)1 << a.level_start;
    const uint64_t tiles = nodes * (a.line_stride / LINES);
    dim3 grid((unsigned)tiles, (unsigned)primes);
    if (canon)
        hipLaunchKernelGGL((k_inv_strided<LOGL, LINES, strided_threads<LOGL>(), true>), grid, dim3(strided_threads<LOGL>()), 0, s, a);
    else
        hipLaunchKernelGGL((k_inv_strided<LOGL, LINES, strided_threads<LOGL>(), false>), grid, dim3(strided_threads<LOGL>()), 0, s, a);
    return NTT_CHECK_LAUNCH();
}

template <int LOGL>
static bool launch_fwd_rows_t(const PassArgs& a, int primes, hipStream_t s) {
    constexpr int LINES = 1;
    const uint64_t rows = (uint64_t)1 << a.level_start;
    dim3 grid((unsigned)(rows / LINES), (unsigned)primes);
    hipLaunchKernelGGL((k_fwd_rows<LOGL, LINES, 256>), grid, dim3(256), 0, s, a);
    return NTT_CHECK_LAUNCH();
}

template <int LOGL>
static bool launch_fused_t(const PassArgs& a, int primes, bool pack, bool square, hipStream_t s) {
    const uint64_t rows = (uint64_t)1 << a.level_start;
    dim3 grid((unsigned)rows, (unsigned)primes);
    constexpr int NT = LOGL <= 8 ? 64 : 256;
    if (pack) {
        if (square)
            hipLaunchKernelGGL((k_fused<LOGL, NT, true, true>), grid, dim3(NT), 0, s, a);
        else
            hipLaunchKernelGGL((k_fused<LOGL, NT, true, false>), grid, dim3(NT), 0, s, a);
    } else {
        if (square)
            hipLaunchKernelGGL((k_fused<LOGL, NT, false, true>), grid, dim3(NT), 0, s, a);
        else
            hipLaunchKernelGGL((k_fused<LOGL, NT, false, false>), grid, dim3(NT), 0, s, a);
    }
    return NTT_CHECK_LAUNCH();
}

bool launch_fwd_strided(int logl, const PassArgs& a, int primes, bool pack, hipStream_t s) {
    switch (logl) {
        case 5: return launch_fwd_strided_t<5>(a, primes, pack, s);
        case 6: return launch_fwd_strided_t<6>(a, primes, pack, s);
        case 7: return launch_fwd_strided_t<7>(a, primes, pack, s);
        case 8: return launch_fwd_strided_t<8>(a, primes, pack, s);
        case 9: return launch_fwd_strided_t<9>(a, primes, pack, s);
        case 10: return launch_fwd_strided_t<10>(a, primes, pack, s);
        case 11: return launch_fwd_strided_t<11>(a, primes, pack, s);
        default: return false;
    }
}

bool launch_inv_strided(int logl, const PassArgs& a, int primes, bool canon, hipStream_t s) {
    switch (logl) {
        case 5: return launch_inv_strided_t<5>(a, primes, canon, s);
        case 6: return launch_inv_strided_t<6>(a, primes, canon, s);
        case 7: return launch_inv_strided_t<7>(a, primes, canon, s);
        case 8: return launch_inv_strided_t<8>(a, primes, canon, s);
        case 9: return launch_inv_strided_t<9>(a, primes, canon, s);
        case 10: return launch_inv_strided_t<10>(a, primes, canon, s);
        case 11: return launch_inv_strided_t<11>(a, primes, canon, s);
        default: return false;
    }
}

bool launch_fwd_rows(int logl, const PassArgs& a, int primes, hipStream_t s) {
    switch (logl) {
        case 5: return launch_fwd_rows_t<5>(a, primes, s);
        case 6: return launch_fwd_rows_t<6>(a, primes, s);
        case 7: return launch_fwd_rows_t<7>(a, primes, s);
        case 8: return launch_fwd_rows_t<8>(a, primes, s);
        case 9: return launch_fwd_rows_t<9>(a, primes, s);
        case 10: return launch_fwd_rows_t<10>(a, primes, s);
        case 11: return launch_fwd_rows_t<11>(a, primes, s);
        case 12: return launch_fwd_rows_t<12>(a, primes, s);
        default: return false;
    }
}

bool launch_fused(int logl, const PassArgs& a, int primes, bool pack, bool square, hipStream_t s) {
    switch (logl) {
        case 5: return launch_fused_t<5>(a, primes, pack, square, s);
        case 6: return launch_fused_t<6>(a, primes, pack, square, s);
        case 7: return launch_fused_t<7>(a, primes, pack, square, s);
        case 8: return launch_fused_t<8>(a, primes, pack, square, s);
        case 9: return launch_fused_t<9>(a, primes, pack, square, s);
        case 10: return launch_fused_t<10>(a, primes, pack, square, s);
        case 11: return launch_fused_t<11>(a, primes, pack, square, s);
        case 12: return launch_fused_t<12>(a, primes, pack, square, s);
        default: return false;
    }
}

bool launch_tables(const TableArgs& a, int primes, hipStream_t s) {
    const unsigned blocks = (unsigned)((a.len + 255) / 256 < 4096 ? (a.len + 255) / 256 : 4096);
    hipLaunchKernelGGL(k_tables, dim3(blocks, primes), dim3(256), 0, s, a);
    return NTT_CHECK_LAUNCH();
}

bool launch_crt_product(const CrtArgs& a, hipStream_t s) {
    dim3 grid(a.blocks);
    switch (a.primes) {
        case 1: hipLaunchKernelGGL(k_crt_product<1>, grid, dim3(kCrtThreads), 0, s, a); break;
        case 2: hipLaunchKernelGGL(k_crt_product<2>, grid, dim3(kCrtThreads), 0, s, a); break;
        case 3: hipLaunchKernelGGL(k_crt_product<3>, grid, dim3(kCrtThreads), 0, s, a); break;
        case 4: hipLaunchKernelGGL(k_crt_product<4>, grid, dim3(kCrtThreads), 0, s, a); break;
        case 5: hipLaunchKernelGGL(k_crt_product<5>, grid, dim3(kCrtThreads), 0, s, a); break;
        case 6: hipLaunchKernelGGL(k_crt_product<6>, grid, dim3(kCrtThreads), 0, s, a); break;
        case 7: hipLaunchKernelGGL(k_crt_product<7>, grid, dim3(kCrtThreads), 0, s, a); break;
        case 8: hipLaunchKernelGGL(k_crt_product<8>, grid, dim3(kCrtThreads), 0, s, a); break;
        default: return false;
    }
    if (!NTT_CHECK_LAUNCH()) return false;
    hipLaunchKernelGGL(k_carry_scan, dim3(1), dim3(kScanThreads), 0, s, a);
    if (!NTT_CHECK_LAUNCH()) return false;
    hipLaunchKernelGGL(k_carry_fix, dim3(a.blocks), dim3(kCrtThreads), 0, s, a);
    return NTT_CHECK_LAUNCH();
}

bool launch_crt_disjoint(const CrtArgs& a, hipStream_t s) {
    dim3 grid((unsigned)((a.coefficients + 255) / 256));
    switch (a.primes) {
        case 1: hipLaunchKernelGGL(k_crt_disjoint<1>, grid, dim3(256), 0, s, a); break;
        case 2: hipLaunchKernelGGL(k_crt_disjoint<2>, grid, dim3(256), 0, s, a); break;
        case 3: hipLaunchKernelGGL(k_crt_disjoint<3>, grid, dim3(256), 0, s, a); break;
        case 4: hipLaunchKernelGGL(k_crt_disjoint<4>, grid, dim3(256), 0, s, a); break;
        default: return false;
    }
    return NTT_CHECK_LAUNCH();
}

}  // namespace nttamd

namespace nttamd {

// Direct column sums for a convolution whose shorter operand has at most 24
// terms (src/convolve.rs:188-203, 303-351): block k receives
// sum a[i]*b[k-i] in `stride` limbs, two's complement when signed.
__global__ __launch_bounds__(256) void k_conv_columns(const uint64_t* a, uint64_t la, const uint64_t* b,
                                                      uint64_t lb, uint64_t* out, uint32_t stride,
                                                      uint32_t is_signed) {
    typedef unsigned __int128 u128;
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k + 1 >= la + lb) return;
    const uint64_t next = k + 1;
    const uint64_t first = next > lb ? next - lb : 0;
    const uint64_t end = next < la ? next : la;
    u128 low = 0;
    uint64_t high = 0;
    for (uint64_t i = first; i < end; ++i) {
        const uint64_t x = a[i], y = b[k - i];
        u128 product = (u128)x * y;
        if (is_signed) {
            uint64_t hi = (uint64_t)(product >> 64);
            if ((int64_t)x < 0) hi -= y;
            if ((int64_t)y < 0) hi -= x;
            product = ((u128)hi << 64) | (uint64_t)product;
        }
        const u128 sum = low + product;
        high += sum < low ? 1 : 0;
        if (is_signed) high -= (uint64_t)(product >> 127);
        low = sum;
    }
    const uint64_t words[3] = {(uint64_t)low, (uint64_t)(low >> 64), high};
    for (uint32_t q = 0; q < stride; ++q) out[k * stride + q] = q < 3 ? words[q] : 0;
}

bool launch_conv_columns(const uint64_t* a, uint64_t la, const uint64_t* b, uint64_t lb, uint64_t* out,
                         uint32_t stride, uint32_t is_signed, hipStream_t s) {
    const uint64_t n = la + lb - 1;
    hipLaunchKernelGGL(k_conv_columns, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, la, b,
                       lb, out, stride, is_signed);
    return hipGetLastError() == hipSuccess;
}

}  // namespace nttamd
