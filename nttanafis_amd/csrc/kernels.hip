// Root tables, composite-row recombination, CRT reconstruction, carries, schoolbook and convolution columns.
#include "passes.h"

namespace nttamd {

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

// Recombination of a composite row, in place. With c = c_lo + x^N c_hi the
// product (deg c_lo < N, deg c_hi < tail*N/4), the cyclic part holds
// u = c mod (x^N - 1) = c_lo + c_hi and a tail factor f of x^N + 1 holds
// v = c mod f = (u - 2 c_hi) mod f, since x^N = -1 mod f. So
// c_hi mod f = (u mod f - v) / 2, which is c_hi itself when f has degree
// at least tail*N/4 (one tail), and c_lo = u - c_hi. Two tails (7N/4):
// g1 = c_hi mod (x^(N/2) - W1), g2 = c_hi mod (x^(N/4) - W3) and
// c_hi = A0 + x^(N/4) A1 + x^(N/2) B give A1 = g1[N/4..], A0 + W1 B = g1[..N/4]
// and A0 + W3 A1 - W1 B = g2 (W3^2 = -W1), so A0 = (g2 + g1[..N/4] - W3 A1)/2
// and B = W1 (A0 - g1[..N/4]) (W1^-1 = -W1). Thread k handles coefficient k
// of each quarter (of each half for tail 2); every value read is canonical,
// sums stay below 8p before each reduction and products below p^2.
__global__ __launch_bounds__(256) void k_recombine(RecombineArgs a) {
    const int y = blockIdx.y;
    const double p = a.field.p[y], pinv = a.field.pinv[y];
    const uint64_t N = a.n, Q = N >> 2;
    const uint64_t count = a.tail == 2 ? N >> 1 : Q;
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    double* r = a.rows + (uint64_t)y * a.row_stride;
    const double half = a.half[y], w1 = a.w1[y];
    auto fix = [&](double x) { return canonical(reduce(x, p, pinv), p); };
    if (a.tail == 2) {
        const double u0 = r[k], u1 = r[k + 2 * Q], v = r[N + k];
        const double h = mulmod(u0 + mulmod(u1, w1, p, pinv) - v, half, p, pinv);
        r[k] = fix(u0 - h);
        r[N + k] = fix(h);
        return;
    }
    const double u0 = r[k], u1 = r[k + Q], u2 = r[k + 2 * Q], u3 = r[k + 3 * Q];
    if (a.tail == 1) {
        const double* s = a.s[y];
        const double x = u0 + mulmod(u1, s[0], p, pinv) + mulmod(u2, s[1], p, pinv) + mulmod(u3, s[2], p, pinv) -
                         r[N + k];
        const double h = mulmod(reduce(x, p, pinv), half, p, pinv);
        r[k] = fix(u0 - h);
        r[N + k] = fix(h);
        return;
    }
    const double* t = a.t[y];
    const double v1a = r[N + k], v1b = r[N + Q + k], v2 = r[N + 2 * Q + k];
    const double g1a = mulmod(reduce(u0 + mulmod(u2, w1, p, pinv) - v1a, p, pinv), half, p, pinv);
    const double g1b = mulmod(reduce(u1 + mulmod(u3, w1, p, pinv) - v1b, p, pinv), half, p, pinv);
    const double g2 = mulmod(reduce(u0 + mulmod(u1, t[0], p, pinv) + mulmod(u2, t[1], p, pinv) +
                                        mulmod(u3, t[2], p, pinv) - v2,
                                    p, pinv),
                             half, p, pinv);
    const double A1 = g1b;
    const double A0 = mulmod(reduce(g2 + g1a - mulmod(t[0], A1, p, pinv), p, pinv), half, p, pinv);
    const double B = mulmod(A0 - g1a, w1, p, pinv);
    r[k] = fix(u0 - A0);
    r[k + Q] = fix(u1 - A1);
    r[k + 2 * Q] = fix(u2 - B);
    r[N + k] = fix(A0);
    r[N + Q + k] = fix(A1);
    r[N + 2 * Q + k] = fix(B);
}

// Product reconstruction, one block of kCrtBlockLimbs output limbs: every
// coefficient X_i touching the block is rebuilt once, shifted to its bit
// offset and added word by word into LDS limb accumulators with atomic adds
// that return the previous value, so each limb's sum (mod 2^64) and its
// number of wrap-arounds (the carries it sends on) come out exact in any
// order. Carries inside the block then resolve by a scan of carry functions
// (per wave with lane shuffles, then across the block's waves). Blocks
// publish their carry function for k_carry_apply (the block merge of
// src/limb/reconstruction.rs:211-247). A workgroup of NT threads covers
// 4*NT limbs (kCrtBlockLimbs for the 256-thread form); the 512-thread form
// reconstructs a whole output of up to 2048 limbs in one block, which needs
// no carry pass after it.
template <int NT>
struct CrtShape {
    static constexpr int BLOCK = NT * kCrtLimbsPerThread;  // limbs per block
    static constexpr int LDS_WORDS = 16 * NT;               // coefficient words per block
};

__device__ __forceinline__ CarryFn shfl_up_fn(CarryFn f, int d) {
    CarryFn r;
    r.base = __shfl_up(f.base, d, 64);
    r.thresh = __shfl_up(f.thresh, d, 64);
    return r;
}

// Reconstruction parameters of one workgroup: the single operation's, or
// those of the batched product its item names.
// Its pointers are typed as global memory (NTT_GLOBAL): the reconstruction's
// atomic accumulation waits on the LDS counter, which generic accesses
// would hold up as well.
struct CrtJob {
    const NTT_GLOBAL double* residues;
    uint64_t row_stride, coefficients, out_len;
    NTT_GLOBAL uint64_t* out;
    uint32_t width, magic, value_bits, block, blocks;
    NTT_GLOBAL uint32_t *agg_base, *agg_thresh, *first_thresh, *ripple_end;
    uint32_t split_blocks;  // 0: no split
    uint64_t split_digits, block_stride;
    const NTT_GLOBAL uint64_t *op_a, *op_b;  // schoolbook operands
    uint64_t len_a, len_b;
};


__device__ __forceinline__ CrtJob crt_job(const CrtArgs& a) {
    CrtJob j;
    if (a.batch == nullptr) {
        j.residues = global_ptr(a.residues);
        j.row_stride = a.row_stride;
        j.coefficients = a.coefficients;
        j.out_len = a.out_len;
        j.out = global_ptr(a.out);
        j.width = a.width;
        j.magic = a.width_magic;
        j.value_bits = a.value_bits;
        j.block = blockIdx.x;
        j.blocks = a.blocks;
        j.agg_base = global_ptr(a.agg_base);
        j.agg_thresh = global_ptr(a.agg_thresh);
        j.first_thresh = global_ptr(a.first_thresh);
        j.ripple_end = global_ptr(a.ripple_end);
        j.split_blocks = a.split_blocks;
        j.split_digits = a.split_digits;
        j.block_stride = a.block_stride;
        j.op_a = global_ptr(a.op_a);
        j.op_b = global_ptr(a.op_b);
        j.len_a = a.len_a;
        j.len_b = a.len_b;
        return j;
    }
    const BatchItem it = a.items[blockIdx.x];
    const BatchProduct& P = a.batch[it.product];
    j.residues = global_ptr((P.square || P.split_blocks) ? P.spec_a : P.spec_b);
    j.split_blocks = P.split_blocks;
    j.split_digits = P.block_digits;
    j.block_stride = P.block_stride;
    j.row_stride = P.row_stride;
    j.coefficients = P.coefficients;
    j.out_len = P.out_len;
    j.out = global_ptr(P.out);
    j.width = P.width;
    j.magic = P.width_magic;
    j.value_bits = P.value_bits;
    j.block = it.sub;
    j.blocks = P.blocks;
    j.agg_base = global_ptr(P.carry);
    j.agg_thresh = global_ptr(P.carry + P.blocks);
    j.first_thresh = global_ptr(P.carry + 2 * (uint64_t)P.blocks);
    j.ripple_end = global_ptr(P.carry + 3 * (uint64_t)P.blocks);
    j.op_a = global_ptr(P.src_a.limbs);
    j.op_b = global_ptr(P.src_b.limbs);
    j.len_a = P.src_a.limb_len;
    j.len_b = P.src_b.limb_len;
    return j;
}

// Second half of a reconstruction block: acc[j] holds limb j's sum mod 2^64
// and wraps[j] the carries it sends on. Resolves the carries inside the
// block, stores the limbs and publishes the block's carry function.
template <int NT>
__device__ __forceinline__ void block_finish(const CrtJob& J, unsigned long long* acc, uint32_t* wraps,
                                             CarryFn* wave_total, uint32_t* ripple_s) {
    const uint64_t L0 = (uint64_t)J.block * CrtShape<NT>::BLOCK;
    uint64_t v[kCrtLimbsPerThread];
    CarryFn f[kCrtLimbsPerThread];
#pragma unroll
    for (int q = 0; q < kCrtLimbsPerThread; ++q) {
        const int jl = threadIdx.x * kCrtLimbsPerThread + q;
        const uint64_t sum = acc[jl];
        v[q] = sum;
        f[q] = CarryFn{wraps[jl], limb_thresh(sum)};
    }
    // Inclusive scan of the threads' carry functions within each wave.
    CarryFn agg = f[0];
#pragma unroll
    for (int q = 1; q < kCrtLimbsPerThread; ++q) agg = compose(agg, f[q]);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const CarryFn prev = shfl_up_fn(agg, d);
        if (lane >= d) agg = compose(prev, agg);
    }
    if (lane == 63) wave_total[wave] = agg;
    const CarryFn lane_excl = shfl_up_fn(agg, 1);
    __syncthreads();
    // Exclusive prefix of this thread: earlier waves, then earlier lanes.
    bool have = false;
    CarryFn g{0, kInf};
    for (int u = 0; u < wave; ++u) {
        g = have ? compose(g, wave_total[u]) : wave_total[u];
        have = true;
    }
    if (lane > 0) {
        g = have ? compose(g, lane_excl) : lane_excl;
        have = true;
    }
    const uint32_t first_t = f[0].thresh;
    uint32_t my_ripple = 0;
    uint64_t outv[kCrtLimbsPerThread];
#pragma unroll
    for (int q = 0; q < kCrtLimbsPerThread; ++q) {
        const uint32_t rel = threadIdx.x * kCrtLimbsPerThread + q;
        if (have && g.thresh != kInf) my_ripple = rel;
        outv[q] = v[q] + (have ? g.base : 0u);
        g = have ? compose(g, f[q]) : f[q];
        have = true;
    }
    if (my_ripple) atomicMax(ripple_s, my_ripple);
    __syncthreads();  // every accumulator read is done; reuse them to coalesce the stores
#pragma unroll
    for (int q = 0; q < kCrtLimbsPerThread; ++q) acc[threadIdx.x * kCrtLimbsPerThread + q] = outv[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kCrtLimbsPerThread; ++q) {
        const uint64_t j = L0 + threadIdx.x + q * NT;
        if (j < J.out_len) J.out[j] = acc[threadIdx.x + q * NT];
    }
    if (threadIdx.x == NT - 1) {
        J.agg_base[J.block] = g.base;
        J.agg_thresh[J.block] = g.thresh;
    }
    if (threadIdx.x == 0) {
        J.first_thresh[J.block] = first_t;
        J.ripple_end[J.block] = *ripple_s;
    }
}

template <int NP, int NT>
__global__ __launch_bounds__(NT) void k_crt_product(CrtArgs a) {
    constexpr int BLOCK = CrtShape<NT>::BLOCK;
    constexpr int NW = CrtWords<NP>::value;
    constexpr int RMAX = (CrtShape<NT>::LDS_WORDS / NW + NT - 1) / NT;
    constexpr int WAVES = NT / 64;
    __shared__ unsigned long long acc[BLOCK];
    __shared__ uint32_t wraps[BLOCK];
    __shared__ CarryFn wave_total[WAVES];
    __shared__ uint32_t ripple;
    const CrtJob J = crt_job(a);
    const uint64_t L0 = (uint64_t)J.block * BLOCK;
    const uint64_t w = J.width;
    const uint64_t xb = J.value_bits;
    const uint64_t lo_bit = L0 * 64;
    const uint64_t i_lo = lo_bit > xb ? (lo_bit - xb) / w : 0;
    uint64_t i_hi = (lo_bit + 64 * (uint64_t)BLOCK - 1) / w + 1;
    if (i_hi > J.coefficients) i_hi = J.coefficients;
    const int count = i_hi > i_lo ? (int)(i_hi - i_lo) : 0;
    // Residues of this thread's coefficients, every load issued first.
    double r[RMAX][NP];
    if (J.split_blocks == 0) {
#pragma unroll
        for (int q = 0; q < RMAX; ++q) {
            const int idx = threadIdx.x + q * NT;
#pragma unroll
            for (int k = 0; k < NP; ++k)
                r[q][k] = idx < count ? J.residues[(uint64_t)k * J.row_stride + i_lo + idx] : 0.0;
        }
    } else {
        // Coefficient i of a split product: the rows of block z1 = i / D
        // (the last block for i past it) at i - z1*D, plus those of block
        // z1 - 1 at i - z1*D + D while that lies in its row; no third block
        // reaches i, since blocks are at least as long as the shorter operand.
        const uint64_t D = J.split_digits;
#pragma unroll
        for (int q = 0; q < RMAX; ++q) {
            const int idx = threadIdx.x + q * NT;
            const uint64_t i = i_lo + idx;
            uint64_t z1 = i / D;
            if (z1 >= J.split_blocks) z1 = J.split_blocks - 1;
            const uint64_t k1 = i - z1 * D;
            const bool two = z1 > 0 && k1 + D < J.row_stride;
            const NTT_GLOBAL double* r1 = J.residues + z1 * J.block_stride + k1;
            const NTT_GLOBAL double* r0 = two ? r1 - J.block_stride + D : r1;
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                double v = 0.0;
                if (idx < count) {
                    v = r1[(uint64_t)k * J.row_stride];
                    if (two) {
                        v += r0[(uint64_t)k * J.row_stride];
                        if (v >= a.field.p[k]) v -= a.field.p[k];
                    }
                }
                r[q][k] = v;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kCrtLimbsPerThread; ++q) {
        acc[threadIdx.x + q * NT] = 0;
        wraps[threadIdx.x + q * NT] = 0;
    }
    if (threadIdx.x == 0) ripple = 0;
    __syncthreads();
    // Coefficient i starts at bit i*w - 64*L0 of the block, which lies in
    // [-xb, 64*BLOCK); offsets from bit i_lo*w keep the arithmetic
    // in 32 bits. Word m of X_i << s lands in limb qrel + m.
    const uint32_t wr = (uint32_t)w;
    const uint32_t rel0 = (uint32_t)(lo_bit - i_lo * w);  // bit of limb L0 from coefficient i_lo
#pragma unroll
    for (int q = 0; q < RMAX; ++q) {
        const int idx = threadIdx.x + q * NT;
        if (idx < count) {
            uint64_t X[NW];
            garner_r<NP>(a, r[q], X, false);
            const int32_t bit = (int32_t)((uint32_t)idx * wr) - (int32_t)rel0;
            const int32_t qrel = bit >> 6;  // arithmetic shift: floor division
            const uint32_t sh = (uint32_t)bit & 63u;
            uint64_t prev = 0;
#pragma unroll
            for (int m = 0; m <= NW; ++m) {
                const uint64_t cur = m < NW ? X[m] : 0;
                const uint64_t word = sh ? (cur << sh) | (prev >> (64 - sh)) : cur;
                prev = cur;
                const int32_t j = qrel + m;
                if (word != 0 && j >= 0 && j < BLOCK) {
                    const unsigned long long old = atomicAdd(&acc[j], (unsigned long long)word);
                    if (old + word < old) atomicAdd(&wraps[j], 1u);
                }
            }
        }
    }
    __syncthreads();
    block_finish<NT>(J, acc, wraps, wave_total, &ripple);
}

// Schoolbook product, no transform, for a shorter operand of at most
// kSchoolMaxShort limbs. One workgroup per block of kSchoolBlockLimbs output
// limbs, so that even a small product spreads over many CUs. Limbs are cut
// into pieces of 22, 22 and 20 bits held as doubles: a piece product is
// below 2^44 and exact, and the sums of one chunk of kSchoolChunk terms stay
// below 2^53, so nine FMAs per limb product accumulate exactly (the FP64
// pipes issue them at full rate). Column c = i + j of the limb products
// feeds limbs c, c + 1 and c + 2, so the block takes the columns L0 - 2 ..
// L0 + kSchoolBlockLimbs - 1; a work unit is four consecutive columns times
// one chunk of terms, whose window of the longer operand slides by one limb
// per term. Each unit's column values, three words each, go into the limb
// accumulators with LDS atomics that count wrap-arounds; one wave then
// resolves the carries inside the block and publishes the block's carry
// function for k_carry_apply.
constexpr int kSchoolChunk = 16;
constexpr int kSchoolCols = 4;
constexpr int kSchoolThreads = 256;

__device__ __forceinline__ void lds_add(unsigned long long* acc, uint32_t* wraps, int r, uint64_t w) {
    if (w != 0 && r >= 0 && r < kSchoolBlockLimbs) {
        const unsigned long long old = atomicAdd(&acc[r], (unsigned long long)w);
        if (old + w < old) atomicAdd(&wraps[r], 1u);
    }
}

// Pieces of a limb: bits 0-21, 22-43 and 44-63.
__device__ __forceinline__ void limb_pieces(uint64_t x, double* p0, double* p1, double* p2) {
    *p0 = (double)(uint32_t)(x & 0x3FFFFFu);
    *p1 = (double)(uint32_t)((x >> 22) & 0x3FFFFFu);
    *p2 = (double)(uint32_t)(x >> 44);
}

__global__ __launch_bounds__(kSchoolThreads) void k_school(CrtArgs a) {
    constexpr int BL = kSchoolBlockLimbs;
    constexpr int NC = BL + 2;                                    // columns L0-2 .. L0+BL-1
    constexpr int GROUPS = (NC + kSchoolCols - 1) / kSchoolCols;  // column groups
    constexpr int PAD = kSchoolChunk;                             // window slack below
    constexpr int WIN = PAD + GROUPS * kSchoolCols + kSchoolMaxShort;
    static_assert(kSchoolMaxShort % kSchoolChunk == 0 && kSchoolChunk % kSchoolCols == 0, "chunking");
    __shared__ unsigned long long acc[BL];
    __shared__ uint32_t wraps[BL];
    __shared__ double lp0[WIN], lp1[WIN], lp2[WIN];
    __shared__ double sp0[kSchoolMaxShort], sp1[kSchoolMaxShort], sp2[kSchoolMaxShort];
    const CrtJob J = crt_job(a);
    const bool swap = J.len_b > J.len_a;
    const NTT_GLOBAL uint64_t* pl = swap ? J.op_b : J.op_a;
    const NTT_GLOBAL uint64_t* ps = swap ? J.op_a : J.op_b;
    const int64_t ll = (int64_t)(swap ? J.len_b : J.len_a);
    const int ls = (int)(swap ? J.len_a : J.len_b);  // 1 .. kSchoolMaxShort (host-checked)
    const int nch = (ls + kSchoolChunk - 1) / kSchoolChunk;
    const int64_t L0 = (int64_t)J.block * BL;
    const int64_t c0 = L0 - 2;
    // Window: limb c - j of column c = c0 + cr sits at index PAD + cr + ls - 1 - j
    // for cr < GROUPS*kSchoolCols and 0 <= j < nch*kSchoolChunk; limbs outside
    // the operand are zero, and terms past ls multiply zero pieces of the
    // shorter operand.
    const int64_t w0 = c0 - (ls - 1) - PAD;
    const int W = PAD + GROUPS * kSchoolCols + ls - 1;
    for (int t = threadIdx.x; t < W; t += kSchoolThreads) {
        const int64_t i = w0 + t;
        limb_pieces((i >= 0 && i < ll) ? pl[i] : 0, &lp0[t], &lp1[t], &lp2[t]);
    }
    for (int t = threadIdx.x; t < nch * kSchoolChunk; t += kSchoolThreads)
        limb_pieces(t < ls ? ps[t] : 0, &sp0[t], &sp1[t], &sp2[t]);
    for (int t = threadIdx.x; t < BL; t += kSchoolThreads) {
        acc[t] = 0;
        wraps[t] = 0;
    }
    __syncthreads();
    for (int u = threadIdx.x; u < GROUPS * nch; u += kSchoolThreads) {
        const int g = u % GROUPS, ch = u / GROUPS;
        const int cr = g * kSchoolCols;  // first column of the unit, relative to c0
        const int jb = ch * kSchoolChunk;
        // The unit reads limbs c - j in [c0+cr-jb-CH+1, c0+cr+3-jb]: skip it when
        // they all lie outside the longer operand.
        const int64_t i_hi = c0 + cr + kSchoolCols - 1 - jb;
        const int64_t i_lo = c0 + cr - (jb + kSchoolChunk - 1);
        if (i_hi < 0 || i_lo >= ll) continue;
        double s[kSchoolCols][5];
#pragma unroll
        for (int q = 0; q < kSchoolCols; ++q)
#pragma unroll
            for (int k = 0; k < 5; ++k) s[q][k] = 0.0;
        // Ring of the four limbs in use: at term jb + t, column q takes slot
        // (q - t) mod 4, and the new limb of column 0 replaces the one column 3
        // no longer needs.
        const int base = PAD + cr + ls - 1 - jb;  // window index of limb c0 + cr - jb
        double r0[kSchoolCols], r1[kSchoolCols], r2[kSchoolCols];
#pragma unroll
        for (int q = 1; q < kSchoolCols; ++q) {
            r0[q] = lp0[base + q];
            r1[q] = lp1[base + q];
            r2[q] = lp2[base + q];
        }
#pragma unroll 1
        for (int t4 = 0; t4 < kSchoolChunk; t4 += kSchoolCols) {
#pragma unroll
            for (int tt = 0; tt < kSchoolCols; ++tt) {
                const int t = t4 + tt;
                const int in = (kSchoolCols - tt) & (kSchoolCols - 1);
                r0[in] = lp0[base - t];
                r1[in] = lp1[base - t];
                r2[in] = lp2[base - t];
                const double y0 = sp0[jb + t], y1 = sp1[jb + t], y2 = sp2[jb + t];
#pragma unroll
                for (int q = 0; q < kSchoolCols; ++q) {
                    const int sl = (q - tt + kSchoolCols) & (kSchoolCols - 1);
                    const double x0 = r0[sl], x1 = r1[sl], x2 = r2[sl];
                    s[q][0] = fma(x0, y0, s[q][0]);
                    s[q][1] = fma(x0, y1, s[q][1]);
                    s[q][1] = fma(x1, y0, s[q][1]);
                    s[q][2] = fma(x0, y2, s[q][2]);
                    s[q][2] = fma(x1, y1, s[q][2]);
                    s[q][2] = fma(x2, y0, s[q][2]);
                    s[q][3] = fma(x1, y2, s[q][3]);
                    s[q][3] = fma(x2, y1, s[q][3]);
                    s[q][4] = fma(x2, y2, s[q][4]);
                }
            }
        }
        // Column value sum_k s[q][k] * 2^(22k) in three words.
#pragma unroll
        for (int q = 0; q < kSchoolCols; ++q) {
            typedef unsigned __int128 u128;
            const uint64_t a0 = (uint64_t)s[q][0], a1 = (uint64_t)s[q][1], a2 = (uint64_t)s[q][2];
            const uint64_t a3 = (uint64_t)s[q][3], a4 = (uint64_t)s[q][4];
            const u128 t = (u128)a0 + ((u128)a1 << 22) + ((u128)a2 << 44);
            const u128 w = ((u128)a3 << 2) + ((u128)a4 << 24) + (uint64_t)(t >> 64);
            const int r = cr + q - 2;  // limb c relative to L0
            lds_add(acc, wraps, r, (uint64_t)t);
            lds_add(acc, wraps, r + 1, (uint64_t)w);
            lds_add(acc, wraps, r + 2, (uint64_t)(w >> 64));
        }
    }
    __syncthreads();
    // One wave resolves the block's carries: lane = limb.
    if (threadIdx.x < 64) {
        static_assert(BL == 64, "one limb per lane");
        const int lane = threadIdx.x;
        const uint64_t v = acc[lane];
        const CarryFn f{wraps[lane], limb_thresh(v)};
        CarryFn agg = f;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const CarryFn prev = shfl_up_fn(agg, d);
            if (lane >= d) agg = compose(prev, agg);
        }
        const CarryFn excl = shfl_up_fn(agg, 1);
        const uint64_t outv = v + (lane > 0 ? excl.base : 0u);
        uint32_t my_ripple = (lane > 0 && excl.thresh != kInf) ? (uint32_t)lane : 0u;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            const uint32_t o = __shfl_xor(my_ripple, d, 64);
            my_ripple = o > my_ripple ? o : my_ripple;
        }
        const uint64_t jl = (uint64_t)L0 + lane;
        if (jl < J.out_len) J.out[jl] = outv;
        if (lane == 63) {
            J.agg_base[J.block] = agg.base;
            J.agg_thresh[J.block] = agg.thresh;
        }
        if (lane == 0) {
            J.first_thresh[J.block] = f.thresh;
            J.ripple_end[J.block] = my_ripple;
        }
    }
}

// Carry into every block and its application. Block b's carry in is
// F_{b-1}(F_{b-2}(... F_0(0))) for the published carry functions F (the block
// merge of src/limb/reconstruction.rs:211-247).
//
// An output of more than kCarryDirectBlocks blocks first runs k_carry_scan,
// one workgroup per chunk of kCarryChunk blocks (one block per thread): it
// replaces each block's function by the composition of the functions before
// it in its chunk (the first block of a chunk gets the identity, marked by a
// zero threshold, which no carry function has) and publishes the chunk's
// composed function. The apply kernel then composes the functions of the
// chunks before its own (at most a few per lane of its wave) and applies its
// block's prefix: O(blocks) work for any carry pattern, all-ones products
// included. A shorter output skips the scan (and its launch): block b
// composes the functions of blocks 0..b-1 itself, each thread a contiguous
// slice of at most kCarryDirectBlocks / NT of them, the slices in order by a
// scan over each wave and the wave results in order.
//
// Then the first limb of the block takes the carry, and the limbs up to the
// recorded ripple end take one more when the carry reaches the first limb's
// threshold.
// Carry bookkeeping of the product a scan or apply workgroup works on.
struct CarryView {
    uint32_t* words;  // carry_words(blocks) words
    uint32_t blocks;
};

__global__ __launch_bounds__(kCarryChunk) void k_carry_scan(CrtArgs a) {
    constexpr int WAVES = kCarryChunk / 64;
    __shared__ CarryFn wave_total[WAVES];
    CarryView v;
    uint32_t chunk;
    if (a.batch == nullptr) {
        v = CarryView{a.agg_base, a.blocks};
        chunk = blockIdx.x;
    } else {
        const BatchItem it = a.scan_items[blockIdx.x];
        const BatchProduct& P = a.batch[it.product];
        v = CarryView{P.carry, P.blocks};
        chunk = it.sub;
    }
    uint32_t* base = v.words;
    uint32_t* thresh = v.words + v.blocks;
    const uint32_t first = chunk * (uint32_t)kCarryChunk;
    const uint32_t b = first + threadIdx.x;
    const bool have = b < v.blocks;
    // The threads of blocks past the output are a suffix of the workgroup, so
    // their placeholder only reaches threads after them.
    const CarryFn f = have ? CarryFn{base[b], thresh[b]} : CarryFn{0u, kInf};
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    CarryFn inc = f;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const CarryFn prev = shfl_up_fn(inc, d);
        if (lane >= d) inc = compose(prev, inc);
    }
    if (lane == 63) wave_total[wave] = inc;
    const CarryFn lane_excl = shfl_up_fn(inc, 1);
    __syncthreads();
    bool pre = false;
    CarryFn g{0u, kInf};
    for (int u = 0; u < wave; ++u) {
        g = pre ? compose(g, wave_total[u]) : wave_total[u];
        pre = true;
    }
    if (lane > 0) {
        g = pre ? compose(g, lane_excl) : lane_excl;
        pre = true;
    }
    if (!have) return;
    base[b] = pre ? g.base : 0u;
    thresh[b] = pre ? g.thresh : 0u;  // 0: the identity (first block of the chunk)
    const uint32_t count = v.blocks - first < (uint32_t)kCarryChunk ? v.blocks - first : (uint32_t)kCarryChunk;
    if (threadIdx.x == count - 1) {
        const CarryFn whole = pre ? compose(g, f) : f;
        const uint32_t chunks = carry_chunks(v.blocks);
        v.words[4ull * v.blocks + chunk] = whole.base;
        v.words[4ull * v.blocks + chunks + chunk] = whole.thresh;
    }
}

// Threads of an apply workgroup: they find the carry into the block
// together (one wave of them after a scan), then add the ripple (up to
// BL - 1 limbs).
template <int BL>
__host__ __device__ constexpr int apply_threads() {
    return 256;
}

// The apply workgroup of a single operation whose block functions were
// scanned takes BPW consecutive blocks (one chunk never splits them: BPW
// divides kCarryChunk): the 2^24-limb square's 32,768 blocks take 16 us in
// groups of four against 31 us one by one (groups of eight: 14 us, but an
// all-ones product's ripples then queue up behind one another). Shorter
// outputs and batch items take one block per workgroup.
constexpr int kApplyBlocks = 4;
static_assert(kCarryChunk % kApplyBlocks == 0, "a group of blocks lies in one chunk");

template <int BL, int BPW>
__global__ __launch_bounds__(apply_threads<BL>()) void k_carry_apply(CrtArgs a) {
    constexpr int NT = apply_threads<BL>();
    __shared__ CarryFn wave_fn[NT / 64];
    __shared__ uint32_t cin_s[BPW];
    const CrtJob J = crt_job(a);
    const uint32_t b0 = BPW == 1 ? J.block : (uint32_t)blockIdx.x * BPW;
    const uint32_t bn = b0 + BPW < J.blocks ? b0 + BPW : J.blocks;  // blocks b0 .. bn-1
    const int lane = threadIdx.x;  // wave 0 only in the scanned branch
    if (carry_scanned(J.blocks)) {
        if (threadIdx.x < 64) {
            // Carry into this group's chunk: the chunk functions before it,
            // composed in order, lane l taking a contiguous slice (the lanes
            // with a slice are a prefix of the wave).
            const uint32_t j = b0 / (uint32_t)kCarryChunk;
            uint32_t cin_chunk = 0;
            if (j > 0) {
                const uint32_t chunks = carry_chunks(J.blocks);
                const NTT_GLOBAL uint32_t* cb = J.agg_base + 4ull * J.blocks;
                const NTT_GLOBAL uint32_t* ct = cb + chunks;
                const uint32_t per = (j + 63) / 64;
                const uint32_t s0 = lane * per;
                const uint32_t s1 = s0 + per < j ? s0 + per : j;
                CarryFn agg{0u, kInf};
                for (uint32_t q = s0; q < s1; ++q) {
                    const CarryFn fq{cb[q], ct[q]};
                    agg = q == s0 ? fq : compose(agg, fq);
                }
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const CarryFn prev = shfl_up_fn(agg, d);
                    if (lane >= d) agg = compose(prev, agg);
                }
                const int last = (int)((j + per - 1) / per) - 1;
                cin_chunk = __shfl(agg.base, last, 64);  // the composition applied to 0
            }
            // Each block's prefix within its chunk, from k_carry_scan (a zero
            // threshold marks the first block of the chunk: the identity).
            if (lane < BPW && b0 + lane < bn) {
                const uint32_t pb = J.agg_base[b0 + lane], pt = J.agg_thresh[b0 + lane];
                cin_s[lane] = pt == 0 ? cin_chunk : pb + (cin_chunk >= pt ? 1u : 0u);
            }
        }
    } else {
        // b0 <= kCarryDirectBlocks: thread t composes the functions of blocks
        // [t*per, t*per + per) below b0 (the threads with a slice are a
        // prefix of the workgroup), each wave scans its threads' functions
        // in order, and thread 0 applies the waves' results in order to the
        // zero carry into block 0, then the group's own functions in turn.
        // Every load of a slice is issued before the first composition.
        constexpr int PER_MAX = ((int)kCarryDirectBlocks + NT - 1) / NT;
        const uint32_t per = (b0 + NT - 1) / NT;  // 0 when b0 == 0: no thread has a slice
        const uint32_t s0 = threadIdx.x * per;
        const uint32_t s1 = s0 + per < b0 ? s0 + per : b0;
        CarryFn fs[PER_MAX];
#pragma unroll
        for (int i = 0; i < PER_MAX; ++i)
            if (s0 + i < s1) fs[i] = CarryFn{J.agg_base[s0 + i], J.agg_thresh[s0 + i]};
        CarryFn agg = fs[0];
#pragma unroll
        for (int i = 1; i < PER_MAX; ++i)
            if (s0 + i < s1) agg = compose(agg, fs[i]);
        if (s0 >= s1) agg = CarryFn{0u, kInf};
        const int wl = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const CarryFn prev = shfl_up_fn(agg, d);
            if (wl >= d) agg = compose(prev, agg);
        }
        // Threads with a slice: nv = ceil(b0 / per); wave w holds nv - 64w of
        // them (clamped to 0..64), its result at its last such lane.
        const int nv = per ? (int)((b0 + per - 1) / per) : 0;
        const int in_wave = nv - 64 * wave < 0 ? 0 : (nv - 64 * wave > 64 ? 64 : nv - 64 * wave);
        if (in_wave > 0 && wl == in_wave - 1) wave_fn[wave] = agg;
        __syncthreads();
        if (threadIdx.x == 0) {
            const int waves = (nv + 63) / 64;
            uint32_t cc = 0;
            for (int w = 0; w < waves; ++w) cc = wave_fn[w].base + (cc >= wave_fn[w].thresh ? 1u : 0u);
            for (uint32_t k = b0; k < bn; ++k) {
                cin_s[k - b0] = cc;
                cc = J.agg_base[k] + (cc >= J.agg_thresh[k] ? 1u : 0u);
            }
        }
    }
    __syncthreads();
    // Each block's carry: its first limb takes it, and the limbs up to the
    // block's ripple end one more when it reaches the first limb's threshold.
    // The blocks' limbs are disjoint.
    // Ripple limbs are read four per thread before any is written back.
    constexpr int U = 4;
    for (uint32_t k = b0; k < bn; ++k) {
        const uint32_t c = k == 0 ? 0u : cin_s[k - b0];
        if (c == 0) continue;
        const uint64_t L0 = (uint64_t)k * BL;
        if (threadIdx.x == 0 && L0 < J.out_len) J.out[L0] += c;
        if (c >= J.first_thresh[k]) {
            const uint32_t rend = J.ripple_end[k];
            for (uint32_t r0 = 1; r0 <= rend; r0 += U * NT) {
                uint64_t v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t rel = r0 + threadIdx.x + u * NT;
                    if (rel <= rend && L0 + rel < J.out_len) v[u] = J.out[L0 + rel];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t rel = r0 + threadIdx.x + u * NT;
                    if (rel <= rend && L0 + rel < J.out_len) J.out[L0 + rel] = v[u] + 1;
                }
            }
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

bool launch_tables(const TableArgs& a, int primes, hipStream_t s) {
    const unsigned blocks = (unsigned)((a.len + 255) / 256 < 4096 ? (a.len + 255) / 256 : 4096);
    NTT_LAUNCH(nullptr, k_tables, dim3(blocks, primes), dim3(256), s, a);
    return NTT_CHECK_LAUNCH();
}

// Carries across the output blocks of blocks of BL limbs: the scan over
// the block functions of every long output, then one workgroup per block.
// In a batched launch a.blocks counts (product, block) items and
// a.scan_count the products that take the scan.
template <int BL>
static bool launch_carries(const CrtArgs& a, hipStream_t s) {
    if (a.batch ? a.scan_count > 0 : carry_scanned(a.blocks)) {
        NTT_LAUNCH("carry_scan", k_carry_scan, dim3(a.batch ? a.scan_count : carry_chunks(a.blocks)),
                   dim3(kCarryChunk), s, a);
        if (!NTT_CHECK_LAUNCH()) return false;
    }
    if (a.batch)
        NTT_LAUNCH("carry_apply", (k_carry_apply<BL, 1>), dim3(a.blocks), dim3(apply_threads<BL>()), s, a);
    else if (carry_scanned(a.blocks))
        NTT_LAUNCH("carry_apply", (k_carry_apply<BL, kApplyBlocks>), dim3((a.blocks + kApplyBlocks - 1) / kApplyBlocks),
                   dim3(apply_threads<BL>()), s, a);
    else if (a.blocks > 1)
        NTT_LAUNCH("carry_apply", (k_carry_apply<BL, 1>), dim3(a.blocks), dim3(apply_threads<BL>()), s, a);
    return NTT_CHECK_LAUNCH();
}

// In a batched launch a.blocks counts (product, block) items. a.wide: one
// 512-thread block covers the whole output (at most 2048 limbs), so there
// are no carries between blocks.
template <int NT>
static bool launch_crt_t(const CrtArgs& a, dim3 grid, hipStream_t s) {
    switch (a.primes) {
        case 1: NTT_LAUNCH("crt_product", (k_crt_product<1, NT>), grid, dim3(NT), s, a); break;
        case 2: NTT_LAUNCH("crt_product", (k_crt_product<2, NT>), grid, dim3(NT), s, a); break;
        case 3: NTT_LAUNCH("crt_product", (k_crt_product<3, NT>), grid, dim3(NT), s, a); break;
        case 4: NTT_LAUNCH("crt_product", (k_crt_product<4, NT>), grid, dim3(NT), s, a); break;
        case 5: NTT_LAUNCH("crt_product", (k_crt_product<5, NT>), grid, dim3(NT), s, a); break;
        case 6: NTT_LAUNCH("crt_product", (k_crt_product<6, NT>), grid, dim3(NT), s, a); break;
        case 7: NTT_LAUNCH("crt_product", (k_crt_product<7, NT>), grid, dim3(NT), s, a); break;
        case 8: NTT_LAUNCH("crt_product", (k_crt_product<8, NT>), grid, dim3(NT), s, a); break;
        default: return false;
    }
    return NTT_CHECK_LAUNCH();
}

bool launch_crt_product(const CrtArgs& a, hipStream_t s) {
    if (a.wide) {
        if (a.batch || a.blocks != 1 || a.out_len > (uint64_t)CrtShape<2 * kCrtThreads>::BLOCK) return false;
        return launch_crt_t<2 * kCrtThreads>(a, dim3(1), s);
    }
    if (!launch_crt_t<kCrtThreads>(a, dim3(a.blocks), s)) return false;
    return launch_carries<kCrtBlockLimbs>(a, s);
}

// Schoolbook product blocks of kSchoolBlockLimbs limbs, then the carries
// across blocks; in a batched launch a.blocks counts (product, block) items.
bool launch_school(const CrtArgs& a, hipStream_t s) {
    NTT_LAUNCH("school", k_school, dim3(a.blocks), dim3(kSchoolThreads), s, a);
    if (!NTT_CHECK_LAUNCH()) return false;
    return launch_carries<kSchoolBlockLimbs>(a, s);
}

bool launch_recombine(const RecombineArgs& a, int primes, hipStream_t s) {
    const uint64_t count = a.tail == 2 ? a.n >> 1 : a.n >> 2;
    if (a.tail < 1 || a.tail > 3 || count == 0) return false;
    NTT_LAUNCH("recombine", k_recombine, dim3((unsigned)((count + 255) / 256), primes), dim3(256), s, a);
    return NTT_CHECK_LAUNCH();
}

bool launch_crt_disjoint(const CrtArgs& a, hipStream_t s) {
    dim3 grid((unsigned)((a.coefficients + 255) / 256));
    switch (a.primes) {
        case 1: NTT_LAUNCH(nullptr, k_crt_disjoint<1>, grid, dim3(256), s, a); break;
        case 2: NTT_LAUNCH(nullptr, k_crt_disjoint<2>, grid, dim3(256), s, a); break;
        case 3: NTT_LAUNCH(nullptr, k_crt_disjoint<3>, grid, dim3(256), s, a); break;
        case 4: NTT_LAUNCH(nullptr, k_crt_disjoint<4>, grid, dim3(256), s, a); break;
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

// Plain device copy of 16-byte words: the calibration kernel for HBM byte
// counters (reads and writes each byte once).
__global__ __launch_bounds__(256) void k_copy(const uint4* src, uint4* dst, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

bool launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t s) {
    const uint64_t n = bytes / 16;
    NTT_LAUNCH(nullptr, k_copy, dim3(4096), dim3(256), s, (const uint4*)src, (uint4*)dst, n);
    return hipGetLastError() == hipSuccess;
}

bool launch_conv_columns(const uint64_t* a, uint64_t la, const uint64_t* b, uint64_t lb, uint64_t* out,
                         uint32_t stride, uint32_t is_signed, hipStream_t s) {
    const uint64_t n = la + lb - 1;
    NTT_LAUNCH(nullptr, k_conv_columns, dim3((unsigned)((n + 255) / 256)), dim3(256), s, a, la, b,
                       lb, out, stride, is_signed);
    return hipGetLastError() == hipSuccess;
}

}  // namespace nttamd
