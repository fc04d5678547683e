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

#include <hip/hip_ext.h>

namespace nttamd {

// Per-launch timing hooks (engine.hip). While profiling is enabled every
// launch goes through hipExtLaunchKernelGGL, whose start and stop events
// bracket exactly that dispatch on its stream.
bool prof_enabled();
const char* prof_kind();
void prof_record(const char* kind, hipEvent_t start, hipEvent_t stop);
// Keeps the last launch of a kind (kernel, shape, stream, argument block) for
// ntt_profile_replay: kernels whose only argument is one argument block.
void prof_keep(const char* kind, const void* fn, dim3 grid, dim3 block, hipStream_t s, const void* arg,
               size_t size);

template <class... A>
inline void prof_remember(const char* kind, const void* fn, dim3 grid, dim3 block, hipStream_t s,
                          const A&... args) {
    if constexpr (sizeof...(A) == 1) (prof_keep(kind, fn, grid, block, s, &args, sizeof(args)), ...);
}

#define NTT_LAUNCH(KIND, KERNEL, GRID, BLOCK, STREAM, ...)                                 \
    do {                                                                                   \
        hipEvent_t e0_ = nullptr, e1_ = nullptr;                                           \
        if (prof_enabled() && hipEventCreate(&e0_) == hipSuccess &&                        \
            hipEventCreate(&e1_) == hipSuccess) {                                          \
            hipExtLaunchKernelGGL(KERNEL, GRID, BLOCK, 0, STREAM, e0_, e1_, 0, __VA_ARGS__); \
            const char* k_ = (KIND) ? (KIND) : prof_kind();                                \
            prof_record(k_, e0_, e1_);                                                     \
            prof_remember(k_, reinterpret_cast<const void*>(KERNEL), GRID, BLOCK, STREAM, __VA_ARGS__); \
        } else {                                                                           \
            hipLaunchKernelGGL(KERNEL, GRID, BLOCK, 0, STREAM, __VA_ARGS__);              \
        }                                                                                  \
    } while (0)

// Padded LDS index: one pad word after every 16 values. The last round of a
// pass gives thread t the points 16t..16t+15 (or 8t..8t+7 in radix 8), so
// the 16 lanes of one LDS cycle read values 16 (or 8) apart; with a pad per
// 16 values they fall on 16 distinct bank pairs (a pad per 32 values left
// pairs of lanes on one bank pair). Consecutive points and the 2-value
// stripes of row loads stay conflict-free.
__device__ __forceinline__ int pidx(int e) { return e + (e >> 4); }

template <int LOGL>
struct Lay {
    static constexpr int L = 1 << LOGL;
    static constexpr int STRIDE = L + (L >> 4) + 1;  // padded line in LDS (odd: lines fall on distinct banks)
    static constexpr int ROUNDS = (LOGL + 3) / 4;
    static constexpr int FIRST = LOGL - 4 * (ROUNDS - 1);  // levels of round 0
};

// Roots of the tree come from the root table of one prime in global memory.
// Line c of a tile starts at node h0 + c*hstep of the pass's first level, so
// its node x at local level l has the root W[(h0 + c*hstep) << l | x]. From
// local level `split` on, where those indices spread over the whole table,
// the root is formed as W[x] * W[(h0 + c*hstep) << l] instead: the table is
// bit-reversed, so W[j 2^l + x] = W[j 2^l] W[x] for x < 2^l, and W[x] comes
// from the short, cache-resident prefix of the table.
//
// A product root is left unreduced: both factors are centered (at most p/2),
// so the product is at most p^2/4 and its rounded quotient is within
// 1/2 + 1/16 of the exact one, which leaves |root| < 9p/16 + p/16 = 5p/8.
// Product roots serve only the last forward round of a row (entered below
// 31p/16, its multiplied inputs stay below 5.2p and their products below
// 3.3p^2 < 4p^2, and it still ends below 9p/16 + 11p/8 = 31p/16) and the
// inverse levels (differences below 11p/2 times 5p/8 stay below 4p^2, and
// below 2p^2 for differences below 11p/4), so every bound below holds.
struct Tw {
    const double* table;
    uint64_t h0;
    int hstep;  // 1 when every line is its own node (rows), 0 when lines share one
    int cmask;  // line c is node h0 + (c & cmask) * hstep (several tiles of rows)
    int split;  // first local level whose roots are formed as products
    double p, pinv;
    __device__ __forceinline__ double at(int c, int lvl, int node) const {
        const uint64_t h = (h0 + (uint64_t)((c & cmask) * hstep)) << lvl;
        if (lvl >= split) return mulmod(table[node], table[h], p, pinv);
        return table[h + node];
    }
};

// One DIF round: local levels [L0, L0 + RLOG) of every line of the tile.
//
// Reduction schedule, fixed at compile time: the added (upper) input of a
// butterfly is reduced at the last level of each round only. With
// |root| <= p/2, a level maps inputs below B to outputs below B + 15p/16
// while the multiplied input stays below 4p (its product below 2p^2), or
// B + 11p/8 while it stays below 8p (product below 4p^2). A round entered
// below 31p/16 reaches 31p/16 + 3*15p/16 = 19p/4 before its last level,
// which reduces the added input to below 9p/16 and multiplies one below
// 19p/4 (product below 2.4p^2), so it ends below 9p/16 + 11p/8 = 31p/16.
// A short first round of 1-3 levels, entered below 31p/16, ends below 3p/2.
// Values never exceed 19p/4 < 2^53, packed inputs are below p, and spectra
// stay below 31p/16, so a pointwise product stays below 3.76p^2 < 4p^2.
//
// ROOT1: level 0 of this round is the first level of the transform at table
// node 0, whose root W[0] is 1, and its inputs are packed (below p), so the
// multiplied input enters its butterfly as it is: outputs below 2p at level
// 0 and, for a round of one level, below 9p/16 + p < 31p/16.
template <int LOGL, int LINES, int NT, int L0, int RLOG, bool ROOT1 = false>
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
                    const double u = s == RLOG - 1 ? reduce(x[k], p, pinv) : x[k];
                    const double t = (ROOT1 && L0 == 0 && s == 0) ? x[k + half] : mulmod(x[k + half], w, p, pinv);
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
//
// ROOT1: level 0 of this round is the last level of the transform at table
// node 0, whose root is 1: the difference is only reduced (below 9p/16).
template <int LOGL, int LINES, int NT, int L0, int RLOG, bool ROOT1 = false>
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
                    x[k + half] = (ROOT1 && L0 == 0 && s == 0) ? reduce(u - v, p, pinv) : mulmod(u - v, w, p, pinv);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < RADIX; ++k) line[pidx(base + k * D)] = x[k];
    }
}

// Rounds R.. of a forward pass; root1 (uniform): the pass starts the
// transform at table node 0 from packed inputs (fwd_round's ROOT1).
template <int LOGL, int LINES, int NT, int R>
__device__ __forceinline__ void fwd_rounds(double* lds, Tw tw, double p, double pinv, bool root1 = false) {
    if constexpr (R < Lay<LOGL>::ROUNDS) {
        constexpr int L0 = R == 0 ? 0 : Lay<LOGL>::FIRST + 4 * (R - 1);
        constexpr int RL = R == 0 ? Lay<LOGL>::FIRST : 4;
        if (R == 0 && root1)
            fwd_round<LOGL, LINES, NT, L0, RL, true>(lds, tw, p, pinv);
        else
            fwd_round<LOGL, LINES, NT, L0, RL, false>(lds, tw, p, pinv);
        __syncthreads();
        fwd_rounds<LOGL, LINES, NT, R + 1>(lds, tw, p, pinv);
    }
}

// Rounds R..0 of an inverse pass; root1 (uniform): the pass ends the
// transform at table node 0 (inv_round's ROOT1).
template <int LOGL, int LINES, int NT, int R>
__device__ __forceinline__ void inv_rounds(double* lds, Tw tw, double p, double pinv, bool root1 = false) {
    if constexpr (R >= 0) {
        constexpr int L0 = R == 0 ? 0 : Lay<LOGL>::FIRST + 4 * (R - 1);
        constexpr int RL = R == 0 ? Lay<LOGL>::FIRST : 4;
        if (R == 0 && root1)
            inv_round<LOGL, LINES, NT, L0, RL, true>(lds, tw, p, pinv);
        else
            inv_round<LOGL, LINES, NT, L0, RL, false>(lds, tw, p, pinv);
        __syncthreads();
        inv_rounds<LOGL, LINES, NT, R - 1>(lds, tw, p, pinv, root1);
    }
}

// Inverse rounds of radix 8: rounds of 3 levels (the first of LOGL mod 3, or
// 3), twice the groups of radix 16, for a tile whose radix-16 groups would
// leave half of the threads idle. The schedule above holds for rounds of 1,
// 2 and 3 levels: each ends on a reducing level 0 below 11p/8.
template <int LOGL>
struct Lay8 {
    static constexpr int ROUNDS = (LOGL + 2) / 3;
    static constexpr int FIRST = LOGL - 3 * (ROUNDS - 1);
};

template <int LOGL, int LINES, int NT, int R>
__device__ __forceinline__ void inv_rounds8(double* lds, Tw tw, double p, double pinv, bool root1 = false) {
    if constexpr (R >= 0) {
        constexpr int L0 = R == 0 ? 0 : Lay8<LOGL>::FIRST + 3 * (R - 1);
        constexpr int RL = R == 0 ? Lay8<LOGL>::FIRST : 3;
        if (R == 0 && root1)
            inv_round<LOGL, LINES, NT, L0, RL, true>(lds, tw, p, pinv);
        else
            inv_round<LOGL, LINES, NT, L0, RL, false>(lds, tw, p, pinv);
        __syncthreads();
        inv_rounds8<LOGL, LINES, NT, R - 1>(lds, tw, p, pinv, root1);
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
// Tile movement. Every loop below has a compile-time trip count and is fully
// unrolled, so a thread issues all of its global loads before it waits on
// the first one.

// Strided tile: LINES consecutive lines lo0.. of one node; element m of line c
// at g[m*S + c]. Consecutive threads take consecutive lines, so each group of
// LINES threads reads one LINES*8-byte run.
template <int LOGL, int LINES, int NT>
__device__ __forceinline__ void load_strided(double* lds, const double* g, uint64_t S) {
    constexpr int L = 1 << LOGL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    constexpr int TOTAL = LINES * L;
    constexpr int R = (TOTAL + NT - 1) / NT;
    double v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int f = threadIdx.x + r * NT;
        if (TOTAL % NT == 0 || f < TOTAL) v[r] = g[(uint64_t)(f / LINES) * S + (f % LINES)];
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int f = threadIdx.x + r * NT;
        if (TOTAL % NT == 0 || f < TOTAL) lds[(f % LINES) * LS + pidx(f / LINES)] = v[r];
    }
}

template <int LOGL, int LINES, int NT, bool CANON>
__device__ __forceinline__ void store_strided(const double* lds, double* g, uint64_t S, double tinv, double p,
                                              double pinv) {
    constexpr int L = 1 << LOGL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    constexpr int TOTAL = LINES * L;
    constexpr int R = (TOTAL + NT - 1) / NT;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int f = threadIdx.x + r * NT;
        if (TOTAL % NT != 0 && f >= TOTAL) continue;
        double v = lds[(f % LINES) * LS + pidx(f / LINES)];
        if constexpr (CANON) v = canonical(mulmod(v, tinv, p, pinv), p);
        g[(uint64_t)(f / LINES) * S + (f % LINES)] = v;
    }
}

// Strided tile packed from limbs: element m of line c is coefficient
// n0 + m*S + c of the operand.
// Packs the tile's coefficients n0 + m*S + c (m < L, c < LINES) into
// lds[c*LS + pidx(m)]. Each thread takes its points in groups of BATCH whose
// coefficient indices increase; a group entirely at or beyond `digits` is
// zero padding and costs nothing, a group entirely below `safe` loads
// without clamping, and the few groups in between take the clamped form.
template <int LOGL, int LINES, int NT, int KIND>
__device__ __forceinline__ void pack_tile_as(double* lds, const Source& src, uint64_t n0, uint64_t S,
                                             const double* pw, double p, double pinv) {
    constexpr int L = 1 << LOGL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    constexpr int TOTAL = LINES * L;
    constexpr int R = (TOTAL + NT - 1) / NT;
    constexpr int BATCH = (R >= 4 && R % 4 == 0) ? 4 : 1;
    const uint64_t safe = pack_safe<KIND>(src);
    const uint64_t digits = src.digits;
#pragma unroll
    for (int r0 = 0; r0 < R; r0 += BATCH) {
        double v[BATCH];
        uint64_t n[BATCH];
#pragma unroll
        for (int r = 0; r < BATCH; ++r) {
            const int f = threadIdx.x + (r0 + r) * NT;
            n[r] = n0 + (uint64_t)(f / LINES) * S + (f % LINES);
            v[r] = 0.0;
        }
        const bool valid_last = TOTAL % NT == 0 || (int)threadIdx.x + (r0 + BATCH - 1) * NT < TOTAL;
        if (n[0] < digits) {
            if (valid_last && n[BATCH - 1] < safe) {
#pragma unroll
                for (int r = 0; r < BATCH; ++r) v[r] = pack_as<KIND, false>(src, n[r], pw, p, pinv);
            } else {
#pragma unroll
                for (int r = 0; r < BATCH; ++r) v[r] = pack_as<KIND, true>(src, n[r], pw, p, pinv);
            }
        }
#pragma unroll
        for (int r = 0; r < BATCH; ++r) {
            const int f = threadIdx.x + (r0 + r) * NT;
            if (TOTAL % NT == 0 || f < TOTAL) lds[(f % LINES) * LS + pidx(f / LINES)] = v[r];
        }
    }
}

// Fold of the operand onto a part of a composite row: point n gains
// coefficient n + q*span times m[q-1] for 1 <= q < terms.
struct Fold {
    uint64_t span;
    int terms;
    double m[3];
};

__device__ __forceinline__ Fold fold_of(const PassArgs& a, int y) {
    return Fold{a.fold_span, a.fold_terms, {a.fold_mult[y][0], a.fold_mult[y][1], a.fold_mult[y][2]}};
}

// Adds the folded terms to the points of the tile this thread packed (only
// points below digits - span have any). Each added term is below 15p/16 and
// the packed point below p, so the sum stays below 8p before its reduction.
template <int LOGL, int LINES, int NT, int KIND>
__device__ __forceinline__ void fold_tile_as(double* lds, const Source& src, uint64_t n0, uint64_t S,
                                             const Fold& fold, const double* pw, double p, double pinv) {
    constexpr int L = 1 << LOGL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    constexpr int TOTAL = LINES * L;
    constexpr int R = (TOTAL + NT - 1) / NT;
    const uint64_t digits = src.digits;
    if (n0 + fold.span >= digits) return;
#pragma unroll 1
    for (int r = 0; r < R; ++r) {
        const int f = threadIdx.x + r * NT;
        if (TOTAL % NT != 0 && f >= TOTAL) break;
        const uint64_t n = n0 + (uint64_t)(f / LINES) * S + (f % LINES);
        if (n + fold.span >= digits) continue;
        double* x = lds + (f % LINES) * LS + pidx(f / LINES);
        double acc = *x;
#pragma unroll 1
        for (int q = 1; q < fold.terms; ++q) {
            const uint64_t m = n + (uint64_t)q * fold.span;
            if (m >= digits) break;
            const double w = q == 1 ? fold.m[0] : (q == 2 ? fold.m[1] : fold.m[2]);
            acc += mulmod(pack_as<KIND, true>(src, m, pw, p, pinv), w, p, pinv);
        }
        *x = reduce(acc, p, pinv);
    }
}

template <int LOGL, int LINES, int NT, int KIND>
__device__ __forceinline__ void pack_fold_as(double* lds, const Source& src, uint64_t n0, uint64_t S,
                                             const Fold& fold, const double* pw, double p, double pinv) {
    pack_tile_as<LOGL, LINES, NT, KIND>(lds, src, n0, S, pw, p, pinv);
    if (fold.terms > 1) fold_tile_as<LOGL, LINES, NT, KIND>(lds, src, n0, S, fold, pw, p, pinv);
}

template <int LOGL, int LINES, int NT>
__device__ __forceinline__ void pack_strided(double* lds, const Source& src, uint64_t n0, uint64_t S,
                                             const Fold& fold, const double* pw, double p, double pinv) {
    switch (pack_kind(src)) {
        case kPackW64: pack_fold_as<LOGL, LINES, NT, kPackW64>(lds, src, n0, S, fold, pw, p, pinv); break;
        case kPackW64Signed:
            pack_fold_as<LOGL, LINES, NT, kPackW64Signed>(lds, src, n0, S, fold, pw, p, pinv);
            break;
        case kPack1: pack_fold_as<LOGL, LINES, NT, kPack1>(lds, src, n0, S, fold, pw, p, pinv); break;
        case kPack2: pack_fold_as<LOGL, LINES, NT, kPack2>(lds, src, n0, S, fold, pw, p, pinv); break;
        case kPack3: pack_fold_as<LOGL, LINES, NT, kPack3>(lds, src, n0, S, fold, pw, p, pinv); break;
        default: pack_fold_as<LOGL, LINES, NT, kPack4>(lds, src, n0, S, fold, pw, p, pinv); break;
    }
}

// Row tile: LINES rows of L contiguous points; thread t moves points 2t, 2t+1
// of each 2*NT-point stripe with 16-byte accesses.
template <int LOGL, int LINES, int NT>
__device__ __forceinline__ void load_rows(double* lds, const double* g) {
    constexpr int L = 1 << LOGL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    constexpr int PAIRS = LINES * L / 2;
    constexpr int R = (PAIRS + NT - 1) / NT;
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (PAIRS % NT == 0 || threadIdx.x + r * NT < PAIRS)
            v[r] = reinterpret_cast<const double2*>(g)[threadIdx.x + r * NT];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (PAIRS % NT != 0 && threadIdx.x + r * NT >= PAIRS) continue;
        const int f = 2 * (threadIdx.x + r * NT);
        const int c = f / L, m = f % L;
        lds[c * LS + pidx(m)] = v[r].x;
        lds[c * LS + pidx(m + 1)] = v[r].y;
    }
}

template <int LOGL, int LINES, int NT>
__device__ __forceinline__ void store_rows(const double* lds, double* g) {
    constexpr int L = 1 << LOGL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    constexpr int PAIRS = LINES * L / 2;
    constexpr int R = (PAIRS + NT - 1) / NT;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (PAIRS % NT != 0 && threadIdx.x + r * NT >= PAIRS) continue;
        const int f = 2 * (threadIdx.x + r * NT);
        const int c = f / L, m = f % L;
        double2 v;
        v.x = lds[c * LS + pidx(m)];
        v.y = lds[c * LS + pidx(m + 1)];
        reinterpret_cast<double2*>(g)[threadIdx.x + r * NT] = v;
    }
}

// ---------------------------------------------------------------------------
// Work resolution. A single operation takes its prime from blockIdx.y and its
// tile from blockIdx.x; a batched launch reads the product and prime of its
// workgroup from the item table, so one launch covers every product of a
// size bucket.

// Consecutive tiles kept on one XCD: enough to cover a 128-byte line.
__host__ __device__ constexpr int xcd_group(int lines) { return lines >= 16 ? 1 : 16 / lines; }
// Strided passes deal tiles to XCDs when neighbouring tiles share lines or
// the units of a tile read the same limbs (packing); otherwise the plain
// (tile, unit) grid is used.
__host__ __device__ constexpr bool strided_xcd(bool pack, int lines) { return pack || xcd_group(lines) > 1; }

struct Job {
    double* data;          // rows of the operand this pass transforms
    const double* other;   // fused passes: the rows of operand a
    Source src_a;          // packing sources, held by value so that the
    Source src_b;          // packing loops never reload them
    int y;                 // prime
    uint64_t tile;
};

// Workgroup b of a single-operation strided pass runs on XCD b % 8
// (round-robin dispatch). Tiles are dealt to XCDs in groups of G consecutive
// tiles, and all units (primes) of a tile follow each other on one XCD: tiles
// of a group share cache lines (G*LINES*8 >= 128 bytes), and the primes of a
// packing tile read the same limbs, which the first one brings into that
// XCD's L2.
template <int G>
__device__ __forceinline__ bool xcd_tile(int units, uint64_t tiles, uint64_t* tile, int* y) {
    const uint64_t b = blockIdx.x;
    const uint64_t slot = b >> 3;
    const uint64_t kk = slot / (uint64_t)units;
    *y = (int)(slot % (uint64_t)units);
    const uint64_t t = (kk / G) * (8 * G) + (b & 7) * G + (kk % G);
    *tile = t;
    return t < tiles;
}

// Coefficients first.. (at most count of them) of a source, as a source of
// their own; first*width is a multiple of 64, so they start at a limb.
__device__ __forceinline__ Source block_source(Source s, uint64_t first, uint64_t count) {
    const uint64_t limb0 = first * s.width / 64;
    s.limbs += limb0;
    s.limb_len -= limb0;
    s.digits = s.digits - first < count ? s.digits - first : count;
    return s;
}

template <bool BATCH, int G = 1>
__device__ __forceinline__ bool resolve(const PassArgs& a, uint64_t tiles, bool xcd, Job* j) {
    if constexpr (!BATCH) {
        const uint32_t z = blockIdx.z;
        if (a.split_blocks) {
            // Block z of the longer operand, or (z == split_blocks, forward
            // passes only) the shorter one.
            const bool block = z < a.split_blocks;
            j->data = block ? a.data + (uint64_t)z * a.block_stride : a.other;
            j->src_a = block ? block_source(a.src_a, (uint64_t)z * a.block_digits, a.block_digits) : a.src_b;
        } else {
            // Forward passes over both operands in one launch: z = 1 is
            // operand b, whose rows are a.other and whose source is a.src_b.
            j->data = z != 0 ? a.other : a.data;
            j->src_a = z != 0 ? a.src_b : a.src_a;
        }
        j->other = a.other;
        j->src_b = a.src_b;
        if (xcd) return xcd_tile<G>(a.primes, tiles, &j->tile, &j->y);
        j->tile = blockIdx.x;
        j->y = blockIdx.y;
        return true;
    }
    const BatchItem it = a.items[blockIdx.x / a.tiles_per_item];
    const BatchProduct& P = a.batch[it.product];
    j->tile = blockIdx.x % a.tiles_per_item;
    j->y = (int)(it.sub & kItemPrimeMask);
    j->src_a = P.src_a;
    j->src_b = P.src_b;
    if (P.split_blocks) {
        // Block z of the longer operand, or (forward items flagged as the
        // second operand) the shorter one; a block's product lands in its rows.
        const uint32_t z = (it.sub >> kItemBlockShift) & kItemBlockMask;
        j->other = P.spec_b;
        if (a.operand == kOperandForward && (it.sub & kItemOperandB)) {
            j->data = P.spec_b;
            j->src_a = P.src_b;
        } else {
            j->data = P.spec_a + (uint64_t)z * P.block_stride;
            j->src_a = block_source(P.src_a, (uint64_t)z * P.block_digits, P.block_digits);
        }
        return true;
    }
    j->other = P.spec_a;
    if (a.operand == kOperandProduct) {
        j->data = P.square ? P.spec_a : P.spec_b;
    } else if (it.sub & kItemOperandB) {
        j->data = P.spec_b;
        j->src_a = P.src_b;
    } else {
        j->data = P.spec_a;
    }
    return true;
}

// ---------------------------------------------------------------------------
// Strided passes: the tile is LINES consecutive lines lo0..lo0+LINES-1 of
// first node H; element m of line lo sits at H*(T >> level_start) + m*S + lo.

template <int LOGL, int LINES, int NT, bool PACK, bool BATCH>
__global__ __launch_bounds__(NT) void k_fwd_strided(PassArgs a) {
    constexpr int LS = Lay<LOGL>::STRIDE;
    __shared__ double lds[LINES * LS];
    const uint64_t S = a.line_stride;
    const uint64_t tiles_per_node = S / LINES;
    Job j;
    if (!resolve<BATCH, xcd_group(LINES)>(a, tiles_per_node << a.level_start, strided_xcd(PACK, LINES), &j))
        return;
    const int y = j.y;
    const double p = a.field.p[y], pinv = a.field.pinv[y];
    const uint64_t H = j.tile / tiles_per_node;
    const uint64_t lo0 = (j.tile % tiles_per_node) * LINES;
    const uint64_t span = (uint64_t)1 << (a.log_t - a.level_start);
    const uint64_t base = H * span + lo0;
    double* row = j.data + (uint64_t)y * a.row_stride;
    if constexpr (PACK) {
        pack_strided<LOGL, LINES, NT>(lds, j.src_a, base, S, fold_of(a, y), a.powers[y], p, pinv);
    } else {
        load_strided<LOGL, LINES, NT>(lds, row + base, S);
    }
    __syncthreads();
    const Tw tw{a.fwd + (uint64_t)y * a.table_stride, H + (a.node_base << a.level_start), 0, 0, 64, p, pinv};
    // A packing pass starts the transform (level_start 0); at node 0 its first
    // root is 1.
    fwd_rounds<LOGL, LINES, NT, 0>(lds, tw, p, pinv, PACK && a.node_base == 0);
    store_strided<LOGL, LINES, NT, false>(lds, row + base, S, 0.0, p, pinv);
}

// Forward rounds from local level L0 to the end of the line, MAXR levels per
// round (the last one shorter).
template <int LOGL, int LINES, int NT, int L0, int MAXR>
__device__ __forceinline__ void fwd_rounds_from(double* lds, Tw tw, double p, double pinv) {
    if constexpr (L0 < LOGL) {
        constexpr int RL = LOGL - L0 < MAXR ? LOGL - L0 : MAXR;
        fwd_round<LOGL, LINES, NT, L0, RL, false>(lds, tw, p, pinv);
        __syncthreads();
        fwd_rounds_from<LOGL, LINES, NT, L0 + RL, MAXR>(lds, tw, p, pinv);
    }
}

// Levels S0..RL-1 of the transform's first round (radix 2^RL) on the points
// of one group, in registers (fwd_round's butterflies with ROOT1: the group
// sits at table node 0, whose level-0 root is 1, and holds packed inputs).
// Levels below S0 were trivial: their multiplied inputs were all zero.
template <int RL, int S0>
__device__ __forceinline__ void first_round_regs(double* x, const Tw& tw, double p, double pinv) {
    constexpr int RADIX = 1 << RL;
#pragma unroll
    for (int s = S0; s < RL; ++s) {
        const int half = RADIX >> (s + 1);
#pragma unroll
        for (int blk = 0; blk < (1 << s); ++blk) {
            const double w = tw.at(0, s, blk);
#pragma unroll
            for (int j = 0; j < half; ++j) {
                const int k = blk * 2 * half + j;
                const double u = s == RL - 1 ? reduce(x[k], p, pinv) : x[k];
                const double t = s == 0 ? x[k + half] : mulmod(x[k + half], w, p, pinv);
                x[k] = u + t;
                x[k + half] = u - t;
            }
        }
    }
}

// Tiles of the shared packing pass: 4096 points (its extracted parts stay in
// registers), eight points per thread.
template <int LOGL>
static constexpr int pack_primes_lines() {
    return 4096 >> LOGL;
}
template <int LOGL>
static constexpr int pack_primes_threads() {
    return 4096 / 8;
}

// First forward pass with packing shared across primes: one workgroup packs
// one tile of LINES lines of 2^LOGL points (4096 points) for every prime of
// the operation. Each thread extracts the parts of its 8 coefficients once
// (elements m0 + r*L/8 of line c: exactly one radix-8 group of the first
// three levels), then per prime combines them into residues, runs those
// levels in registers, stores the group to LDS, runs the remaining rounds,
// and writes the tile to that prime's row. The rows start at table node 0
// and are not folded (single operations on power-of-two rows or cyclic
// parts).
//
// Zero padding past `digits` prunes the first levels: a group whose upper
// half is zero has a trivial level 0 (both outputs are the lower input),
// and one whose upper three quarters are zero also a trivial level 1 (the
// pruned first stages of src/arch/schedule/expression/stages.rs:109-163).
// Bounds: packed inputs are below p, level 0 (root 1) gives 2p, level 1 adds
// 15p/16 (multiplied inputs below 4p, products below 2p^2), level 2 reduces
// the added input and ends below 9p/16 + 15p/16 = 3p/2, and every later round
// is entered below 31p/16 (fwd_round's schedule).
template <int LOGL, int LINES, int NT, int KIND>
__device__ __forceinline__ void pack_primes_tile(const PassArgs& a, const Job& j, int primes, double* lds) {
    constexpr int L = 1 << LOGL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    constexpr int D = L / 8;
    constexpr int P = PackParts<KIND>::value;
    static_assert(LINES * L == 8 * NT, "one radix-8 group of the first levels per thread");
    const uint64_t S = a.line_stride;
    const int c = threadIdx.x % LINES, m0 = threadIdx.x / LINES;
    const uint64_t base = j.tile * LINES;
    const uint64_t n0 = base + c + (uint64_t)m0 * S;  // element m0 + r*D at n0 + r*D*S
    const uint64_t step = (uint64_t)D * S;
    const Source& src = j.src_a;
    const uint64_t digits = src.digits;
    const uint64_t safe = pack_safe<KIND>(src);
    double d[8][P];
    int fill = 0;  // leading elements below digits (the rest are zero)
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const uint64_t n = n0 + (uint64_t)r * step;
        if (n < safe) {
            extract_as<KIND, false>(src, n, d[r]);
        } else {
            extract_as<KIND, true>(src, n, d[r]);
        }
        fill += n < digits ? 1 : 0;
    }
#pragma unroll 1
    for (int y = 0; y < primes; ++y) {
        const double p = a.field.p[y], pinv = a.field.pinv[y];
        const double* pw = a.powers[y];
        double x[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) x[r] = combine_as<KIND>(src, d[r], pw, p, pinv);
        const Tw tw{a.fwd + (uint64_t)y * a.table_stride, 0, 0, 0, 64, p, pinv};
        if (fill <= 2) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                x[k + 2] = x[k];
                x[k + 4] = x[k];
                x[k + 6] = x[k];
            }
            first_round_regs<3, 2>(x, tw, p, pinv);
        } else if (fill <= 4) {
#pragma unroll
            for (int k = 0; k < 4; ++k) x[k + 4] = x[k];
            first_round_regs<3, 1>(x, tw, p, pinv);
        } else {
            first_round_regs<3, 0>(x, tw, p, pinv);
        }
        double* line = lds + c * LS;
#pragma unroll
        for (int k = 0; k < 8; ++k) line[pidx(m0 + k * D)] = x[k];
        __syncthreads();
        // Radix-8 rounds: one group per thread, and eight values in flight
        // beside the extracted parts.
        fwd_rounds_from<LOGL, LINES, NT, 3, 3>(lds, tw, p, pinv);
        store_strided<LOGL, LINES, NT, false>(lds, j.data + (uint64_t)y * a.row_stride + base, S, 0.0, p, pinv);
        __syncthreads();  // the tile is read out before the next prime overwrites it
    }
}

template <int LOGL, int LINES, int NT, int KIND>
__global__ __launch_bounds__(NT) void k_pack_primes(PassArgs a) {
    __shared__ double lds[LINES * Lay<LOGL>::STRIDE];
    Job j;
    resolve<false>(a, a.line_stride / LINES, false, &j);
    pack_primes_tile<LOGL, LINES, NT, KIND>(a, j, a.primes, lds);
}

// Batched form: workgroup b packs tile b % tiles_per_item of the operand (or
// block) that item b / tiles_per_item names, for every prime of its product.
// The items are (product, block, operand) with prime 0; the host sends no
// product whose coefficients need four parts.
template <int LOGL, int LINES, int NT>
__global__ __launch_bounds__(NT) void k_pack_primes_batch(PassArgs a) {
    __shared__ double lds[LINES * Lay<LOGL>::STRIDE];
    Job j;
    resolve<true>(a, 0, false, &j);
    const int primes = (int)a.batch[a.items[blockIdx.x / a.tiles_per_item].product].primes;
    switch (pack_kind(j.src_a)) {
        case kPackW64: pack_primes_tile<LOGL, LINES, NT, kPackW64>(a, j, primes, lds); break;
        case kPackW64Signed: pack_primes_tile<LOGL, LINES, NT, kPackW64Signed>(a, j, primes, lds); break;
        case kPack1: pack_primes_tile<LOGL, LINES, NT, kPack1>(a, j, primes, lds); break;
        case kPack2: pack_primes_tile<LOGL, LINES, NT, kPack2>(a, j, primes, lds); break;
        case kPack3: pack_primes_tile<LOGL, LINES, NT, kPack3>(a, j, primes, lds); break;
        default: break;
    }
}

// Inverse strided pass; CANON (level_start == 0) multiplies by T^-1 and stores
// canonical residues in [0, p).
template <int LOGL, int LINES, int NT, bool CANON, bool BATCH>
__global__ __launch_bounds__(NT) void k_inv_strided(PassArgs a) {
    constexpr int LS = Lay<LOGL>::STRIDE;
    __shared__ double lds[LINES * LS];
    const uint64_t S = a.line_stride;
    const uint64_t tiles_per_node = S / LINES;
    Job j;
    if (!resolve<BATCH, xcd_group(LINES)>(a, tiles_per_node << a.level_start, strided_xcd(false, LINES), &j))
        return;
    const int y = j.y;
    const double p = a.field.p[y], pinv = a.field.pinv[y];
    const uint64_t H = j.tile / tiles_per_node;
    const uint64_t lo0 = (j.tile % tiles_per_node) * LINES;
    const uint64_t span = (uint64_t)1 << (a.log_t - a.level_start);
    const uint64_t base = H * span + lo0;
    double* row = j.data + (uint64_t)y * a.row_stride;
    load_strided<LOGL, LINES, NT>(lds, row + base, S);
    __syncthreads();
    const Tw tw{a.inv + (uint64_t)y * a.table_stride, H + (a.node_base << a.level_start), 0, 0, 64, p, pinv};
    // The canonical pass ends the transform (level_start 0); at node 0 its
    // last root is 1.
    inv_rounds<LOGL, LINES, NT, Lay<LOGL>::ROUNDS - 1>(lds, tw, p, pinv, CANON && a.node_base == 0);
    store_strided<LOGL, LINES, NT, CANON>(lds, row + base, S, a.tinv[y], p, pinv);
}

// Row passes form the roots of their last four levels as products (every
// row is its own node, so reading them would cost one table entry per point
// at the last level; the products read W[x] from the short, cache-resident
// prefix of the table and one W[H 2^l] per row and level).
__host__ __device__ constexpr int row_split(int logl) { return logl > 4 ? logl - 4 : 0; }

// ---------------------------------------------------------------------------
// Row pass: row H of 2^LOGL contiguous points holds the last levels of the
// transform.
//
// Fused row pass over LINES consecutive rows H0.. of each operand: forward
// levels of the rows of both operands, pointwise product (or a square),
// inverse levels, store. The rows of b are j.data and those of a are
// j.other; the product replaces the rows of b. Tile t, line c sits at
// lds + (t*LINES + c)*LS and is node H0 + c. With PACK the row is the whole
// transform (level_start == 0, LINES == 1): both operands are packed from
// limbs into LDS and the result is stored as canonical residues times T^-1.
// FORM 0: radix-16 forward rounds, inverse rounds in radix 16 for a square
// and radix 8 for a product; FORM 1: radix 16 both ways; FORM 2: radix 8
// both ways, with one thread per radix-8 group (twice the threads of the
// radix-16 shapes over the same tile: more waves per CU, one more round).
template <int LOGL, int LINES, int NT, bool PACK, bool SQUARE, bool BATCH, int FORM = 0>
__global__ __launch_bounds__(NT) void k_fused(PassArgs a) {
    static_assert(!PACK || LINES == 1, "a packed transform is one row");
    constexpr int L = 1 << LOGL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    constexpr int TILES = SQUARE ? 1 : 2;
    constexpr int PTS = LINES * L;
    constexpr int R = (PTS + NT - 1) / NT;
    constexpr bool FULL = PTS % NT == 0;
    __shared__ double lds[TILES * LINES * LS];
    Job j;
    resolve<BATCH>(a, 0, false, &j);
    const int y = j.y;
    const double p = a.field.p[y], pinv = a.field.pinv[y];
    const uint64_t H0 = j.tile * LINES;
    const uint64_t HR = H0 + (a.node_base << a.level_start);
    double* row = j.data + (uint64_t)y * a.row_stride + H0 * L;
    if constexpr (PACK) {
        const Fold fold = fold_of(a, y);
        pack_strided<LOGL, 1, NT>(lds, j.src_a, 0, 1, fold, a.powers[y], p, pinv);
        if constexpr (!SQUARE) pack_strided<LOGL, 1, NT>(lds + LS, j.src_b, 0, 1, fold, a.powers[y], p, pinv);
    } else {
        load_rows<LOGL, LINES, NT>(lds, row);
        if constexpr (!SQUARE)
            load_rows<LOGL, LINES, NT>(lds + LINES * LS, j.other + (uint64_t)y * a.row_stride + H0 * L);
    }
    __syncthreads();
    const int split = PACK ? 64 : row_split(LOGL);
    const Tw twF{a.fwd + (uint64_t)y * a.table_stride, HR, 1, LINES - 1, split, p, pinv};
    const Tw twI{a.inv + (uint64_t)y * a.table_stride, HR, 1, LINES - 1, split, p, pinv};
    // The lines of both tiles go through one set of rounds; line c of either
    // tile is node H0 + (c mod LINES).
    if constexpr (FORM == 2)
        fwd_rounds_from<LOGL, TILES * LINES, NT, 0, 3>(lds, twF, p, pinv);
    else
        fwd_rounds<LOGL, TILES * LINES, NT, 0>(lds, twF, p, pinv, PACK && a.node_base == 0);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int e = threadIdx.x + r * NT;
        if (!FULL && e >= PTS) continue;
        const int c = e >> LOGL, m = e & (L - 1);
        double* x = lds + c * LS + pidx(m);
        if constexpr (SQUARE)
            *x = mulmod(*x, *x, p, pinv);
        else
            *x = mulmod(*x, x[LINES * LS], p, pinv);
    }
    __syncthreads();
    // The product is one tile: radix-16 groups would occupy half of the
    // threads, so a product (not a square) runs its inverse in radix 8
    // unless the autotuner chose radix 16 (FORM 1) for this shape.
    if constexpr (FORM != 2 && (SQUARE || FORM == 1))
        inv_rounds<LOGL, LINES, NT, Lay<LOGL>::ROUNDS - 1>(lds, twI, p, pinv, PACK && a.node_base == 0);
    else
        inv_rounds8<LOGL, LINES, NT, Lay8<LOGL>::ROUNDS - 1>(lds, twI, p, pinv, PACK && a.node_base == 0);
    if constexpr (PACK) {
        const double tinv = a.tinv[y];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int m = threadIdx.x + r * NT;
            if (!FULL && m >= L) continue;
            row[m] = canonical(mulmod(lds[pidx(m)], tinv, p, pinv), p);
        }
    } else {
        store_rows<LOGL, LINES, NT>(lds, row);
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
struct CrtJob {
    const double* residues;
    uint64_t row_stride, coefficients, out_len;
    uint64_t* out;
    uint32_t width, magic, value_bits, block, blocks;
    uint32_t *agg_base, *agg_thresh, *first_thresh, *ripple_end;
    uint32_t split_blocks;  // 0: no split
    uint64_t split_digits, block_stride;
    const uint64_t *op_a, *op_b;  // schoolbook operands
    uint64_t len_a, len_b;
};

__device__ __forceinline__ CrtJob crt_job(const CrtArgs& a) {
    CrtJob j;
    if (a.batch == nullptr) {
        j.residues = a.residues;
        j.row_stride = a.row_stride;
        j.coefficients = a.coefficients;
        j.out_len = a.out_len;
        j.out = a.out;
        j.width = a.width;
        j.magic = a.width_magic;
        j.value_bits = a.value_bits;
        j.block = blockIdx.x;
        j.blocks = a.blocks;
        j.agg_base = a.agg_base;
        j.agg_thresh = a.agg_thresh;
        j.first_thresh = a.first_thresh;
        j.ripple_end = a.ripple_end;
        j.split_blocks = a.split_blocks;
        j.split_digits = a.split_digits;
        j.block_stride = a.block_stride;
        j.op_a = a.op_a;
        j.op_b = a.op_b;
        j.len_a = a.len_a;
        j.len_b = a.len_b;
        return j;
    }
    const BatchItem it = a.items[blockIdx.x];
    const BatchProduct& P = a.batch[it.product];
    j.residues = (P.square || P.split_blocks) ? P.spec_a : P.spec_b;
    j.split_blocks = P.split_blocks;
    j.split_digits = P.block_digits;
    j.block_stride = P.block_stride;
    j.row_stride = P.row_stride;
    j.coefficients = P.coefficients;
    j.out_len = P.out_len;
    j.out = P.out;
    j.width = P.width;
    j.magic = P.width_magic;
    j.value_bits = P.value_bits;
    j.block = it.sub;
    j.blocks = P.blocks;
    j.agg_base = P.carry;
    j.agg_thresh = P.carry + P.blocks;
    j.first_thresh = P.carry + 2 * (uint64_t)P.blocks;
    j.ripple_end = P.carry + 3 * (uint64_t)P.blocks;
    j.op_a = P.src_a.limbs;
    j.op_b = P.src_b.limbs;
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
            const double* r1 = J.residues + z1 * J.block_stride + k1;
            const double* r0 = two ? r1 - J.block_stride + D : r1;
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
    const uint64_t* pl = swap ? J.op_b : J.op_a;
    const uint64_t* ps = swap ? J.op_a : J.op_b;
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
// An output of more than kCarryLookBackBlocks blocks first runs k_carry_scan,
// one workgroup per chunk of kCarryChunk blocks (one block per thread): it
// replaces each block's function by the composition of the functions before
// it in its chunk (the first block of a chunk gets the identity, marked by a
// zero threshold, which no carry function has) and publishes the chunk's
// composed function. The apply kernel then composes the functions of the
// chunks before its own (at most a few per lane of its wave) and applies its
// block's prefix: O(blocks) work for any carry pattern, all-ones products
// included. A shorter output skips the scan: each block looks back to the
// nearest block whose function has no threshold (its output is fixed whatever
// comes in) and applies the functions after it, at most kCarryLookBackBlocks
// steps.
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

// Threads of an apply workgroup: one wave finds the carry into the block,
// then all of them add the ripple (up to BL - 1 limbs).
template <int BL>
__host__ __device__ constexpr int apply_threads() {
    return BL >= 1024 ? 256 : 64;
}

template <int BL>
__global__ __launch_bounds__(apply_threads<BL>()) void k_carry_apply(CrtArgs a) {
    constexpr int NT = apply_threads<BL>();
    __shared__ uint32_t found;
    __shared__ uint32_t cin_s;
    const CrtJob J = crt_job(a);
    const int b = (int)J.block;
    if (b == 0) return;
    const int lane = threadIdx.x;  // wave 0 only below
    if (carry_scanned(J.blocks)) {
        if (threadIdx.x < 64) {
            // Carry into this block's chunk: the chunk functions before it,
            // composed in order, lane l taking a contiguous slice (the lanes
            // with a slice are a prefix of the wave).
            const uint32_t j = (uint32_t)b / (uint32_t)kCarryChunk;
            uint32_t cin_chunk = 0;
            if (j > 0) {
                const uint32_t chunks = carry_chunks(J.blocks);
                const uint32_t* cb = J.agg_base + 4ull * J.blocks;
                const uint32_t* ct = cb + chunks;
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
            // The block's prefix within its chunk, from k_carry_scan.
            const uint32_t pb = J.agg_base[b], pt = J.agg_thresh[b];
            if (lane == 0) cin_s = pt == 0 ? cin_chunk : pb + (cin_chunk >= pt ? 1u : 0u);
        }
    } else {
        // b <= kCarryLookBackBlocks = 64: one look-back step by wave 0.
        if (threadIdx.x == 0) found = 0xFFFFFFFFu;
        __syncthreads();
        const int k = b - 1 - lane;
        if (lane < 64 && k >= 0 && J.agg_thresh[k] == kInf) atomicMin(&found, (uint32_t)lane);
        __syncthreads();
        const uint32_t t = found;
        // Block whose output is fixed, or -1: begin with a zero carry into block 0.
        const int start = t != 0xFFFFFFFFu ? b - 1 - (int)t : -1;
        if (threadIdx.x == 0) {
            uint32_t cc = start >= 0 ? J.agg_base[start] : 0u;
            for (int q = start + 1; q < b; ++q) cc = J.agg_base[q] + (cc >= J.agg_thresh[q] ? 1u : 0u);
            cin_s = cc;
        }
    }
    __syncthreads();
    const uint32_t c = cin_s;
    if (c == 0) return;
    const uint64_t L0 = (uint64_t)b * BL;
    if (threadIdx.x == 0 && L0 < J.out_len) J.out[L0] += c;
    if (c >= J.first_thresh[b]) {
        const uint32_t rend = J.ripple_end[b];
        for (uint32_t rel = 1 + threadIdx.x; rel <= rend; rel += NT) {
            const uint64_t j = L0 + rel;
            if (j < J.out_len) J.out[j] += 1;
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

// Lines per strided tile (strided_lines, device.h). Measured with a plain
// tile copy at 2^24 points x 6 rows: 64 x 128 moves 5.6 TB/s at a first
// pass's line stride and 6.0 TB/s at a middle pass's, against 5.1 and 5.35
// for 32 x 128; 32 x 256 moves 5.3-5.5 TB/s against 4.6-4.8 for 16 x 256.

// Threads of a strided tile: one radix-16 group per thread and round.
template <int LOGL, int LINES>
static constexpr int strided_threads() {
    return LINES * (1 << LOGL) / 16 >= 512 ? 512 : 256;
}

// In a batched launch (a.batch set) `units` counts work items (product,
// prime); otherwise it counts primes.
template <int LOGL, int LINES>
static bool launch_fwd_strided_l(const PassArgs& a, int units, int operands, bool pack, hipStream_t s) {
    constexpr int NT = strided_threads<LOGL, LINES>();
    if (a.line_stride < (uint64_t)LINES) return false;
    const uint64_t nodes = (uint64_t)1 << a.level_start;
    const uint64_t tiles = nodes * (a.line_stride / LINES);
    PassArgs b = a;
    b.primes = units;
    b.tiles_per_item = tiles;
    constexpr uint64_t G8 = 8 * (uint64_t)xcd_group(LINES);
    dim3 grid((unsigned)tiles, (unsigned)units, (unsigned)operands);
    if (strided_xcd(pack, LINES)) grid = dim3((unsigned)(((tiles + G8 - 1) / G8) * G8 * units), 1, (unsigned)operands);
    if (a.batch) grid = dim3((unsigned)(tiles * units), 1);
    if (a.batch) {
        if (pack)
            NTT_LAUNCH(nullptr, (k_fwd_strided<LOGL, LINES, NT, true, true>), grid, dim3(NT), s, b);
        else
            NTT_LAUNCH(nullptr, (k_fwd_strided<LOGL, LINES, NT, false, true>), grid, dim3(NT), s, b);
    } else {
        if (pack)
            NTT_LAUNCH(nullptr, (k_fwd_strided<LOGL, LINES, NT, true, false>), grid, dim3(NT), s, b);
        else
            NTT_LAUNCH(nullptr, (k_fwd_strided<LOGL, LINES, NT, false, false>), grid, dim3(NT), s, b);
    }
    return NTT_CHECK_LAUNCH();
}

template <int LOGL>
static bool launch_fwd_strided_t(const PassArgs& a, int units, int operands, bool pack, bool alt, hipStream_t s) {
    return alt ? launch_fwd_strided_l<LOGL, alt_strided_lines(LOGL)>(a, units, operands, pack, s)
               : launch_fwd_strided_l<LOGL, strided_lines(LOGL)>(a, units, operands, pack, s);
}

// First forward pass of a single operation with packing shared across its
// `primes` primes: one workgroup per tile and operand (grid z as in
// k_fwd_strided). Whole-limb and one- to three-part coefficients.
template <int LOGL>
static bool launch_pack_primes_t(const PassArgs& a, int primes, int operands, hipStream_t s) {
    constexpr int LINES = pack_primes_lines<LOGL>();
    constexpr int NT = pack_primes_threads<LOGL>();
    static_assert(LINES * (1 << LOGL) == 8 * NT, "one radix-8 group per thread");
    if (a.line_stride < (uint64_t)LINES || a.level_start != 0 || a.fold_terms != 1 || a.node_base != 0 || a.batch)
        return false;
    PassArgs b = a;
    b.primes = primes;
    const dim3 grid((unsigned)(a.line_stride / LINES), 1, (unsigned)operands);
    switch (pack_kind(a.src_a)) {
        case kPackW64: NTT_LAUNCH(nullptr, (k_pack_primes<LOGL, LINES, NT, kPackW64>), grid, dim3(NT), s, b); break;
        case kPackW64Signed:
            NTT_LAUNCH(nullptr, (k_pack_primes<LOGL, LINES, NT, kPackW64Signed>), grid, dim3(NT), s, b);
            break;
        case kPack1: NTT_LAUNCH(nullptr, (k_pack_primes<LOGL, LINES, NT, kPack1>), grid, dim3(NT), s, b); break;
        case kPack2: NTT_LAUNCH(nullptr, (k_pack_primes<LOGL, LINES, NT, kPack2>), grid, dim3(NT), s, b); break;
        case kPack3: NTT_LAUNCH(nullptr, (k_pack_primes<LOGL, LINES, NT, kPack3>), grid, dim3(NT), s, b); break;
        default: return false;
    }
    return NTT_CHECK_LAUNCH();
}

template <int LOGL, int LINES>
static bool launch_inv_strided_l(const PassArgs& a, int units, bool canon, hipStream_t s) {
    constexpr int NT = strided_threads<LOGL, LINES>();
    if (a.line_stride < (uint64_t)LINES) return false;
    const uint64_t nodes = (uint64_t)1 << a.level_start;
    const uint64_t tiles = nodes * (a.line_stride / LINES);
    PassArgs b = a;
    b.primes = units;
    b.tiles_per_item = tiles;
    constexpr uint64_t G8 = 8 * (uint64_t)xcd_group(LINES);
    const unsigned blocks = a.split_blocks ? a.split_blocks : 1;
    dim3 grid((unsigned)tiles, (unsigned)units, blocks);
    if (strided_xcd(false, LINES)) grid = dim3((unsigned)(((tiles + G8 - 1) / G8) * G8 * units), 1, blocks);
    if (a.batch) grid = dim3((unsigned)(tiles * units), 1);
    if (a.batch) {
        if (canon)
            NTT_LAUNCH(nullptr, (k_inv_strided<LOGL, LINES, NT, true, true>), grid, dim3(NT), s, b);
        else
            NTT_LAUNCH(nullptr, (k_inv_strided<LOGL, LINES, NT, false, true>), grid, dim3(NT), s, b);
    } else {
        if (canon)
            NTT_LAUNCH(nullptr, (k_inv_strided<LOGL, LINES, NT, true, false>), grid, dim3(NT), s, b);
        else
            NTT_LAUNCH(nullptr, (k_inv_strided<LOGL, LINES, NT, false, false>), grid, dim3(NT), s, b);
    }
    return NTT_CHECK_LAUNCH();
}

template <int LOGL>
static bool launch_inv_strided_t(const PassArgs& a, int units, bool canon, bool alt, hipStream_t s) {
    return alt ? launch_inv_strided_l<LOGL, alt_strided_lines(LOGL)>(a, units, canon, s)
               : launch_inv_strided_l<LOGL, strided_lines(LOGL)>(a, units, canon, s);
}

// Shapes of the fused row pass: about 256 radix-16 groups per workgroup
// round (enough rows per workgroup to fill it), one thread per group; the
// radix-8 form (FORM 2) keeps the tile and doubles the threads.
template <int LOGL, bool PACK, bool SQUARE, int FORM = 0>
struct FusedShape {
    static constexpr int TILES = SQUARE ? 1 : 2;
    static constexpr int GROUPS = TILES * (1 << LOGL) / 16;  // per row, per round
    static constexpr int LINES = PACK ? 1 : (GROUPS >= 256 ? 1 : (256 / GROUPS > 32 ? 32 : 256 / GROUPS));
    static constexpr int NT_RAW = GROUPS * LINES * (FORM == 2 ? 2 : 1);
    static constexpr int NT = NT_RAW < 64 ? 64 : (NT_RAW > 512 ? 512 : NT_RAW);
};

template <int LOGL, bool PACK, bool SQUARE, int FORM>
static bool launch_fused_v(const PassArgs& a, int units, hipStream_t s) {
    using F = FusedShape<LOGL, PACK, SQUARE, FORM>;
    const uint64_t rows = (uint64_t)1 << a.level_start;
    if (rows % F::LINES != 0) return false;
    const uint64_t tiles = rows / F::LINES;
    PassArgs b = a;
    b.tiles_per_item = tiles;
    dim3 grid((unsigned)tiles, (unsigned)units, a.split_blocks ? a.split_blocks : 1);
    if (a.batch) {
        grid = dim3((unsigned)(tiles * units), 1);
        NTT_LAUNCH(nullptr, (k_fused<LOGL, F::LINES, F::NT, PACK, SQUARE, true, FORM>), grid, dim3(F::NT), s, b);
    } else {
        NTT_LAUNCH(nullptr, (k_fused<LOGL, F::LINES, F::NT, PACK, SQUARE, false, FORM>), grid, dim3(F::NT), s, b);
    }
    return NTT_CHECK_LAUNCH();
}

// form: k_fused's FORM (1, radix-16 inverse rows, applies to products only).
template <int LOGL, bool SQUARE>
static bool launch_fused_f(const PassArgs& a, int units, bool pack, int form, hipStream_t s) {
    if (form == 2)
        return pack ? launch_fused_v<LOGL, true, SQUARE, 2>(a, units, s)
                    : launch_fused_v<LOGL, false, SQUARE, 2>(a, units, s);
    if (form == 1 && !SQUARE)
        return pack ? launch_fused_v<LOGL, true, SQUARE, 1>(a, units, s)
                    : launch_fused_v<LOGL, false, SQUARE, 1>(a, units, s);
    return pack ? launch_fused_v<LOGL, true, SQUARE, 0>(a, units, s)
                : launch_fused_v<LOGL, false, SQUARE, 0>(a, units, s);
}

template <int LOGL>
static bool launch_fused_t(const PassArgs& a, int units, bool pack, bool square, int form, hipStream_t s) {
    return square ? launch_fused_f<LOGL, true>(a, units, pack, form, s)
                  : launch_fused_f<LOGL, false>(a, units, pack, form, s);
}

bool launch_fwd_strided(int logl, const PassArgs& a, int primes, int operands, bool pack, bool alt, hipStream_t s) {
    switch (logl) {
        case 5: return launch_fwd_strided_t<5>(a, primes, operands, pack, alt, s);
        case 6: return launch_fwd_strided_t<6>(a, primes, operands, pack, alt, s);
        case 7: return launch_fwd_strided_t<7>(a, primes, operands, pack, alt, s);
        case 8: return launch_fwd_strided_t<8>(a, primes, operands, pack, alt, s);
        case 9: return launch_fwd_strided_t<9>(a, primes, operands, pack, alt, s);
        case 10: return launch_fwd_strided_t<10>(a, primes, operands, pack, alt, s);
        case 11: return launch_fwd_strided_t<11>(a, primes, operands, pack, alt, s);
        default: return false;
    }
}

// False when the pass cannot take the shared form (the caller then packs
// per prime): a batch, a folded or tail part, four-part coefficients, a
// source kind other than the second operand's, or lines of other than 2^6 to
// 2^9 points (4096-point tiles).
bool launch_pack_primes(int logl, const PassArgs& a, int primes, int operands, hipStream_t s) {
    if (pack_kind(a.src_a) != pack_kind(a.src_b) || pack_kind(a.src_a) == kPack4) return false;
    switch (logl) {
        case 5: return launch_pack_primes_t<5>(a, primes, operands, s);
        case 6: return launch_pack_primes_t<6>(a, primes, operands, s);
        case 7: return launch_pack_primes_t<7>(a, primes, operands, s);
        case 8: return launch_pack_primes_t<8>(a, primes, operands, s);
        case 9: return launch_pack_primes_t<9>(a, primes, operands, s);
        default: return false;
    }
}

// Batched shared packing: `items` (product, block, operand) units, every
// product of kind below kPack4 (pack_primes_batch_fits checks the shape).
template <int LOGL>
static bool launch_pack_primes_batch_t(const PassArgs& a, int items, hipStream_t s) {
    constexpr int LINES = pack_primes_lines<LOGL>();
    constexpr int NT = pack_primes_threads<LOGL>();
    if (a.line_stride < (uint64_t)LINES || a.level_start != 0 || !a.batch) return false;
    PassArgs b = a;
    b.tiles_per_item = a.line_stride / LINES;
    const dim3 grid((unsigned)(b.tiles_per_item * (uint64_t)items), 1);
    NTT_LAUNCH(nullptr, (k_pack_primes_batch<LOGL, LINES, NT>), grid, dim3(NT), s, b);
    return NTT_CHECK_LAUNCH();
}

bool pack_primes_batch_fits(int logl, uint64_t line_stride) {
    return logl >= 5 && logl <= 9 && line_stride >= (uint64_t)(4096 >> logl);
}

bool launch_pack_primes_batch(int logl, const PassArgs& a, int items, hipStream_t s) {
    switch (logl) {
        case 5: return launch_pack_primes_batch_t<5>(a, items, s);
        case 6: return launch_pack_primes_batch_t<6>(a, items, s);
        case 7: return launch_pack_primes_batch_t<7>(a, items, s);
        case 8: return launch_pack_primes_batch_t<8>(a, items, s);
        case 9: return launch_pack_primes_batch_t<9>(a, items, s);
        default: return false;
    }
}

bool launch_inv_strided(int logl, const PassArgs& a, int primes, bool canon, bool alt, hipStream_t s) {
    switch (logl) {
        case 5: return launch_inv_strided_t<5>(a, primes, canon, alt, s);
        case 6: return launch_inv_strided_t<6>(a, primes, canon, alt, s);
        case 7: return launch_inv_strided_t<7>(a, primes, canon, alt, s);
        case 8: return launch_inv_strided_t<8>(a, primes, canon, alt, s);
        case 9: return launch_inv_strided_t<9>(a, primes, canon, alt, s);
        case 10: return launch_inv_strided_t<10>(a, primes, canon, alt, s);
        case 11: return launch_inv_strided_t<11>(a, primes, canon, alt, s);
        default: return false;
    }
}


bool launch_fused(int logl, const PassArgs& a, int primes, bool pack, bool square, int form, hipStream_t s) {
    switch (logl) {
        case 5: return launch_fused_t<5>(a, primes, pack, square, form, s);
        case 6: return launch_fused_t<6>(a, primes, pack, square, form, s);
        case 7: return launch_fused_t<7>(a, primes, pack, square, form, s);
        case 8: return launch_fused_t<8>(a, primes, pack, square, form, s);
        case 9: return launch_fused_t<9>(a, primes, pack, square, form, s);
        case 10: return launch_fused_t<10>(a, primes, pack, square, form, s);
        case 11: return launch_fused_t<11>(a, primes, pack, square, form, s);
        case 12: return launch_fused_t<12>(a, primes, pack, square, form, s);
        default: return false;
    }
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
    if (a.blocks > 1 || a.batch)
        NTT_LAUNCH("carry_apply", k_carry_apply<BL>, dim3(a.blocks), dim3(apply_threads<BL>()), s, a);
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
