// Device code shared by the gfx950 kernels of the multi-prime NTT engine.
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
#pragma once

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

// Global memory. Pointers that reach a kernel inside structures (a batch's
// products, a reconstruction's job) are generic to the compiler, and generic
// (flat) accesses count against the LDS counter as well as the vector memory
// counter, so each LDS wait after them would also wait for them (a store
// included). Every pointer the kernels read or write rows through addresses
// global memory, so the tile movements below cast to it.
#define NTT_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ NTT_GLOBAL T* global_ptr(T* p) {
    return (NTT_GLOBAL T*)p;
}

// Padded LDS index: one pad word after every 16 values. The last round of a
// pass gives thread t the points 16t..16t+15 (or 8t..8t+7 in radix 8), so
// the 16 lanes of one LDS cycle read values 16 (or 8) apart; with a pad per
// 16 values they fall on 16 distinct bank pairs (a pad per 32 values left
// pairs of lanes on one bank pair). Consecutive points and the 2-value
// stripes of row loads stay conflict-free.
__host__ __device__ constexpr int pidx(int e) { return e + (e >> 4); }

template <int LOGL>
struct Lay {
    static constexpr int L = 1 << LOGL;
    static constexpr int STRIDE = L + (L >> 4) + 1;  // padded line in LDS (odd: lines fall on distinct banks)
    static constexpr int ROUNDS = (LOGL + 3) / 4;
    static constexpr int FIRST = LOGL - 4 * (ROUNDS - 1);  // levels of round 0
};

// pidx(m0 + m) = pidx(m0) + pidx(m) when m is a multiple of a power of two
// M and m0 < M, or when m0 is a multiple of 16: the pads before m0 + m are
// then those before m0 plus m >> 4. The LDS walks split every index that
// way, so a thread computes the variable part once and the constant part of
// each access folds into the instruction's offset (the compiler does not
// see through the shift of a sum on its own).
//
// Offset of point f of a tile of rows: row f >> LOGL, point f mod 2^LOGL.
// row_off(f0 + s) = row_off(f0) + row_off(s) for s a multiple of a power of
// two S > f0.
template <int LOGL>
__host__ __device__ constexpr int row_off(int f) {
    return (f >> LOGL) * Lay<LOGL>::STRIDE + pidx(f & ((1 << LOGL) - 1));
}
// Offset of element f of a strided tile: line f mod LINES, point f / LINES.
// col_off(f0 + s) = col_off(f0) + col_off(s) for s a multiple of a power of
// two S > f0.
template <int LOGL, int LINES>
__host__ __device__ constexpr int col_off(int f) {
    return (f % LINES) * Lay<LOGL>::STRIDE + pidx(f / LINES);
}

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
    // Tables are read through the constant address space: no pass writes
    // them, so a root every lane of a wave shares becomes one scalar load,
    // which waits on its own counter instead of behind the wave's earlier
    // vector loads and stores (those complete in issue order).
    __device__ __forceinline__ double ld(uint64_t i) const {
        return ((const __attribute__((address_space(4))) double*)table)[i];
    }
    __device__ __forceinline__ double at(int c, int lvl, int node) const {
        const uint64_t h = (h0 + (uint64_t)((c & cmask) * hstep)) << lvl;
        if (lvl >= split) return mulmod(ld(node), ld(h), p, pinv);
        return ld(h + node);
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
// DIF butterflies of one radix-2^RLOG group in registers: x[k] is point
// base + k*D of line c, whose node at local level l0 + s is (b << s) + blk.
// NV groups (x, then y) of the same nodes share every root. R1: level 0 of
// the group is the transform's first level at table node 0 (root 1).
template <int RLOG, bool R1, int NV = 1>
__device__ __forceinline__ void fwd_group(double* x, double* y, const Tw& tw, int c, int l0, int b, double p,
                                          double pinv) {
    constexpr int RADIX = 1 << RLOG;
#pragma unroll
    for (int s = 0; s < RLOG; ++s) {
        const int half = RADIX >> (s + 1);
#pragma unroll
        for (int blk = 0; blk < (1 << s); ++blk) {
            const double w = (R1 && s == 0) ? 1.0 : tw.at(c, l0 + s, (b << s) + blk);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                double* z = v == 0 ? x : y;
#pragma unroll
                for (int j = 0; j < half; ++j) {
                    const int k = blk * 2 * half + j;
                    const double u = s == RLOG - 1 ? reduce(z[k], p, pinv) : z[k];
                    const double t = (R1 && s == 0) ? z[k + half] : mulmod(z[k + half], w, p, pinv);
                    z[k] = u + t;
                    z[k + half] = u - t;
                }
            }
        }
    }
}

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
        // base = b * (L >> L0) + i with i < D: point base + k*D sits at
        // pidx(base) + pidx(k*D).
        double* line = lds + c * LS + pidx(b * (L >> L0) + i);
        double x[RADIX];
#pragma unroll
        for (int k = 0; k < RADIX; ++k) x[k] = line[pidx(k * D)];
        fwd_group<RLOG, ROOT1 && L0 == 0>(x, nullptr, tw, c, L0, b, p, pinv);
#pragma unroll
        for (int k = 0; k < RADIX; ++k) line[pidx(k * D)] = x[k];
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
template <int RLOG, bool R1>
__device__ __forceinline__ void inv_group(double* x, const Tw& tw, int c, int l0, int b, double p, double pinv) {
    constexpr int RADIX = 1 << RLOG;
#pragma unroll
    for (int s = RLOG - 1; s >= 0; --s) {
        const int half = RADIX >> (s + 1);
#pragma unroll
        for (int blk = 0; blk < (1 << s); ++blk) {
            const double w = (R1 && s == 0) ? 1.0 : tw.at(c, l0 + s, (b << s) + blk);
#pragma unroll
            for (int j = 0; j < half; ++j) {
                const int k = blk * 2 * half + j;
                const double u = x[k], v = x[k + half];
                const double sum = u + v;
                x[k] = (s & 1) ? sum : reduce(sum, p, pinv);
                x[k + half] = (R1 && s == 0) ? reduce(u - v, p, pinv) : mulmod(u - v, w, p, pinv);
            }
        }
    }
}

// Roots of one radix-16 group loaded ahead of its round (so that their
// latency overlaps the work before it): w[(1 << s) - 1 + blk] is the root of
// node (b << s) + blk at local level l0 + s of line c.
__device__ __forceinline__ void group_roots16(double* w, const Tw& tw, int c, int l0, int b) {
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int blk = 0; blk < (1 << s); ++blk) w[(1 << s) - 1 + blk] = tw.at(c, l0 + s, (b << s) + blk);
}

// fwd_group / inv_group of a radix-16 group whose roots were loaded ahead.
__device__ __forceinline__ void fwd_group16_w(double* x, const double* w, double p, double pinv) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int half = 16 >> (s + 1);
#pragma unroll
        for (int blk = 0; blk < (1 << s); ++blk) {
#pragma unroll
            for (int j = 0; j < half; ++j) {
                const int k = blk * 2 * half + j;
                const double u = s == 3 ? reduce(x[k], p, pinv) : x[k];
                const double t = mulmod(x[k + half], w[(1 << s) - 1 + blk], p, pinv);
                x[k] = u + t;
                x[k + half] = u - t;
            }
        }
    }
}

__device__ __forceinline__ void inv_group16_w(double* x, const double* w, double p, double pinv) {
#pragma unroll
    for (int s = 3; s >= 0; --s) {
        const int half = 16 >> (s + 1);
#pragma unroll
        for (int blk = 0; blk < (1 << s); ++blk) {
#pragma unroll
            for (int j = 0; j < half; ++j) {
                const int k = blk * 2 * half + j;
                const double u = x[k], v = x[k + half];
                const double sum = u + v;
                x[k] = (s & 1) ? sum : reduce(sum, p, pinv);
                x[k + half] = mulmod(u - v, w[(1 << s) - 1 + blk], p, pinv);
            }
        }
    }
}

// A radix-16 round (local levels L0..L0+3, not the transform's first or
// last level) whose roots the thread loaded ahead for its first group; every
// further group the thread takes (g + NT, g + 2 NT, ...) must sit at the same
// nodes, which holds when NT is a multiple of GPL * TILE_LINES (lines of one
// tile of rows share nodes with the same lines of the other tile).
template <int LOGL, int LINES, int NT, int L0, int TILE_LINES, bool INV>
__device__ __forceinline__ void round16_w(double* lds, const double* w, double p, double pinv) {
    constexpr int L = 1 << LOGL;
    constexpr int D = L >> (L0 + 4);
    constexpr int GPL = L / 16;
    constexpr int GROUPS = LINES * GPL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    static_assert(NT >= GROUPS || NT % (GPL * TILE_LINES) == 0, "later groups of a thread sit at other nodes");
    for (int g = threadIdx.x; g < GROUPS; g += NT) {
        const int c = g / GPL;
        const int gl = g - c * GPL;
        const int b = gl / D;
        const int i = gl - b * D;
        // base = b * (L >> L0) + i with i < D: point base + k*D sits at
        // pidx(base) + pidx(k*D).
        double* line = lds + c * LS + pidx(b * (L >> L0) + i);
        double x[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) x[k] = line[pidx(k * D)];
        if constexpr (INV)
            inv_group16_w(x, w, p, pinv);
        else
            fwd_group16_w(x, w, p, pinv);
#pragma unroll
        for (int k = 0; k < 16; ++k) line[pidx(k * D)] = x[k];
    }
}

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
        // base = b * (L >> L0) + i with i < D: point base + k*D sits at
        // pidx(base) + pidx(k*D).
        double* line = lds + c * LS + pidx(b * (L >> L0) + i);
        double x[RADIX];
#pragma unroll
        for (int k = 0; k < RADIX; ++k) x[k] = line[pidx(k * D)];
        inv_group<RLOG, ROOT1 && L0 == 0>(x, tw, c, L0, b, p, pinv);
#pragma unroll
        for (int k = 0; k < RADIX; ++k) line[pidx(k * D)] = x[k];
    }
}

// Rounds R.. of a forward pass; root1 (uniform): the pass starts the
// transform at table node 0 from packed inputs (fwd_round's ROOT1).
// Rounds R..END-1 of a forward pass (END: all of them).
template <int LOGL, int LINES, int NT, int R, int END = Lay<LOGL>::ROUNDS>
__device__ __forceinline__ void fwd_rounds(double* lds, Tw tw, double p, double pinv, bool root1 = false) {
    if constexpr (R < END) {
        constexpr int L0 = R == 0 ? 0 : Lay<LOGL>::FIRST + 4 * (R - 1);
        constexpr int RL = R == 0 ? Lay<LOGL>::FIRST : 4;
        if (R == 0 && root1)
            fwd_round<LOGL, LINES, NT, L0, RL, true>(lds, tw, p, pinv);
        else
            fwd_round<LOGL, LINES, NT, L0, RL, false>(lds, tw, p, pinv);
        __syncthreads();
        fwd_rounds<LOGL, LINES, NT, R + 1, END>(lds, tw, p, pinv);
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
// Tile movement. Every loop below has a compile-time trip count and is fully
// unrolled, so a thread issues all of its global loads before it waits on
// the first one.

// Strided tile: LINES consecutive lines lo0.. of one node; element m of line c
// at g[m*S + c]. Consecutive threads take consecutive lines, so each group of
// LINES threads reads one LINES*8-byte run.
template <int LOGL, int LINES, int NT>
__device__ __forceinline__ void load_strided(double* lds, const double* g0, uint64_t S) {
    const NTT_GLOBAL double* g = global_ptr(g0);
    constexpr int L = 1 << LOGL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    constexpr int TOTAL = LINES * L;
    constexpr int R = (TOTAL + NT - 1) / NT;
    const int t0 = col_off<LOGL, LINES>(threadIdx.x);
    double v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int f = threadIdx.x + r * NT;
        if (TOTAL % NT == 0 || f < TOTAL) v[r] = g[(uint64_t)(f / LINES) * S + (f % LINES)];
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int f = threadIdx.x + r * NT;
        if (TOTAL % NT == 0 || f < TOTAL) lds[t0 + col_off<LOGL, LINES>(r * NT)] = v[r];
    }
}

template <int LOGL, int LINES, int NT, bool CANON>
__device__ __forceinline__ void store_strided(const double* lds, double* g0, uint64_t S, double tinv, double p,
                                              double pinv) {
    NTT_GLOBAL double* g = global_ptr(g0);
    constexpr int L = 1 << LOGL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    constexpr int TOTAL = LINES * L;
    constexpr int R = (TOTAL + NT - 1) / NT;
    const int t0 = col_off<LOGL, LINES>(threadIdx.x);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int f = threadIdx.x + r * NT;
        if (TOTAL % NT != 0 && f >= TOTAL) continue;
        double v = lds[t0 + col_off<LOGL, LINES>(r * NT)];
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
    const int t0 = col_off<LOGL, LINES>(threadIdx.x);
    if (n0 + (uint64_t)(L - 1) * S + (LINES - 1) < safe) {
        // The whole tile lies inside the operand (one test for the tile):
        // every load is issued before the first result is needed.
        double v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int f = threadIdx.x + r * NT;
            if (TOTAL % NT == 0 || f < TOTAL)
                v[r] = pack_as<KIND, false>(src, n0 + (uint64_t)(f / LINES) * S + (f % LINES), pw, p, pinv);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int f = threadIdx.x + r * NT;
            if (TOTAL % NT == 0 || f < TOTAL) lds[t0 + col_off<LOGL, LINES>(r * NT)] = v[r];
        }
        return;
    }
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
            if (TOTAL % NT == 0 || f < TOTAL) lds[t0 + col_off<LOGL, LINES>((r0 + r) * NT)] = v[r];
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

// Two doubles moved with one 16-byte access (a native vector type, which
// loads and stores through a global pointer as one value).
typedef double Pair __attribute__((ext_vector_type(2)));

// Row tile: LINES rows of L contiguous points; thread t moves points 2t, 2t+1
// of each 2*NT-point stripe with 16-byte accesses.
template <int LOGL, int LINES, int NT>
__device__ __forceinline__ void load_rows(double* lds, const double* g) {
    constexpr int L = 1 << LOGL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    constexpr int PAIRS = LINES * L / 2;
    constexpr int R = (PAIRS + NT - 1) / NT;
    const NTT_GLOBAL Pair* gp = global_ptr(reinterpret_cast<const Pair*>(g));
    Pair v[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (PAIRS % NT == 0 || threadIdx.x + r * NT < PAIRS) v[r] = gp[threadIdx.x + r * NT];
    // Point 2t + 2rNT at row_off(2t) + row_off(2rNT); 2t is even, so 2t + 1
    // sits right after it.
    double* t0 = lds + row_off<LOGL>(2 * threadIdx.x);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (PAIRS % NT != 0 && threadIdx.x + r * NT >= PAIRS) continue;
        t0[row_off<LOGL>(2 * r * NT)] = v[r].x;
        t0[row_off<LOGL>(2 * r * NT) + 1] = v[r].y;
    }
}

template <int LOGL, int LINES, int NT>
__device__ __forceinline__ void store_rows(const double* lds, double* g) {
    constexpr int L = 1 << LOGL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    constexpr int PAIRS = LINES * L / 2;
    constexpr int R = (PAIRS + NT - 1) / NT;
    const double* t0 = lds + row_off<LOGL>(2 * threadIdx.x);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (PAIRS % NT != 0 && threadIdx.x + r * NT >= PAIRS) continue;
        Pair v;
        v.x = t0[row_off<LOGL>(2 * r * NT)];
        v.y = t0[row_off<LOGL>(2 * r * NT) + 1];
        global_ptr(reinterpret_cast<Pair*>(g))[threadIdx.x + r * NT] = v;
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

// Forward rounds from local level L0 to the end of the line, MAXR levels per
// round (the last one shorter).
// root1 (uniform, L0 == 0 only): as fwd_rounds.
template <int LOGL, int LINES, int NT, int L0, int MAXR>
__device__ __forceinline__ void fwd_rounds_from(double* lds, Tw tw, double p, double pinv, bool root1 = false) {
    if constexpr (L0 < LOGL) {
        constexpr int RL = LOGL - L0 < MAXR ? LOGL - L0 : MAXR;
        if (L0 == 0 && root1)
            fwd_round<LOGL, LINES, NT, L0, RL, true>(lds, tw, p, pinv);
        else
            fwd_round<LOGL, LINES, NT, L0, RL, false>(lds, tw, p, pinv);
        __syncthreads();
        fwd_rounds_from<LOGL, LINES, NT, L0 + RL, MAXR>(lds, tw, p, pinv);
    }
}

// Row passes form the roots of their last four levels as products (every
// row is its own node, so reading them would cost one table entry per point
// at the last level; the products read W[x] from the short, cache-resident
// prefix of the table and one W[H 2^l] per row and level).
__host__ __device__ constexpr int row_split(int logl) { return logl > 4 ? logl - 4 : 0; }

#define NTT_CHECK_LAUNCH() (hipGetLastError() == hipSuccess)

}  // namespace nttamd
