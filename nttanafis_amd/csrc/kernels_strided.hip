// Strided passes of the transforms (packing, forward, inverse) and their launchers.
#include "passes.h"

namespace nttamd {

// ---------------------------------------------------------------------------
// Strided passes: the tile is LINES consecutive lines lo0..lo0+LINES-1 of
// first node H; element m of line lo sits at H*(T >> level_start) + m*S + lo.

// R8: radix-8 rounds with one thread per radix-8 group (twice the threads of
// the radix-16 form, half the values in flight per thread).
template <int LOGL, int LINES, int NT, bool PACK, bool BATCH, bool R8 = false>
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
    if constexpr (R8)
        fwd_rounds_from<LOGL, LINES, NT, 0, 3>(lds, tw, p, pinv, PACK && a.node_base == 0);
    else
        fwd_rounds<LOGL, LINES, NT, 0>(lds, tw, p, pinv, PACK && a.node_base == 0);
    store_strided<LOGL, LINES, NT, false>(lds, row + base, S, 0.0, p, pinv);
}

// Strided pass of 2^5 or 2^6-point lines in wave-owned lines: each wave
// takes WL = 1024 / L whole lines of the tile (16 values per lane), so the
// two radix rounds exchange values through LDS within the wave only and the
// pass runs without a workgroup barrier. Lane (c, q) of a wave (line c of
// its WL lines, q < L/16) holds in the round of the first FIRST = LOGL - 4
// levels the groups i = q + QW*jj (points i + 16k) and in the round of the
// last four levels the group of points 16q..16q+15; both rounds read or
// write global memory directly, WL consecutive lines (WL * 8 bytes) per
// row of the tile per instruction, and the LDS only carries the transpose
// between them. MODE 0: forward (no packing), 2: inverse, 3: inverse ending
// the transform (times T^-1, canonical residues). Bounds as fwd_round and
// inv_round (a forward round entered below 31p/16, an inverse round below
// 11p/8).
// Orders a wave's LDS stores before its later LDS loads for every lane: the
// hardware runs one wave's LDS instructions in order, and the fences keep the
// compiler from moving the loads above the stores.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int LOGL, int WAVES, int MODE>
__global__ __launch_bounds__(64 * WAVES) void k_strided_waves(PassArgs a) {
    constexpr int L = 1 << LOGL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    constexpr int FIRST = LOGL - 4;
    constexpr int RA = 1 << FIRST;       // points of a first-round group
    constexpr int WL = 1024 / L;         // lines per wave
    constexpr int QW = L / 16;           // lanes per line
    constexpr int GA = 16 / QW;          // first-round groups per lane
    constexpr int LINES = WL * WAVES;
    static_assert(LOGL == 5 || LOGL == 6, "two rounds with 16 first-round groups per line");
    constexpr bool INV = MODE >= 2, CANON = MODE == 3;
    __shared__ double lds[LINES * LS];
    const uint64_t S = a.line_stride;
    const uint64_t tiles_per_node = S / LINES;
    Job j;
    if (!resolve<false, xcd_group(LINES)>(a, tiles_per_node << a.level_start, strided_xcd(false, LINES), &j))
        return;
    const int y = j.y;
    const double p = a.field.p[y], pinv = a.field.pinv[y];
    const uint64_t H = j.tile / tiles_per_node;
    const uint64_t lo0 = (j.tile % tiles_per_node) * LINES;
    const uint64_t span = (uint64_t)1 << (a.log_t - a.level_start);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int cl = lane % WL, q = lane / WL;
    const int c = wave * WL + cl;  // line of the tile
    NTT_GLOBAL double* g = global_ptr(j.data + (uint64_t)y * a.row_stride + H * span + lo0 + c);  // point m at g[m * S]
    double* line = lds + c * LS;
    const Tw tw{(INV ? a.inv : a.fwd) + (uint64_t)y * a.table_stride, H + (a.node_base << a.level_start), 0, 0, 64, p,
                pinv};
    double x[16];
    if constexpr (!INV) {
#pragma unroll
        for (int jj = 0; jj < GA; ++jj)
#pragma unroll
            for (int k = 0; k < RA; ++k) x[jj * RA + k] = g[(uint64_t)(q + QW * jj + 16 * k) * S];
#pragma unroll
        for (int jj = 0; jj < GA; ++jj) {
            fwd_group<FIRST, false>(x + jj * RA, nullptr, tw, c, 0, 0, p, pinv);
#pragma unroll
            for (int k = 0; k < RA; ++k) line[q + QW * jj + 17 * k] = x[jj * RA + k];
        }
        wave_lds_sync();
#pragma unroll
        for (int k = 0; k < 16; ++k) x[k] = line[17 * q + k];
        fwd_group<4, false>(x, nullptr, tw, c, FIRST, q, p, pinv);
#pragma unroll
        for (int k = 0; k < 16; ++k) g[(uint64_t)(16 * q + k) * S] = x[k];
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k) x[k] = g[(uint64_t)(16 * q + k) * S];
        inv_group<4, false>(x, tw, c, FIRST, q, p, pinv);
#pragma unroll
        for (int k = 0; k < 16; ++k) line[17 * q + k] = x[k];
        wave_lds_sync();
        const bool root1 = CANON && a.node_base == 0;
        const double tinv = a.tinv[y];
#pragma unroll
        for (int jj = 0; jj < GA; ++jj) {
            double* z = x + jj * RA;
#pragma unroll
            for (int k = 0; k < RA; ++k) z[k] = line[q + QW * jj + 17 * k];
            if (root1)
                inv_group<FIRST, true>(z, tw, c, 0, 0, p, pinv);
            else
                inv_group<FIRST, false>(z, tw, c, 0, 0, p, pinv);
#pragma unroll
            for (int k = 0; k < RA; ++k) {
                double v = z[k];
                if constexpr (CANON) v = canonical(mulmod(v, tinv, p, pinv), p);
                g[(uint64_t)(q + QW * jj + 16 * k) * S] = v;
            }
        }
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
// Last forward round of a strided tile (local levels L0..LOGL-1) read from
// LDS and written straight to global memory (point m of line c at
// g[m * S + c]): thread t takes line t % LINES, so each store instruction
// writes LINES consecutive points of one row of the tile.
template <int LOGL, int LINES, int NT, int L0>
__device__ __forceinline__ void last_round_direct(const double* lds, const Tw& tw, double* g0, uint64_t S, double p,
                                                  double pinv) {
    NTT_GLOBAL double* g = global_ptr(g0);
    constexpr int L = 1 << LOGL;
    constexpr int RL = LOGL - L0;
    constexpr int RADIX = 1 << RL;
    constexpr int GPL = L / RADIX;  // groups per line; D = 1
    constexpr int LS = Lay<LOGL>::STRIDE;
    constexpr int TPL = NT / LINES;  // threads per line
    static_assert(NT % LINES == 0, "whole lines per thread set");
    const int c = threadIdx.x % LINES;
    const double* line = lds + c * LS;
#pragma unroll
    for (int b0 = threadIdx.x / LINES; b0 < GPL; b0 += TPL) {
        // With LINES a multiple of the wave size every lane of a wave takes
        // the same group, so its roots are one scalar load: they then wait
        // on the scalar counter, not behind the vector stores of the tile's
        // earlier primes (vector memory operations complete in issue order).
        const int b = LINES % 64 == 0 ? __builtin_amdgcn_readfirstlane(b0) : b0;
        double x[RADIX];
#pragma unroll
        for (int k = 0; k < RADIX; ++k) x[k] = line[pidx(b * RADIX) + k];
        fwd_group<RL, false>(x, nullptr, tw, c, L0, b, p, pinv);
#pragma unroll
        for (int k = 0; k < RADIX; ++k) g[(uint64_t)(b * RADIX + k) * S + c] = x[k];
    }
}

// Levels L0..LOGL-1 of a packing tile after its first round: radix-8 rounds
// through LDS, the last one straight to global memory.
template <int LOGL, int LINES, int NT, int L0>
__device__ __forceinline__ void pack_rest(double* lds, const Tw& tw, double* g, uint64_t S, double p, double pinv) {
    if constexpr (LOGL - L0 > 3) {
        fwd_round<LOGL, LINES, NT, L0, 3, false>(lds, tw, p, pinv);
        __syncthreads();
        pack_rest<LOGL, LINES, NT, L0 + 3>(lds, tw, g, S, p, pinv);
    } else {
        last_round_direct<LOGL, LINES, NT, L0>(lds, tw, g, S, p, pinv);
    }
}

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
        // Primes alternate between two LDS tiles: a prime's first barrier
        // (after its stores below) also tells that every thread has read the
        // tile of the prime before last, so no barrier closes a prime.
        double* buf = lds + (y & 1) * LINES * LS;
        double* line = buf + c * LS;
#pragma unroll
        for (int k = 0; k < 8; ++k) line[pidx(m0) + pidx(k * D)] = x[k];
        __syncthreads();
        // Radix-8 rounds, one group per thread, eight values in flight beside
        // the extracted parts; the last one stores to global memory directly.
        pack_rest<LOGL, LINES, NT, 3>(buf, tw, j.data + (uint64_t)y * a.row_stride + base, S, p, pinv);
    }
}

template <int LOGL, int LINES, int NT, int KIND>
__global__ __launch_bounds__(NT) void k_pack_primes(PassArgs a) {
    __shared__ double lds[2 * LINES * Lay<LOGL>::STRIDE];  // two tiles (pack_primes_tile)
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
    __shared__ double lds[2 * LINES * Lay<LOGL>::STRIDE];  // two tiles (pack_primes_tile)
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
template <int LOGL, int LINES, int NT, bool CANON, bool BATCH, bool R8 = false>
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
    if constexpr (R8)
        inv_rounds8<LOGL, LINES, NT, Lay8<LOGL>::ROUNDS - 1>(lds, tw, p, pinv, CANON && a.node_base == 0);
    else
        inv_rounds<LOGL, LINES, NT, Lay<LOGL>::ROUNDS - 1>(lds, tw, p, pinv, CANON && a.node_base == 0);
    store_strided<LOGL, LINES, NT, CANON>(lds, row + base, S, a.tinv[y], p, pinv);
}

// Lines per strided tile (strided_lines, device.h). Measured with a plain
// tile copy at 2^24 points x 6 rows: 64 x 128 moves 5.6 TB/s at a first
// pass's line stride and 6.0 TB/s at a middle pass's, against 5.1 and 5.35
// for 32 x 128; 32 x 256 moves 5.3-5.5 TB/s against 4.6-4.8 for 16 x 256.

// Threads of a strided tile: one radix-16 group per thread and round.
template <int LOGL, int LINES>
static constexpr int strided_threads() {
    return LINES * (1 << LOGL) / 16 >= 512 ? 512 : 256;
}

// Threads of the radix-8 form: one radix-8 group per thread and round.
template <int LOGL, int LINES>
static constexpr int strided_threads8() {
    return LINES * (1 << LOGL) / 8 >= 1024 ? 1024 : (LINES * (1 << LOGL) / 8 <= 256 ? 256 : LINES * (1 << LOGL) / 8);
}
// Line lengths whose radix-8 rounds are as many as the radix-16 rounds.
__host__ __device__ constexpr bool r8_fits(int logl) { return logl == 5 || logl == 6 || logl == 9; }

// In a batched launch (a.batch set) `units` counts work items (product,
// prime); otherwise it counts primes.
template <int LOGL, int LINES>
static bool launch_fwd_strided_l(const PassArgs& a, int units, int operands, bool pack, hipStream_t s,
                                 bool r8 = false) {
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
    } else if (r8 && !pack) {
        if constexpr (r8_fits(LOGL)) {
            constexpr int NT8 = strided_threads8<LOGL, LINES>();
            NTT_LAUNCH(nullptr, (k_fwd_strided<LOGL, LINES, NT8, false, false, true>), grid, dim3(NT8), s, b);
        }
    } else {
        if (pack)
            NTT_LAUNCH(nullptr, (k_fwd_strided<LOGL, LINES, NT, true, false>), grid, dim3(NT), s, b);
        else
            NTT_LAUNCH(nullptr, (k_fwd_strided<LOGL, LINES, NT, false, false>), grid, dim3(NT), s, b);
    }
    return NTT_CHECK_LAUNCH();
}

// Wave-owned lines (k_strided_waves, four waves per workgroup) for lines of
// 2^5 and 2^6 points; mode as the kernel's.
__host__ __device__ constexpr bool waves_fit(int logl) { return logl == 5 || logl == 6; }
template <int LOGL, int WAVES>
static bool launch_waves_w(const PassArgs& a, int units, int z, int mode, hipStream_t s) {
    if constexpr (waves_fit(LOGL)) {
        constexpr int LINES = 1024 / (1 << LOGL) * WAVES;
        if (a.line_stride < (uint64_t)LINES || a.batch) return false;
        const uint64_t tiles = ((uint64_t)1 << a.level_start) * (a.line_stride / LINES);
        PassArgs b = a;
        b.primes = units;
        b.tiles_per_item = tiles;
        const dim3 grid((unsigned)tiles, (unsigned)units, (unsigned)z);
        if (mode == 0)
            NTT_LAUNCH(nullptr, (k_strided_waves<LOGL, WAVES, 0>), grid, dim3(64 * WAVES), s, b);
        else if (mode == 2)
            NTT_LAUNCH(nullptr, (k_strided_waves<LOGL, WAVES, 2>), grid, dim3(64 * WAVES), s, b);
        else
            NTT_LAUNCH(nullptr, (k_strided_waves<LOGL, WAVES, 3>), grid, dim3(64 * WAVES), s, b);
        return NTT_CHECK_LAUNCH();
    }
    return false;
}

// Four waves per workgroup (64 lines of 2^6 points, 512-byte runs per row of
// the tile); measured at 2^24 points x 2 primes: 1, 2, 8 or 16 waves per
// workgroup move the same or less.
template <int LOGL>
static bool launch_waves(const PassArgs& a, int units, int z, int mode, hipStream_t s) {
    return launch_waves_w<LOGL, 4>(a, units, z, mode, s);
}

template <int LOGL>
static bool launch_fwd_strided_t(const PassArgs& a, int units, int operands, bool pack, bool alt, bool r8,
                                 hipStream_t s, bool waves = false) {
    if (waves && !pack && !a.batch && waves_fit(LOGL) && launch_waves<LOGL>(a, units, operands, 0, s)) return true;
    r8 = r8 && r8_fits(LOGL);
    return alt ? launch_fwd_strided_l<LOGL, alt_strided_lines(LOGL)>(a, units, operands, pack, s, r8)
               : launch_fwd_strided_l<LOGL, strided_lines(LOGL)>(a, units, operands, pack, s, r8);
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
static bool launch_inv_strided_l(const PassArgs& a, int units, bool canon, hipStream_t s, bool r8 = false) {
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
    } else if (r8) {
        if constexpr (r8_fits(LOGL)) {
            constexpr int NT8 = strided_threads8<LOGL, LINES>();
            if (canon)
                NTT_LAUNCH(nullptr, (k_inv_strided<LOGL, LINES, NT8, true, false, true>), grid, dim3(NT8), s, b);
            else
                NTT_LAUNCH(nullptr, (k_inv_strided<LOGL, LINES, NT8, false, false, true>), grid, dim3(NT8), s, b);
        }
    } else {
        if (canon)
            NTT_LAUNCH(nullptr, (k_inv_strided<LOGL, LINES, NT, true, false>), grid, dim3(NT), s, b);
        else
            NTT_LAUNCH(nullptr, (k_inv_strided<LOGL, LINES, NT, false, false>), grid, dim3(NT), s, b);
    }
    return NTT_CHECK_LAUNCH();
}

template <int LOGL>
static bool launch_inv_strided_t(const PassArgs& a, int units, bool canon, bool alt, bool r8, hipStream_t s,
                                 bool waves = false) {
    if (waves && !a.batch && waves_fit(LOGL) &&
        launch_waves<LOGL>(a, units, a.split_blocks ? a.split_blocks : 1, canon ? 3 : 2, s))
        return true;
    r8 = r8 && r8_fits(LOGL);
    return alt ? launch_inv_strided_l<LOGL, alt_strided_lines(LOGL)>(a, units, canon, s, r8)
               : launch_inv_strided_l<LOGL, strided_lines(LOGL)>(a, units, canon, s, r8);
}

bool launch_fwd_strided(int logl, const PassArgs& a, int primes, int operands, bool pack, bool alt, bool r8,
                        hipStream_t s, bool waves) {
    switch (logl) {
        case 5: return launch_fwd_strided_t<5>(a, primes, operands, pack, alt, r8, s, waves);
        case 6: return launch_fwd_strided_t<6>(a, primes, operands, pack, alt, r8, s, waves);
        case 7: return launch_fwd_strided_t<7>(a, primes, operands, pack, alt, r8, s, waves);
        case 8: return launch_fwd_strided_t<8>(a, primes, operands, pack, alt, r8, s, waves);
        case 9: return launch_fwd_strided_t<9>(a, primes, operands, pack, alt, r8, s, waves);
        case 10: return launch_fwd_strided_t<10>(a, primes, operands, pack, alt, r8, s, waves);
        case 11: return launch_fwd_strided_t<11>(a, primes, operands, pack, alt, r8, s, waves);
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

bool launch_inv_strided(int logl, const PassArgs& a, int primes, bool canon, bool alt, bool r8, hipStream_t s,
                        bool waves) {
    switch (logl) {
        case 5: return launch_inv_strided_t<5>(a, primes, canon, alt, r8, s, waves);
        case 6: return launch_inv_strided_t<6>(a, primes, canon, alt, r8, s, waves);
        case 7: return launch_inv_strided_t<7>(a, primes, canon, alt, r8, s, waves);
        case 8: return launch_inv_strided_t<8>(a, primes, canon, alt, r8, s, waves);
        case 9: return launch_inv_strided_t<9>(a, primes, canon, alt, r8, s, waves);
        case 10: return launch_inv_strided_t<10>(a, primes, canon, alt, r8, s, waves);
        case 11: return launch_inv_strided_t<11>(a, primes, canon, alt, r8, s, waves);
        default: return false;
    }
}


}  // namespace nttamd
