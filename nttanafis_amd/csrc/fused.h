// Row passes: forward rows of both operands, pointwise product, inverse rows.
#pragma once

#include "passes.h"

namespace nttamd {

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
// radix-16 shapes over the same tile: more waves per CU, one more round);
// FORM 3 (products): radix 16 both ways with one thread per group of the
// product tile, each thread taking both operands' groups in the forward
// rounds. Squares (FORM 0 and 1) and FORM 3 run the last forward round, the
// pointwise product and the first inverse round on one group in registers
// (fused_middle).

// Last forward round, pointwise product and first inverse round of a row
// pass in registers: both rounds hold the same groups (16 consecutive points
// of one line, the last four levels), so the product needs no LDS round trip
// and no barrier of its own. For a product, the groups of both operands at
// the same points (tile 1 at lds + LINES*LS) go through the forward levels
// with shared roots; the result replaces tile 0. Bounds: the forward round
// ends below 31p/16, products stay below 3.76p^2 < 4p^2 (below 11p/8), which
// is what an inverse round takes.
template <int LOGL, int LINES, int NT, bool SQUARE>
__device__ __forceinline__ void fused_middle(double* lds, const Tw& twF, const Tw& twI, double p, double pinv) {
    constexpr int L = 1 << LOGL;
    constexpr int L0 = LOGL - 4;
    constexpr int GPL = L / 16;
    constexpr int GROUPS = LINES * GPL;
    constexpr int LS = Lay<LOGL>::STRIDE;
    for (int g = threadIdx.x; g < GROUPS; g += NT) {
        const int c = g / GPL;
        const int b = g - c * GPL;
        double* line = lds + c * LS + pidx(16 * b);  // 16 consecutive points, no pad inside
        double x[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) x[k] = line[k];
        if constexpr (SQUARE) {
            fwd_group<4, false>(x, nullptr, twF, c, L0, b, p, pinv);
#pragma unroll
            for (int k = 0; k < 16; ++k) x[k] = mulmod(x[k], x[k], p, pinv);
        } else {
            double y[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) y[k] = line[LINES * LS + k];
            fwd_group<4, false, 2>(x, y, twF, c, L0, b, p, pinv);
#pragma unroll
            for (int k = 0; k < 16; ++k) x[k] = mulmod(x[k], y[k], p, pinv);
        }
        inv_group<4, false>(x, twI, c, L0, b, p, pinv);
#pragma unroll
        for (int k = 0; k < 16; ++k) line[k] = x[k];
    }
}

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
    const bool root1 = PACK && a.node_base == 0;
    constexpr int ROUNDS = Lay<LOGL>::ROUNDS;
    // The lines of both tiles go through one set of rounds; line c of either
    // tile is node H0 + (c mod LINES).
    if constexpr ((SQUARE && FORM != 2) || FORM == 3) {
        constexpr int L1 = Lay<LOGL>::FIRST;  // first level of round 1
        constexpr int GPL = L / 16;
        if constexpr (ROUNDS == 3 && (NT >= TILES * LINES * GPL || NT % (GPL * LINES) == 0)) {
            // Round 1 reads its roots from the table at nodes spread over
            // the table (rows of long transforms miss the L2): each thread
            // loads them for its group ahead, the forward ones before round
            // 0 and the inverse ones before the middle rounds.
            const int g = threadIdx.x, c1 = g / GPL, b1 = (g - c1 * GPL) / (L >> (L1 + 4));
            double w1[15], v1[15];
            group_roots16(w1, twF, c1, L1, b1);
            fwd_rounds<LOGL, TILES * LINES, NT, 0, 1>(lds, twF, p, pinv, root1);
            round16_w<LOGL, TILES * LINES, NT, L1, LINES, false>(lds, w1, p, pinv);
            __syncthreads();
            group_roots16(v1, twI, c1, L1, b1);
            fused_middle<LOGL, LINES, NT, SQUARE>(lds, twF, twI, p, pinv);
            __syncthreads();
            round16_w<LOGL, LINES, NT, L1, LINES, true>(lds, v1, p, pinv);
            __syncthreads();
            inv_rounds<LOGL, LINES, NT, 0>(lds, twI, p, pinv, root1);
        } else {
            fwd_rounds<LOGL, TILES * LINES, NT, 0, ROUNDS - 1>(lds, twF, p, pinv, root1);
            fused_middle<LOGL, LINES, NT, SQUARE>(lds, twF, twI, p, pinv);
            __syncthreads();
            inv_rounds<LOGL, LINES, NT, ROUNDS - 2>(lds, twI, p, pinv, root1);
        }
    } else {
        if constexpr (FORM == 2)
            fwd_rounds_from<LOGL, TILES * LINES, NT, 0, 3>(lds, twF, p, pinv);
        else
            fwd_rounds<LOGL, TILES * LINES, NT, 0>(lds, twF, p, pinv, root1);
        // Pointwise product: every value loaded before the first store (the
        // stores could alias later loads, which would serialise them).
        double v[R], w[R];
        const int t0 = row_off<LOGL>(threadIdx.x);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int e = threadIdx.x + r * NT;
            if (!FULL && e >= PTS) continue;
            const double* x = lds + t0 + row_off<LOGL>(r * NT);
            v[r] = x[0];
            w[r] = SQUARE ? x[0] : x[LINES * LS];
        }
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = mulmod(v[r], w[r], p, pinv);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int e = threadIdx.x + r * NT;
            if (!FULL && e >= PTS) continue;
            lds[t0 + row_off<LOGL>(r * NT)] = v[r];
        }
        __syncthreads();
        // The product is one tile: radix-16 groups would occupy half of the
        // threads, so a product runs its inverse in radix 8 unless the
        // autotuner chose radix 16 (FORM 1) for this shape.
        if constexpr (FORM == 1)
            inv_rounds<LOGL, LINES, NT, ROUNDS - 1>(lds, twI, p, pinv, root1);
        else
            inv_rounds8<LOGL, LINES, NT, Lay8<LOGL>::ROUNDS - 1>(lds, twI, p, pinv, root1);
    }
    if constexpr (PACK) {
        const double tinv = a.tinv[y];
        const int t0 = row_off<LOGL>(threadIdx.x);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int m = threadIdx.x + r * NT;
            if (!FULL && m >= L) continue;
            global_ptr(row)[m] = canonical(mulmod(lds[t0 + row_off<LOGL>(r * NT)], tinv, p, pinv), p);
        }
    } else {
        store_rows<LOGL, LINES, NT>(lds, row);
    }
}

// Shapes of the fused row pass: about 256 radix-16 groups per workgroup
// round (enough rows per workgroup to fill it), one thread per group; the
// radix-8 form (FORM 2) keeps the tile and doubles the threads.
template <int LOGL, bool PACK, bool SQUARE, int FORM = 0>
struct FusedShape {
    static constexpr int TILES = SQUARE ? 1 : 2;
    // Radix-16 groups one thread set covers per row and round (FORM 3: the
    // product tile's; the forward rounds then take two groups per thread).
    static constexpr int GROUPS = (FORM == 3 ? 1 : TILES) * (1 << LOGL) / 16;
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

// form: k_fused's FORM (1 and 3 apply to products only).
template <int LOGL, bool SQUARE>
static bool launch_fused_f(const PassArgs& a, int units, bool pack, int form, hipStream_t s) {
    if (form == 2)
        return pack ? launch_fused_v<LOGL, true, SQUARE, 2>(a, units, s)
                    : launch_fused_v<LOGL, false, SQUARE, 2>(a, units, s);
    if constexpr (!SQUARE) {
        if (form == 1)
            return pack ? launch_fused_v<LOGL, true, SQUARE, 1>(a, units, s)
                        : launch_fused_v<LOGL, false, SQUARE, 1>(a, units, s);
        if (form == 3)
            return pack ? launch_fused_v<LOGL, true, SQUARE, 3>(a, units, s)
                        : launch_fused_v<LOGL, false, SQUARE, 3>(a, units, s);
    }
    return pack ? launch_fused_v<LOGL, true, SQUARE, 0>(a, units, s)
                : launch_fused_v<LOGL, false, SQUARE, 0>(a, units, s);
}

}  // namespace nttamd
