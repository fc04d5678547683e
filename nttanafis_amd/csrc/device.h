// Device-side arithmetic and kernel argument blocks of the gfx950 engine.
//
// Residues are integer-valued binary64 values. A product of two residues is
// reduced exactly with one FMA recovering the rounding error of the product
// and one FMA subtracting the rounded quotient times p, the scheme of the
// reference's src/arch/scalar/product.rs:96-109. Every kernel here is compiled
// with -ffp-contract=off, so only the fma() calls below fuse.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace nttamd {

constexpr int kMaxPrimes = 8;
constexpr int kMaxWords = 7;     // ceil(8 * 50 / 64): words of one CRT value
constexpr int kPartBits = 49;    // bits of one packing part, below every prime
constexpr int kMaxParts = 4;     // parts of the widest coefficient (196 bits)
constexpr int kMinLog = 5;       // shortest transform: 32 points
constexpr int kSinglePassMaxLog = 12;

// Per-prime constants, passed by value to every kernel.
struct Field {
    double p[kMaxPrimes];
    double pinv[kMaxPrimes];
};

// Exact residue product: |a*b| < 2p^2 gives |r| < 15p/16, |a*b| < 4p^2 gives
// |r| < 11p/8 for every catalog prime.
// `low` is the exact rounding error a*b - high, so a*b - q*p = (high - q*p) + low.
__host__ __device__ __forceinline__ double mulmod(double a, double b, double p, double pinv) {
    const double high = a * b;
    const double low = fma(a, b, -high);
    const double q = rint(high * pinv);
    return fma(-q, p, high) + low;
}

// Centered reduction of |x| < 8p to |r| < 9p/16 (|r| <= p/2 for |x| < p).
__host__ __device__ __forceinline__ double reduce(double x, double p, double pinv) {
    return fma(-rint(x * pinv), p, x);
}

// Canonical residue in [0, p) of |x| < p.
__host__ __device__ __forceinline__ double canonical(double x, double p) { return x < 0.0 ? x + p : x; }

// Operand description for packing coefficients straight from limbs.
struct Source {
    const uint64_t* limbs;  // little-endian limbs
    uint64_t limb_len;
    uint64_t digits;        // coefficients; later columns are zero
    uint32_t width;         // bits between coefficients (64: one limb each)
    uint32_t value_bits;    // significant bits per coefficient (<= width)
    uint32_t is_signed;     // two's complement coefficients (width 64 only)
    uint32_t pad;
};

// One product of a batch, as the batched kernels read it from device memory.
struct BatchProduct {
    Source src_a, src_b;
    double* spec_a;          // primes rows of T doubles
    double* spec_b;          // unused for a square
    uint64_t* out;           // result limbs
    uint64_t out_len;
    uint64_t coefficients;   // convolution coefficients reconstructed
    uint64_t row_stride;     // T
    uint32_t primes;
    uint32_t square;
    uint32_t width;
    uint32_t width_magic;
    uint32_t value_bits;
    uint32_t blocks;         // carry blocks of the output
    uint32_t* carry;         // carry_words(blocks): base, thresh, first thresh, ripple end; chunk functions
    // Unbalanced split (as in PassArgs): block z of the longer operand (src_a)
    // owns rows at spec_a + z*block_stride; the shorter operand (src_b) is
    // transformed once into spec_b. 0: no split.
    uint32_t split_blocks;
    uint32_t pad;
    uint64_t block_digits;
    uint64_t block_stride;
};

// Work items of a batched launch: product index and sub = prime | block << 4
// (passes; block 0 unless the product is split) or product index and carry
// block (reconstruction). In a forward pass, sub carries kItemOperandB when
// the item transforms the second (shorter) operand.
struct BatchItem {
    uint32_t product;
    uint32_t sub;
};
constexpr uint32_t kItemOperandB = 0x80000000u;
constexpr uint32_t kItemPrimeMask = 0xFu;
constexpr int kItemBlockShift = 4;
constexpr uint32_t kItemBlockMask = 0x07FFFFFFu;  // after the shift

// Which rows of a batched product a pass works on: forward passes take the
// operand each item names; product passes take the product rows.
enum BatchOperand : int32_t { kOperandForward = 0, kOperandProduct = 2 };

// Arguments of every transform pass kernel.
struct PassArgs {
    double* data;            // rows of one operand, prime y at data + y * row_stride
    double* other;           // rows of operand b in forward passes over both operands
                             // (grid z = 1); rows of operand a in a fused pass
    const double* fwd;       // forward root table, prime y at fwd + y * table_stride
    const double* inv;       // inverse root table
    uint64_t row_stride;     // T
    uint64_t table_stride;   // entries per prime in the root tables
    int32_t log_t;           // K = log2(T)
    int32_t level_start;     // global level of this pass's first level
    int32_t primes;          // work units (primes) of a single-operation strided pass
    int32_t pad0;
    uint64_t line_stride;    // element stride of a strided pass, 1 for rows
    Source src_a;            // packing sources (first forward pass)
    Source src_b;
    Field field;
    double tinv[kMaxPrimes];                 // centered T^-1 mod p
    double powers[kMaxPrimes][kMaxParts];    // centered 2^(49k) mod p
    // Part of a composite row: the transform of T points runs below root-table
    // node node_base (0 for x^T - 1; 2, 4, 6 for the factors x^T - W[node]^2
    // of x^(2T') + 1 that a composite row adds), so a line of depth d that
    // is node h of the part is node (node_base << d) + h of the table.
    uint64_t node_base;
    // Packing into a part: point n of the part is the sum over q < fold_terms
    // of coefficient n + q*fold_span times fold_mult[y][q-1] (q >= 1), the
    // residue of the operand modulo x^fold_span - W[node_base]^2.
    uint64_t fold_span;
    int32_t fold_terms;                      // 1: no fold
    // Unbalanced split of a product (src/multiply/execute.rs:131-411): grid z
    // below split_blocks is block z of the longer operand, whose rows sit at
    // data + z*block_stride and whose coefficients are z*block_digits.. of
    // src_a; z == split_blocks is the shorter operand (rows at other, source
    // src_b), transformed once for every block.
    uint32_t split_blocks;
    double fold_mult[kMaxPrimes][3];         // centered
    uint64_t block_digits;
    uint64_t block_stride;
    // Batched launches: items[blockIdx.x / tiles_per_item] names the product
    // and prime of a workgroup; batch == nullptr for a single operation.
    const BatchProduct* batch;
    const BatchItem* items;
    uint64_t tiles_per_item;
    int32_t operand;                         // BatchOperand
    int32_t pad1;
};

// Arguments of the CRT and carry kernels.
struct CrtArgs {
    const double* residues;  // canonical residues, prime y at residues + y * row_stride
    uint64_t row_stride;
    uint64_t coefficients;   // number of convolution coefficients
    uint64_t out_len;        // output limbs
    uint64_t* out;           // output limbs (device)
    uint32_t width;          // bits between coefficients (product) or stride*64 (conv)
    uint32_t width_magic;    // ceil(2^32 / width): x / width == mulhi(x, magic) for x < 2^24
    uint32_t value_words;    // words of one reconstructed value
    uint32_t value_bits;     // bits of one reconstructed value (bound)
    uint32_t stride;         // conv: limbs per coefficient
    uint32_t is_signed;
    uint32_t primes;
    // Unbalanced split: residues of coefficient i are the sum of the rows of
    // the (at most two) blocks covering it, block z at residues +
    // z*block_stride holding coefficients z*split_digits.. (0: no split).
    uint32_t split_blocks;
    uint64_t split_digits;
    uint64_t block_stride;
    Field field;
    double pmod[kMaxPrimes][kMaxPrimes];  // centered p_j mod p_k (j < k)
    double ginv[kMaxPrimes];              // centered (p_0..p_{k-1})^-1 mod p_k
    uint64_t pint[kMaxPrimes];
    // carry bookkeeping, one entry per carry block
    uint32_t* agg_base;      // B_end of each block
    uint32_t* agg_thresh;    // T_end of each block
    uint32_t* first_thresh;  // t of the block's first limb
    uint32_t* ripple_end;    // last limb index (relative) reached by a ripple
    uint32_t blocks;
    uint32_t pad;
    // Batched reconstruction: items[blockIdx.x] names product and block;
    // scan_items (scan_count of them) name the (product, chunk) pairs of the
    // products whose carries go through k_carry_scan (those of more than
    // kCarryDirectBlocks blocks).
    const BatchProduct* batch;
    const BatchItem* items;
    const BatchItem* scan_items;
    uint32_t scan_count;
    uint32_t wide;  // single operation: one 512-thread block covers the whole output
    // Schoolbook products: the operand limbs (single operation; a batch
    // reads them from src_a and src_b of each product).
    const uint64_t* op_a;
    const uint64_t* op_b;
    uint64_t len_a, len_b;
};

// Root-table generation arguments.
struct TableArgs {
    double* fwd;
    double* inv;
    uint64_t table_stride;
    uint64_t len;
    Field field;
    double roots[kMaxPrimes][48];
    double inv_roots[kMaxPrimes][48];
};

// Recombination of a composite row of T = N + tail*N/4 points (the
// recombine<5|3|7> step of src/transform/finish.rs:106-158). The cyclic
// part u holds the product modulo x^N - 1 and the tails hold it modulo the
// factors of x^N + 1 at nodes 4 (tail 1), 2 (tail 2), or 2 and 6 (tail 3),
// all canonical; the kernel rewrites the row in place as the T canonical
// coefficients. W1 = W[1] (a fourth root of unity), W2 = W[2], W3 = W[3].
struct RecombineArgs {
    double* rows;  // prime y at rows + y * row_stride
    uint64_t row_stride;
    uint64_t n;  // cyclic length N
    uint32_t tail;
    uint32_t pad;
    Field field;
    double w1[kMaxPrimes];
    double s[kMaxPrimes][3];  // W2, W2^2, W2^3
    double t[kMaxPrimes][3];  // W3, W3^2, W3^3
    double half[kMaxPrimes];  // 2^-1
};

// Lines per strided tile of lines of 2^logl points: up to 2^7 points 64 at a
// time (512-byte runs), 2^8 points 32 at a time (256-byte runs); longer
// lines fill 8192-point tiles. A pass needs its line stride to be at least
// this many points.
__host__ __device__ constexpr int strided_lines(int logl) {
    return logl <= 7 ? 64 : logl == 8 ? 32 : (logl <= 10 ? 8 : 4);
}
// The alternative tiles the autotuner may choose instead: half the lines up
// to 2^8 points and for 2^10-2^11 points (4096-point tiles), twice the lines
// for 2^9 points (8192 points).
__host__ __device__ constexpr int alt_strided_lines(int logl) {
    return logl <= 7 ? 32 : logl == 8 ? 16 : logl == 9 ? 16 : logl == 10 ? 4 : 2;
}
__host__ __device__ constexpr int tile_lines(int logl, bool alt) { return alt ? alt_strided_lines(logl) : strided_lines(logl); }

constexpr uint32_t kInf = 0xFFFFFFFFu;
// Carries across the blocks of an output: up to kCarryDirectBlocks blocks,
// each block composes the published functions of the blocks before it
// itself, spread over its threads (at most kCarryDirectBlocks / 256 per
// thread); longer outputs first scan their block functions in chunks of
// kCarryChunk blocks, one workgroup per chunk.
constexpr uint32_t kCarryDirectBlocks = 2048;
constexpr int kCarryChunk = 1024;
__host__ __device__ __forceinline__ bool carry_scanned(uint32_t blocks) { return blocks > kCarryDirectBlocks; }
__host__ __device__ __forceinline__ uint32_t carry_chunks(uint32_t blocks) {
    return (blocks + kCarryChunk - 1) / kCarryChunk;
}
// Carry bookkeeping words of an output of `blocks` carry blocks: four per
// block (function base and threshold, first-limb threshold, ripple end), then
// two per chunk (the chunk's composed function).
__host__ __device__ __forceinline__ uint64_t carry_words(uint32_t blocks) {
    return 4ull * blocks + 2ull * carry_chunks(blocks);
}
constexpr int kCrtThreads = 256;
constexpr int kCrtLimbsPerThread = 4;
constexpr int kCrtBlockLimbs = kCrtThreads * kCrtLimbsPerThread;
// Schoolbook products: the shorter operand has at most kSchoolMaxShort limbs;
// carry blocks of kSchoolBlockLimbs output limbs.
constexpr int kSchoolMaxShort = 512;
constexpr int kSchoolBlockLimbs = 64;

}  // namespace nttamd
