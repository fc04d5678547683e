/*
 * ntt_oracle.h -- CPU restatement of the ntt_anafis multiplication path.
 *
 * TEST INFRASTRUCTURE. This library is the checker for the HIP engine in
 * nttanafis_amd/. Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it, and only to check or to time the CPU path.
 * The product never links it and never falls back to it.
 *
 * The restatement follows the reference crate's algorithm: multi-prime
 * integer-valued binary64 residues (src/arch/scalar/product.rs:96-109),
 * bit-reversed root tables (src/transform/twiddle.rs:116-156), radix-4
 * remainder-tree stages (src/arch/schedule/expression/stages.rs:332-383),
 * cofactor CRT with one approximate quotient (src/limb/cofactor.rs:59-111)
 * and carry reconstruction (src/limb/reconstruction.rs:77-165).
 *
 * Geometry restated from src/profile/planner.rs with power-of-two rows only
 * (no 5N/4, 3N/2, 7N/4 rows) and without the unbalanced block split: both
 * change the cost of a product, never its exact limbs.
 */
#ifndef NTT_ORACLE_H
#define NTT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_geometry {
    uint32_t primes;       /* catalog prefix length, 1..8 */
    uint32_t chunk_bits;   /* bits between operand coefficients */
    uint64_t digits_a;     /* coefficients of the left operand */
    uint64_t digits_b;     /* coefficients of the right operand */
    uint32_t value_bits_a; /* significant bits per left coefficient */
    uint32_t value_bits_b;
    uint64_t conv_len;     /* digits_a + digits_b - 1 */
    uint64_t transform_len;/* power-of-two row length */
    uint32_t disjoint;     /* 1 for convolution blocks of whole limbs */
    int32_t sign_bit;      /* b-1 for signed b-bit coefficients, else -1 */
    uint64_t output_stride;/* bits between result coefficients */
    uint32_t capacity;     /* CRT capacity of the prefix in bits */
} oracle_geometry;

typedef struct oracle_mul_plan {
    uint64_t len_a, len_b;
    oracle_geometry geometry;
} oracle_mul_plan;

typedef struct oracle_conv_plan {
    uint64_t len_a, len_b;
    uint64_t max_a, max_b;
    uint32_t is_signed;
    uint32_t schoolbook;
    uint64_t stride;
    uint64_t output_len;
    oracle_geometry geometry;
} oracle_conv_plan;

/* Catalog and field arithmetic (src/field/catalog.rs, src/arch/scalar/product.rs). */
uint64_t oracle_prime(int index);
uint64_t oracle_generator(int index);
double oracle_mulmod(double a, double b, int prime);
double oracle_reduce(double a, int prime);
uint32_t oracle_capacity(int primes);
/* Fills `out` (np*table_len) with W[j] = omega^rev(j) per prime. */
void oracle_root_table(int primes, uint64_t table_len, int inverse, double *out);

/* One forward remainder-tree transform of a row of `len` points for `prime`,
 * output in the tree's (bit-reversed) order. */
void oracle_forward_row(double *row, uint64_t len, int prime);
/* Inverse of oracle_forward_row, unnormalized (scaled by len). */
void oracle_inverse_row(double *row, uint64_t len, int prime);

/* MulPlan (src/plan.rs:53-114). Return 1 on success, 0 for None/false. */
int oracle_mul_plan_new(uint64_t len_a, uint64_t len_b, oracle_mul_plan *out);
int oracle_mul(const oracle_mul_plan *plan, uint64_t *dst, uint64_t dst_len,
               const uint64_t *a, uint64_t a_len, const uint64_t *b, uint64_t b_len,
               int threads);
int oracle_sqr(const oracle_mul_plan *plan, uint64_t *dst, uint64_t dst_len,
               const uint64_t *a, uint64_t a_len, int threads);
/* MulWorkspace::mul semantics (src/plan.rs:182-205): empty operand writes zero. */
int oracle_workspace_mul(uint64_t *dst, uint64_t dst_len, const uint64_t *a, uint64_t a_len,
                         const uint64_t *b, uint64_t b_len, int square, int threads);

/* ConvPlan (src/convolve.rs:75-184). */
int oracle_conv_plan_new(uint64_t len_a, uint64_t len_b, uint64_t max_a, uint64_t max_b,
                         int is_signed, oracle_conv_plan *out);
int oracle_convolve(const oracle_conv_plan *plan, uint64_t *dst, uint64_t dst_len,
                    const uint64_t *a, uint64_t a_len, const uint64_t *b, uint64_t b_len,
                    int threads);

/* Schoolbook product (src/tests/schoolbook.rs:13-45). */
void oracle_schoolbook(uint64_t *dst, const uint64_t *a, uint64_t a_len, const uint64_t *b,
                       uint64_t b_len);

/* SplitMix64 operand (bench/src/lib.rs:13-29). */
void oracle_operand(uint64_t *out, uint64_t len, uint64_t seed);

/* Residue of a little-endian limb array modulo q < 2^63 (Horner, u128). */
uint64_t oracle_mod_limbs(const uint64_t *limbs, uint64_t len, uint64_t q);

#ifdef __cplusplus
}
#endif

#endif
