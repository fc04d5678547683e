/*
 * nttanafis.h -- C ABI of the MI355X (gfx950) multi-prime NTT engine.
 *
 * Drop-in boundary for the reference crate's public multiplication API. Each
 * entry point cites the reference item it replaces (paths relative to the
 * reference crate root). Limbs are little-endian uint64_t words, least
 * significant first (src/lib.rs:3-6).
 *
 * Shared semantics (src/lib.rs:21-26):
 *   - plan constructors return false where the reference returns None;
 *   - calls return false on rejected shapes or short buffers, and then leave
 *     dst unchanged;
 *   - success writes exactly the product span (a_len+b_len limbs, 2*a_len for a
 *     square, output_len for a convolution) and preserves the rest of dst;
 *   - the same pointer and length for both operands runs the square path
 *     (core::ptr::eq on slices, src/plan.rs:96, src/convolve.rs:156);
 *   - dst must not overlap an operand (guaranteed by borrows in the
 *     reference, required here);
 *   - device failures (allocation, launch) return false with dst unchanged:
 *     results are staged in device memory and copied out last.
 *
 * Deviation: the GPU engine keeps no work in caller scratch. Device arenas
 * belong to a process-wide context (plan path) or to the workspace, so the
 * scratch lengths below are 0 and any scratch buffer, including NULL, passes.
 */
#ifndef NTTANAFIS_H
#define NTTANAFIS_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Products                                                                  */

/* MulPlan (src/plan.rs:33-36): plain data, copyable, no heap. */
typedef struct ntt_mul_plan {
    uint64_t len_a;
    uint64_t len_b;
    uint32_t primes;     /* catalog prefix of the device geometry */
    uint32_t chunk_bits; /* bits per coefficient */
    uint32_t log_len;    /* log2 of the cyclic transform length N (of one block's row when the
                            product is split, see ntt_mul_split) */
    uint32_t valid;      /* 1 for a plan from ntt_mul_plan_new */
    uint32_t row_tail;   /* transform row of N + row_tail * N/4 points (0..3) */
    uint32_t schoolbook; /* 1: the product runs as a schoolbook (ntt_mul_schoolbook) */
} ntt_mul_plan;

/* MulPlan::new, src/plan.rs:53-69. false <=> None: a zero length, an
 * overflowing output length, or a transform order beyond the catalog. */
bool ntt_mul_plan_new(size_t len_a, size_t len_b, ntt_mul_plan *out);

/* MulPlan::scratch_len, src/plan.rs:73-76 (0, see the deviation above). */
size_t ntt_mul_plan_scratch_len(const ntt_mul_plan *plan);

/* MulPlan::sqr_scratch_len, src/plan.rs:81-83 (0). */
size_t ntt_mul_plan_sqr_scratch_len(const ntt_mul_plan *plan);

/* MulPlan::mul, src/plan.rs:94-102. Thread-safe. */
bool ntt_mul_plan_mul(const ntt_mul_plan *plan, uint64_t *dst, size_t dst_len, const uint64_t *a,
                      size_t a_len, const uint64_t *b, size_t b_len, uint64_t *scratch,
                      size_t scratch_len);

/* MulPlan::sqr, src/plan.rs:112-114. Thread-safe. */
bool ntt_mul_plan_sqr(const ntt_mul_plan *plan, uint64_t *dst, size_t dst_len, const uint64_t *a,
                      size_t a_len, uint64_t *scratch, size_t scratch_len);

/* MulWorkspace (src/plan.rs:134-138): retained device tables and arenas.
 * One handle must not be used by two threads at once. */
typedef struct ntt_mul_workspace ntt_mul_workspace;

/* MulWorkspace::new, src/plan.rs:143-148. NULL only when out of host memory. */
ntt_mul_workspace *ntt_mul_workspace_new(void);

/* Drop of MulWorkspace. NULL is ignored. */
void ntt_mul_workspace_free(ntt_mul_workspace *ws);

/* MulWorkspace::prepare, src/plan.rs:151-156: retains the plan and grows the
 * device tables and arenas for it. */
void ntt_mul_workspace_prepare(ntt_mul_workspace *ws, const ntt_mul_plan *plan);

/* MulWorkspace::mul, src/plan.rs:164-170. An empty operand writes zero. */
bool ntt_mul_workspace_mul(ntt_mul_workspace *ws, uint64_t *dst, size_t dst_len, const uint64_t *a,
                           size_t a_len, const uint64_t *b, size_t b_len);

/* MulWorkspace::sqr, src/plan.rs:176-179. */
bool ntt_mul_workspace_sqr(ntt_mul_workspace *ws, uint64_t *dst, size_t dst_len, const uint64_t *a,
                           size_t a_len);

/* ------------------------------------------------------------------------ */
/* Convolutions                                                              */

/* ConvPlan (src/convolve.rs:51-57). */
typedef struct ntt_conv_plan {
    uint64_t len_a;
    uint64_t len_b;
    uint64_t max_a;
    uint64_t max_b;
    uint64_t stride;     /* limbs per result coefficient */
    uint64_t output_len; /* (len_a + len_b - 1) * stride */
    uint32_t is_signed;
    uint32_t schoolbook; /* shorter operand of at most 24 terms: direct sums */
    uint32_t primes;
    uint32_t log_len;    /* log2 of the cyclic transform length N */
    uint32_t bound_bits; /* bits of t*max_a*max_b, plus a sign bit if signed */
    uint32_t value_bits_a;
    uint32_t value_bits_b;
    uint32_t valid;
    uint32_t row_tail;     /* transform row of N + row_tail * N/4 points (0..3) */
    uint32_t split_blocks; /* unbalanced split: blocks of the longer operand (0: none) */
    uint64_t block_len;    /* coefficients of the longer operand per block */
} ntt_conv_plan;

/* ConvPlan::new, src/convolve.rs:75-113. false <=> None: an empty operand,
 * an unaddressable shape, a bound beyond the catalog, or a signed bound of
 * 2^63 or more. */
bool ntt_conv_plan_new(size_t len_a, size_t len_b, uint64_t max_a, uint64_t max_b, bool is_signed,
                       ntt_conv_plan *out);

/* ConvPlan::stride, src/convolve.rs:118-120. */
size_t ntt_conv_plan_stride(const ntt_conv_plan *plan);

/* ConvPlan::output_len, src/convolve.rs:124-126. */
size_t ntt_conv_plan_output_len(const ntt_conv_plan *plan);

/* ConvPlan::scratch_len, src/convolve.rs:132-138 (0, see above). */
size_t ntt_conv_plan_scratch_len(const ntt_conv_plan *plan);

/* ConvPlan::convolve, src/convolve.rs:148-162. Returns false for a length
 * mismatch, a coefficient above its bound, or a short dst. Thread-safe. */
bool ntt_conv_plan_convolve(const ntt_conv_plan *plan, uint64_t *dst, size_t dst_len,
                            const uint64_t *a, size_t a_len, const uint64_t *b, size_t b_len,
                            uint64_t *scratch, size_t scratch_len);

/* ConvWorkspace (src/convolve.rs:228-232). */
typedef struct ntt_conv_workspace ntt_conv_workspace;

/* ConvWorkspace::new, src/convolve.rs:237-243. */
ntt_conv_workspace *ntt_conv_workspace_new(void);

/* Drop of ConvWorkspace. */
void ntt_conv_workspace_free(ntt_conv_workspace *ws);

/* ConvWorkspace::prepare, src/convolve.rs:247-255. */
void ntt_conv_workspace_prepare(ntt_conv_workspace *ws, const ntt_conv_plan *plan);

/* ConvWorkspace::convolve, src/convolve.rs:264-280. false when unprepared. */
bool ntt_conv_workspace_convolve(ntt_conv_workspace *ws, uint64_t *dst, size_t dst_len,
                                 const uint64_t *a, size_t a_len, const uint64_t *b, size_t b_len);

/* ------------------------------------------------------------------------ */
/* Device-resident operands (no counterpart in the reference, whose operands
 * are host slices). Pointers are device addresses; `stream` is a hipStream_t
 * or NULL for the workspace's own stream. The call returns once the work is
 * queued on that stream; synchronize the stream before reading dst.
 * Coefficient bounds of a convolution are not checked on this path. */

bool ntt_mul_workspace_mul_device(ntt_mul_workspace *ws, uint64_t *d_dst, size_t dst_len,
                                  const uint64_t *d_a, size_t a_len, const uint64_t *d_b,
                                  size_t b_len, void *stream);

bool ntt_mul_workspace_sqr_device(ntt_mul_workspace *ws, uint64_t *d_dst, size_t dst_len,
                                  const uint64_t *d_a, size_t a_len, void *stream);

bool ntt_conv_workspace_convolve_device(ntt_conv_workspace *ws, uint64_t *d_dst, size_t dst_len,
                                        const uint64_t *d_a, size_t a_len, const uint64_t *d_b,
                                        size_t b_len, void *stream);

/* Allocation-free product of device operands. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It computes exactly what ntt_mul_workspace_mul_device computes (the same
 * slice twice squares) and calls no hipMalloc, hipFree or other allocating
 * routine: its spectra, carry words and root tables live in one fixed device
 * array of 32 MiB per GPU (8 MiB of root tables, 24 MiB of working memory),
 * held from library load to process end. The first call on a device takes
 * the array's address and generates the root tables into it (synchronizing
 * `stream` once). Limit: transforms of at most 2^17 points and a working set
 * within 24 MiB (every product of two operands up to 65,536 limbs fits); a
 * larger product returns false with a message in ntt_last_error and dst
 * unchanged. Every call on one GPU shares that GPU's array, so calls that
 * overlap on one GPU (concurrent streams or threads) race: serialize them. */
bool ntt_mul_device_noalloc(uint64_t *d_dst, size_t dst_len, const uint64_t *d_a, size_t a_len,
                            const uint64_t *d_b, size_t b_len, void *stream);

/* ------------------------------------------------------------------------ */
/* Batched products (no counterpart in the reference, which multiplies one
 * pair per call: MulWorkspace::mul, src/plan.rs:164-170, is what each batch
 * entry computes). A batch is planned once for a list of shapes; operands are
 * bound as device pointers and may be re-bound for every run. A run queues
 * one hipGraph replay that computes every product of the batch: one launch
 * per transform pass and length bucket over all (product, prime) items, and
 * one reconstruction launch per prime count over all (product, block) items.
 * The first run times candidate pass splits for each transform length
 * (autotuning, process-wide) and captures the graph. One batch handle must not
 * be used by two threads at once. */

typedef struct ntt_batch ntt_batch;

ntt_batch *ntt_batch_new(void);
void ntt_batch_free(ntt_batch *batch);

/* Plans `count` products: product i is a_i * b_i with len_a[i] and len_b[i]
 * limbs, or a_i squared when square != NULL and square[i] != 0 (len_b may
 * then be NULL or hold anything at i). Allocates the device arenas. false on
 * a zero length or a shape beyond the catalog (ntt_last_error says which). */
bool ntt_batch_plan(ntt_batch *batch, size_t count, const size_t *len_a, const size_t *len_b,
                    const uint8_t *square);

/* Number of planned products, and the limbs product i writes (len_a+len_b). */
size_t ntt_batch_count(const ntt_batch *batch);
size_t ntt_batch_output_len(const ntt_batch *batch, size_t i);

/* Binds device operands and outputs: arrays of `count` device pointers.
 * d_b may be NULL when every product is a square; d_b[i] is ignored for a
 * square. Output i receives exactly ntt_batch_output_len(i) limbs and must
 * not overlap any operand or other output. */
bool ntt_batch_bind(ntt_batch *batch, const uint64_t *const *d_a, const uint64_t *const *d_b,
                    uint64_t *const *d_out);

/* Queues every product of the batch on `stream` (NULL: the batch's own
 * stream) as one graph launch. Synchronize the stream before reading. */
bool ntt_batch_run(ntt_batch *batch, void *stream);

/* The same products as a plain loop of one ntt_mul_workspace_mul_device (or
 * _sqr_device) call per product on `stream`: the per-call path a batch
 * replaces, for comparison. */
bool ntt_batch_run_plain(ntt_batch *batch, void *stream);

/* Kernel launches in the batch's graph. */
int ntt_batch_launches(const ntt_batch *batch);

/* Host-operand convenience: copies every a[i] (and b[i] unless a square) to
 * device staging, binds, runs, and writes each product to out[i] (exactly
 * ntt_batch_output_len(i) limbs). Synchronous. */
bool ntt_batch_mul_host(ntt_batch *batch, const uint64_t *const *a, const uint64_t *const *b,
                        uint64_t *const *out);

/* ------------------------------------------------------------------------ */
/* Request server (no counterpart in the reference). Serves a stream of
 * products, squares and convolutions from host threads with continuous
 * batching: worker slots with their own streams take work whenever they are
 * free, in arrival order: the small products queued ahead of the oldest
 * other request run as one batched launch sequence, and that request is
 * admitted as soon as its device working set fits in the memory budget
 * (small products behind it run only while it waits for memory, so it is
 * never starved). Root tables stay resident. Each request
 * computes what MulWorkspace::mul / sqr (src/plan.rs:164-179) or
 * ConvPlan::convolve (src/convolve.rs:148-162) computes for it alone. */

enum {
    NTT_REQ_MUL = 0,  /* dst = a * b (len_a + len_b limbs) */
    NTT_REQ_SQR = 1,  /* dst = a * a (2 len_a limbs); b ignored */
    NTT_REQ_CONV = 2, /* dst = ConvPlan(len_a, len_b, max_a, max_b, signed).convolve(a, b) */
};

enum {
    NTT_REQ_SIGNED = 1, /* convolution of two's complement coefficients */
    NTT_REQ_DEVICE = 2, /* a, b and dst are device addresses */
};

typedef struct ntt_request {
    uint32_t kind;
    uint32_t flags;
    const uint64_t *a;
    uint64_t len_a;
    const uint64_t *b;
    uint64_t len_b;
    uint64_t max_a; /* convolution coefficient bounds */
    uint64_t max_b;
    uint64_t *dst;
    uint64_t dst_len;
} ntt_request;

typedef struct ntt_server_stats {
    uint64_t submitted;
    uint64_t completed;
    uint64_t failed;
    uint64_t groups;           /* batched groups of small products run */
    uint64_t grouped;          /* requests that ran inside groups */
    uint64_t budget_bytes;     /* device memory budget of large requests */
    uint64_t peak_bytes;       /* most of it in use at once */
    uint64_t waits_for_memory; /* admissions deferred for lack of memory */
} ntt_server_stats;

typedef struct ntt_server ntt_server;

/* Starts a server on the current device with `slots` worker streams (0: 4)
 * and a budget of `hbm_budget_bytes` of device memory for large requests (0:
 * half of the free memory), allocated once at start. Each slot also keeps up
 * to about 0.3 GiB for its groups of small products. */
ntt_server *ntt_server_new(uint64_t hbm_budget_bytes, uint32_t slots);

/* Finishes every submitted request, then stops the server. */
void ntt_server_free(ntt_server *server);

/* Queues a request and returns its ticket, or -1 when the request is
 * malformed (lengths, bounds, dst_len) or can never fit the budget. Buffers
 * must stay valid and unchanged until the request completes. A host dst is
 * written only by a request that completes successfully; a convolution whose
 * host coefficients exceed their bounds fails and leaves dst unchanged. With
 * NTT_REQ_DEVICE the kernels write dst directly and coefficient bounds are
 * not checked (as for the *_device calls). Thread-safe. */
int64_t ntt_server_submit(ntt_server *server, const ntt_request *request);

/* 0 while queued or running, 1 when done, -1 when failed or unknown. */
int ntt_server_poll(ntt_server *server, int64_t ticket);

/* Blocks until the request ends; returns 1 (done) or -1 (failed). */
int ntt_server_wait(ntt_server *server, int64_t ticket);

/* Blocks until every submitted request has ended; true when none failed. */
bool ntt_server_drain(ntt_server *server);

/* Submission, start and completion times of a request on ntt_clock_s's
 * clock. */
bool ntt_server_times(ntt_server *server, int64_t ticket, double *submit_s, double *start_s,
                      double *done_s);

/* Forgets an ended request (its ticket becomes unknown). */
void ntt_server_release(ntt_server *server, int64_t ticket);

void ntt_server_get_stats(ntt_server *server, ntt_server_stats *out);

/* Monotonic clock of the server's time stamps, in seconds. */
double ntt_clock_s(void);

/* Pinned host memory (hipHostMalloc) for request buffers: copies from it run
 * at full link speed. */
void *ntt_host_alloc(size_t bytes);
void ntt_host_free(void *ptr);

/* ------------------------------------------------------------------------ */
/* Engine utilities (no counterpart in the reference).                       */

/* 1 when a gfx950 device is usable, 0 otherwise. Never initializes more than
 * the HIP runtime. */
int ntt_device_available(void);

/* Device memory for callers without their own allocator. */
void *ntt_device_alloc(size_t bytes);
void ntt_device_free(void *ptr);
bool ntt_copy_to_device(void *d_dst, const void *src, size_t bytes);
bool ntt_copy_to_host(void *dst, const void *d_src, size_t bytes);
bool ntt_device_synchronize(void);

/* Last error message of the calling thread ("" if none). */
const char *ntt_last_error(void);

/* Per-kernel timing. While enabled, every kernel launch of the engine is
 * bracketed by HIP events on the stream it runs on; ntt_profile_read returns
 * per-kernel-kind totals. `names` receives up to `cap` NUL-separated kind
 * names; `ms` and `counts` receive totals. Returns the number of kinds. */
void ntt_profile_enable(int on);
int ntt_profile_read(char *names, size_t names_len, double *ms, uint64_t *counts, int cap);
void ntt_profile_reset(void);

/* Re-issues the last launch of kernel kind `kind` recorded while profiling
 * (same kernel, grid and argument block, on the stream it ran on) `reps`
 * times back to back between two HIP events on that stream, after one warm
 * launch, and writes the average duration per launch in milliseconds: the
 * kernel's own time without per-launch event overhead. The replays rewrite
 * that launch's output buffers, so run the operation again before reading
 * its result. false when no launch of that kind was recorded. No
 * counterpart in the reference (measurement). */
bool ntt_profile_replay(const char *kind, int reps, double *avg_ms);

/* Pass split the engine runs for a transform of 2^log_len points as a
 * single operation's product (use 0), its square (use 1) or a batch bucket
 * (use 2): returns the number of passes and writes their level counts
 * (strided passes first, the row pass last) to pass_logs[0..n) and, when
 * `flags` is not NULL, the kernel forms chosen with it (NTT_FORM_*); the split
 * is the one the autotuner chose for this length and use, the one
 * NTTAMD_ROW_LOG / NTTAMD_PASSES (and NTTAMD_ALT_TILES, NTTAMD_INV16,
 * NTTAMD_PER_PRIME_PACK, NTTAMD_ROWS8, NTTAMD_STRIDED8, NTTAMD_PAIRED,
 * NTTAMD_WAVES set to 0 or 1) force, or the default. 0 for lengths out of
 * range. No counterpart in the reference (the reference's split lives in its
 * Geometry). */
#define NTT_FORM_ALT_TILES 1u        /* strided passes use the alternative tiles */
#define NTT_FORM_INV16 2u            /* product rows run the inverse in radix 16 (else 8) */
#define NTT_FORM_PER_PRIME_PACK 4u   /* first pass packs per (tile, prime), not shared across primes */
#define NTT_FORM_ROWS8 8u            /* row kernel runs radix-8 rounds with twice the threads */
#define NTT_FORM_STRIDED8 16u        /* strided passes run radix-8 rounds with twice the threads */
#define NTT_FORM_PAIRED 32u          /* product rows: one thread per product group, both operands' groups */
#define NTT_FORM_WAVES 64u           /* 2^5- and 2^6-point strided passes in wave-owned lines, no barrier */
#define NTT_FORM_TUNED 256u          /* the autotuner has timed this length and use */
int ntt_transform_split(uint32_t log_len, int use, uint32_t *pass_logs, uint32_t *flags);

/* ---------------------------------------------------------------------------
 * Prime fields and two-prime CRT: the public scalar API of the reference
 * (src/field/prime_field.rs:91-442). Host functions (no GPU). A residue is a
 * canonical uint64_t in [0, p); a field's operations expect its own residues.
 */
typedef struct ntt_prime_field {
    uint64_t modulus;  /* prime, 2^49 < modulus < 2^50 */
    double reciprocal; /* RN(1/modulus) */
} ntt_prime_field;

/* PrimeField::new (prime_field.rs:99-116): false unless modulus is a prime
 * with 2^49 < modulus < 2^50. */
bool ntt_field_new(uint64_t modulus, ntt_prime_field *out);
/* PrimeField::below (118-129): the largest prime below bound and 2^50,
 * false when none exceeds 2^49. */
bool ntt_field_below(uint64_t bound, ntt_prime_field *out);
/* PrimeField::catalog (131-137): the eight transform primes in
 * reconstruction order; writes min(cap, 8), returns 8. */
size_t ntt_field_catalog(ntt_prime_field *out, size_t cap);
/* PrimeField::residue (147-153): false unless value < modulus. */
bool ntt_field_residue(const ntt_prime_field *f, uint64_t value, uint64_t *out);
/* PrimeField::reduce, reduce_u64, signed (158-199): canonical residue of a
 * signed or unsigned integer; the centered representative in
 * -(p-1)/2..=(p-1)/2. */
uint64_t ntt_field_reduce(const ntt_prime_field *f, int64_t value);
uint64_t ntt_field_reduce_u64(const ntt_prime_field *f, uint64_t value);
int64_t ntt_field_signed(const ntt_prime_field *f, uint64_t residue);
/* PrimeField::add, sub, neg, mul, pow (205-265). */
uint64_t ntt_field_add(const ntt_prime_field *f, uint64_t a, uint64_t b);
uint64_t ntt_field_sub(const ntt_prime_field *f, uint64_t a, uint64_t b);
uint64_t ntt_field_neg(const ntt_prime_field *f, uint64_t a);
uint64_t ntt_field_mul(const ntt_prime_field *f, uint64_t a, uint64_t b);
uint64_t ntt_field_pow(const ntt_prime_field *f, uint64_t base, uint64_t exponent);
/* PrimeField::inv (271-290): false for zero. */
bool ntt_field_inv(const ntt_prime_field *f, uint64_t a, uint64_t *out);
/* PrimeField::mul_sub, scale, evaluate (296-320): target[i] -= source[i] *
 * factor; target[i] *= factor; Horner evaluation of coefficients (constant
 * term first) at point. */
void ntt_field_mul_sub(const ntt_prime_field *f, uint64_t *target, const uint64_t *source, size_t n,
                       uint64_t factor);
void ntt_field_scale(const ntt_prime_field *f, uint64_t *target, size_t n, uint64_t factor);
uint64_t ntt_field_evaluate(const ntt_prime_field *f, const uint64_t *coefficients, size_t n,
                            uint64_t point);

typedef struct ntt_crt_pair {
    ntt_prime_field first;
    ntt_prime_field second;
    uint64_t inverse; /* first modulus^-1 in the second field */
} ntt_crt_pair;

/* CrtPair::new (404-416): false when the fields coincide. */
bool ntt_crt_pair_new(const ntt_prime_field *first, const ntt_prime_field *second, ntt_crt_pair *out);
/* CrtPair::modulus (418-422): the 128-bit product of the moduli as two words. */
void ntt_crt_pair_modulus(const ntt_crt_pair *pair, uint64_t *lo, uint64_t *hi);
/* CrtPair::combine (424-441): the integer in -(M-1)/2..=(M-1)/2 congruent to
 * both residues, as a 128-bit two's complement value (lo word, signed hi word). */
void ntt_crt_pair_combine(const ntt_crt_pair *pair, uint64_t first, uint64_t second, uint64_t *lo,
                          int64_t *hi);

/* Unbalanced split the engine chose for a product shape
 * (src/profile/planner.rs:283-393): the longer operand runs in `blocks`
 * blocks of block_digits coefficients, each on a row of 2^log_len points
 * (ntt_mul_geometry) against the shorter operand's spectrum; blocks == 0
 * for a product that is not split. NTTAMD_SPLIT=0 disables splits, =1 takes
 * the cheapest admissible one (tests); results are the same limbs. */
bool ntt_mul_split(size_t len_a, size_t len_b, uint32_t *blocks, uint64_t *block_digits);

/* 1 when the engine runs a product of this shape as a schoolbook instead
 * of transforms: direct sums of the 128-bit limb products, one launch of
 * 1024-limb output blocks plus the carry pass, for products whose shorter
 * operand has at most 512 limbs and whose transforms would cost more (the
 * basecase below the transform crossover that README.md:6-8 leaves to
 * callers). NTTAMD_SCHOOL=0 disables it, =1 takes it wherever admissible
 * (tests); results are the same limbs. 0 for an unsupported shape. */
int ntt_mul_schoolbook(size_t len_a, size_t len_b);

/* Row shape of src/profile/planner.rs:565-591 (row_shape): the shortest
 * row of 2^k, 5N/4, 3N/2 or 7N/4 points (N = 2^(k-1) holding both operands,
 * every tail of at least 32 points) for a convolution of conv_len
 * coefficients whose operands have at most `longest`, with its cyclic
 * length and root-table length. A single operation runs this composite row
 * when it prices below the power-of-two row (host-only; needs no GPU). */
bool ntt_row_shape(uint64_t conv_len, uint64_t longest, uint64_t *row_len, uint64_t *cyclic_len,
                   uint64_t *table_len);

/* Geometry the engine chose for a product shape (for reports and tests):
 * the transform row has N + row_tail * N/4 points, N = 2^log_len. */
bool ntt_mul_geometry(size_t len_a, size_t len_b, uint32_t *primes, uint32_t *chunk_bits,
                      uint32_t *log_len, uint32_t *row_tail);

/* Diagnostics (no counterpart in the reference; used by tests). */

/* Checks the arithmetic the kernels share with the host (residue products,
 * packing, Garner recombination, carry functions) against exact integer
 * arithmetic on the host. Returns the number of failures; needs no GPU. */
int ntt_selftest_host(uint64_t seed);

/* Copies `bytes` (a multiple of 16) between device buffers with a plain copy
 * kernel, the calibration reference for hardware byte counters. */
bool ntt_debug_copy(void *d_dst, const void *d_src, uint64_t bytes);

/* Copies the device root tables of all eight primes at `len` entries each. */
bool ntt_debug_tables(uint64_t len, double *fwd, double *inv);

/* Runs the transforms of a product and copies the canonical residue rows
 * (primes x row length) to `out`; `geometry` (4 entries) receives primes,
 * chunk bits, log2 of the cyclic length N and the row tail (row length
 * N + tail * N/4). With out == NULL only the geometry is reported. */
bool ntt_debug_residues(const uint64_t *a, size_t la, const uint64_t *b, size_t lb, int square,
                        double *out, size_t out_len, uint32_t *geometry);

/* Runs the CRT and carry kernels of a product shape on host-given residue
 * rows and copies the la+lb result limbs to `out`. */
bool ntt_debug_crt(size_t la, size_t lb, const double *rows, uint64_t *out);

#ifdef __cplusplus
}
#endif

#endif
