/*
 * ntt_oracle.c -- CPU restatement of the ntt_anafis multiplication path.
 *
 * TEST INFRASTRUCTURE: the checker for the HIP engine. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library; the product never links it. See ntt_oracle.h for the scope of the
 * restatement (power-of-two rows, no block split).
 *
 * Every function cites the reference file:line it restates. Paths are
 * relative to the reference crate root.
 */
#include "ntt_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

#define MAX_PRIMES 8
#define WORDS 7                 /* ceil(8*50/64), src/limb/cofactor.rs:19 */
#define PART_BITS 49            /* src/limb/packing.rs:23 */
#define MAX_PARTS 4             /* src/limb/packing.rs:31-37 */
#define MAX_CHUNK_BITS (PART_BITS * MAX_PARTS) /* src/profile/catalog.rs:8 */
#define LEVEL_ROOTS 43          /* src/field/catalog.rs:33 */
#define MIN_TRANSFORM_LEN 32    /* src/arch/contract.rs:62-66 */
#define MAX_ROOT_ORDER 41       /* src/field/catalog.rs:100-112 for this catalog */
#define SCHOOLBOOK_MAX_TERMS 24 /* src/convolve.rs:24 */
#define COLUMN_ALIGN 8          /* src/arch/contract.rs:55 */

/* PRIME_CATALOG, src/field/catalog.rs:67-76: (modulus, generator). */
static const uint64_t CATALOG[MAX_PRIMES][2] = {
    {1108307720798209ULL, 11}, {1086317488242689ULL, 3}, {910395627798529ULL, 7},
    {659706976665601ULL, 11},  {699289395265537ULL, 5},  {1022545813831681ULL, 7},
    {1013749720809473ULL, 3},  {868614185943041ULL, 3},
};

typedef struct {
    uint64_t inverses[MAX_PRIMES];
    uint64_t factors[MAX_PRIMES][WORDS];
    uint64_t multiples[MAX_PRIMES][WORDS];
    uint32_t capacity;
    uint64_t quotient_factor;
} cofactors_t;

static struct {
    double p[MAX_PRIMES], pinv[MAX_PRIMES];
    double level_root[MAX_PRIMES][LEVEL_ROOTS];
    double powers[MAX_PRIMES][MAX_PARTS];
    cofactors_t cof[MAX_PRIMES]; /* index = prefix count - 1 */
} K;

static pthread_once_t k_once = PTHREAD_ONCE_INIT;

/* Prime::pow_mod, src/field/catalog.rs:370-388. */
static uint64_t pow_mod(uint64_t base, uint64_t e, uint64_t m) {
    u128 v = 1, f = base % m;
    while (e) {
        if (e & 1) v = (v * f) % m;
        e >>= 1;
        if (e) f = (f * f) % m;
    }
    return (uint64_t)v;
}

/* nearest: round half to even, |x| < 2^52 (src/arch/scalar/product.rs:186-221). */
static inline double nearest(double x) { return nearbyint(x); }

/* Scalar::reduce, src/arch/scalar/product.rs:99-101: |x| < 8p -> < 9p/16. */
static inline double reduce_p(double x, double p, double pinv) {
    return fma(-nearest(x * pinv), p, x);
}

/* Scalar::mulmod, src/arch/scalar/product.rs:105-108: exact for |ab| < 4p^2. */
static inline double mulmod_p(double a, double b, double p, double pinv) {
    double high = a * b;
    return fma(a, b, -high) + fma(-nearest(high * pinv), p, high);
}

/* capacity = floor(log2(2P/3)), src/limb/cofactor.rs:239-255. */
static uint32_t capacity_of(const uint64_t *product, int words) {
    uint64_t top = product[words - 1];
    int bits = (words - 1) * 64 + (64 - __builtin_clzll(top));
    int second = bits - 2;
    if ((product[second >> 6] >> (second % 64)) & 1) return (uint32_t)(bits - 1);
    return (uint32_t)second;
}

/* quotient_factor, src/limb/cofactor.rs:269-300. */
static uint64_t quotient_factor_of(const uint64_t *product, uint32_t capacity) {
    int words = (int)((capacity + 63) / 64);
    int top = words - 1;
    uint64_t low = top == 0 ? 0 : product[top - 1];
    u128 leading = ((u128)product[top] << 64) | low;
    u128 factor = (~(u128)0) / (leading + 1);
    return (uint64_t)factor;
}

/* cofactor_tables, src/limb/cofactor.rs:136-228; catalog_prime level roots,
 * src/field/catalog.rs:314-331; packing powers, src/limb/packing.rs:371-405. */
static void init_constants(void) {
    for (int i = 0; i < MAX_PRIMES; i++) {
        uint64_t m = CATALOG[i][0], g = CATALOG[i][1];
        K.p[i] = (double)m;
        K.pinv[i] = 1.0 / (double)m;
        int v = __builtin_ctzll(m - 1);
        uint64_t half = m >> 1;
        for (int level = 0; level < LEVEL_ROOTS; level++) {
            if (level + 2 > v) { K.level_root[i][level] = 0.0; continue; }
            uint64_t power = pow_mod(g, (m - 1) >> (level + 2), m);
            K.level_root[i][level] = power > half ? -(double)(m - power) : (double)power;
        }
        u128 power = 1;
        for (int part = 0; part < MAX_PARTS; part++) {
            K.powers[i][part] = power > (m >> 1) ? (double)power - (double)m : (double)power;
            power = (power << PART_BITS) % m;
        }
    }
    uint64_t product[WORDS] = {1};
    int words = 1;
    for (int count = 0; count < MAX_PRIMES; count++) {
        uint64_t pm = CATALOG[count][0];
        u128 carry = 0;
        for (int w = 0; w < words; w++) {
            u128 full = (u128)product[w] * pm + carry;
            product[w] = (uint64_t)full;
            carry = full >> 64;
        }
        if (carry) product[words++] = (uint64_t)carry;
        cofactors_t *t = &K.cof[count];
        memset(t, 0, sizeof *t);
        for (int row = 0; row <= count; row++) {
            uint64_t q = CATALOG[row][0];
            u128 rem = 0;
            for (int w = words - 1; w >= 0; w--) {
                u128 dividend = (rem << 64) | product[w];
                t->factors[row][w] = (uint64_t)(dividend / q);
                rem = dividend % q;
            }
            rem = 0;
            for (int w = words - 1; w >= 0; w--) rem = ((rem << 64) | t->factors[row][w]) % q;
            t->inverses[row] = pow_mod((uint64_t)rem, q - 2, q);
        }
        uint64_t multiple[WORDS] = {0};
        for (int quotient = 0; quotient <= count; quotient++) {
            memcpy(t->multiples[quotient], multiple, sizeof multiple);
            unsigned overflow = 0;
            for (int w = 0; w < WORDS; w++) {
                uint64_t s = multiple[w] + product[w];
                unsigned c1 = s < multiple[w];
                uint64_t s2 = s + overflow;
                unsigned c2 = s2 < s;
                multiple[w] = s2;
                overflow = c1 | c2;
            }
        }
        t->capacity = capacity_of(product, words);
        t->quotient_factor = quotient_factor_of(product, t->capacity);
    }
}

static void ensure_init(void) { pthread_once(&k_once, init_constants); }

uint64_t oracle_prime(int index) { return (index >= 0 && index < MAX_PRIMES) ? CATALOG[index][0] : 0; }
uint64_t oracle_generator(int index) { return (index >= 0 && index < MAX_PRIMES) ? CATALOG[index][1] : 0; }

double oracle_mulmod(double a, double b, int prime) {
    ensure_init();
    return mulmod_p(a, b, K.p[prime], K.pinv[prime]);
}

double oracle_reduce(double a, int prime) {
    ensure_init();
    return reduce_p(a, K.p[prime], K.pinv[prime]);
}

uint32_t oracle_capacity(int primes) {
    ensure_init();
    if (primes < 1 || primes > MAX_PRIMES) return 0;
    return K.cof[primes - 1].capacity;
}

/* generate_tables, src/transform/twiddle.rs:116-156: W[j] = omega^rev(j),
 * W[2^m + k] = W[2^m] * W[k], inverse Winv[2^(m+1)-1-k] = -W[2^m+k]. */
static void generate_tables(double *fwd, double *inv, uint64_t len, int prime) {
    double p = K.p[prime], pinv = K.pinv[prime];
    fwd[0] = 1.0;
    if (inv) inv[0] = 1.0;
    uint64_t width = 1;
    while (width < len) {
        double root = K.level_root[prime][__builtin_ctzll(width)];
        for (uint64_t index = 0; index < width; index++)
            fwd[width + index] = reduce_p(mulmod_p(fwd[index], root, p, pinv), p, pinv);
        uint64_t end = width * 2;
        if (inv)
            for (uint64_t index = 0; index < width; index++) inv[end - 1 - index] = -fwd[width + index];
        width = end;
    }
}

void oracle_root_table(int primes, uint64_t table_len, int inverse, double *out) {
    ensure_init();
    double *scratch = malloc(sizeof(double) * table_len);
    for (int i = 0; i < primes; i++) {
        if (inverse) {
            generate_tables(scratch, out + (uint64_t)i * table_len, table_len, i);
        } else {
            generate_tables(out + (uint64_t)i * table_len, NULL, table_len, i);
        }
    }
    free(scratch);
}

/* Forward butterfly, src/arch/schedule/expression/stages.rs:365-381 (non-unit
 * node; W[0] = 1 covers the unit node). Inputs below 3p, outputs below 5p/2,
 * or below 3p/2 for the LAST stage (src/arch/contract.rs:20-44). */
static inline void forward4(double *a, double *b, double *c, double *d, double w, double left,
                            double right, int last, double p, double pinv) {
    double c1 = mulmod_p(*c, w, p, pinv);
    double d1 = mulmod_p(*d, w, p, pinv);
    double u, s;
    if (last) {
        u = reduce_p(*a + c1, p, pinv);
        s = reduce_p(*a - c1, p, pinv);
    } else {
        double centered = reduce_p(*a, p, pinv);
        u = centered + c1;
        s = centered - c1;
    }
    double v = mulmod_p(*b + d1, left, p, pinv);
    double t = mulmod_p(*b - d1, right, p, pinv);
    *a = u + v;
    *b = u - v;
    *c = s + t;
    *d = s - t;
}

/* Inverse butterfly, src/arch/schedule/expression/stages.rs:348-362. Inputs and
 * outputs below 2p. */
static inline void inverse4(double *a, double *b, double *c, double *d, double w, double left,
                            double right, double p, double pinv) {
    double u = *a + *b;
    double s = *c + *d;
    double v = mulmod_p(*a - *b, left, p, pinv);
    double t = mulmod_p(*c - *d, right, p, pinv);
    *a = reduce_p(u + s, p, pinv);
    *b = v + t;
    *c = mulmod_p(u - s, w, p, pinv);
    *d = mulmod_p(v - t, w, p, pinv);
}

/* Forward remainder tree of one row (src/transform/forward.rs:200-291): a
 * radix-2 root stage for an odd order (radix2_node, src/arch/schedule/
 * expression/tail.rs:106-149), then radix-4 stages at every node j with
 * W[j], W[2j], W[2j+1] down to single points. */
static void forward_tree(double *x, uint64_t len, const double *W, int prime) {
    double p = K.p[prime], pinv = K.pinv[prime];
    int logn = __builtin_ctzll(len);
    uint64_t n = len, nodes = 1;
    if (logn & 1) {
        uint64_t half = len / 2;
        double root = W[0];
        for (uint64_t i = 0; i < half; i++) {
            double rotated = mulmod_p(x[half + i], root, p, pinv);
            double centered = reduce_p(x[i], p, pinv);
            x[i] = centered + rotated;
            x[half + i] = centered - rotated;
        }
        n = half;
        nodes = 2;
    }
    while (n >= 4) {
        uint64_t q = n / 4;
        int last = (n == 4);
        for (uint64_t j = 0; j < nodes; j++) {
            double w = W[j], left = W[2 * j], right = W[2 * j + 1];
            double *base = x + j * n;
            for (uint64_t i = 0; i < q; i++)
                forward4(base + i, base + q + i, base + 2 * q + i, base + 3 * q + i, w, left, right,
                         last, p, pinv);
        }
        n = q;
        nodes *= 4;
    }
}

/* Inverse remainder tree (src/transform/inverse.rs:218-394): radix-4 DIT
 * stages from the leaves up, and a radix-2 root stage for an odd order. */
static void inverse_tree(double *x, uint64_t len, const double *Winv, int prime) {
    double p = K.p[prime], pinv = K.pinv[prime];
    int logn = __builtin_ctzll(len);
    uint64_t top = (logn & 1) ? len / 2 : len;
    for (uint64_t n = 4; n <= top; n *= 4) {
        uint64_t q = n / 4, nodes = len / n;
        for (uint64_t j = 0; j < nodes; j++) {
            double w = Winv[j], left = Winv[2 * j], right = Winv[2 * j + 1];
            double *base = x + j * n;
            for (uint64_t i = 0; i < q; i++)
                inverse4(base + i, base + q + i, base + 2 * q + i, base + 3 * q + i, w, left, right,
                         p, pinv);
        }
    }
    if (logn & 1) {
        uint64_t half = len / 2;
        double root = Winv[0];
        for (uint64_t i = 0; i < half; i++) {
            double a = x[i], b = x[half + i];
            x[i] = reduce_p(a + b, p, pinv);
            x[half + i] = mulmod_p(a - b, root, p, pinv);
        }
    }
}

void oracle_forward_row(double *row, uint64_t len, int prime) {
    ensure_init();
    double *W = malloc(sizeof(double) * (len / 2));
    generate_tables(W, NULL, len / 2, prime);
    forward_tree(row, len, W, prime);
    free(W);
}

void oracle_inverse_row(double *row, uint64_t len, int prime) {
    ensure_init();
    double *W = malloc(sizeof(double) * (len / 2));
    double *Winv = malloc(sizeof(double) * (len / 2));
    generate_tables(W, Winv, len / 2, prime);
    inverse_tree(row, len, Winv, prime);
    free(W);
    free(Winv);
}

/* ------------------------------------------------------------------------ */
/* Geometry, src/profile/planner.rs.                                         */

static uint64_t next_pow2(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

static int bitlen64(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

/* product_bits, src/profile/planner.rs:558-563. */
static uint64_t product_bits(uint64_t width, uint64_t terms) {
    return 2 * width + (uint64_t)bitlen64(terms - 1);
}

/* admit, src/profile/planner.rs:211-281, with row_shape (573-591) restricted
 * to power-of-two rows. */
static int admit(int primes, uint64_t chunk, uint64_t da, uint32_t vba, uint64_t db, uint32_t vbb,
                 uint64_t coefficient_bits, uint64_t output_stride, int blocks, oracle_geometry *g) {
    uint32_t capacity = K.cof[primes - 1].capacity;
    int disjoint = blocks >= 0;
    if (chunk > MAX_CHUNK_BITS || vba > chunk || vbb > chunk || coefficient_bits > capacity) return 0;
    if (disjoint && (output_stride % 64 != 0 || output_stride < coefficient_bits)) return 0;
    uint64_t conv_len = da + db - 1;
    uint64_t len = next_pow2(conv_len);
    if (len < MIN_TRANSFORM_LEN) len = MIN_TRANSFORM_LEN;
    uint64_t table_len = len / 2;
    if ((uint64_t)__builtin_ctzll(table_len * 2) > MAX_ROOT_ORDER) return 0;
    g->primes = (uint32_t)primes;
    g->chunk_bits = (uint32_t)chunk;
    g->digits_a = da;
    g->digits_b = db;
    g->value_bits_a = vba;
    g->value_bits_b = vbb;
    g->conv_len = conv_len;
    g->transform_len = len;
    g->disjoint = (uint32_t)disjoint;
    g->sign_bit = (blocks == 1) ? (int32_t)(coefficient_bits - 1) : -1;
    g->output_stride = output_stride;
    g->capacity = capacity;
    return 1;
}

/* under, src/profile/planner.rs:190-202. */
static int under(uint64_t bits_a, uint64_t bits_b, int primes, uint64_t width, oracle_geometry *g) {
    uint64_t da = (bits_a + width - 1) / width, db = (bits_b + width - 1) / width;
    uint64_t terms = da < db ? da : db;
    return admit(primes, width, da, (uint32_t)width, db, (uint32_t)width, product_bits(width, terms),
                 width, -1, g);
}

/* widest_under, src/profile/planner.rs:168-183. */
static int widest_under(uint64_t bits_a, uint64_t bits_b, int primes, oracle_geometry *g) {
    uint32_t capacity = K.cof[primes - 1].capacity;
    uint64_t width = capacity / 2 < MAX_CHUNK_BITS ? capacity / 2 : MAX_CHUNK_BITS;
    for (;;) {
        uint64_t ta = (bits_a + width - 1) / width, tb = (bits_b + width - 1) / width;
        uint64_t terms = ta < tb ? ta : tb;
        uint64_t pb = product_bits(width, terms);
        uint64_t excess = pb > capacity ? pb - capacity : 0;
        if (excess == 0) return under(bits_a, bits_b, primes, width, g);
        uint64_t cut = (excess + 1) / 2;
        if (cut >= width) return 0;
        width -= cut;
    }
}

/* work, src/profile/planner.rs:401-422 (power-of-two rows have no tails). */
static double work_score(const oracle_geometry *g, int crt_cost) {
    double n = (double)g->transform_len;
    double transform = (double)g->primes * n * (double)__builtin_ctzll(g->transform_len);
    double crt = (double)crt_cost * g->primes * (g->primes - 1) * (double)g->conv_len;
    return 3.0 * transform + crt;
}

/* Geometry::select, src/profile/planner.rs:147-161. */
static int select_geometry(uint64_t la, uint64_t lb, oracle_geometry *out) {
    if (la > (UINT64_MAX / 64) || lb > (UINT64_MAX / 64)) return 0;
    uint64_t bits_a = la * 64, bits_b = lb * 64;
    int found = 0;
    double best = 0;
    for (int primes = 1; primes <= MAX_PRIMES; primes++) {
        oracle_geometry g;
        if (!widest_under(bits_a, bits_b, primes, &g)) continue;
        double score = work_score(&g, 1);
        if (!found || score < best) {
            best = score;
            *out = g;
            found = 1;
        }
    }
    return found;
}

/* MulPlan::new, src/plan.rs:53-69. */
int oracle_mul_plan_new(uint64_t len_a, uint64_t len_b, oracle_mul_plan *out) {
    ensure_init();
    if (len_a == 0 || len_b == 0) return 0;
    if (len_a + len_b < len_a) return 0;
    oracle_mul_plan plan;
    if (!select_geometry(len_a, len_b, &plan.geometry)) return 0;
    plan.len_a = len_a;
    plan.len_b = len_b;
    *out = plan;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Packing, src/limb/packing.rs:113-292 and src/arch/schedule/expression/
 * packing.rs:68-109.                                                        */

/* checked_window, src/limb/packing.rs:338-364: 49 bits at a bit position,
 * zero beyond the input. */
static uint64_t checked_window(const uint64_t *limbs, uint64_t len, uint64_t position) {
    uint64_t index = position / 64, offset = position % 64;
    uint64_t value = index < len ? limbs[index] >> offset : 0;
    if (offset + PART_BITS > 64 && index + 1 < len) value |= limbs[index + 1] << (64 - offset);
    return value;
}

typedef struct {
    const uint64_t *input;
    uint64_t input_len;
    uint64_t width;
    uint32_t value_bits;
    int is_signed;
    double *rows; /* primes rows of `stride` */
    uint64_t stride;
    int primes;
    uint64_t first, end;
} pack_job;

static void pack_range(const pack_job *job) {
    uint32_t parts = (job->value_bits + PART_BITS - 1) / PART_BITS;
    uint32_t high_bits = job->value_bits - (parts - 1) * PART_BITS;
    uint64_t high_mask = (1ULL << high_bits) - 1;
    uint64_t part_mask = (1ULL << PART_BITS) - 1;
    uint64_t flip = job->is_signed ? (1ULL << (high_bits - 1)) : 0;
    int short_high = parts > 1 && high_bits <= 3;
    for (uint64_t i = job->first; i < job->end; i++) {
        double digit[MAX_PARTS];
        for (uint32_t part = 0; part < parts; part++) {
            uint64_t word;
            if (job->width == 64) {
                word = i < job->input_len ? job->input[i] >> (part * PART_BITS) : 0;
            } else {
                word = checked_window(job->input, job->input_len, i * job->width + part * PART_BITS);
            }
            if (part + 1 == parts) {
                word &= high_mask;
                digit[part] = (double)(int64_t)(word ^ flip) - (double)(int64_t)flip;
            } else {
                digit[part] = (double)(word & part_mask);
            }
        }
        for (int prime = 0; prime < job->primes; prime++) {
            double p = K.p[prime], pinv = K.pinv[prime];
            double sum = digit[0];
            if (parts > 1) {
                for (uint32_t part = 1; part + 1 < parts; part++)
                    sum += mulmod_p(digit[part], K.powers[prime][part], p, pinv);
                double power = K.powers[prime][parts - 1];
                double product = short_high ? digit[parts - 1] * power
                                            : mulmod_p(digit[parts - 1], power, p, pinv);
                sum = reduce_p(sum + product, p, pinv);
            }
            job->rows[(uint64_t)prime * job->stride + i] = sum;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* A fork-join helper standing in for the reference's executor joins.        */

typedef void (*task_fn)(void *ctx, uint64_t index);

typedef struct {
    task_fn fn;
    void *ctx;
    uint64_t count;
    uint64_t next;
    pthread_mutex_t lock;
} pool_t;

static void *pool_worker(void *arg) {
    pool_t *pool = arg;
    for (;;) {
        pthread_mutex_lock(&pool->lock);
        uint64_t index = pool->next++;
        pthread_mutex_unlock(&pool->lock);
        if (index >= pool->count) break;
        pool->fn(pool->ctx, index);
    }
    return NULL;
}

static void parallel_for(int threads, uint64_t count, task_fn fn, void *ctx) {
    if (threads <= 1 || count <= 1) {
        for (uint64_t i = 0; i < count; i++) fn(ctx, i);
        return;
    }
    pool_t pool = {fn, ctx, count, 0, PTHREAD_MUTEX_INITIALIZER};
    int n = threads < (int)count ? threads : (int)count;
    pthread_t tid[256];
    if (n > 256) n = 256;
    for (int t = 1; t < n; t++) pthread_create(&tid[t], NULL, pool_worker, &pool);
    pool_worker(&pool);
    for (int t = 1; t < n; t++) pthread_join(tid[t], NULL);
}

/* ------------------------------------------------------------------------ */
/* The product pipeline, src/multiply/execute.rs:419-459 and
 * src/multiply/schedule.rs:160-281.                                          */

typedef struct {
    const oracle_geometry *g;
    const uint64_t *a, *b;
    uint64_t a_len, b_len;
    int square;
    double *left, *right; /* primes rows of transform_len */
    double *fwd, *inv;    /* primes tables of table_len */
    uint64_t table_len;
    uint64_t chunk; /* pack coefficients per task */
} pipe_ctx;

static void task_tables(void *arg, uint64_t prime) {
    pipe_ctx *c = arg;
    generate_tables(c->fwd + prime * c->table_len, c->inv + prime * c->table_len, c->table_len,
                    (int)prime);
}

static void task_pack(void *arg, uint64_t index) {
    pipe_ctx *c = arg;
    const oracle_geometry *g = c->g;
    uint64_t n = g->transform_len;
    int side = c->square ? 0 : (int)(index & 1);
    uint64_t tile = c->square ? index : index >> 1;
    const uint64_t *input = side ? c->b : c->a;
    uint64_t input_len = side ? c->b_len : c->a_len;
    uint64_t digits = side ? g->digits_b : g->digits_a;
    uint64_t first = tile * c->chunk, end = first + c->chunk;
    if (first >= n) return;
    if (end > n) end = n;
    uint64_t active_end = end < digits ? end : digits;
    double *rows = side ? c->right : c->left;
    if (first < active_end) {
        pack_job job = {input, input_len, g->chunk_bits,
                        side ? g->value_bits_b : g->value_bits_a,
                        g->sign_bit >= 0, rows, n, (int)g->primes, first, active_end};
        pack_range(&job);
    }
    uint64_t zero_from = first > active_end ? first : active_end;
    for (uint32_t prime = 0; prime < g->primes; prime++)
        for (uint64_t i = zero_from; i < end; i++) rows[(uint64_t)prime * n + i] = 0.0;
}

static void task_forward(void *arg, uint64_t index) {
    pipe_ctx *c = arg;
    uint64_t n = c->g->transform_len;
    int prime = (int)(c->square ? index : index >> 1);
    int side = c->square ? 0 : (int)(index & 1);
    double *row = (side ? c->right : c->left) + (uint64_t)prime * n;
    forward_tree(row, n, c->fwd + (uint64_t)prime * c->table_len, prime);
}

/* inverse_prime, src/multiply/schedule.rs:225-281: pointwise product, inverse
 * tree, then normalization by N^-1 (P/p_i)^-1 and the canonical store of
 * src/arch/contract.rs:46-51 (here as an integer-valued float in 0..p). */
static void task_inverse(void *arg, uint64_t prime_index) {
    pipe_ctx *c = arg;
    const oracle_geometry *g = c->g;
    int prime = (int)prime_index;
    uint64_t n = g->transform_len;
    double p = K.p[prime], pinv = K.pinv[prime];
    double *left = c->left + (uint64_t)prime * n;
    const double *right = c->square ? left : c->right + (uint64_t)prime * n;
    for (uint64_t i = 0; i < n; i++) left[i] = mulmod_p(left[i], right[i], p, pinv);
    inverse_tree(left, n, c->inv + (uint64_t)prime * c->table_len, prime);
    uint64_t m = CATALOG[prime][0];
    uint64_t inverse_scale = m - ((m - 1) >> __builtin_ctzll(n));
    uint64_t factor = (uint64_t)(((u128)inverse_scale * K.cof[g->primes - 1].inverses[prime]) % m);
    double centered = factor > m / 2 ? (double)factor - p : (double)factor;
    for (uint64_t i = 0; i < g->conv_len; i++) {
        double r = mulmod_p(left[i], centered, p, pinv);
        left[i] = r < 0 ? r + p : r;
    }
}

/* Cofactors::bounded, src/limb/cofactor.rs:59-111. `r` holds the weighted
 * canonical residues; writes ceil(capacity/64) words. */
static void cofactor_value(const cofactors_t *t, int count, const uint64_t *r, uint64_t *value) {
    int words = (int)((t->capacity + 63) / 64);
    int factor_words = count == 1 ? 1 : (int)((50 * (count - 1) + 63) / 64);
    u128 carry = 0;
    for (int w = 0; w < words; w++) {
        u128 sum = carry;
        if (w < factor_words)
            for (int row = 0; row < count; row++) sum += (u128)t->factors[row][w] * r[row];
        carry = sum >> 64;
        value[w] = (uint64_t)sum;
    }
    uint64_t top = value[words - 1];
    u128 estimate = (u128)top * t->quotient_factor + ((u128)1 << 62);
    const uint64_t *multiple = t->multiples[(int)(estimate >> 64)];
    unsigned borrow = 0;
    for (int w = 0; w < words; w++) {
        uint64_t x = value[w], y = multiple[w];
        uint64_t d = x - y - borrow;
        borrow = (x < y) || (x - y < borrow);
        value[w] = d;
    }
}

typedef struct {
    pipe_ctx *pc;
    uint64_t coefficients;
    uint64_t block;   /* coefficients per task, a multiple of 64 */
    uint64_t **local; /* per task accumulator */
    uint64_t *local_len;
} crt_ctx;

/* add_shifted, src/limb/reconstruction.rs:434-455: adds value*2^bit into acc. */
static void add_shifted(uint64_t *acc, uint64_t acc_len, const uint64_t *value, int words,
                        uint64_t bit) {
    uint64_t limb = bit / 64;
    unsigned shift = bit % 64;
    uint64_t shifted[WORDS + 2];
    int n = words + 1;
    for (int i = 0; i < n; i++) {
        uint64_t lo = i < words ? value[i] : 0;
        uint64_t hi = (i > 0) ? value[i - 1] : 0;
        shifted[i] = shift ? ((lo << shift) | (hi >> (64 - shift))) : lo;
    }
    uint64_t carry = 0;
    uint64_t k = 0;
    for (; k < (uint64_t)n && limb + k < acc_len; k++) {
        uint64_t s = acc[limb + k] + shifted[k];
        uint64_t c1 = s < shifted[k];
        uint64_t s2 = s + carry;
        uint64_t c2 = s2 < s;
        acc[limb + k] = s2;
        carry = c1 | c2;
    }
    for (uint64_t idx = limb + k; carry && idx < acc_len; idx++) {
        acc[idx] += 1;
        carry = acc[idx] == 0;
    }
}

static void task_crt(void *arg, uint64_t task) {
    crt_ctx *c = arg;
    const oracle_geometry *g = c->pc->g;
    uint64_t n = g->transform_len;
    const cofactors_t *t = &K.cof[g->primes - 1];
    int words = (int)((t->capacity + 63) / 64);
    uint64_t first = task * c->block;
    uint64_t end = first + c->block;
    if (end > c->coefficients) end = c->coefficients;
    uint64_t width = g->output_stride;
    uint64_t limbs = ((end - first) * width + 64 * (uint64_t)(words + 2)) / 64 + 1;
    uint64_t *acc = calloc(limbs, sizeof(uint64_t));
    for (uint64_t i = first; i < end; i++) {
        uint64_t r[MAX_PRIMES], value[WORDS + 1] = {0};
        for (uint32_t prime = 0; prime < g->primes; prime++)
            r[prime] = (uint64_t)c->pc->left[(uint64_t)prime * n + i];
        cofactor_value(t, (int)g->primes, r, value);
        add_shifted(acc, limbs, value, words, (i - first) * width);
    }
    c->local[task] = acc;
    c->local_len[task] = limbs;
}

/* Adds every task's block into the output span with carries, the block merge
 * of src/limb/reconstruction.rs:211-247. */
static void merge_blocks(crt_ctx *c, uint64_t tasks, uint64_t *out, uint64_t out_len) {
    uint64_t width = c->pc->g->output_stride;
    memset(out, 0, out_len * sizeof(uint64_t));
    for (uint64_t task = 0; task < tasks; task++) {
        uint64_t start = task * c->block * width / 64; /* block is a multiple of 64 */
        uint64_t carry = 0;
        uint64_t k = 0;
        for (; k < c->local_len[task] && start + k < out_len; k++) {
            uint64_t x = c->local[task][k];
            uint64_t s = out[start + k] + x;
            uint64_t c1 = s < x;
            uint64_t s2 = s + carry;
            uint64_t c2 = s2 < s;
            out[start + k] = s2;
            carry = c1 | c2;
        }
        for (uint64_t idx = start + k; carry && idx < out_len; idx++) {
            out[idx] += 1;
            carry = out[idx] == 0;
        }
        free(c->local[task]);
    }
}

/* Stores disjoint coefficient blocks, src/limb/reconstruction.rs:255-261 and
 * 395-425. */
static void task_disjoint(void *arg, uint64_t task) {
    crt_ctx *c = arg;
    const oracle_geometry *g = c->pc->g;
    uint64_t n = g->transform_len;
    const cofactors_t *t = &K.cof[g->primes - 1];
    uint64_t stride = g->output_stride / 64;
    uint64_t first = task * c->block;
    uint64_t end = first + c->block;
    if (end > c->coefficients) end = c->coefficients;
    uint64_t *out = c->local[0];
    for (uint64_t i = first; i < end; i++) {
        uint64_t r[MAX_PRIMES], value[WORDS + 1] = {0};
        for (uint32_t prime = 0; prime < g->primes; prime++)
            r[prime] = (uint64_t)c->pc->left[(uint64_t)prime * n + i];
        cofactor_value(t, (int)g->primes, r, value);
        for (uint64_t w = 0; w < stride; w++) out[i * stride + w] = w < WORDS ? value[w] : 0;
        if (g->sign_bit >= 0) {
            /* Subtract 2^s inside the block, src/limb/reconstruction.rs:148-163. */
            uint64_t limb = (uint64_t)g->sign_bit / 64;
            uint64_t borrow = 1ULL << (g->sign_bit % 64);
            for (uint64_t w = limb; w < stride; w++) {
                uint64_t x = out[i * stride + w];
                out[i * stride + w] = x - borrow;
                borrow = x < borrow;
            }
        }
    }
}

/* Runs the whole pipeline into `out` (exactly the output span). */
static void run_pipeline(const oracle_geometry *g, uint64_t *out, uint64_t out_len,
                         const uint64_t *a, uint64_t a_len, const uint64_t *b, uint64_t b_len,
                         int square, int threads) {
    uint64_t n = g->transform_len;
    pipe_ctx c;
    memset(&c, 0, sizeof c);
    c.g = g;
    c.a = a;
    c.b = b;
    c.a_len = a_len;
    c.b_len = b_len;
    c.square = square;
    c.table_len = n / 2;
    c.left = malloc(sizeof(double) * n * g->primes);
    c.right = square ? NULL : malloc(sizeof(double) * n * g->primes);
    c.fwd = malloc(sizeof(double) * c.table_len * g->primes);
    c.inv = malloc(sizeof(double) * c.table_len * g->primes);
    parallel_for(threads, g->primes, task_tables, &c);
    c.chunk = 1ULL << 16;
    uint64_t tiles = (n + c.chunk - 1) / c.chunk;
    parallel_for(threads, tiles * (square ? 1 : 2), task_pack, &c);
    parallel_for(threads, (uint64_t)g->primes * (square ? 1 : 2), task_forward, &c);
    parallel_for(threads, g->primes, task_inverse, &c);

    /* Signed offset: add the weighted residue of 2^s to every residue,
     * src/limb/reconstruction.rs:94-113. */
    if (g->sign_bit >= 0) {
        for (uint32_t prime = 0; prime < g->primes; prime++) {
            uint64_t m = CATALOG[prime][0];
            uint64_t power = 1ULL << (g->sign_bit % PART_BITS);
            for (int k = 0; k < g->sign_bit / PART_BITS; k++)
                power = (uint64_t)(((u128)power << PART_BITS) % m);
            uint64_t offset =
                (uint64_t)(((u128)power * K.cof[g->primes - 1].inverses[prime]) % m);
            double *row = c.left + (uint64_t)prime * n;
            for (uint64_t i = 0; i < g->conv_len; i++) {
                uint64_t sum = (uint64_t)row[i] + offset;
                if (sum >= m) sum -= m;
                row[i] = (double)sum;
            }
        }
    }

    crt_ctx cc;
    memset(&cc, 0, sizeof cc);
    cc.pc = &c;
    uint64_t coefficients = (out_len * 64 + g->output_stride - 1) / g->output_stride;
    cc.coefficients = coefficients < g->conv_len ? coefficients : g->conv_len;
    cc.block = 64 * 1024;
    uint64_t tasks = (cc.coefficients + cc.block - 1) / cc.block;
    if (g->disjoint) {
        uint64_t *slot = out;
        cc.local = &slot;
        memset(out, 0, out_len * sizeof(uint64_t));
        parallel_for(threads, tasks, task_disjoint, &cc);
    } else {
        cc.local = calloc(tasks, sizeof(uint64_t *));
        cc.local_len = calloc(tasks, sizeof(uint64_t));
        parallel_for(threads, tasks, task_crt, &cc);
        merge_blocks(&cc, tasks, out, out_len);
        free(cc.local);
        free(cc.local_len);
    }
    free(c.left);
    free(c.right);
    free(c.fwd);
    free(c.inv);
}

/* Layout::run, src/multiply/storage.rs:95-156: lengths must match the plan,
 * dst must hold the product span; failure leaves dst untouched. */
static int mul_common(const oracle_mul_plan *plan, uint64_t *dst, uint64_t dst_len,
                      const uint64_t *a, uint64_t a_len, const uint64_t *b, uint64_t b_len,
                      int square, int threads) {
    ensure_init();
    if (a_len != plan->len_a || b_len != plan->len_b) return 0;
    uint64_t out_len = a_len + b_len;
    if (dst_len < out_len) return 0;
    uint64_t *out = malloc(sizeof(uint64_t) * out_len);
    run_pipeline(&plan->geometry, out, out_len, a, a_len, b, b_len, square, threads);
    memcpy(dst, out, sizeof(uint64_t) * out_len);
    free(out);
    return 1;
}

/* MulPlan::mul, src/plan.rs:94-102: the same slice twice squares. */
int oracle_mul(const oracle_mul_plan *plan, uint64_t *dst, uint64_t dst_len, const uint64_t *a,
               uint64_t a_len, const uint64_t *b, uint64_t b_len, int threads) {
    int square = (a == b) && (a_len == b_len);
    return mul_common(plan, dst, dst_len, a, a_len, b, b_len, square, threads);
}

/* MulPlan::sqr, src/plan.rs:112-114. */
int oracle_sqr(const oracle_mul_plan *plan, uint64_t *dst, uint64_t dst_len, const uint64_t *a,
               uint64_t a_len, int threads) {
    return mul_common(plan, dst, dst_len, a, a_len, a, a_len, 1, threads);
}

/* MulWorkspace::run, src/plan.rs:182-205. */
int oracle_workspace_mul(uint64_t *dst, uint64_t dst_len, const uint64_t *a, uint64_t a_len,
                         const uint64_t *b, uint64_t b_len, int square, int threads) {
    ensure_init();
    if (a_len == 0 || b_len == 0) {
        uint64_t out_len = a_len + b_len;
        if (dst_len < out_len) return 0;
        memset(dst, 0, sizeof(uint64_t) * out_len);
        return 1;
    }
    oracle_mul_plan plan;
    if (!oracle_mul_plan_new(a_len, b_len, &plan)) return 0;
    return mul_common(&plan, dst, dst_len, a, a_len, b, b_len, square, threads);
}

/* ------------------------------------------------------------------------ */
/* Convolution, src/convolve.rs.                                             */

/* bound_bits, src/convolve.rs:358-369. */
static int bound_bits(uint64_t terms, uint64_t max_a, uint64_t max_b, uint64_t *bits) {
    u128 product = (u128)max_a * max_b;
    int plen = product == 0 ? 0 : (product >> 64 ? 128 - __builtin_clzll((uint64_t)(product >> 64))
                                                 : 64 - __builtin_clzll((uint64_t)product));
    /* checked_mul by terms in u128 */
    u128 hi_check = product >> 64;
    u128 lo = (uint64_t)product;
    u128 t_lo = lo * terms;
    u128 t_hi = hi_check * terms + (t_lo >> 64);
    if (t_hi >> 64) {
        *bits = (uint64_t)plen + (uint64_t)bitlen64(terms - 1);
    } else {
        u128 bound = (t_hi << 64) | (uint64_t)t_lo;
        int blen = bound == 0 ? 0 : (bound >> 64 ? 128 - __builtin_clzll((uint64_t)(bound >> 64))
                                                 : 64 - __builtin_clzll((uint64_t)bound));
        *bits = (uint64_t)blen;
    }
    return 1;
}

/* ConvPlan::new, src/convolve.rs:75-113 and Geometry::for_residues,
 * src/profile/planner.rs:120-138. */
int oracle_conv_plan_new(uint64_t len_a, uint64_t len_b, uint64_t max_a, uint64_t max_b,
                         int is_signed, oracle_conv_plan *out) {
    ensure_init();
    if (len_a == 0 || len_b == 0) return 0;
    uint64_t terms = len_a < len_b ? len_a : len_b;
    uint64_t sign = is_signed ? 1 : 0;
    uint64_t bits;
    bound_bits(terms, max_a, max_b, &bits);
    bits += sign;
    uint64_t stride = (bits + 63) / 64;
    if (stride < 1) stride = 1;
    uint64_t vba = (uint64_t)bitlen64(max_a) + sign, vbb = (uint64_t)bitlen64(max_b) + sign;
    if (vba == 0) vba = 1;
    if (vbb == 0) vbb = 1;
    if (vba > 64 || vbb > 64) return 0;
    int primes = 0;
    for (int count = 1; count <= MAX_PRIMES; count++)
        if (K.cof[count - 1].capacity >= bits) { primes = count; break; }
    if (!primes) return 0;
    oracle_conv_plan plan;
    memset(&plan, 0, sizeof plan);
    if (!admit(primes, 64, len_a, (uint32_t)vba, len_b, (uint32_t)vbb, bits, stride * 64,
               is_signed ? 1 : 0, &plan.geometry))
        return 0;
    if (len_a + len_b < len_a) return 0;
    plan.len_a = len_a;
    plan.len_b = len_b;
    plan.max_a = max_a;
    plan.max_b = max_b;
    plan.is_signed = (uint32_t)is_signed;
    plan.schoolbook = terms <= SCHOOLBOOK_MAX_TERMS;
    plan.stride = stride;
    plan.output_len = (len_a + len_b - 1) * stride;
    *out = plan;
    return 1;
}

/* admits, src/convolve.rs:169-184. */
static int conv_admits(const oracle_conv_plan *plan, const uint64_t *a, uint64_t a_len,
                       const uint64_t *b, uint64_t b_len) {
    if (a_len != plan->len_a || b_len != plan->len_b) return 0;
    const uint64_t *ops[2] = {a, b};
    uint64_t lens[2] = {a_len, b_len}, maxes[2] = {plan->max_a, plan->max_b};
    for (int s = 0; s < 2; s++) {
        if (plan->is_signed) {
            for (uint64_t i = 0; i < lens[s]; i++) {
                int64_t v = (int64_t)ops[s][i];
                uint64_t mag = v < 0 ? (uint64_t)0 - (uint64_t)v : (uint64_t)v;
                if (mag > maxes[s]) return 0;
            }
        } else {
            uint64_t u = 0;
            for (uint64_t i = 0; i < lens[s]; i++) u |= ops[s][i];
            if (u > maxes[s])
                for (uint64_t i = 0; i < lens[s]; i++)
                    if (ops[s][i] > maxes[s]) return 0;
        }
    }
    return 1;
}

/* columns, src/convolve.rs:303-351: direct column sums of WORDS limbs. */
static void columns(uint64_t *output, uint64_t stride, int is_signed, const uint64_t *a,
                    uint64_t a_len, const uint64_t *b, uint64_t b_len) {
    uint64_t count = a_len + b_len - 1;
    const u128 half = (u128)1 << 63;
    for (uint64_t index = 0; index < count; index++) {
        uint64_t next = index + 1;
        uint64_t first = next > b_len ? next - b_len : 0;
        uint64_t end = next < a_len ? next : a_len;
        u128 low = 0;
        uint64_t high = 0;
        uint64_t sum1 = 0;
        for (uint64_t i = first; i < end; i++) {
            uint64_t x = a[i], y = b[next - 1 - i];
            if (stride == 1) {
                sum1 += x * y;
            } else {
                u128 wx = is_signed ? (((u128)x ^ half) - half) : (u128)x;
                u128 wy = is_signed ? (((u128)y ^ half) - half) : (u128)y;
                u128 product = wx * wy;
                if (stride == 2) {
                    low += product;
                } else {
                    u128 sum = low + product;
                    high += sum < low;
                    low = sum;
                    if (is_signed) high -= (uint64_t)(product >> 127);
                }
            }
        }
        if (stride == 1) low = sum1;
        uint64_t words[3] = {(uint64_t)low, (uint64_t)(low >> 64), high};
        uint64_t limit = stride < 3 ? stride : 3;
        for (uint64_t w = 0; w < stride; w++) output[index * stride + w] = w < limit ? words[w] : 0;
    }
}

/* ConvPlan::convolve, src/convolve.rs:148-162. */
int oracle_convolve(const oracle_conv_plan *plan, uint64_t *dst, uint64_t dst_len,
                    const uint64_t *a, uint64_t a_len, const uint64_t *b, uint64_t b_len,
                    int threads) {
    ensure_init();
    if (!conv_admits(plan, a, a_len, b, b_len)) return 0;
    if (dst_len < plan->output_len) return 0;
    uint64_t *out = malloc(sizeof(uint64_t) * plan->output_len);
    if (plan->schoolbook) {
        columns(out, plan->stride, (int)plan->is_signed, a, a_len, b, b_len);
    } else {
        int square = (a == b) && (a_len == b_len);
        run_pipeline(&plan->geometry, out, plan->output_len, a, a_len, b, b_len, square, threads);
    }
    memcpy(dst, out, sizeof(uint64_t) * plan->output_len);
    free(out);
    return 1;
}

/* ------------------------------------------------------------------------ */

/* src/tests/schoolbook.rs:13-45. */
void oracle_schoolbook(uint64_t *dst, const uint64_t *a, uint64_t a_len, const uint64_t *b,
                       uint64_t b_len) {
    memset(dst, 0, sizeof(uint64_t) * (a_len + b_len));
    for (uint64_t i = 0; i < a_len; i++) {
        u128 carry = 0;
        for (uint64_t j = 0; j < b_len; j++) {
            u128 v = (u128)a[i] * b[j] + dst[i + j] + carry;
            dst[i + j] = (uint64_t)v;
            carry = v >> 64;
        }
        dst[i + b_len] = (uint64_t)carry;
    }
}

/* operand, bench/src/lib.rs:13-29. */
void oracle_operand(uint64_t *out, uint64_t len, uint64_t seed) {
    uint64_t state = seed;
    for (uint64_t i = 0; i < len; i++) {
        state += 0x9e3779b97f4a7c15ULL;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        out[i] = z ^ (z >> 31);
    }
}

uint64_t oracle_mod_limbs(const uint64_t *limbs, uint64_t len, uint64_t q) {
    u128 r = 0;
    for (uint64_t i = len; i-- > 0;) {
        r = ((r << 64) | limbs[i]) % q;
    }
    return (uint64_t)r;
}
