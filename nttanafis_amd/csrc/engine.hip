// Host side of the gfx950 multi-prime NTT engine: catalog constants, device
// geometry, device context, the product/convolution pipelines and the C ABI
// declared in include/nttanafis.h.
#include "engine.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nttanafis.h"

namespace nttamd {

typedef unsigned __int128 u128;

bool launch_fwd_strided(int logl, const PassArgs& a, int primes, int operands, bool pack, bool alt, bool r8,
                        hipStream_t s, bool waves = false);
bool launch_pack_primes(int logl, const PassArgs& a, int primes, int operands, hipStream_t s);
bool launch_inv_strided(int logl, const PassArgs& a, int primes, bool canon, bool alt, bool r8, hipStream_t s,
                        bool waves = false);
bool launch_fused(int logl, const PassArgs& a, int primes, bool pack, bool square, int form, hipStream_t s);
bool pack_primes_batch_fits(int logl, uint64_t line_stride);
bool launch_pack_primes_batch(int logl, const PassArgs& a, int items, hipStream_t s);
bool launch_crt_product(const CrtArgs& a, hipStream_t s);
bool launch_crt_disjoint(const CrtArgs& a, hipStream_t s);
bool launch_school(const CrtArgs& a, hipStream_t s);
bool launch_recombine(const RecombineArgs& a, int primes, hipStream_t s);
bool launch_conv_columns(const uint64_t* a, uint64_t la, const uint64_t* b, uint64_t lb, uint64_t* out,
                         uint32_t stride, uint32_t is_signed, hipStream_t s);

bool launch_tables(const TableArgs& a, int primes, hipStream_t s);
bool launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t s);

// ---------------------------------------------------------------------------
// Catalog: the reference's eight primes 2^49 < p < 2^50 with generators of
// their two-power subgroups (src/field/catalog.rs:67-76). Everything else is
// derived here at start-up with exact integer arithmetic.

static const uint64_t kCatalog[kMaxPrimes][2] = {
    {1108307720798209ULL, 11}, {1086317488242689ULL, 3}, {910395627798529ULL, 7},
    {659706976665601ULL, 11},  {699289395265537ULL, 5},  {1022545813831681ULL, 7},
    {1013749720809473ULL, 3},  {868614185943041ULL, 3},
};
constexpr int kMaxRootOrder = 41;  // min 2-adic valuation of p-1 over the catalog
// Device packing reads one four-limb window per coefficient, so device
// coefficients are at most 192 bits wide (the reference admits 196).
constexpr int kMaxChunkBits = 192;
constexpr int kMaxLog = 30;
constexpr uint64_t kSchoolbookMaxTerms = 24;  // src/convolve.rs:24

static uint64_t pow_mod(uint64_t b, uint64_t e, uint64_t m) {
    u128 v = 1, f = b % m;
    while (e) {
        if (e & 1) v = v * f % m;
        f = f * f % m;
        e >>= 1;
    }
    return (uint64_t)v;
}

static double centered(uint64_t v, uint64_t m) { return v > m / 2 ? (double)v - (double)m : (double)v; }

struct Catalog {
    uint64_t p[kMaxPrimes];
    Field field;
    double roots[kMaxPrimes][48];
    double inv_roots[kMaxPrimes][48];
    double powers[kMaxPrimes][kMaxParts];
    uint32_t capacity[kMaxPrimes + 1];  // by prime count
    // Garner constants for a prefix of `count` primes do not depend on count.
    double pmod[kMaxPrimes][kMaxPrimes];
    double ginv[kMaxPrimes];
};

static Catalog g_cat;
static std::once_flag g_cat_once;

static void init_catalog() {
    Catalog& c = g_cat;
    memset(&c, 0, sizeof c);
    for (int i = 0; i < kMaxPrimes; ++i) {
        const uint64_t m = kCatalog[i][0], g = kCatalog[i][1];
        c.p[i] = m;
        c.field.p[i] = (double)m;
        c.field.pinv[i] = 1.0 / (double)m;
        const int v = __builtin_ctzll(m - 1);
        for (int level = 0; level < 48; ++level) {
            if (level + 2 > v) continue;
            const uint64_t r = pow_mod(g, (m - 1) >> (level + 2), m);
            c.roots[i][level] = centered(r, m);
            c.inv_roots[i][level] = centered(pow_mod(r, m - 2, m), m);
        }
        u128 power = 1;
        for (int part = 0; part < kMaxParts; ++part) {
            c.powers[i][part] = centered((uint64_t)power, m);
            power = (power << kPartBits) % m;
        }
    }
    // Capacity floor(log2(2P/3)) of each prefix (src/limb/cofactor.rs:239-255).
    uint64_t prod[kMaxWords + 1] = {1};
    int words = 1;
    for (int count = 1; count <= kMaxPrimes; ++count) {
        u128 carry = 0;
        for (int w = 0; w < words; ++w) {
            const u128 full = (u128)prod[w] * c.p[count - 1] + carry;
            prod[w] = (uint64_t)full;
            carry = full >> 64;
        }
        if (carry) prod[words++] = (uint64_t)carry;
        const int bits = (words - 1) * 64 + (64 - __builtin_clzll(prod[words - 1]));
        const int second = bits - 2;
        c.capacity[count] = ((prod[second >> 6] >> (second & 63)) & 1) ? bits - 1 : second;
    }
    for (int k = 0; k < kMaxPrimes; ++k) {
        const uint64_t pk = c.p[k];
        u128 prefix = 1;
        for (int j = 0; j < k; ++j) {
            c.pmod[j][k] = centered(c.p[j] % pk, pk);
            prefix = prefix * (c.p[j] % pk) % pk;
        }
        c.ginv[k] = k == 0 ? 1.0 : centered(pow_mod((uint64_t)prefix, pk - 2, pk), pk);
    }
}

const Catalog& cat() {
    std::call_once(g_cat_once, init_catalog);
    return g_cat;
}

// ---------------------------------------------------------------------------
// Errors and profiling.

static thread_local std::string t_error;
void set_error(const std::string& e) { t_error = e; }
void clear_error() { t_error.clear(); }

struct ProfEntry {
    const char* name;
    hipEvent_t start, stop;
};
static std::mutex g_prof_lock;
static std::atomic<int> g_prof_on{0};
static std::vector<ProfEntry> g_prof_pending;
static std::vector<std::pair<std::string, std::pair<double, uint64_t>>> g_prof_totals;

static thread_local const char* t_kind = "kernel";

// Names the launches of a scope for per-launch timing (the events themselves
// are recorded around each dispatch by the launchers in kernels.hip).
Timed::Timed(const char* n, hipStream_t) : saved(t_kind) { t_kind = n; }
Timed::~Timed() { t_kind = saved; }

static thread_local bool t_capturing = false;

bool prof_enabled() { return g_prof_on.load() != 0 && !t_capturing; }
const char* prof_kind() { return t_kind; }
void prof_record(const char* kind, hipEvent_t start, hipEvent_t stop) {
    std::lock_guard<std::mutex> g(g_prof_lock);
    g_prof_pending.push_back({kind, start, stop});
}

// The last launch of each kind while profiling, for replays.
struct ProfLaunch {
    std::string kind;
    const void* fn;
    dim3 grid, block;
    hipStream_t stream;
    std::vector<unsigned char> arg;
};
static std::vector<ProfLaunch> g_prof_last;

void prof_keep(const char* kind, const void* fn, dim3 grid, dim3 block, hipStream_t s, const void* arg,
               size_t size) {
    std::lock_guard<std::mutex> g(g_prof_lock);
    ProfLaunch* slot = nullptr;
    for (auto& l : g_prof_last)
        if (l.kind == kind) slot = &l;
    if (!slot) {
        g_prof_last.push_back(ProfLaunch{kind, nullptr, dim3(), dim3(), nullptr, {}});
        slot = &g_prof_last.back();
    }
    slot->fn = fn;
    slot->grid = grid;
    slot->block = block;
    slot->stream = s;
    const unsigned char* b = static_cast<const unsigned char*>(arg);
    slot->arg.assign(b, b + size);
}

static void prof_drain() {
    std::lock_guard<std::mutex> g(g_prof_lock);
    for (auto& e : g_prof_pending) {
        float ms = 0;
        hipEventSynchronize(e.stop);
        hipEventElapsedTime(&ms, e.start, e.stop);
        bool found = false;
        for (auto& t : g_prof_totals) {
            if (t.first == e.name) {
                t.second.first += ms;
                t.second.second += 1;
                found = true;
                break;
            }
        }
        if (!found) g_prof_totals.push_back({e.name, {ms, 1}});
        hipEventDestroy(e.start);
        hipEventDestroy(e.stop);
    }
    g_prof_pending.clear();
}

// ---------------------------------------------------------------------------
// Geometry of one device operation.


static int bitlen(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

static int log2_ceil(uint64_t x) {
    int k = 0;
    while (((uint64_t)1 << k) < x) ++k;
    return k;
}

static int crt_words(int np) {
    static const int w[9] = {0, 1, 2, 3, 4, 4, 5, 6, 7};
    return w[np];
}

static int passes_of(int K) { return K <= kSinglePassMaxLog ? 1 : (K <= 23 ? 2 : 3); }

// Device price of transforms, in units of about 0.6 ps of kernel time: HBM
// passes and levels over the points of each part, a recombination pass over
// a composite row, and a fixed price per launch.
static double transform_price(int np, int log_n) {
    return np * (double)((uint64_t)1 << log_n) * (8.0 * passes_of(log_n) + log_n);
}
constexpr double kLaunchPrice = 5e6;
// Price of one 64 x 64-bit limb product of a schoolbook block on one CU,
// in the units of transform_price (the whole device), set at the measured
// crossover of single squares: 384^2 limbs take 13.0-13.3 us either way
// (schoolbook 192^2: 9.2 us, 512^2: 16.8 us; transforms 11.5 and 13.8 us).
constexpr double kSchoolProductPrice = 414.0;
constexpr int kDeviceCUs = 256;

int row_parts(const Geometry& g, RowPart* parts) {
    const uint64_t n = (uint64_t)1 << g.log_t;
    parts[0] = RowPart{g.log_t, 0, 0};
    switch (g.tail) {
        case 1: parts[1] = RowPart{g.log_t - 2, 4, n}; return 2;
        case 2: parts[1] = RowPart{g.log_t - 1, 2, n}; return 2;
        case 3:
            parts[1] = RowPart{g.log_t - 1, 2, n};
            parts[2] = RowPart{g.log_t - 2, 6, n + n / 2};
            return 3;
        default: return 1;
    }
}

static double row_price(const Geometry& g, double launch_price) {
    RowPart parts[3];
    const int n = row_parts(g, parts);
    double price = 0;
    int launches = 0;
    for (int i = 0; i < n; ++i) {
        price += transform_price((int)g.primes, parts[i].log_n);
        launches += 2 * passes_of(parts[i].log_n) - 1;
    }
    if (g.tail) {
        price += 8.0 * g.primes * (double)row_len(g);
        ++launches;
    }
    return price + launch_price * launches;
}

// The composite row of row_shape (src/profile/planner.rs:565-591) for a
// convolution of conv coefficients whose power-of-two row has 2^K points
// (K >= kMinLog): with N = 2^(K-1) holding both operands, the shortest of
// 5N/4 (N/4 >= 2^kMinLog), 3N/2 (N/2 >= 2^kMinLog), 7N/4 (N/4 >= 2^kMinLog)
// that holds conv; 0 for none.
static uint32_t composite_tail(uint64_t conv, uint64_t longest, int K) {
    const uint64_t n = (uint64_t)1 << (K - 1), half = n >> 1, quarter = n >> 2;
    const uint64_t min = (uint64_t)1 << kMinLog;
    if (longest > n) return 0;
    if (quarter >= min && conv <= n + quarter) return 1;
    if (half >= min && conv <= n + half) return 2;
    if (quarter >= min && conv <= n + half + quarter) return 3;
    return 0;
}

// Row of a single operation's convolution: the cheaper of the power-of-two
// row and the composite row, when one is admitted. NTTAMD_COMPOSITE=0 keeps
// power-of-two rows, =1 takes every admitted composite row (both give the
// same results; tests use them to reach every row form).
static void choose_row(Geometry* g, uint64_t longest, int K) {
    g->log_t = K;
    g->tail = 0;
    const char* env = getenv("NTTAMD_COMPOSITE");
    const int mode = env ? atoi(env) : -1;
    const uint32_t t = mode == 0 ? 0 : composite_tail(g->conv_len, longest, K);
    if (!t) return;
    Geometry c = *g;
    c.log_t = K - 1;
    c.tail = t;
    if (mode == 1 || row_price(c, kLaunchPrice) < row_price(*g, kLaunchPrice)) *g = c;
}

// Unbalanced split (Geometry::cheapest_blocks and search_blocks,
// src/profile/planner.rs:283-393): blocks of D coefficients of the longer
// operand, D at least the shorter operand's count, a multiple of 64 (so a
// block starts at a limb) and grown to fill a row of 2^K points. Priced as
// one forward transform of the shorter operand plus a forward and an
// inverse per block (a one-pass row repeats the shorter operand's forward
// in every block); the reconstruction covers the same coefficients either
// way. Taken when it prices at most three quarters of the whole product;
// NTTAMD_SPLIT=0 never splits, =1 takes the cheapest admissible split.
constexpr uint64_t kMaxBlocks = 8192;

static double split_price(const Geometry& g, double launch_price) {
    const int passes = passes_of(g.log_t);
    const double tp = transform_price((int)g.primes, g.log_t);
    const double work = passes == 1 ? tp * g.blocks : tp * (1.0 + 2.0 * g.blocks) / 3.0;
    return work + launch_price * (2 * passes - 1);
}

// launch_price: 0 inside a batch, whose launches every product shares.
static void choose_split(Geometry* g, double launch_price) {
    const char* env = getenv("NTTAMD_SPLIT");
    const int mode = env ? atoi(env) : -1;
    if (mode == 0) return;
    const uint64_t dl = std::max(g->digits_a, g->digits_b), ds = std::min(g->digits_a, g->digits_b);
    if (dl < 2 * ds) return;
    double best = mode == 1 ? 1e300 : 0.75 * row_price(*g, launch_price);
    bool found = false;
    Geometry pick = *g;
    for (int K = std::max(kMinLog, log2_ceil(2 * ds - 1)); K <= g->log_t; ++K) {
        const uint64_t T = (uint64_t)1 << K;
        const uint64_t D = ((T - ds + 1) / 64) * 64;
        if (D < ds) continue;
        const uint64_t blocks = (dl + D - 1) / D;
        if (blocks < 2) break;
        if (blocks > kMaxBlocks) continue;
        Geometry c = *g;
        c.log_t = K;
        c.tail = 0;
        c.blocks = (uint32_t)blocks;
        c.block_digits = D;
        const double price = split_price(c, launch_price);
        if (price < best) {
            best = price;
            pick = c;
            found = true;
        }
    }
    if (found) *g = pick;
}

// Schoolbook or transforms. A schoolbook block costs its share of the
// la*lb limb products on one CU and needs no transform launches; it is
// admissible while the shorter operand fits the kernel's LDS window.
// NTTAMD_SCHOOL=0|1 forces the choice (tests).
static bool choose_school(uint64_t la, uint64_t lb, const Geometry& g, double ntt_price, double launch_price) {
    if (std::min(la, lb) > (uint64_t)kSchoolMaxShort) return false;
    const char* env = getenv("NTTAMD_SCHOOL");
    if (env && env[0] == '0') return false;
    if (env && env[0] == '1') return true;
    (void)g;
    // Blocks run one per CU: a single operation waits for its waves of
    // blocks; a batch shares the device, so it pays its share of all CUs.
    const uint64_t blocks = carry_blocks(la + lb, true);
    const double products = (double)la * (double)lb;
    const double price = launch_price == 0.0
                             ? kSchoolProductPrice * products / kDeviceCUs
                             : kSchoolProductPrice * products / (double)blocks *
                                       (double)((blocks + kDeviceCUs - 1) / kDeviceCUs) +
                                   launch_price;
    return price < ntt_price;
}

// Product geometry: for every prefix, the widest coefficient whose products
// fit its capacity (the narrowing of src/profile/planner.rs:168-183), priced
// by device work: HBM passes and transform levels over the row's points.
bool select_product(uint64_t la, uint64_t lb, Geometry* out, bool batched) {
    const Catalog& c = cat();
    if (la == 0 || lb == 0 || la > (UINT64_MAX >> 7) || lb > (UINT64_MAX >> 7)) return false;
    const uint64_t bits_a = la * 64, bits_b = lb * 64;
    bool found = false;
    double best = 0;
    const char* forced_env = getenv("NTTAMD_PRIMES");  // a prime count for tests and sweeps
    const int forced = forced_env ? atoi(forced_env) : 0;
    for (int np = 1; np <= kMaxPrimes; ++np) {
        if (forced && np != forced) continue;
        const uint32_t cap = c.capacity[np];
        uint64_t w = std::min<uint64_t>(cap / 2, kMaxChunkBits);
        uint64_t da = 0, db = 0, pb = 0;
        bool ok = false;
        for (;;) {
            da = (bits_a + w - 1) / w;
            db = (bits_b + w - 1) / w;
            const uint64_t terms = std::min(da, db);
            pb = 2 * w + bitlen(terms - 1);
            if (pb <= cap) {
                ok = true;
                break;
            }
            const uint64_t cut = (pb - cap + 1) / 2;
            if (cut >= w) break;
            w -= cut;
        }
        if (!ok) continue;
        const uint64_t conv = da + db - 1;
        const int K = std::max(kMinLog, log2_ceil(conv));
        if (K > kMaxLog || K > kMaxRootOrder - 1) continue;
        const int nw = crt_words(np);
        const uint64_t span = (64 * (uint64_t)kCrtBlockLimbs + pb) / w + 2;
        if (span * nw > 4096) continue;
        Geometry g;
        g.primes = np;
        g.width = (uint32_t)w;
        g.value_bits_a = g.value_bits_b = (uint32_t)w;
        g.digits_a = da;
        g.digits_b = db;
        g.conv_len = conv;
        g.bound_bits = (uint32_t)pb;
        // A batch shares every launch among its products.
        g.log_t = K;
        const double cost = row_price(g, batched ? 0.0 : kLaunchPrice) + 4.0 * np * np * (double)conv;
        if (!found || cost < best) {
            found = true;
            best = cost;
            *out = g;
        }
    }
    // The prime count is chosen on power-of-two rows; a single operation
    // then takes the composite row of that geometry when it prices lower
    // (same primes and coefficients, so the reconstruction is unchanged).
    // A batch runs power-of-two rows; both forms split unbalanced products.
    if (found) {
        const double launch_price = batched ? 0.0 : kLaunchPrice;
        if (!batched) choose_row(out, std::max(out->digits_a, out->digits_b), out->log_t);
        choose_split(out, launch_price);
        // Both forms end with the carry pass across output blocks; the
        // transforms also launch the reconstruction.
        const double ntt = (out->blocks ? split_price(*out, launch_price) : row_price(*out, launch_price)) +
                           4.0 * out->primes * out->primes * (double)out->conv_len + launch_price;
        out->school = choose_school(la, lb, *out, ntt, launch_price);
    }
    return found;
}

// bound_bits of src/convolve.rs:358-369.
bool conv_bound_bits(uint64_t terms, uint64_t max_a, uint64_t max_b, uint64_t* bits) {
    const u128 product = (u128)max_a * max_b;
    const uint64_t hi = (uint64_t)(product >> 64), lo = (uint64_t)product;
    const u128 t_lo = (u128)lo * terms;
    const u128 t_hi = (u128)hi * terms + (t_lo >> 64);
    auto bitlen128 = [](u128 v) { return v >> 64 ? 64 + bitlen((uint64_t)(v >> 64)) : bitlen((uint64_t)v); };
    if (t_hi >> 64) {
        *bits = (uint64_t)bitlen128(product) + (uint64_t)bitlen(terms - 1);
    } else {
        *bits = (uint64_t)bitlen128((t_hi << 64) | (uint64_t)t_lo);
    }
    return true;
}

// ---------------------------------------------------------------------------
// Device context: root tables shared by every caller (grow-only; a table for
// T serves every shorter transform) and device buffers of one executor.

static std::mutex g_table_lock;
static double* g_fwd = nullptr;
static double* g_inv = nullptr;
static uint64_t g_table_len = 0;  // entries per prime
static std::vector<void*> g_retired;

bool ensure_tables(uint64_t need, hipStream_t s, const double** fwd, const double** inv,
                          uint64_t* stride) {
    std::lock_guard<std::mutex> g(g_table_lock);
    if (need > g_table_len) {
        uint64_t len = std::max<uint64_t>(need, 1 << 12);
        double *f = nullptr, *v = nullptr;
        if (hipMalloc(&f, len * kMaxPrimes * sizeof(double)) != hipSuccess) return false;
        if (hipMalloc(&v, len * kMaxPrimes * sizeof(double)) != hipSuccess) {
            hipFree(f);
            return false;
        }
        const Catalog& c = cat();
        TableArgs a;
        memset(&a, 0, sizeof a);
        a.fwd = f;
        a.inv = v;
        a.table_stride = len;
        a.len = len;
        a.field = c.field;
        memcpy(a.roots, c.roots, sizeof a.roots);
        memcpy(a.inv_roots, c.inv_roots, sizeof a.inv_roots);
        {
            Timed t("tables", s);
            if (!launch_tables(a, kMaxPrimes, s)) return false;
        }
        if (hipStreamSynchronize(s) != hipSuccess) return false;
        // Older tables may still be read by queued work of other executors.
        if (g_fwd) g_retired.push_back(g_fwd);
        if (g_inv) g_retired.push_back(g_inv);
        g_fwd = f;
        g_inv = v;
        g_table_len = len;
    }
    *fwd = g_fwd;
    *inv = g_inv;
    *stride = g_table_len;
    return true;
}



// Executors of the plan path (MulPlan::mul / sqr, ConvPlan::convolve take
// &self and run concurrently in the reference, src/plan.rs:94-114): each
// calling thread gets its own executor (stream and device arenas) the first
// time it calls, from a process-wide pool, so calls from different threads
// overlap on the device. Executors live until the process ends: work queued
// from a thread that has since exited stays valid.
Executor* plan_executor() {
    static std::mutex pool_lock;
    static std::vector<Executor*>* pool = new std::vector<Executor*>();
    thread_local Executor* mine = nullptr;
    if (!mine) {
        Executor* ex = new (std::nothrow) Executor();
        if (!ex) return nullptr;
        std::lock_guard<std::mutex> g(pool_lock);
        pool->push_back(ex);
        mine = ex;
    }
    return mine;
}

// ---------------------------------------------------------------------------
// Pipelines.


static std::mutex g_tune_lock;

// Splits of a transform of 2^K points, K > 12: a row pass of 2^rows points
// (5 <= rows <= 12) after one strided pass (2^13 .. 2^23 points) or two
// (from 2^18 points on), each of 5 to 11 levels. A strided pass needs a line
// stride of at least its tile's lines (tile_lines: the default or the
// alternative tiles).
static Split split_rows(int K, int rows, int passes) {
    Split s{};
    if (passes == 2) {
        s.n = 2;
        s.logs[0] = K - rows;
        s.logs[1] = rows;
    } else {
        const int k1 = (K - rows + 1) / 2;
        s.n = 3;
        s.logs[0] = k1;
        s.logs[1] = K - rows - k1;
        s.logs[2] = rows;
    }
    return s;
}

static bool split_ok(int K, const Split& s) {
    int level = 0;
    for (int i = 0; i + 1 < s.n; ++i) {
        if (s.logs[i] < kMinLog || s.logs[i] > 11) return false;
        const int stride_log = K - level - s.logs[i];
        if (stride_log < 0 || ((uint64_t)1 << stride_log) < (uint64_t)tile_lines(s.logs[i], s.alt_tiles != 0))
            return false;
        level += s.logs[i];
    }
    const int rows = s.logs[s.n - 1];
    return rows >= kMinLog && rows <= kSinglePassMaxLog && level + rows == K;
}

static bool same_split(const Split& a, const Split& b) {
    if (a.n != b.n || a.alt_tiles != b.alt_tiles || a.inv16 != b.inv16 || a.per_prime_pack != b.per_prime_pack ||
        a.rows8 != b.rows8 || a.strided8 != b.strided8 || a.paired != b.paired || a.waves != b.waves)
        return false;
    for (int i = 0; i < a.n; ++i)
        if (a.logs[i] != b.logs[i]) return false;
    return true;
}

// Kernel-form flags forced from the environment (tests, A/B runs):
// NTTAMD_ALT_TILES, NTTAMD_INV16, NTTAMD_PER_PRIME_PACK, NTTAMD_ROWS8,
// NTTAMD_STRIDED8, NTTAMD_PAIRED and NTTAMD_WAVES set to 0 or 1. -1: not forced.
static int env_flag(const char* name) {
    const char* e = getenv(name);
    return e && *e ? (atoi(e) != 0) : -1;
}
static bool flags_forced() {
    return env_flag("NTTAMD_ALT_TILES") >= 0 || env_flag("NTTAMD_INV16") >= 0 ||
           env_flag("NTTAMD_PER_PRIME_PACK") >= 0 || env_flag("NTTAMD_ROWS8") >= 0 ||
           env_flag("NTTAMD_STRIDED8") >= 0 || env_flag("NTTAMD_PAIRED") >= 0 || env_flag("NTTAMD_WAVES") >= 0;
}
static Split with_env_flags(int K, Split s) {
    const int alt = env_flag("NTTAMD_ALT_TILES"), inv = env_flag("NTTAMD_INV16"),
              ppp = env_flag("NTTAMD_PER_PRIME_PACK"), r8 = env_flag("NTTAMD_ROWS8"),
              s8 = env_flag("NTTAMD_STRIDED8"), pr = env_flag("NTTAMD_PAIRED"), wv = env_flag("NTTAMD_WAVES");
    if (alt >= 0 && s.n > 1) {
        Split t = s;
        t.alt_tiles = (uint8_t)alt;
        if (split_ok(K, t)) s = t;
    }
    if (inv >= 0) s.inv16 = (uint8_t)inv;
    if (ppp >= 0 && s.n > 1) s.per_prime_pack = (uint8_t)ppp;
    if (r8 >= 0) s.rows8 = (uint8_t)r8;
    if (s8 >= 0 && s.n > 1) s.strided8 = (uint8_t)s8;
    if (pr >= 0) s.paired = (uint8_t)pr;
    if (wv >= 0 && s.n > 1) s.waves = (uint8_t)wv;
    return s;
}

// The flag variants the autotuner tries after choosing the levels, in order,
// each on top of the flags kept so far and kept when it beats the best so
// far: alternative strided tiles, radix-16 inverse rows (products; a
// square's rows already run radix 16), radix-8 rows with twice the threads,
// packing per prime instead of the shared first packing pass, radix-8
// strided passes (single operations whose strided passes have 2^5, 2^6 or
// 2^9 points), paired product rows (products), strided passes in
// wave-owned lines (single operations whose strided passes have 2^5 or 2^6
// points).
constexpr int kMaxVariants = 7;
static int flag_variants(int K, int use, const Split& best, Split out[kMaxVariants]) {
    int n = 0;
    Split t = best;
    t.alt_tiles = 1;
    if (best.n > 1 && !best.alt_tiles && split_ok(K, t)) out[n++] = t;
    t = best;
    t.inv16 = 1;
    if (use != kUseSquare && !best.inv16) out[n++] = t;
    t = best;
    t.rows8 = 1;
    if (!best.rows8) out[n++] = t;
    t = best;
    t.per_prime_pack = 1;
    if (best.n > 1 && !best.per_prime_pack) out[n++] = t;
    t = best;
    t.strided8 = 1;
    bool r8_pass = false;
    for (int i = 0; i + 1 < best.n; ++i) r8_pass |= best.logs[i] == 5 || best.logs[i] == 6 || best.logs[i] == 9;
    if (use != kUseBatch && r8_pass && !best.strided8) out[n++] = t;
    t = best;
    t.paired = 1;
    if (use != kUseSquare && !best.paired) out[n++] = t;
    t = best;
    t.waves = 1;
    bool w_pass = false;
    for (int i = 0; i + 1 < best.n; ++i) w_pass |= best.logs[i] == 5 || best.logs[i] == 6;
    if (use != kUseBatch && w_pass && !best.waves) out[n++] = t;
    return n;
}

// best with the flag variant v sets on top.
static Split with_variant(Split best, const Split& v) {
    best.alt_tiles |= v.alt_tiles;
    best.inv16 |= v.inv16;
    best.per_prime_pack |= v.per_prime_pack;
    best.rows8 |= v.rows8;
    best.strided8 |= v.strided8;
    best.paired |= v.paired;
    best.waves |= v.waves;
    return best;
}

// Admissible splits, in a fixed array (split_of runs on the allocation-free
// path too): two-pass splits first, each form from the largest row down.
struct SplitCandidates {
    int n = 0;
    Split s[16];
    const Split* begin() const { return s; }
    const Split* end() const { return s + n; }
    bool empty() const { return n == 0; }
    size_t size() const { return (size_t)n; }
    const Split& operator[](int i) const { return s[i]; }
};

static SplitCandidates split_candidates(int K) {
    SplitCandidates c;
    if (K <= kSinglePassMaxLog) return c;
    for (int passes = 2; passes <= 3; ++passes) {
        if ((passes == 2 && K > 23) || (passes == 3 && K < 18)) continue;
        for (int r = kSinglePassMaxLog; r >= kMinLog && c.n < 16; --r) {
            const Split sp = split_rows(K, r, passes);
            if (split_ok(K, sp)) c.s[c.n++] = sp;
        }
    }
    return c;
}

// Tuned splits per use and length (n == 0: not tuned yet).
static Split g_tuned[3][kMaxLog + 1];

// Split of a transform of 2^K points: the autotuner's candidate while it
// times one; otherwise a single pass up to 2^12, the candidate that
// NTTAMD_ROW_LOG (and NTTAMD_PASSES, 2 or 3) name when set and admissible,
// the tuned split of this use once the autotuner timed this length, else
// the first candidate (two passes with the largest row below 2^24 points,
// three from there on); then the flags the environment forces.
static thread_local Split t_force{};  // set while the autotuner times a candidate

static Split split_default(int K, int use) {
    const SplitCandidates cand = split_candidates(K);
    const char* er = getenv("NTTAMD_ROW_LOG");
    const char* ep = getenv("NTTAMD_PASSES");
    if (K > kSinglePassMaxLog && (er || ep)) {
        const int rows = er ? atoi(er) : 0, passes = ep ? atoi(ep) : 0;
        for (const Split& s : cand)
            if ((!rows || s.logs[s.n - 1] == rows) && (!passes || s.n == passes)) return s;
    }
    if (K >= 0 && K <= kMaxLog && use >= 0 && use < 3) {
        std::lock_guard<std::mutex> g(g_tune_lock);
        if (g_tuned[use][K].n) return g_tuned[use][K];
    }
    if (K <= kSinglePassMaxLog) {
        Split s{};
        s.n = 1;
        s.logs[0] = K;
        return s;
    }
    if (!cand.empty()) return cand[0];
    return split_rows(K, kSinglePassMaxLog, K <= 23 ? 2 : 3);
}

Split split_of(int K, int use) {
    if (t_force.n) return t_force;
    return with_env_flags(K, split_default(K, use));
}

// Pass arguments of a transform of 2^log_n points over rows of row_len(g)
// points (a whole power-of-two row, or one part of a composite row).
static PassArgs base_args(const Geometry& g, double* data, const double* fwd, const double* inv,
                          uint64_t table_stride, int log_n) {
    const Catalog& c = cat();
    PassArgs a;
    memset(&a, 0, sizeof a);
    a.data = data;
    a.fwd = fwd;
    a.inv = inv;
    a.row_stride = row_len(g);
    a.table_stride = table_stride;
    a.log_t = log_n;
    a.field = c.field;
    a.fold_terms = 1;
    memcpy(a.powers, c.powers, sizeof a.powers);
    for (uint32_t y = 0; y < g.primes; ++y) {
        const uint64_t m = c.p[y];
        const uint64_t tinv = m - ((m - 1) >> log_n);  // 2^-log_n mod p, 2^log_n | p-1
        a.tinv[y] = centered(tinv, m);
    }
    return a;
}

static PassArgs base_args(const Geometry& g, double* data, const double* fwd, const double* inv,
                          uint64_t table_stride) {
    return base_args(g, data, fwd, inv, table_stride, g.log_t);
}

static uint64_t mul_mod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)a * b % m); }

// W[j] of the root table for j < 4 (W[1] a fourth root of unity, W[2] an
// eighth root, W[3] = W[1] W[2]), exact.
static uint64_t low_root(int y, int j) {
    const Catalog& c = cat();
    const uint64_t m = c.p[y], g = kCatalog[y][1];
    const uint64_t w1 = pow_mod(g, (m - 1) >> 2, m), w2 = pow_mod(g, (m - 1) >> 3, m);
    return j == 0 ? 1 : (j == 1 ? w1 : (j == 2 ? w2 : mul_mod(w1, w2, m)));
}

// Packing fold onto a part at table node `node` of span points: with
// operands of at most 2*span (node 2) or 4*span (nodes 4, 6) coefficients,
// coefficient n + q*span meets point n times W[node/2]^q, since the part's
// modulus is x^span - W[node]^2 = x^span - W[node/2].
static void set_fold(PassArgs* a, const Geometry& g, uint32_t node, uint64_t span) {
    const Catalog& c = cat();
    a->node_base = node;
    a->fold_span = span;
    a->fold_terms = node == 0 ? 1 : (node == 2 ? 2 : 4);
    for (uint32_t y = 0; y < g.primes; ++y) {
        const uint64_t m = c.p[y], w = low_root((int)y, (int)node / 2);
        uint64_t v = 1;
        for (int q = 0; q < 3; ++q) {
            v = mul_mod(v, w, m);
            a->fold_mult[y][q] = centered(v, m);
        }
    }
}

static Source source_of(const uint64_t* limbs, uint64_t len, uint64_t digits, uint32_t width,
                        uint32_t value_bits, bool is_signed) {
    Source s;
    s.limbs = limbs;
    s.limb_len = len;
    s.digits = digits;
    s.width = width;
    s.value_bits = value_bits;
    s.is_signed = is_signed ? 1 : 0;
    s.pad = 0;
    return s;
}

// Forward transform of operand `src` into `rows` through all passes but the
// row pass; returns the args of the row pass.
// Strided forward passes of the transform (all but the row pass). With
// operands == 2 every launch covers both operands: a0.data with source sa and
// a0.other with source sb.
// The first pass packs the limbs of each tile once for every prime
// (k_pack_primes) unless NTTAMD_PACK_SHARED=0 or the pass cannot take that
// form; then one workgroup per tile and prime packs its own residues.
static bool pack_shared() {
    const char* env = getenv("NTTAMD_PACK_SHARED");
    return !(env && env[0] == '0');
}

static bool forward_head(const Split& sp, const PassArgs& a0, const Source& sa, const Source& sb, int operands,
                         int primes, hipStream_t s, int first = 0, int last = -1) {
    PassArgs a = a0;
    a.src_a = sa;
    a.src_b = sb;
    int level = 0;
    if (last < 0) last = sp.n - 2;
    for (int i = 0; i + 1 < sp.n; ++i) {
        if (i < first || i > last) {
            level += sp.logs[i];
            continue;
        }
        a.level_start = level;
        a.line_stride = (uint64_t)1 << (a.log_t - level - sp.logs[i]);
        Timed t(i == 0 ? "fwd_pack_strided" : "fwd_strided", s);
        const bool shared = i == 0 && pack_shared() && !sp.per_prime_pack &&
                            launch_pack_primes(sp.logs[i], a, primes, operands, s);
        if (!shared && !launch_fwd_strided(sp.logs[i], a, primes, operands, i == 0, sp.alt_tiles, sp.strided8, s, sp.waves)) return false;
        level += sp.logs[i];
    }
    return true;
}

static bool inverse_tail(const Split& sp, const PassArgs& a0, int primes, hipStream_t s) {
    PassArgs a = a0;
    int level = a.log_t - sp.logs[sp.n - 1];
    for (int i = sp.n - 2; i >= 0; --i) {
        level -= sp.logs[i];
        a.level_start = level;
        a.line_stride = (uint64_t)1 << (a.log_t - level - sp.logs[i]);
        Timed t(i == 0 ? "inv_strided_canon" : "inv_strided", s);
        if (!launch_inv_strided(sp.logs[i], a, primes, i == 0, sp.alt_tiles, sp.strided8, s, sp.waves)) return false;
    }
    return true;
}

// Device working memory of one product or convolution: the spectra rows
// (primes x row_len doubles each; spec_b unused for a square) and the carry
// words of the reconstruction (4 per output block: carry function base and
// threshold, first-limb threshold, ripple end).
uint64_t op_spectra_bytes(const Geometry& g) {
    return g.school ? 0 : (uint64_t)g.primes * row_len(g) * 8 * (g.blocks ? g.blocks : 1);
}
uint64_t op_spectra_bytes_b(const Geometry& g) { return g.school ? 0 : (uint64_t)g.primes * row_len(g) * 8; }
uint64_t carry_blocks(uint64_t out_len, bool school) {
    const uint64_t bl = school ? kSchoolBlockLimbs : kCrtBlockLimbs;
    return (out_len + bl - 1) / bl;
}
uint64_t op_carry_bytes(uint64_t out_len, bool school) { return carry_words((uint32_t)carry_blocks(out_len, school)) * 4; }

// Root-table entries a geometry reads: T/2 for a row of 2^K points; a
// composite row reaches nodes below 7N/8 of the table of 2N points.
static uint64_t table_need(const Geometry& g) {
    const uint64_t n = (uint64_t)1 << g.log_t;
    return g.tail ? n : n / 2;
}

// Primes per chunk of a multi-pass transform. The passes of a chunk run back
// to back, so when its rows (every operand's, all primes of the chunk) fit the
// 256 MiB memory-side cache (MALL, "Infinity Cache") of the MI355X, each pass
// after the chunk's first reads what the previous one wrote from that cache
// rather than from HBM: an in-place tile pass over 256 MiB of rows runs at
// 8.5 TB/s against 5.2-5.8 TB/s over 768 MiB. Chunks split the primes as
// evenly as the budget allows; NTTAMD_CHUNK_PRIMES forces a chunk size.
constexpr uint64_t kMallBytes = 256ull << 20;

static int prime_chunk(int primes, int log_n, int operands) {
    const char* env = getenv("NTTAMD_CHUNK_PRIMES");
    int chunk;
    if (env && *env) {
        chunk = atoi(env);
    } else {
        const uint64_t per_prime = (uint64_t)operands * 8 << log_n;
        chunk = (int)std::max<uint64_t>(1, kMallBytes / per_prime);
    }
    chunk = std::max(1, std::min(chunk, primes));
    const int chunks = (primes + chunk - 1) / chunk;
    return (primes + chunks - 1) / chunks;
}

// Pass arguments for primes y0.. of a: rows, root tables and every per-prime
// constant start at prime y0, so a launch over n primes covers y0..y0+n-1.
static PassArgs chunk_args(const PassArgs& a, int y0) {
    PassArgs c = a;
    if (y0 == 0) return c;
    c.data = a.data + (uint64_t)y0 * a.row_stride;
    if (a.other) c.other = a.other + (uint64_t)y0 * a.row_stride;
    c.fwd = a.fwd + (uint64_t)y0 * a.table_stride;
    c.inv = a.inv + (uint64_t)y0 * a.table_stride;
    for (int y = 0; y + y0 < kMaxPrimes; ++y) {
        c.field.p[y] = a.field.p[y + y0];
        c.field.pinv[y] = a.field.pinv[y + y0];
        c.tinv[y] = a.tinv[y + y0];
        for (int k = 0; k < kMaxParts; ++k) c.powers[y][k] = a.powers[y + y0][k];
        for (int k = 0; k < 3; ++k) c.fold_mult[y][k] = a.fold_mult[y + y0][k];
    }
    return c;
}

// Canonical residues of the convolution modulo x^(2^log_n) - W[node]^2 into
// the part of the rows at `offset`: forward transforms of both operands
// (packing folded onto the part), pointwise product and inverse. The result
// lands in buf.spec_b (spec_a for a square).
static bool part_residues(hipStream_t s, const Geometry& g, const RowPart& part, const Source& sa,
                          const Source& sb, bool square, const OpBuffers& buf, const double* fwd,
                          const double* inv, uint64_t tstride) {
    const int np = (int)g.primes;
    double* A = buf.spec_a + part.offset;
    double* B = (square ? buf.spec_a : buf.spec_b) + part.offset;
    const Split sp = split_of(part.log_n, square ? kUseSquare : kUseProduct);
    if (sp.n == 1) {
        // One workgroup per prime packs both operands, transforms, multiplies
        // and stores the product rows into B.
        PassArgs a = base_args(g, B, fwd, inv, tstride, part.log_n);
        set_fold(&a, g, part.node, (uint64_t)1 << part.log_n);
        a.src_a = sa;
        a.src_b = sb;
        a.level_start = 0;
        Timed t("fused_single", s);
        return launch_fused(part.log_n, a, np, true, square, sp.fused_form(), s);
    }
    const int row_log = sp.logs[sp.n - 1];
    const int row_level = part.log_n - row_log;
    // Per chunk of primes: strided forward passes of both operands, then one
    // fused row pass (forward rows of both, pointwise product, inverse rows),
    // then the inverse strided passes.
    PassArgs pa = base_args(g, A, fwd, inv, tstride, part.log_n);
    set_fold(&pa, g, part.node, (uint64_t)1 << part.log_n);
    pa.other = B;
    PassArgs pf = base_args(g, B, fwd, inv, tstride, part.log_n);
    pf.node_base = part.node;
    pf.other = A;
    pf.level_start = row_level;
    pf.line_stride = 1;
    // The packing pass runs over every prime at once (it extracts each
    // coefficient's parts once for all of them); the chunks start after it.
    const int chunk = prime_chunk(np, part.log_n, square ? 1 : 2);
    if (chunk < np && !forward_head(sp, pa, sa, sb, square ? 1 : 2, np, s, 0, 0)) return false;
    for (int y0 = 0; y0 < np; y0 += chunk) {
        const int nc = std::min(chunk, np - y0);
        if (!forward_head(sp, chunk_args(pa, y0), sa, sb, square ? 1 : 2, nc, s, chunk < np ? 1 : 0)) return false;
        const PassArgs pc = chunk_args(pf, y0);
        {
            Timed t("fused_rows", s);
            if (!launch_fused(row_log, pc, nc, false, square, sp.fused_form(), s)) return false;
        }
        if (!inverse_tail(sp, pc, nc, s)) return false;
    }
    return true;
}

static bool recombine(hipStream_t s, const Geometry& g, double* rows) {
    const Catalog& c = cat();
    RecombineArgs a;
    memset(&a, 0, sizeof a);
    a.rows = rows;
    a.row_stride = row_len(g);
    a.n = (uint64_t)1 << g.log_t;
    a.tail = g.tail;
    a.field = c.field;
    for (uint32_t y = 0; y < g.primes; ++y) {
        const uint64_t m = c.p[y];
        a.w1[y] = centered(low_root((int)y, 1), m);
        uint64_t sv = 1, tv = 1;
        for (int q = 0; q < 3; ++q) {
            sv = mul_mod(sv, low_root((int)y, 2), m);
            tv = mul_mod(tv, low_root((int)y, 3), m);
            a.s[y][q] = centered(sv, m);
            a.t[y][q] = centered(tv, m);
        }
        a.half[y] = centered((m + 1) / 2, m);
    }
    Timed t("recombine", s);
    return launch_recombine(a, (int)g.primes, s);
}

// Residues of a split product: every block of the longer operand (grid z)
// times the shorter operand's spectrum, transformed once (z == blocks in the
// forward passes), into buf.spec_a, block z at z * primes * T.
static bool split_residues(hipStream_t s, const Geometry& g, const Source& sa, const Source& sb,
                           const OpBuffers& buf, const double* fwd, const double* inv, uint64_t tstride,
                           double** rows_out) {
    const bool swapped = g.digits_a < g.digits_b;
    const Source& lng = swapped ? sb : sa;
    const Source& sht = swapped ? sa : sb;
    const int np = (int)g.primes;
    const uint64_t T = (uint64_t)1 << g.log_t;
    PassArgs a = base_args(g, buf.spec_a, fwd, inv, tstride);
    a.split_blocks = g.blocks;
    a.block_digits = g.block_digits;
    a.block_stride = (uint64_t)np * T;
    a.other = buf.spec_b;
    a.src_a = lng;
    a.src_b = sht;
    const Split sp = split_of(g.log_t, kUseProduct);
    *rows_out = buf.spec_a;
    if (sp.n == 1) {
        Timed t("fused_single", s);
        return launch_fused(g.log_t, a, np, true, false, sp.fused_form(), s);
    }
    if (!forward_head(sp, a, lng, sht, (int)g.blocks + 1, np, s)) return false;
    const int row_log = sp.logs[sp.n - 1];
    a.level_start = g.log_t - row_log;
    a.line_stride = 1;
    {
        Timed t("fused_rows", s);
        if (!launch_fused(row_log, a, np, false, false, sp.fused_form(), s)) return false;
    }
    return inverse_tail(sp, a, np, s);
}

// Computes the canonical residues of the convolution into buf.spec_b (or
// spec_a for a square or a split) and returns their pointer.
static bool residues_with(hipStream_t s, const Geometry& g, const uint64_t* d_a, uint64_t la, const uint64_t* d_b,
                          uint64_t lb, bool square, const OpBuffers& buf, const double* fwd, const double* inv,
                          uint64_t tstride, double** rows_out) {
    const Source sa = source_of(d_a, la, g.digits_a, g.width, g.value_bits_a, g.is_signed);
    const Source sb = source_of(d_b, lb, g.digits_b, g.width, g.value_bits_b, g.is_signed);
    if (g.blocks && !square) return split_residues(s, g, sa, sb, buf, fwd, inv, tstride, rows_out);
    RowPart parts[3];
    const int n = row_parts(g, parts);
    for (int i = 0; i < n; ++i)
        if (!part_residues(s, g, parts[i], sa, sb, square, buf, fwd, inv, tstride)) return false;
    double* rows = square ? buf.spec_a : buf.spec_b;
    if (g.tail && !recombine(s, g, rows)) return false;
    *rows_out = rows;
    return true;
}

// The same with the shared root tables, grown as the geometry needs.
static bool residues(hipStream_t s, const Geometry& g, const uint64_t* d_a, uint64_t la, const uint64_t* d_b,
                     uint64_t lb, bool square, const OpBuffers& buf, double** rows_out) {
    const double *fwd = nullptr, *inv = nullptr;
    uint64_t tstride = 0;
    if (!ensure_tables(table_need(g), s, &fwd, &inv, &tstride)) {
        set_error("root table allocation failed");
        return false;
    }
    return residues_with(s, g, d_a, la, d_b, lb, square, buf, fwd, inv, tstride, rows_out);
}

static CrtArgs crt_args(const Geometry& g, const double* rows, uint64_t* out, uint64_t out_len) {
    const Catalog& c = cat();
    CrtArgs a;
    memset(&a, 0, sizeof a);
    a.residues = rows;
    a.row_stride = row_len(g);
    a.out_len = out_len;
    a.out = out;
    a.primes = g.primes;
    a.field = c.field;
    memcpy(a.pmod, c.pmod, sizeof a.pmod);
    memcpy(a.ginv, c.ginv, sizeof a.ginv);
    memcpy(a.pint, c.p, sizeof a.pint);
    a.value_bits = g.bound_bits;
    a.value_words = (uint32_t)crt_words((int)g.primes);
    const uint32_t wbits = g.disjoint ? g.stride * 64 : g.width;
    a.width_magic = (uint32_t)(((1ull << 32) + wbits - 1) / wbits);
    a.is_signed = g.is_signed ? 1 : 0;
    if (g.blocks) {
        a.split_blocks = g.blocks;
        a.split_digits = g.block_digits;
        a.block_stride = (uint64_t)g.primes * row_len(g);
    }
    return a;
}

static bool crt_product_from(hipStream_t s, const Geometry& g, const double* rows, uint64_t out_len,
                             uint32_t* carry, uint64_t* d_out) {
    CrtArgs a = crt_args(g, rows, d_out, out_len);
    a.width = g.width;
    a.coefficients = std::min<uint64_t>(g.conv_len, (out_len * 64 + g.width - 1) / g.width);
    const uint64_t blocks = (out_len + kCrtBlockLimbs - 1) / kCrtBlockLimbs;
    a.blocks = (uint32_t)blocks;
    // An output of at most two blocks whose coefficients fit one 512-thread
    // block's words is reconstructed by that one block, carries included.
    const uint64_t wide_span = (64 * 2 * (uint64_t)kCrtBlockLimbs + g.bound_bits) / g.width + 2;
    if (blocks == 2 && wide_span * (uint64_t)crt_words((int)g.primes) <= 2 * 4096) {
        a.wide = 1;
        a.blocks = 1;
    }
    a.agg_base = carry;
    a.agg_thresh = carry + blocks;
    a.first_thresh = carry + 2 * blocks;
    a.ripple_end = carry + 3 * blocks;
    return launch_crt_product(a, s);
}

// Schoolbook product of device operands: one launch of output blocks, then
// the carries across blocks.
static bool school_on(hipStream_t s, const uint64_t* d_a, uint64_t la, const uint64_t* d_b, uint64_t lb,
                      uint32_t* carry, uint64_t* d_out) {
    CrtArgs a;
    memset(&a, 0, sizeof a);
    a.out = d_out;
    a.out_len = la + lb;
    const uint64_t blocks = carry_blocks(la + lb, true);
    a.blocks = (uint32_t)blocks;
    a.agg_base = carry;
    a.agg_thresh = carry + blocks;
    a.first_thresh = carry + 2 * blocks;
    a.ripple_end = carry + 3 * blocks;
    a.op_a = d_a;
    a.op_b = d_b;
    a.len_a = la;
    a.len_b = lb;
    return launch_school(a, s);
}

// tables: root tables of at least table_need(g) entries per prime, or null
// for the shared tables.
static bool product_with(hipStream_t s, const Geometry& g, const uint64_t* d_a, uint64_t la, const uint64_t* d_b,
                         uint64_t lb, bool square, const OpBuffers& buf, const double* const* tables,
                         uint64_t tstride, uint64_t* d_out) {
    if (g.school) {
        if (std::min(la, lb) > (uint64_t)kSchoolMaxShort) return false;
        return school_on(s, d_a, la, square ? d_a : d_b, square ? la : lb, buf.carry, d_out);
    }
    double* rows = nullptr;
    const bool ok = tables ? residues_with(s, g, d_a, la, d_b, lb, square, buf, tables[0], tables[1], tstride, &rows)
                           : residues(s, g, d_a, la, d_b, lb, square, buf, &rows);
    if (!ok) return false;
    return crt_product_from(s, g, rows, la + lb, buf.carry, d_out);
}

bool product_on(hipStream_t s, const Geometry& g, const uint64_t* d_a, uint64_t la, const uint64_t* d_b,
                uint64_t lb, bool square, const OpBuffers& buf, uint64_t* d_out) {
    return product_with(s, g, d_a, la, d_b, lb, square, buf, nullptr, 0, d_out);
}

bool convolution_on(hipStream_t s, const Geometry& g, const ntt_conv_plan* p, const uint64_t* d_a,
                    const uint64_t* d_b, bool square, const OpBuffers& buf, uint64_t* d_out) {
    if (p->schoolbook) {
        Timed t("conv_columns", s);
        return launch_conv_columns(d_a, p->len_a, d_b, p->len_b, d_out, (uint32_t)p->stride, p->is_signed, s);
    }
    double* rows = nullptr;
    if (!residues(s, g, d_a, p->len_a, d_b, p->len_b, square, buf, &rows)) return false;
    CrtArgs a = crt_args(g, rows, d_out, (p->len_a + p->len_b - 1) * g.stride);
    a.coefficients = p->len_a + p->len_b - 1;
    a.stride = g.stride;
    a.width = g.stride * 64;
    Timed t("crt_disjoint", s);
    return launch_crt_disjoint(a, s);
}

// Executor forms: grow the executor's own buffers, then run on its stream.
static bool exec_buffers(Executor& ex, const Geometry& g, bool square, uint64_t out_len, OpBuffers* buf) {
    if (!ex.spec_a.ensure(std::max<uint64_t>(op_spectra_bytes(g), 8)) ||
        (!square && !ex.spec_b.ensure(std::max<uint64_t>(op_spectra_bytes_b(g), 8))) ||
        !ex.carry.ensure(op_carry_bytes(out_len, g.school))) {
        set_error("device working memory allocation failed");
        return false;
    }
    buf->spec_a = ex.spec_a.as<double>();
    buf->spec_b = square ? buf->spec_a : ex.spec_b.as<double>();
    buf->carry = ex.carry.as<uint32_t>();
    return true;
}

// Product of device operands into d_out (exactly la+lb limbs).
static bool device_product(Executor& ex, const Geometry& g, const uint64_t* d_a, uint64_t la,
                           const uint64_t* d_b, uint64_t lb, bool square, uint64_t* d_out) {
    OpBuffers buf;
    if (!exec_buffers(ex, g, square, la + lb, &buf)) return false;
    return product_on(ex.stream, g, d_a, la, d_b, lb, square, buf, d_out);
}

// Plan-time autotuning of a single operation's split: the first prepare of a
// transform length times every admissible split on the executor's own
// buffers (operands of the plan's shape, contents irrelevant to the timing),
// for the product and, when the operands are as long, for the square, then
// tries the kernel-form flags on the fastest (flag_variants), and records the
// result for every later call of that length and use. A single-pass length
// (up to 2^12 points) has one split and tries the flags only.
static void tune_split(Executor& ex, int K, int use, const std::function<bool()>& run) {
    if (K < kMinLog || K > kMaxLog || use < 0 || use > 2) return;
    if (getenv("NTTAMD_ROW_LOG") || getenv("NTTAMD_PASSES") || flags_forced()) return;  // forced: no trials
    {
        std::lock_guard<std::mutex> g(g_tune_lock);
        if (g_tuned[use][K].n) return;
    }
    SplitCandidates cand = split_candidates(K);
    if (cand.empty()) {
        cand.n = 1;
        cand.s[0] = Split{};
        cand.s[0].n = 1;
        cand.s[0].logs[0] = K;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) hipEventDestroy(e0);
        return;
    }
    // Time per run of a candidate: one warm run (timed), then two rounds of
    // enough runs for about 2 ms of work each (3 to 20), the faster round
    // kept, so that short products are not decided by launch jitter nor
    // long ones by a single slow round.
    auto time_of = [&](const Split& sp) {
        t_force = sp;
        hipEventRecord(e0, ex.stream);
        bool ok = run();
        hipEventRecord(e1, ex.stream);
        float warm = 1e30f;
        if (ok && hipEventSynchronize(e1) == hipSuccess) hipEventElapsedTime(&warm, e0, e1);
        const int reps = warm <= 0.1f ? 20 : std::max(3, std::min(20, (int)(2.0f / warm)));
        float best_round = 1e30f;
        for (int round = 0; round < 2 && ok; ++round) {
            hipEventRecord(e0, ex.stream);
            for (int r = 0; r < reps && ok; ++r) ok = run();
            hipEventRecord(e1, ex.stream);
            float ms = 1e30f;
            if (ok && hipEventSynchronize(e1) == hipSuccess) hipEventElapsedTime(&ms, e0, e1);
            best_round = std::min(best_round, ms / reps);
        }
        return ok ? best_round : 1e30f;
    };
    // A later candidate replaces the best only when it is faster by the
    // margin (2 %), so near-equal candidates resolve to the first one (the
    // largest row, no flags) instead of to timing noise.
    constexpr float kMargin = 0.98f;
    Split best = cand[0];
    float best_ms = 1e30f;
    if (cand.size() > 1)
        for (const Split& sp : cand) {
            const float ms = time_of(sp);
            if (ms < best_ms * kMargin) {
                best_ms = ms;
                best = sp;
            }
        }
    Split var[kMaxVariants];
    const int nv = flag_variants(K, use, best, var);
    if (nv && cand.size() == 1) best_ms = time_of(best);
    for (int i = 0; i < nv; ++i) {
        const Split t = with_variant(best, var[i]);
        if (!split_ok(K, t) && t.n > 1) continue;
        const float ms = time_of(t);
        if (ms < best_ms * kMargin) {
            best_ms = ms;
            best = t;
        }
    }
    t_force = Split{};
    hipStreamSynchronize(ex.stream);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    clear_error();
    std::lock_guard<std::mutex> g(g_tune_lock);
    if (!g_tuned[use][K].n) g_tuned[use][K] = best;
}

// ---------------------------------------------------------------------------
// Host-buffer calls: operands to device, run, result back, dst written last.

bool host_pinned(const void* p) {
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();  // pageable memory is not an error here
        return false;
    }
    return at.type == hipMemoryTypeHost;
}

bool HostStage::init() {
    if (ev[kBuffers - 1]) return true;
    for (int i = 0; i < kBuffers; ++i) {
        if (!buf[i].ensure(kChunk)) {
            release();
            return false;
        }
        // The host maps pinned pages on their first touch; doing it here
        // keeps that cost (milliseconds for a chunk) out of the copies.
        memset(buf[i].ptr, 0, kChunk);
    }
    for (int i = 0; i < kBuffers; ++i) {
        if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) {
            ev[i] = nullptr;
            release();
            return false;
        }
    }
    return true;
}

void HostStage::release() {
    for (int i = 0; i < kBuffers; ++i) {
        if (ev[i]) hipEventDestroy(ev[i]);
        ev[i] = nullptr;
        buf[i].release();
    }
}

bool HostStage::to_device(hipStream_t s, void* dev, const void* host, size_t bytes) {
    const char* src = static_cast<const char*>(host);
    char* dst = static_cast<char*>(dev);
    // The ring continues where the previous copy left it, so a second
    // operand's first chunks go to buffers whose DMAs are not the latest.
    for (size_t o = 0; o < bytes; o += kChunk, next = (next + 1) % kBuffers) {
        const int k = next;
        const size_t m = bytes - o < kChunk ? bytes - o : kChunk;
        // The buffer's previous DMA (this call's or an earlier one's) must
        // have finished before the buffer is refilled.
        if (hipEventSynchronize(ev[k]) != hipSuccess) return false;
        memcpy(buf[k].ptr, src + o, m);
        if (hipMemcpyAsync(dst + o, buf[k].ptr, m, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipEventRecord(ev[k], s) != hipSuccess)
            return false;
    }
    return true;
}

bool HostStage::to_host(hipStream_t s, void* host, const void* dev, size_t bytes) {
    char* dst = static_cast<char*>(host);
    const char* src = static_cast<const char*>(dev);
    const size_t chunks = (bytes + kChunk - 1) / kChunk;
    // HIP calls are thread-safe; the DMAs of different chunks touch
    // disjoint memory, so the order in which threads queue them is free.
    auto queue = [&](size_t c) {
        const size_t o = c * kChunk, m = bytes - o < kChunk ? bytes - o : kChunk;
        const int k = (int)(c % kBuffers);
        return hipMemcpyAsync(buf[k].ptr, src + o, m, hipMemcpyDeviceToHost, s) == hipSuccess &&
               hipEventRecord(ev[k], s) == hipSuccess;
    };
    for (size_t c = 0; c < chunks && c < (size_t)kBuffers; ++c)
        if (!queue(c)) return false;
    // Buffer k belongs to one thread: it waits for the buffer's DMA, copies
    // the chunk out (the host copy out of DMA-written memory, not the link,
    // is the slow side) and queues the buffer's next chunk.
    std::atomic<bool> ok{true};
    auto drain = [&](int k) {
        for (size_t c = (size_t)k; c < chunks && ok.load(std::memory_order_relaxed); c += kBuffers) {
            const size_t o = c * kChunk, m = bytes - o < kChunk ? bytes - o : kChunk;
            if (hipEventSynchronize(ev[k]) != hipSuccess) {
                ok = false;
                return;
            }
            memcpy(dst + o, buf[k].ptr, m);
            if (c + kBuffers < chunks && !queue(c + kBuffers)) {
                ok = false;
                return;
            }
        }
    };
    const int used = (int)std::min<size_t>(chunks, (size_t)kBuffers);
    std::thread helpers[kBuffers];
    int started = 1;
    try {
        for (; started < used; ++started) helpers[started] = std::thread(drain, started);
    } catch (...) {
        // No more threads: the caller drains the rest of the buffers itself.
    }
    drain(0);
    for (int k = started; k < used; ++k) drain(k);
    for (int k = 1; k < started; ++k) helpers[k].join();
    return ok.load();
}

bool copy_to_device(HostStage& st, hipStream_t s, void* dev, const void* host, size_t bytes) {
    if (bytes == 0) return true;
    if (host_pinned(host) || !st.init())
        return hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, s) == hipSuccess;
    return st.to_device(s, dev, host, bytes);
}

bool copy_to_host(HostStage& st, hipStream_t s, void* host, const void* dev, size_t bytes) {
    if (bytes == 0) return hipStreamSynchronize(s) == hipSuccess;
    if (host_pinned(host) || !st.init())
        return hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s) == hipSuccess &&
               hipStreamSynchronize(s) == hipSuccess;
    return st.to_host(s, host, dev, bytes);
}

// Host-buffer calls moving at most this many bytes (operands and result)
// run on mapped host memory; NTTAMD_MAPPED_BYTES sets another limit (tests
// and sweeps). Measured per call from pageable arrays against the staged
// copies: n x n limbs, n = 1,024: 43 / 76 us, 16,384: 72 / 132 us, 131,072
// (4 MiB moved): 251 / 277 us, 262,144: 511 / 442 us.
constexpr uint64_t kMappedCallBytes = (uint64_t)4 << 20;
static uint64_t mapped_call_bytes() {
    static const uint64_t limit = [] {
        const char* e = getenv("NTTAMD_MAPPED_BYTES");
        return e ? (uint64_t)strtoull(e, nullptr, 10) : kMappedCallBytes;
    }();
    return limit;
}

static bool stage_and_run(Executor& ex, uint64_t out_len, const uint64_t* a, uint64_t la,
                          const uint64_t* b, uint64_t lb, bool same,
                          bool (*run)(Executor&, void*, const uint64_t*, const uint64_t*, uint64_t*),
                          void* ctx, uint64_t* dst) {
    std::lock_guard<std::mutex> guard(ex.lock);
    if (!ex.init()) {
        set_error("stream creation failed");
        return false;
    }
    const uint64_t in_limbs = la + (same ? 0 : lb);
    if ((in_limbs + out_len) * 8 <= mapped_call_bytes() && ex.mapped.ensure((in_limbs + out_len) * 8)) {
        // A small call: the kernels read the operands from, and write the
        // result to, mapped pinned host memory, so no DMA round trip (each
        // ~10 us, against ~15 us of kernels for a 1,024-limb product) sits
        // in the call.
        uint64_t* h = ex.mapped.host.as<uint64_t>();
        uint64_t* d = static_cast<uint64_t*>(ex.mapped.dev);
        memcpy(h, a, la * 8);
        if (!same) memcpy(h + la, b, lb * 8);
        if (!run(ex, ctx, d, same ? d : d + la, d + in_limbs)) {
            hipStreamSynchronize(ex.stream);
            if (t_error.empty()) set_error("kernel launch failed");
            return false;
        }
        if (hipStreamSynchronize(ex.stream) != hipSuccess) {
            set_error("device execution failed");
            return false;
        }
        memcpy(dst, h + in_limbs, out_len * 8);
        return true;
    }
    if (!ex.in_a.ensure(la * 8) || (!same && !ex.in_b.ensure(lb * 8)) || !ex.out.ensure(out_len * 8)) {
        set_error("operand allocation failed");
        return false;
    }
    const hipStream_t s = ex.stream;
    uint64_t* da = ex.in_a.as<uint64_t>();
    uint64_t* db = same ? da : ex.in_b.as<uint64_t>();
    if (!copy_to_device(ex.stage, s, da, a, la * 8) || (!same && !copy_to_device(ex.stage, s, db, b, lb * 8))) {
        set_error("operand copy failed");
        return false;
    }
    uint64_t* dout = ex.out.as<uint64_t>();
    if (!run(ex, ctx, da, db, dout)) {
        hipStreamSynchronize(s);
        if (t_error.empty()) set_error("kernel launch failed");
        return false;
    }
    // Every kernel has completed before dst is touched: a failure up to here
    // leaves dst unchanged.
    if (hipStreamSynchronize(s) != hipSuccess) {
        set_error("device execution failed");
        return false;
    }
    if (!copy_to_host(ex.stage, s, dst, dout, out_len * 8)) {
        set_error("result copy failed");
        return false;
    }
    return true;
}

struct MulCtx {
    const Geometry* g;
    uint64_t la, lb;
    bool square;
};

static bool run_mul(Executor& ex, void* ctx, const uint64_t* da, const uint64_t* db, uint64_t* dout) {
    MulCtx* m = static_cast<MulCtx*>(ctx);
    return device_product(ex, *m->g, da, m->la, db, m->lb, m->square, dout);
}

struct ConvCtx {
    const Geometry* g;
    const ntt_conv_plan* plan;
    bool square;
};

static bool run_conv(Executor& ex, void* ctx, const uint64_t* da, const uint64_t* db, uint64_t* dout) {
    ConvCtx* c = static_cast<ConvCtx*>(ctx);
    OpBuffers buf{nullptr, nullptr, nullptr};
    if (!c->plan->schoolbook && !exec_buffers(ex, *c->g, c->square, 0, &buf)) return false;
    return convolution_on(ex.stream, *c->g, c->plan, da, db, c->square, buf, dout);
}

bool product_geometry(const ntt_mul_plan* plan, Geometry* g) {
    if (!plan || !plan->valid) return false;
    return select_product(plan->len_a, plan->len_b, g);
}

bool conv_geometry(const ntt_conv_plan* p, Geometry* g) {
    if (!p || !p->valid) return false;
    g->primes = p->primes;
    g->width = 64;
    g->value_bits_a = p->value_bits_a;
    g->value_bits_b = p->value_bits_b;
    g->digits_a = p->len_a;
    g->digits_b = p->len_b;
    g->conv_len = p->len_a + p->len_b - 1;
    g->log_t = (int)p->log_len;
    g->tail = p->row_tail;
    g->blocks = p->split_blocks;
    g->block_digits = p->block_len;
    g->bound_bits = p->bound_bits;
    g->disjoint = true;
    g->is_signed = p->is_signed != 0;
    g->stride = (uint32_t)p->stride;
    return true;
}

static bool mul_host(Executor& ex, const ntt_mul_plan* plan, uint64_t* dst, size_t dst_len,
                     const uint64_t* a, size_t a_len, const uint64_t* b, size_t b_len, bool square) {
    t_error.clear();
    Geometry g;
    if (!product_geometry(plan, &g)) return false;
    if (a_len != plan->len_a || b_len != plan->len_b) return false;
    const uint64_t out_len = (uint64_t)a_len + b_len;
    if (!dst || dst_len < out_len || !a || (!square && !b)) return false;
    MulCtx ctx{&g, a_len, b_len, square};
    return stage_and_run(ex, out_len, a, a_len, b, b_len, square, run_mul, &ctx, dst);
}

// ConvPlan::admits, src/convolve.rs:169-184.
bool conv_admits(const ntt_conv_plan* p, const uint64_t* a, size_t a_len, const uint64_t* b,
                        size_t b_len) {
    if (a_len != p->len_a || b_len != p->len_b || !a || !b) return false;
    const uint64_t* ops[2] = {a, b};
    const size_t lens[2] = {a_len, b_len};
    const uint64_t maxes[2] = {p->max_a, p->max_b};
    for (int s = 0; s < 2; ++s) {
        if (p->is_signed) {
            for (size_t i = 0; i < lens[s]; ++i) {
                const int64_t v = (int64_t)ops[s][i];
                const uint64_t mag = v < 0 ? 0 - (uint64_t)v : (uint64_t)v;
                if (mag > maxes[s]) return false;
            }
        } else {
            uint64_t u = 0;
            for (size_t i = 0; i < lens[s]; ++i) u |= ops[s][i];
            if (u > maxes[s])
                for (size_t i = 0; i < lens[s]; ++i)
                    if (ops[s][i] > maxes[s]) return false;
        }
    }
    return true;
}

static bool conv_host(Executor& ex, const ntt_conv_plan* plan, uint64_t* dst, size_t dst_len,
                      const uint64_t* a, size_t a_len, const uint64_t* b, size_t b_len) {
    t_error.clear();
    if (!plan || !plan->valid) return false;
    if (!conv_admits(plan, a, a_len, b, b_len)) return false;
    if (!dst || dst_len < plan->output_len) return false;
    Geometry g;
    if (!conv_geometry(plan, &g)) return false;
    const bool square = a == b && a_len == b_len;
    ConvCtx ctx{&g, plan, square && !plan->schoolbook};
    return stage_and_run(ex, plan->output_len, a, a_len, b, b_len, square, run_conv, &ctx, dst);
}

}  // namespace nttamd

namespace nttamd {

// ---------------------------------------------------------------------------
// Batched products: many uneven products in one graph launch.
//
// Products are bucketed by transform length. Every pass of a bucket is one
// launch over all (product, prime) items of the bucket, and the
// reconstruction is one launch per prime count over all (product, block)
// items; products the pricing gives to the schoolbook skip the transforms
// and take one schoolbook launch over their (product, block) items. Pointers and shapes of the products live in a device array that the
// kernels read, so the captured graph serves any operands bound later. The
// split of each bucket's transform into passes comes from an autotuner that
// times the candidate splits on the batch's own data the first time a
// transform length is run.








static bool issue_launch(const BatchLaunch& L, hipStream_t s) {
    Timed t(L.name, s);
    switch (L.kind) {
        case BatchLaunch::kFwdStrided: return launch_fwd_strided(L.logl, L.pass, L.units, 1, L.flag, L.alt, false, s);
        case BatchLaunch::kPackPrimes: return launch_pack_primes_batch(L.logl, L.pass, L.units, s);
        case BatchLaunch::kFused: return launch_fused(L.logl, L.pass, L.units, L.pack, L.flag, L.form, s);
        case BatchLaunch::kInvStrided: return launch_inv_strided(L.logl, L.pass, L.units, L.flag, L.alt, false, s);
        case BatchLaunch::kCrt: return launch_crt_product(L.crt, s);
        case BatchLaunch::kSchool: return launch_school(L.crt, s);
    }
    return false;
}

static bool issue_program(const std::vector<BatchLaunch>& prog, hipStream_t s) {
    for (const BatchLaunch& L : prog)
        if (!issue_launch(L, s)) return false;
    return true;
}

// Appends the transform launches of one bucket for the given split: the
// strided forward passes over both operands of every product, the fused row
// pass (forward rows of both operands, pointwise product, inverse rows) and
// the inverse strided passes, the last of which leaves canonical residues.
static void bucket_program(std::vector<BatchLaunch>& prog, const PassArgs& base, int K, const Split& sp,
                           const BatchItem* I, const BatchBucket& r) {
    auto pass = [&](BatchLaunch::Kind kind, int logl, bool flag, bool pack, int level, uint32_t first,
                    uint32_t units, int operand, const char* name) {
        if (units == 0) return;
        BatchLaunch L;
        memset(&L, 0, sizeof L);
        L.kind = kind;
        L.logl = logl;
        L.flag = flag;
        L.pack = pack;
        L.alt = sp.alt_tiles != 0;
        L.form = sp.fused_form();
        L.units = (int)units;
        L.name = name;
        L.pass = base;
        L.pass.items = I + first;
        L.pass.operand = operand;
        L.pass.level_start = level;
        L.pass.line_stride = kind == BatchLaunch::kFused ? 1 : (uint64_t)1 << (K - level - logl);
        prog.push_back(L);
    };
    if (sp.n == 1) {
        pass(BatchLaunch::kFused, K, false, true, 0, r.ns, r.n_ns, kOperandProduct, "batch_fused_single");
        pass(BatchLaunch::kFused, K, true, true, 0, r.sq, r.n_sq, kOperandProduct, "batch_fused_single");
        return;
    }
    const int row_log = sp.logs[sp.n - 1];
    const int row_level = K - row_log;
    int level = 0;
    for (int i = 0; i + 1 < sp.n; ++i) {
        if (i == 0 && !sp.per_prime_pack && r.shared_ok && pack_shared() &&
            pack_primes_batch_fits(sp.logs[0], (uint64_t)1 << (K - sp.logs[0])))
            pass(BatchLaunch::kPackPrimes, sp.logs[0], true, false, 0, r.pk, r.n_pk, kOperandForward,
                 "batch_pack_primes");
        else
            pass(BatchLaunch::kFwdStrided, sp.logs[i], i == 0, false, level, r.fwd, r.n_fwd, kOperandForward,
                 i == 0 ? "batch_fwd_pack_strided" : "batch_fwd_strided");
        level += sp.logs[i];
    }
    pass(BatchLaunch::kFused, row_log, false, false, row_level, r.ns, r.n_ns, kOperandProduct, "batch_fused_rows");
    pass(BatchLaunch::kFused, row_log, true, false, row_level, r.sq, r.n_sq, kOperandProduct, "batch_fused_rows");
    level = row_level;
    for (int i = sp.n - 2; i >= 0; --i) {
        level -= sp.logs[i];
        pass(BatchLaunch::kInvStrided, sp.logs[i], i == 0, false, level, r.ns, r.n_ns + r.n_sq, kOperandProduct,
             i == 0 ? "batch_inv_strided_canon" : "batch_inv_strided");
    }
}

static PassArgs bucket_base(const Batch& B, int K) {
    Geometry gK;
    gK.log_t = K;
    gK.primes = kMaxPrimes;
    PassArgs base = base_args(gK, nullptr, B.fwd, B.inv, B.tstride);
    base.batch = B.products.as<BatchProduct>();
    return base;
}

static void build_program(Batch& B) {
    B.program.clear();
    B.seg.clear();
    const BatchItem* I = B.items.as<BatchItem>();
    for (const BatchBucket& r : B.buckets) {
        B.seg.push_back(B.program.size());
        bucket_program(B.program, bucket_base(B, r.K), r.K, split_of(r.K, kUseBatch), I, r);
    }
    // The schoolbook launch (np == 0) depends on no transform: its own
    // segment; the reconstructions follow every bucket.
    for (int np = 0; np <= kMaxPrimes; ++np) {
        if (np <= 1) B.seg.push_back(B.program.size());
        if (!B.crt_count[np]) continue;
        BatchLaunch L;
        memset(&L, 0, sizeof L);
        L.kind = np ? BatchLaunch::kCrt : BatchLaunch::kSchool;
        L.name = np ? "batch_crt" : "batch_school";
        if (np) {
            Geometry gN;
            gN.primes = np;
            L.crt = crt_args(gN, nullptr, nullptr, 0);
        }
        L.crt.batch = B.products.as<BatchProduct>();
        L.crt.items = I + B.crt_begin[np];
        L.crt.blocks = B.crt_count[np];
        L.crt.scan_items = I + B.scan_begin[np];
        L.crt.scan_count = B.scan_count[np];
        B.program.push_back(L);
    }
}

bool batch_plan(Batch& B, size_t count, const size_t* la, const size_t* lb, const uint8_t* sq) {
    B.drop_graph();
    B.bound = false;
    B.tuned = false;
    B.count = 0;
    B.program.clear();
    if (!B.ex.init()) {
        set_error("batch: stream creation failed");
        return false;
    }
    B.la.assign(count, 0);
    B.lb.assign(count, 0);
    B.square.assign(count, 0);
    B.swapped.assign(count, 0);
    B.geo.assign(count, Geometry());
    B.host.assign(count, BatchProduct());
    uint64_t arena = 0, carry = 0, max_half = 0;
    std::vector<uint64_t> off_a(count), off_b(count), off_c(count);
    for (size_t i = 0; i < count; ++i) {
        const bool square = sq && sq[i];
        B.square[i] = square ? 1 : 0;
        B.la[i] = la[i];
        B.lb[i] = square ? la[i] : lb[i];
        if (B.la[i] == 0 || B.lb[i] == 0 || !select_product(B.la[i], B.lb[i], &B.geo[i], true)) {
            set_error("batch: product " + std::to_string(i) + " has an unsupported shape");
            return false;
        }
        const Geometry& g = B.geo[i];
        const uint64_t T = (uint64_t)1 << g.log_t;
        max_half = std::max(max_half, T / 2);
        off_a[i] = arena;
        arena += op_spectra_bytes(g) / 8;
        off_b[i] = arena;
        if (!square) arena += op_spectra_bytes_b(g) / 8;
        const uint64_t blocks = carry_blocks(B.la[i] + B.lb[i], g.school);
        off_c[i] = carry;
        carry += carry_words((uint32_t)blocks);
    }
    if (!ensure_tables(max_half, B.ex.stream, &B.fwd, &B.inv, &B.tstride)) {
        set_error("batch: root table allocation failed");
        return false;
    }
    if (!B.arena.ensure(std::max<uint64_t>(arena, 1) * 8) || !B.carry.ensure(std::max<uint64_t>(carry, 1) * 4) ||
        !B.products.ensure(std::max<size_t>(count, 1) * sizeof(BatchProduct))) {
        set_error("batch: device allocation failed");
        return false;
    }
    double* A = B.arena.as<double>();
    uint32_t* C = B.carry.as<uint32_t>();
    for (size_t i = 0; i < count; ++i) {
        const Geometry& g = B.geo[i];
        const bool square = B.square[i] != 0;
        BatchProduct& P = B.host[i];
        memset(&P, 0, sizeof P);
        P.src_a = source_of(nullptr, B.la[i], g.digits_a, g.width, g.value_bits_a, false);
        P.src_b = source_of(nullptr, B.lb[i], g.digits_b, g.width, g.value_bits_b, false);
        if (g.blocks) {
            // src_a is the longer operand (batch_bind keeps the swap).
            B.swapped[i] = g.digits_a < g.digits_b ? 1 : 0;
            if (B.swapped[i]) std::swap(P.src_a, P.src_b);
            P.split_blocks = g.blocks;
            P.block_digits = g.block_digits;
            P.block_stride = (uint64_t)g.primes * ((uint64_t)1 << g.log_t);
        }
        P.spec_a = A + off_a[i];
        P.spec_b = square ? P.spec_a : A + off_b[i];
        P.out_len = B.la[i] + B.lb[i];
        P.coefficients = std::min<uint64_t>(g.conv_len, (P.out_len * 64 + g.width - 1) / g.width);
        P.row_stride = (uint64_t)1 << g.log_t;
        P.primes = g.primes;
        P.square = square ? 1 : 0;
        P.width = g.width;
        P.width_magic = (uint32_t)(((1ull << 32) + g.width - 1) / g.width);
        P.value_bits = g.bound_bits;
        P.blocks = (uint32_t)carry_blocks(P.out_len, g.school);
        P.carry = C + off_c[i];
    }
    // Work items: per transform length every (product, prime) pair, then the
    // pairs of products with two operands and of squares; per prime count
    // (0: the schoolbook products) every (product, output block) pair.
    std::vector<BatchItem> items;
    B.buckets.clear();
    // which: 0 first operands, 1 second operands (forward items); 2 products
    // with two operands, 3 squares (product items).
    // A split product has one first-operand item and one product item per
    // block and prime (sub = prime | block << kItemBlockShift) and one
    // second-operand item per prime for its shorter operand.
    auto add = [&](int K, int which) {
        for (size_t i = 0; i < count; ++i) {
            if (B.geo[i].log_t != K || B.geo[i].school) continue;
            const bool sq = B.square[i] != 0;
            if ((which == 1 || which == 2) && sq) continue;
            if (which == 3 && !sq) continue;
            const uint32_t flag = which == 1 ? kItemOperandB : 0u;
            const uint32_t blocks = which == 1 || B.geo[i].blocks == 0 ? 1 : B.geo[i].blocks;
            for (uint32_t z = 0; z < blocks; ++z)
                for (uint32_t y = 0; y < B.geo[i].primes; ++y)
                    items.push_back({(uint32_t)i, y | (z << kItemBlockShift) | flag});
        }
    };
    for (int K = kMinLog; K <= kMaxLog; ++K) {
        BatchBucket r;
        r.K = K;
        r.fwd = (uint32_t)items.size();
        add(K, 0);
        if ((uint32_t)items.size() == r.fwd) continue;
        add(K, 1);
        r.n_fwd = (uint32_t)items.size() - r.fwd;
        // Shared packing items: the forward items of prime 0.
        r.pk = (uint32_t)items.size();
        for (uint32_t k = r.fwd; k < r.fwd + r.n_fwd; ++k)
            if ((items[k].sub & kItemPrimeMask) == 0) items.push_back(items[k]);
        r.n_pk = (uint32_t)items.size() - r.pk;
        r.shared_ok = true;
        for (size_t i = 0; i < count; ++i)
            if (B.geo[i].log_t == K && !B.geo[i].school &&
                (pack_kind(B.host[i].src_a) == kPack4 || pack_kind(B.host[i].src_b) == kPack4))
                r.shared_ok = false;
        r.ns = (uint32_t)items.size();
        add(K, 2);
        r.n_ns = (uint32_t)items.size() - r.ns;
        r.sq = (uint32_t)items.size();
        add(K, 3);
        r.n_sq = (uint32_t)items.size() - r.sq;
        B.buckets.push_back(r);
    }
    for (int np = 0; np <= kMaxPrimes; ++np) {
        B.crt_begin[np] = (uint32_t)items.size();
        for (size_t i = 0; i < count; ++i)
            if (B.geo[i].school ? np == 0 : (int)B.geo[i].primes == np)
                for (uint32_t blk = 0; blk < B.host[i].blocks; ++blk) items.push_back({(uint32_t)i, blk});
        B.crt_count[np] = (uint32_t)items.size() - B.crt_begin[np];
    }
    for (int np = 0; np <= kMaxPrimes; ++np) {
        B.scan_begin[np] = (uint32_t)items.size();
        for (size_t i = 0; i < count; ++i)
            if ((B.geo[i].school ? np == 0 : (int)B.geo[i].primes == np) && carry_scanned(B.host[i].blocks))
                for (uint32_t ch = 0; ch < carry_chunks(B.host[i].blocks); ++ch) items.push_back({(uint32_t)i, ch});
        B.scan_count[np] = (uint32_t)items.size() - B.scan_begin[np];
    }
    if (!B.items.ensure(std::max<size_t>(items.size(), 1) * sizeof(BatchItem)) ||
        (!items.empty() && hipMemcpy(B.items.ptr, items.data(), items.size() * sizeof(BatchItem),
                                     hipMemcpyHostToDevice) != hipSuccess)) {
        set_error("batch: work item upload failed");
        return false;
    }
    B.count = count;
    build_program(B);
    return true;
}

bool batch_bind(Batch& B, const uint64_t* const* d_a, const uint64_t* const* d_b, uint64_t* const* d_out) {
    for (size_t i = 0; i < B.count; ++i) {
        const uint64_t* pb = B.square[i] ? d_a[i] : d_b[i];
        if (!d_a[i] || !pb || !d_out[i]) {
            set_error("batch: null operand or output pointer at product " + std::to_string(i));
            return false;
        }
        // A split product keeps its longer operand in src_a.
        const bool swap = B.swapped[i] != 0;
        B.host[i].src_a.limbs = swap ? pb : d_a[i];
        B.host[i].src_b.limbs = swap ? d_a[i] : pb;
        B.host[i].out = d_out[i];
    }
    if (B.count && hipMemcpyAsync(B.products.ptr, B.host.data(), B.count * sizeof(BatchProduct),
                                  hipMemcpyHostToDevice, B.ex.stream) != hipSuccess) {
        set_error("batch: descriptor upload failed");
        return false;
    }
    if (hipStreamSynchronize(B.ex.stream) != hipSuccess) {
        set_error("batch: descriptor upload failed");
        return false;
    }
    B.bound = true;
    return true;
}

static float time_launches(const std::vector<BatchLaunch>& prog, hipStream_t s, int reps) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 1e30f;
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess && issue_program(prog, s)) {
        hipEventRecord(e0, s);
        bool ok = true;
        for (int r = 0; r < reps && ok; ++r) ok = issue_program(prog, s);
        hipEventRecord(e1, s);
        if (ok && hipEventSynchronize(e1) == hipSuccess) hipEventElapsedTime(&ms, e0, e1);
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    return ms;
}

static int batch_branches();

// Times the transform work of each untuned bucket under each candidate split,
// then the kernel-form flags on the fastest (flag_variants), and keeps the
// fastest. The operands must be bound: the timed passes read them.
static bool batch_autotune(Batch& B) {
    const BatchItem* I = B.items.as<BatchItem>();
    const bool forced = getenv("NTTAMD_ROW_LOG") || getenv("NTTAMD_PASSES") || flags_forced();
    bool changed = false;
    for (const BatchBucket& r : B.buckets) {
        if (forced) break;
        {
            std::lock_guard<std::mutex> g(g_tune_lock);
            if (g_tuned[kUseBatch][r.K].n) continue;
        }
        SplitCandidates cand = split_candidates(r.K);
        if (cand.empty()) {
            cand.n = 1;
            cand.s[0] = Split{};
            cand.s[0].n = 1;
            cand.s[0].logs[0] = r.K;
        }
        // The bucket's whole transform work under a split (it recomputes
        // everything from the operands, so repeats are harmless).
        auto time_of = [&](const Split& sp) {
            std::vector<BatchLaunch> prog;
            bucket_program(prog, bucket_base(B, r.K), r.K, sp, I, r);
            return time_launches(prog, B.ex.stream, 3);
        };
        Split best = cand[0];
        float best_ms = 1e30f;
        if (cand.size() > 1)
            for (const Split& sp : cand) {
                const float ms = time_of(sp);
                if (ms < best_ms) {
                    best_ms = ms;
                    best = sp;
                }
            }
        Split var[kMaxVariants];
        const int nv = flag_variants(r.K, kUseBatch, best, var);
        if (nv && cand.size() == 1) best_ms = time_of(best);
        for (int i = 0; i < nv; ++i) {
            const Split t = with_variant(best, var[i]);
            if (!split_ok(r.K, t) && t.n > 1) continue;
            const float ms = time_of(t);
            if (ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        }
        if (hipStreamSynchronize(B.ex.stream) != hipSuccess) {
            set_error("batch: autotuning run failed");
            return false;
        }
        std::lock_guard<std::mutex> g(g_tune_lock);
        if (!g_tuned[kUseBatch][r.K].n) {
            g_tuned[kUseBatch][r.K] = best;
            changed = true;
        }
    }
    B.tuned = true;
    if (changed) {
        B.drop_graph();
        build_program(B);
    }
    // Each independent segment timed alone, for the branch assignment of the
    // captured graph.
    B.seg_ms.clear();
    if (batch_branches() > 1) {
        for (size_t i = 0; i + 1 < B.seg.size(); ++i) {
            const std::vector<BatchLaunch> part(B.program.begin() + B.seg[i], B.program.begin() + B.seg[i + 1]);
            B.seg_ms.push_back(part.empty() ? 0.0f : time_launches(part, B.ex.stream, 2));
        }
        if (hipStreamSynchronize(B.ex.stream) != hipSuccess) {
            set_error("batch: segment timing failed");
            return false;
        }
    }
    return true;
}

// Branches of the captured graph: the independent segments (buckets and the
// schoolbook launch) are dealt to kBatchBranches side streams forked from
// the capturing stream, so that small launches and the tails of large ones
// overlap; the reconstructions run after the join, one branch per prime
// count. Streams keep the default priority.
constexpr int kBatchBranches = 4;

// Branch count (NTTAMD_BATCH_BRANCHES=1 captures one chain: tests, A/B).
static int batch_branches() {
    const char* env = getenv("NTTAMD_BATCH_BRANCHES");
    return env ? std::max(1, std::min(atoi(env), 8)) : kBatchBranches;
}

// Side streams and the events of one capture, made before it starts.
static bool prepare_branches(Batch& B, int branches, std::vector<hipEvent_t>* ev) {
    if ((int)B.side.size() != branches) {
        for (hipStream_t x : B.side) hipStreamDestroy(x);
        B.side.clear();
        for (int k = 0; k < branches; ++k) {
            hipStream_t x = nullptr;
            if (hipStreamCreateWithFlags(&x, hipStreamNonBlocking) != hipSuccess) return false;
            B.side.push_back(x);
        }
    }
    // Two fork events and a join event per branch and fork.
    for (int k = 0; k < 2 * (branches + 1); ++k) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return false;
        ev->push_back(e);
    }
    return true;
}

static bool issue_branched(Batch& B, hipStream_t s, const std::vector<hipEvent_t>& ev) {
    const size_t nseg = B.seg.empty() ? 0 : B.seg.size() - 1;
    size_t next = 0;
    // fork: every branch waits for what s has queued so far, runs its
    // launches, and s waits for every branch.
    auto fork_join = [&](const std::vector<std::vector<const BatchLaunch*>>& work) -> bool {
        const hipEvent_t f = ev[next++];
        if (hipEventRecord(f, s) != hipSuccess) return false;
        for (size_t k = 0; k < work.size(); ++k) {
            const hipEvent_t j = ev[next++];
            if (work[k].empty()) continue;
            if (hipStreamWaitEvent(B.side[k], f, 0) != hipSuccess) return false;
            for (const BatchLaunch* L : work[k])
                if (!issue_launch(*L, B.side[k])) return false;
            if (hipEventRecord(j, B.side[k]) != hipSuccess || hipStreamWaitEvent(s, j, 0) != hipSuccess) return false;
        }
        return true;
    };
    // Longest segment first onto the least-loaded branch, by the durations
    // measured at tuning time (round robin in transform-length order when
    // none were measured); the reconstructions likewise, by blocks times
    // primes.
    auto deal = [&](const std::vector<double>& cost, const std::vector<std::pair<size_t, size_t>>& range,
                    std::vector<std::vector<const BatchLaunch*>>& out) {
        std::vector<size_t> order(cost.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
        std::vector<double> load(out.size(), 0.0);
        for (size_t i : order) {
            const size_t k = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
            load[k] += cost[i];
            for (size_t l = range[i].first; l < range[i].second; ++l) out[k].push_back(&B.program[l]);
        }
    };
    std::vector<std::vector<const BatchLaunch*>> work(B.side.size());
    const char* lpt = getenv("NTTAMD_BATCH_LPT");  // 0: round robin (A/B runs)
    if (B.seg_ms.size() == nseg && !(lpt && lpt[0] == '0')) {
        std::vector<double> cost(nseg);
        std::vector<std::pair<size_t, size_t>> range(nseg);
        for (size_t i = 0; i < nseg; ++i) {
            cost[i] = B.seg_ms[i];
            range[i] = {B.seg[i], B.seg[i + 1]};
        }
        deal(cost, range, work);
    } else {
        for (size_t i = 0; i < nseg; ++i)
            for (size_t l = B.seg[i]; l < B.seg[i + 1]; ++l) work[i % work.size()].push_back(&B.program[l]);
    }
    if (!fork_join(work)) return false;
    std::vector<std::vector<const BatchLaunch*>> crt(B.side.size());
    const size_t first = B.seg.empty() ? 0 : B.seg.back();
    std::vector<double> cost;
    std::vector<std::pair<size_t, size_t>> range;
    for (size_t l = first; l < B.program.size(); ++l) {
        const BatchLaunch& L = B.program[l];
        // A reconstruction launch: blocks times primes (the schoolbook one
        // is not among them); its carry launches follow it on its branch.
        if (L.kind == BatchLaunch::kCrt || L.kind == BatchLaunch::kSchool || range.empty()) {
            cost.push_back((double)L.crt.blocks * (L.kind == BatchLaunch::kCrt ? std::max<uint32_t>(1u, L.crt.primes) : 1u));
            range.push_back({l, l + 1});
        } else {
            range.back().second = l + 1;
        }
    }
    deal(cost, range, crt);
    return fork_join(crt);
}

static bool batch_capture(Batch& B) {
    const hipStream_t s = B.ex.stream;
    const int branches = batch_branches();
    std::vector<hipEvent_t> ev;
    if (branches > 1 && !prepare_branches(B, branches, &ev)) {
        for (hipEvent_t e : ev) hipEventDestroy(e);
        set_error("batch: stream or event creation failed");
        return false;
    }
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        for (hipEvent_t e : ev) hipEventDestroy(e);
        set_error("batch: graph capture failed to start");
        return false;
    }
    t_capturing = true;
    const bool ok = branches > 1 ? issue_branched(B, s, ev) : issue_program(B.program, s);
    t_capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    for (hipEvent_t x : ev) hipEventDestroy(x);
    if (!ok || e != hipSuccess || !g) {
        if (g) hipGraphDestroy(g);
        set_error("batch: graph capture failed");
        return false;
    }
    hipGraphExec_t x = nullptr;
    if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
        hipGraphDestroy(g);
        set_error("batch: graph instantiation failed");
        return false;
    }
    B.graph = g;
    B.exec = x;
    return true;
}

// Queues the whole batch on `stream` (the batch's own stream when null):
// autotunes on the first run, captures the program once, then replays it.
bool batch_run(Batch& B, hipStream_t stream, bool use_graph) {
    if (!B.bound) {
        set_error("batch: operands are not bound");
        return false;
    }
    if (B.count == 0) return true;
    if (!B.tuned && !batch_autotune(B)) return false;
    const hipStream_t s = stream ? stream : B.ex.stream;
    if (!use_graph || prof_enabled()) {
        if (!issue_program(B.program, s)) {
            set_error("batch: kernel launch failed");
            return false;
        }
        return true;
    }
    if (!B.exec && !batch_capture(B)) return false;
    if (hipGraphLaunch(B.exec, s) != hipSuccess) {
        set_error("batch: graph launch failed");
        return false;
    }
    return true;
}

}  // namespace nttamd

namespace nttamd {
// Host self-test of the arithmetic the kernels share with the host
// (mulmod, reduce, packing, Garner, carry functions) against exact integer
// arithmetic. Returns the number of failed checks.

namespace {
struct Rng {
    uint64_t s;
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    int64_t range(int64_t lo, int64_t hi) {  // [lo, hi)
        return lo + (int64_t)(next() % (uint64_t)(hi - lo));
    }
};

typedef __int128 i128;

bool check_mod(double r, i128 expect, uint64_t p) {
    if (r != (double)(int64_t)r) return false;
    i128 d = ((i128)(int64_t)r - expect) % (i128)p;
    return d == 0;
}

// X = P - 1 (every residue p_k - 1), the largest value the recombination
// can meet, against P computed by exact multiword products.
template <int NP>
int garner_extreme(const Catalog& c) {
    constexpr int NW = CrtWords<NP>::value;
    uint64_t P[kMaxWords + 2] = {1};
    for (int k = 0; k < NP; ++k) {
        unsigned __int128 carry = 0;
        for (int w = 0; w < kMaxWords + 2; ++w) {
            const unsigned __int128 t = (unsigned __int128)P[w] * c.p[k] + carry;
            P[w] = (uint64_t)t;
            carry = t >> 64;
        }
    }
    for (int w = 0; w < kMaxWords + 2; ++w) {  // P - 1
        if (P[w]--) break;
    }
    double residues[NP];
    for (int k = 0; k < NP; ++k) residues[k] = (double)(c.p[k] - 1);
    CrtArgs a;
    memset(&a, 0, sizeof a);
    a.residues = residues;
    a.row_stride = 1;
    a.field = c.field;
    memcpy(a.pmod, c.pmod, sizeof a.pmod);
    memcpy(a.ginv, c.ginv, sizeof a.ginv);
    memcpy(a.pint, c.p, sizeof a.pint);
    uint64_t got[NW];
    garner<NP>(a, 0, got, false);
    int fails = 0;
    for (int w = 0; w < NW; ++w) fails += got[w] != P[w];
    for (int w = NW; w < kMaxWords + 2; ++w) fails += P[w] != 0;  // P - 1 fits the words
    return fails ? 1 : 0;
}

template <int NP>
int garner_case(const Catalog& c, Rng& rng, bool balanced) {
    constexpr int NW = CrtWords<NP>::value;
    const uint32_t cap = c.capacity[NP];
    uint64_t X[kMaxWords + 1] = {0};
    // Random value below 2^(cap - balanced), then negated half the time when balanced.
    const uint32_t bits = cap - (balanced ? 1 : 0);
    for (uint32_t w = 0; w * 64 < bits; ++w) X[w] = rng.next();
    const uint32_t top = bits % 64;
    if (top) X[(bits - 1) / 64] &= (1ull << top) - 1;
    for (int w = (bits + 63) / 64; w < NW; ++w) X[w] = 0;
    bool neg = balanced && (rng.next() & 1);
    if (neg) {  // two's complement over NW words
        uint64_t carry = 1;
        for (int w = 0; w < NW; ++w) {
            X[w] = ~X[w] + carry;
            carry = carry && X[w] == 0;
        }
    }
    double residues[NP];
    for (int k = 0; k < NP; ++k) {
        // X mod p_k from the words, most significant first.
        const uint64_t pk = c.p[k];
        unsigned __int128 r = 0;
        for (int w = NW - 1; w >= 0; --w) r = ((r << 64) | X[w]) % pk;
        // Signed X: subtract 2^(64 NW) mod p_k.
        if (neg) {
            unsigned __int128 m = 1;
            for (int w = 0; w < NW; ++w) m = (m << 64) % pk;
            r = (r + pk - (uint64_t)m) % pk;
        }
        residues[k] = (double)(uint64_t)r;
    }
    CrtArgs a;
    memset(&a, 0, sizeof a);
    a.residues = residues;
    a.row_stride = 1;
    a.field = c.field;
    memcpy(a.pmod, c.pmod, sizeof a.pmod);
    memcpy(a.ginv, c.ginv, sizeof a.ginv);
    memcpy(a.pint, c.p, sizeof a.pint);
    uint64_t got[NW];
    garner<NP>(a, 0, got, balanced);
    int fails = 0;
    for (int w = 0; w < NW; ++w) fails += got[w] != X[w];
    return fails ? 1 : 0;
}
}  // namespace
}  // namespace nttamd

namespace nttamd {

// ---------------------------------------------------------------------------
// Allocation-free products (ntt_mul_device_noalloc): every device byte the
// call needs beyond its operands and result lives in one file-scope device
// array. Each GPU holds its own copy from the moment the library's code
// object is loaded until the process ends. Its first kNoAllocTableBytes hold
// the forward and inverse root tables of all eight primes (kNoAllocTableLen
// entries each: transforms of up to 2^17 points), generated once per device
// by the first call there; the rest holds one product's spectra and carry
// words. Every word of that working part is written before it is read
// (packing writes whole rows, the reconstruction writes every carry word),
// so nothing has to be cleared between calls.
constexpr uint64_t kNoAllocTableLen = (uint64_t)1 << 16;
constexpr uint64_t kNoAllocTableBytes = 2ull * kMaxPrimes * kNoAllocTableLen * sizeof(double);  // 8 MiB
constexpr uint64_t kNoAllocWorkBytes = 24ull << 20;
constexpr uint64_t kNoAllocScratchBytes = kNoAllocTableBytes + kNoAllocWorkBytes;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) uint64_t scratch[kNoAllocScratchBytes / 8];

// Device address of `scratch` on each device (the host name of the array is
// only a stand-in), taken once per device with that device current, and
// whether that device's root tables are in place.
struct NoAllocDevice {
    uint8_t* base = nullptr;
    bool tables = false;
};
static std::mutex g_noalloc_lock;
static NoAllocDevice g_noalloc[kNoAllocMaxDevices];

static uint64_t align256(uint64_t n) { return (n + 255) & ~(uint64_t)255; }

// Set-up on the calling thread's current device: the scratch address, and
// the root tables generated into it (synchronously, so that calls on any
// stream find them complete).
static bool noalloc_device(hipStream_t s, uint8_t** base) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices) {
        set_error("ntt_mul_device_noalloc: no usable current device");
        return false;
    }
    std::lock_guard<std::mutex> guard(g_noalloc_lock);
    NoAllocDevice& d = g_noalloc[dev];
    if (!d.base) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            set_error("ntt_mul_device_noalloc: the scratch array has no device address");
            return false;
        }
        d.base = static_cast<uint8_t*>(p);
    }
    if (!d.tables) {
        const Catalog& c = cat();
        TableArgs a;
        memset(&a, 0, sizeof a);
        a.fwd = reinterpret_cast<double*>(d.base);
        a.inv = a.fwd + kMaxPrimes * kNoAllocTableLen;
        a.table_stride = kNoAllocTableLen;
        a.len = kNoAllocTableLen;
        a.field = c.field;
        memcpy(a.roots, c.roots, sizeof a.roots);
        memcpy(a.inv_roots, c.inv_roots, sizeof a.inv_roots);
        {
            Timed t("tables", s);
            if (!launch_tables(a, kMaxPrimes, s) || hipStreamSynchronize(s) != hipSuccess) {
                set_error("ntt_mul_device_noalloc: root table generation failed");
                return false;
            }
        }
        d.tables = true;
    }
    *base = d.base;
    return true;
}

}  // namespace nttamd

// ---------------------------------------------------------------------------
// C ABI.

using namespace nttamd;

struct ntt_mul_workspace {
    Executor ex;
    ntt_mul_plan plan;
    bool has_plan = false;
};

struct ntt_conv_workspace {
    Executor ex;
    ntt_conv_plan plan;
    bool has_plan = false;
};

struct ntt_batch {
    Batch batch;
    std::vector<const uint64_t*> d_a, d_b;
    std::vector<uint64_t*> d_out;
    DevBuf staging_in, staging_out;  // host-operand convenience call
    HostBuf host_in, host_out;       // its operands and results, gathered in pinned memory
    ntt_mul_workspace* plain = nullptr;
};

extern "C" {

bool ntt_mul_plan_new(size_t len_a, size_t len_b, ntt_mul_plan* out) {
    if (!out || len_a == 0 || len_b == 0) return false;
    if (len_a + len_b < len_a) return false;
    Geometry g;
    if (!select_product(len_a, len_b, &g)) return false;
    ntt_mul_plan p;
    memset(&p, 0, sizeof p);
    p.len_a = len_a;
    p.len_b = len_b;
    p.primes = g.primes;
    p.chunk_bits = g.width;
    p.log_len = (uint32_t)g.log_t;
    p.row_tail = g.tail;
    p.schoolbook = g.school ? 1 : 0;
    p.valid = 1;
    *out = p;
    return true;
}

size_t ntt_mul_plan_scratch_len(const ntt_mul_plan*) { return 0; }
size_t ntt_mul_plan_sqr_scratch_len(const ntt_mul_plan*) { return 0; }

bool ntt_mul_plan_mul(const ntt_mul_plan* plan, uint64_t* dst, size_t dst_len, const uint64_t* a,
                      size_t a_len, const uint64_t* b, size_t b_len, uint64_t*, size_t) {
    const bool square = a == b && a_len == b_len;
    Executor* ex = plan_executor();
    if (!ex) return false;
    return mul_host(*ex, plan, dst, dst_len, a, a_len, b, b_len, square);
}

bool ntt_mul_plan_sqr(const ntt_mul_plan* plan, uint64_t* dst, size_t dst_len, const uint64_t* a,
                      size_t a_len, uint64_t*, size_t) {
    if (!plan || plan->len_a != plan->len_b) return false;
    Executor* ex = plan_executor();
    if (!ex) return false;
    return mul_host(*ex, plan, dst, dst_len, a, a_len, a, a_len, true);
}

ntt_mul_workspace* ntt_mul_workspace_new(void) { return new (std::nothrow) ntt_mul_workspace(); }

void ntt_mul_workspace_free(ntt_mul_workspace* ws) { delete ws; }

void ntt_mul_workspace_prepare(ntt_mul_workspace* ws, const ntt_mul_plan* plan) {
    if (!ws || !plan || !plan->valid) return;
    ws->plan = *plan;
    ws->has_plan = true;
    Geometry g;
    if (!select_product(plan->len_a, plan->len_b, &g)) return;
    std::lock_guard<std::mutex> guard(ws->ex.lock);
    if (!ws->ex.init()) return;
    if (g.school) {
        ws->ex.carry.ensure(op_carry_bytes(plan->len_a + plan->len_b, true));
        return;
    }
    const double *f, *i;
    uint64_t st;
    ensure_tables(table_need(g), ws->ex.stream, &f, &i, &st);
    ws->ex.spec_a.ensure(op_spectra_bytes(g));
    ws->ex.spec_b.ensure(op_spectra_bytes_b(g));
    // Time the admissible splits of this length on operands of this shape.
    const uint64_t la = plan->len_a, lb = plan->len_b;
    Executor& ex = ws->ex;
    if (ex.in_a.ensure(la * 8) && ex.in_b.ensure(lb * 8) &&
        ex.out.ensure((la + lb) * 8) && hipMemsetAsync(ex.in_a.ptr, 0x5a, la * 8, ex.stream) == hipSuccess &&
        hipMemsetAsync(ex.in_b.ptr, 0xa5, lb * 8, ex.stream) == hipSuccess) {
        tune_split(ex, g.log_t, kUseProduct, [&] {
            return device_product(ex, g, ex.in_a.as<uint64_t>(), la, ex.in_b.as<uint64_t>(), lb, false,
                                  ex.out.as<uint64_t>());
        });
        if (la == lb && !g.blocks)
            tune_split(ex, g.log_t, kUseSquare, [&] {
                return device_product(ex, g, ex.in_a.as<uint64_t>(), la, ex.in_a.as<uint64_t>(), la, true,
                                      ex.out.as<uint64_t>());
            });
    }
}

static bool ws_run(ntt_mul_workspace* ws, uint64_t* dst, size_t dst_len, const uint64_t* a, size_t a_len,
                   const uint64_t* b, size_t b_len, bool square) {
    if (!ws) return false;
    // MulWorkspace::run, src/plan.rs:182-205: an empty operand writes zero.
    if (a_len == 0 || b_len == 0) {
        const uint64_t out_len = (uint64_t)a_len + b_len;
        if (dst_len < out_len || (out_len && !dst)) return false;
        if (out_len) memset(dst, 0, out_len * 8);
        return true;
    }
    if (!ws->has_plan || ws->plan.len_a != a_len || ws->plan.len_b != b_len) {
        ntt_mul_plan p;
        if (!ntt_mul_plan_new(a_len, b_len, &p)) return false;
        ws->plan = p;
        ws->has_plan = true;
    }
    return mul_host(ws->ex, &ws->plan, dst, dst_len, a, a_len, b, b_len, square);
}

bool ntt_mul_workspace_mul(ntt_mul_workspace* ws, uint64_t* dst, size_t dst_len, const uint64_t* a,
                           size_t a_len, const uint64_t* b, size_t b_len) {
    const bool square = a == b && a_len == b_len;
    return ws_run(ws, dst, dst_len, a, a_len, b, b_len, square);
}

bool ntt_mul_workspace_sqr(ntt_mul_workspace* ws, uint64_t* dst, size_t dst_len, const uint64_t* a,
                           size_t a_len) {
    return ws_run(ws, dst, dst_len, a, a_len, a, a_len, true);
}

bool ntt_conv_plan_new(size_t len_a, size_t len_b, uint64_t max_a, uint64_t max_b, bool is_signed,
                       ntt_conv_plan* out) {
    if (!out || len_a == 0 || len_b == 0) return false;
    if (len_a + len_b < len_a) return false;
    const Catalog& c = cat();
    const uint64_t terms = std::min<uint64_t>(len_a, len_b);
    const uint64_t sign = is_signed ? 1 : 0;
    uint64_t bits = 0;
    conv_bound_bits(terms, max_a, max_b, &bits);
    bits += sign;
    const uint64_t stride = std::max<uint64_t>((bits + 63) / 64, 1);
    uint64_t vba = (uint64_t)bitlen(max_a) + sign, vbb = (uint64_t)bitlen(max_b) + sign;
    if (vba == 0) vba = 1;
    if (vbb == 0) vbb = 1;
    if (vba > 64 || vbb > 64) return false;
    int np = 0;
    for (int count = 1; count <= kMaxPrimes; ++count)
        if (c.capacity[count] >= bits) {
            np = count;
            break;
        }
    if (!np) return false;
    const uint64_t conv = (uint64_t)len_a + len_b - 1;
    if (conv > ((uint64_t)1 << kMaxRootOrder)) return false;
    const int K = std::max(kMinLog, log2_ceil(conv));
    if (conv > UINT64_MAX / stride) return false;
    Geometry row;
    row.primes = (uint32_t)np;
    row.conv_len = conv;
    row.digits_a = len_a;
    row.digits_b = len_b;
    choose_row(&row, std::max<uint64_t>(len_a, len_b), K);
    // Unbalanced convolutions split like products (the reference adds the
    // disjoint blocks modulo their width; here the residues of the blocks
    // covering a coefficient are summed before the CRT).
    choose_split(&row, kLaunchPrice);
    // Transforms of at most 2^kMaxLog points (the pass kernels and the
    // per-length split table cover no more); a composite row of 2^kMaxLog
    // cyclic points still serves a slightly longer convolution.
    if (row.log_t > kMaxLog) {
        set_error("convolution longer than the largest transform");
        return false;
    }
    ntt_conv_plan p;
    memset(&p, 0, sizeof p);
    p.len_a = len_a;
    p.len_b = len_b;
    p.max_a = max_a;
    p.max_b = max_b;
    p.stride = stride;
    p.output_len = conv * stride;
    p.is_signed = is_signed ? 1 : 0;
    p.schoolbook = terms <= kSchoolbookMaxTerms ? 1 : 0;
    p.primes = (uint32_t)np;
    p.log_len = (uint32_t)row.log_t;
    p.row_tail = row.tail;
    p.split_blocks = row.blocks;
    p.block_len = row.block_digits;
    p.bound_bits = (uint32_t)bits;
    p.value_bits_a = (uint32_t)vba;
    p.value_bits_b = (uint32_t)vbb;
    p.valid = 1;
    *out = p;
    return true;
}

size_t ntt_conv_plan_stride(const ntt_conv_plan* plan) { return plan ? plan->stride : 0; }
size_t ntt_conv_plan_output_len(const ntt_conv_plan* plan) { return plan ? plan->output_len : 0; }
size_t ntt_conv_plan_scratch_len(const ntt_conv_plan*) { return 0; }

bool ntt_conv_plan_convolve(const ntt_conv_plan* plan, uint64_t* dst, size_t dst_len, const uint64_t* a,
                            size_t a_len, const uint64_t* b, size_t b_len, uint64_t*, size_t) {
    Executor* ex = plan_executor();
    if (!ex) return false;
    return conv_host(*ex, plan, dst, dst_len, a, a_len, b, b_len);
}

ntt_conv_workspace* ntt_conv_workspace_new(void) { return new (std::nothrow) ntt_conv_workspace(); }

void ntt_conv_workspace_free(ntt_conv_workspace* ws) { delete ws; }

void ntt_conv_workspace_prepare(ntt_conv_workspace* ws, const ntt_conv_plan* plan) {
    if (!ws || !plan || !plan->valid) return;
    ws->plan = *plan;
    ws->has_plan = true;
    if (plan->schoolbook) return;
    std::lock_guard<std::mutex> guard(ws->ex.lock);
    if (!ws->ex.init()) return;
    Geometry g;
    if (!conv_geometry(plan, &g)) return;
    const double *f, *i;
    uint64_t st;
    ensure_tables(table_need(g), ws->ex.stream, &f, &i, &st);
    ws->ex.spec_a.ensure(op_spectra_bytes(g));
    ws->ex.spec_b.ensure(op_spectra_bytes_b(g));
    // Time the admissible splits of this length on zero coefficients of this
    // shape (every coefficient within any bound).
    Executor& ex = ws->ex;
    const uint64_t la = plan->len_a, lb = plan->len_b;
    if (ex.in_a.ensure(la * 8) && ex.in_b.ensure(lb * 8) && ex.out.ensure(plan->output_len * 8) &&
        hipMemsetAsync(ex.in_a.ptr, 0, la * 8, ex.stream) == hipSuccess &&
        hipMemsetAsync(ex.in_b.ptr, 0, lb * 8, ex.stream) == hipSuccess) {
        OpBuffers buf;
        if (exec_buffers(ex, g, false, 0, &buf)) {
            tune_split(ex, g.log_t, kUseProduct, [&] {
                return convolution_on(ex.stream, g, plan, ex.in_a.as<uint64_t>(), ex.in_b.as<uint64_t>(), false, buf,
                                      ex.out.as<uint64_t>());
            });
            if (la == lb && !g.blocks)
                tune_split(ex, g.log_t, kUseSquare, [&] {
                    return convolution_on(ex.stream, g, plan, ex.in_a.as<uint64_t>(), ex.in_a.as<uint64_t>(), true,
                                          buf, ex.out.as<uint64_t>());
                });
        }
    }
}

bool ntt_conv_workspace_convolve(ntt_conv_workspace* ws, uint64_t* dst, size_t dst_len, const uint64_t* a,
                                 size_t a_len, const uint64_t* b, size_t b_len) {
    if (!ws || !ws->has_plan) return false;
    return conv_host(ws->ex, &ws->plan, dst, dst_len, a, a_len, b, b_len);
}

// Device-resident operands.

bool ntt_mul_workspace_mul_device(ntt_mul_workspace* ws, uint64_t* d_dst, size_t dst_len, const uint64_t* d_a,
                                  size_t a_len, const uint64_t* d_b, size_t b_len, void* stream) {
    t_error.clear();
    if (!ws || a_len == 0 || b_len == 0 || !d_dst || !d_a || !d_b) return false;
    const bool square = d_a == d_b && a_len == b_len;
    if (dst_len < (uint64_t)a_len + b_len) return false;
    if (!ws->has_plan || ws->plan.len_a != a_len || ws->plan.len_b != b_len) {
        ntt_mul_plan p;
        if (!ntt_mul_plan_new(a_len, b_len, &p)) return false;
        ws->plan = p;
        ws->has_plan = true;
    }
    Geometry g;
    if (!product_geometry(&ws->plan, &g)) return false;
    std::lock_guard<std::mutex> guard(ws->ex.lock);
    if (!ws->ex.init()) return false;
    hipStream_t saved = ws->ex.stream;
    if (stream) ws->ex.stream = (hipStream_t)stream;
    const bool ok = device_product(ws->ex, g, d_a, a_len, d_b, b_len, square, d_dst);
    ws->ex.stream = saved;
    return ok;
}

bool ntt_mul_workspace_sqr_device(ntt_mul_workspace* ws, uint64_t* d_dst, size_t dst_len, const uint64_t* d_a,
                                  size_t a_len, void* stream) {
    return ntt_mul_workspace_mul_device(ws, d_dst, dst_len, d_a, a_len, d_a, a_len, stream);
}

bool ntt_conv_workspace_convolve_device(ntt_conv_workspace* ws, uint64_t* d_dst, size_t dst_len,
                                        const uint64_t* d_a, size_t a_len, const uint64_t* d_b, size_t b_len,
                                        void* stream) {
    t_error.clear();
    if (!ws || !ws->has_plan || !d_dst || !d_a || !d_b) return false;
    const ntt_conv_plan* p = &ws->plan;
    if (a_len != p->len_a || b_len != p->len_b || dst_len < p->output_len) return false;
    Geometry g;
    if (!conv_geometry(p, &g)) return false;
    std::lock_guard<std::mutex> guard(ws->ex.lock);
    if (!ws->ex.init()) return false;
    hipStream_t saved = ws->ex.stream;
    if (stream) ws->ex.stream = (hipStream_t)stream;
    ConvCtx ctx{&g, p, d_a == d_b && a_len == b_len && !p->schoolbook};
    const bool ok = run_conv(ws->ex, &ctx, d_a, d_b, d_dst);
    ws->ex.stream = saved;
    return ok;
}

bool ntt_mul_device_noalloc(uint64_t* d_dst, size_t dst_len, const uint64_t* d_a, size_t a_len,
                            const uint64_t* d_b, size_t b_len, void* stream) {
    t_error.clear();
    if (!d_dst || !d_a || !d_b || a_len == 0 || b_len == 0) return false;
    if ((uint64_t)a_len + b_len < a_len || dst_len < (uint64_t)a_len + b_len) return false;
    const bool square = d_a == d_b && a_len == b_len;
    Geometry g;
    if (!select_product(a_len, b_len, &g)) return false;
    const uint64_t spec_a = g.school ? 0 : align256(op_spectra_bytes(g));
    const uint64_t spec_b = g.school || square ? 0 : align256(op_spectra_bytes_b(g));
    const uint64_t carry = align256(op_carry_bytes((uint64_t)a_len + b_len, g.school));
    if ((!g.school && table_need(g) > kNoAllocTableLen) || spec_a + spec_b + carry > kNoAllocWorkBytes) {
        set_error("ntt_mul_device_noalloc: a product of " + std::to_string(a_len) + " x " + std::to_string(b_len) +
                  " limbs needs " + std::to_string(spec_a + spec_b + carry) + " bytes of working memory and " +
                  std::to_string(g.school ? 0 : table_need(g)) + " root-table entries; the fixed scratch holds " +
                  std::to_string(kNoAllocWorkBytes) + " bytes and " + std::to_string(kNoAllocTableLen) + " entries");
        return false;
    }
    const hipStream_t s = (hipStream_t)stream;
    uint8_t* base = nullptr;
    if (!noalloc_device(s, &base)) return false;
    const double* tables[2] = {reinterpret_cast<const double*>(base),
                               reinterpret_cast<const double*>(base) + kMaxPrimes * kNoAllocTableLen};
    uint8_t* work = base + kNoAllocTableBytes;
    OpBuffers buf;
    buf.spec_a = reinterpret_cast<double*>(work);
    buf.spec_b = square ? buf.spec_a : reinterpret_cast<double*>(work + spec_a);
    buf.carry = reinterpret_cast<uint32_t*>(work + spec_a + spec_b);
    return product_with(s, g, d_a, a_len, d_b, b_len, square, buf, tables, kNoAllocTableLen, d_dst);
}

// Batched products.

ntt_batch* ntt_batch_new(void) { return new (std::nothrow) ntt_batch(); }

void ntt_batch_free(ntt_batch* b) {
    if (!b) return;
    if (b->plain) ntt_mul_workspace_free(b->plain);
    b->staging_in.release();
    b->staging_out.release();
    b->host_in.release();
    b->host_out.release();
    delete b;
}

bool ntt_batch_plan(ntt_batch* b, size_t count, const size_t* len_a, const size_t* len_b, const uint8_t* square) {
    t_error.clear();
    if (!b || (count && (!len_a || (!len_b && !square)))) return false;
    if (!len_b)
        for (size_t i = 0; i < count; ++i)
            if (!square[i]) return false;
    std::vector<size_t> lb(count);
    for (size_t i = 0; i < count; ++i) lb[i] = len_b ? len_b[i] : len_a[i];
    std::lock_guard<std::mutex> guard(b->batch.ex.lock);
    b->d_a.clear();
    return batch_plan(b->batch, count, len_a, lb.data(), square);
}

size_t ntt_batch_count(const ntt_batch* b) { return b ? b->batch.count : 0; }

size_t ntt_batch_output_len(const ntt_batch* b, size_t i) {
    if (!b || i >= b->batch.count) return 0;
    return b->batch.la[i] + b->batch.lb[i];
}

bool ntt_batch_bind(ntt_batch* b, const uint64_t* const* d_a, const uint64_t* const* d_b, uint64_t* const* d_out) {
    t_error.clear();
    if (!b || (b->batch.count && (!d_a || !d_out))) return false;
    std::lock_guard<std::mutex> guard(b->batch.ex.lock);
    const size_t n = b->batch.count;
    std::vector<const uint64_t*> bb(n);
    for (size_t i = 0; i < n; ++i) {
        bb[i] = b->batch.square[i] ? d_a[i] : (d_b ? d_b[i] : nullptr);
        if (!bb[i]) {
            set_error("batch: missing second operand at product " + std::to_string(i));
            return false;
        }
    }
    if (!batch_bind(b->batch, d_a, bb.data(), d_out)) return false;
    b->d_a.assign(d_a, d_a + n);
    b->d_b = bb;
    b->d_out.assign(d_out, d_out + n);
    return true;
}

bool ntt_batch_run(ntt_batch* b, void* stream) {
    t_error.clear();
    if (!b) return false;
    std::lock_guard<std::mutex> guard(b->batch.ex.lock);
    return batch_run(b->batch, (hipStream_t)stream, true);
}

bool ntt_batch_run_plain(ntt_batch* b, void* stream) {
    t_error.clear();
    if (!b || !b->batch.bound) return false;
    if (!b->plain) b->plain = ntt_mul_workspace_new();
    if (!b->plain) return false;
    for (size_t i = 0; i < b->d_a.size(); ++i) {
        const uint64_t la = b->batch.la[i], lb = b->batch.lb[i];
        const bool ok = b->batch.square[i]
                            ? ntt_mul_workspace_sqr_device(b->plain, b->d_out[i], la + lb, b->d_a[i], la, stream)
                            : ntt_mul_workspace_mul_device(b->plain, b->d_out[i], la + lb, b->d_a[i], la,
                                                           b->d_b[i], lb, stream);
        if (!ok) return false;
    }
    return true;
}

int ntt_batch_launches(const ntt_batch* b) { return b ? (int)b->batch.program.size() : 0; }

bool ntt_batch_mul_host(ntt_batch* b, const uint64_t* const* a, const uint64_t* const* bops, uint64_t* const* out) {
    t_error.clear();
    if (!b) return false;
    const size_t n = b->batch.count;
    if (n == 0) return true;
    if (!a || !out) return false;
    uint64_t in_limbs = 0, out_limbs = 0;
    for (size_t i = 0; i < n; ++i) {
        if (!a[i] || !out[i] || (!b->batch.square[i] && (!bops || !bops[i]))) return false;
        in_limbs += b->batch.la[i] + (b->batch.square[i] ? 0 : b->batch.lb[i]);
        out_limbs += b->batch.la[i] + b->batch.lb[i];
    }
    if (!b->staging_in.ensure(in_limbs * 8) || !b->staging_out.ensure(out_limbs * 8)) {
        set_error("batch: staging allocation failed");
        return false;
    }
    if (!b->host_in.ensure(in_limbs * 8) || !b->host_out.ensure(out_limbs * 8)) {
        set_error("batch: pinned staging allocation failed");
        return false;
    }
    std::vector<const uint64_t*> da(n), db(n);
    std::vector<uint64_t*> dout(n);
    uint64_t* host_in = b->host_in.as<uint64_t>();
    uint64_t* din = b->staging_in.as<uint64_t>();
    uint64_t* dst = b->staging_out.as<uint64_t>();
    uint64_t ofs = 0, oofs = 0;
    for (size_t i = 0; i < n; ++i) {
        const uint64_t la = b->batch.la[i], lb = b->batch.lb[i];
        memcpy(&host_in[ofs], a[i], la * 8);
        da[i] = din + ofs;
        ofs += la;
        if (b->batch.square[i]) {
            db[i] = da[i];
        } else {
            memcpy(&host_in[ofs], bops[i], lb * 8);
            db[i] = din + ofs;
            ofs += lb;
        }
        dout[i] = dst + oofs;
        oofs += la + lb;
    }
    if (hipMemcpy(din, host_in, in_limbs * 8, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("batch: operand upload failed");
        return false;
    }
    if (!ntt_batch_bind(b, da.data(), db.data(), dout.data())) return false;
    if (!ntt_batch_run(b, nullptr)) return false;
    uint64_t* host_out = b->host_out.as<uint64_t>();
    if (hipStreamSynchronize(b->batch.ex.stream) != hipSuccess ||
        hipMemcpy(host_out, dst, out_limbs * 8, hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("batch: device execution failed");
        return false;
    }
    oofs = 0;
    for (size_t i = 0; i < n; ++i) {
        const uint64_t len = b->batch.la[i] + b->batch.lb[i];
        memcpy(out[i], &host_out[oofs], len * 8);
        oofs += len;
    }
    return true;
}

// Utilities.

int ntt_device_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
    return strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

void* ntt_device_alloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return nullptr;
    return p;
}

void ntt_device_free(void* ptr) {
    if (ptr) hipFree(ptr);
}

bool ntt_copy_to_device(void* d_dst, const void* src, size_t bytes) {
    return hipMemcpy(d_dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
}

bool ntt_copy_to_host(void* dst, const void* d_src, size_t bytes) {
    return hipMemcpy(dst, d_src, bytes, hipMemcpyDeviceToHost) == hipSuccess;
}

bool ntt_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess; }

const char* ntt_last_error(void) { return t_error.c_str(); }

void ntt_profile_enable(int on) { g_prof_on.store(on ? 1 : 0); }

void ntt_profile_reset(void) {
    prof_drain();
    std::lock_guard<std::mutex> g(g_prof_lock);
    g_prof_totals.clear();
    g_prof_last.clear();
}

bool ntt_profile_replay(const char* kind, int reps, double* avg_ms) {
    if (!kind || reps < 1 || !avg_ms) return false;
    ProfLaunch l;
    {
        std::lock_guard<std::mutex> g(g_prof_lock);
        bool found = false;
        for (auto& x : g_prof_last)
            if (x.kind == kind) {
                l = x;
                found = true;
            }
        if (!found) return false;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) hipEventDestroy(e0);
        return false;
    }
    void* args[1] = {l.arg.data()};
    bool ok = hipLaunchKernel(l.fn, l.grid, l.block, args, 0, l.stream) == hipSuccess &&  // warm
              hipEventRecord(e0, l.stream) == hipSuccess;
    for (int r = 0; r < reps && ok; ++r) ok = hipLaunchKernel(l.fn, l.grid, l.block, args, 0, l.stream) == hipSuccess;
    ok = ok && hipEventRecord(e1, l.stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
    float ms = 0;
    if (ok) ok = hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (ok) *avg_ms = ms / reps;
    return ok;
}

int ntt_transform_split(uint32_t log_len, int use, uint32_t* pass_logs, uint32_t* flags) {
    if (log_len < (uint32_t)kMinLog || log_len > (uint32_t)kMaxLog || use < 0 || use > 2) return 0;
    const Split sp = split_of((int)log_len, use);
    for (int i = 0; i < sp.n && pass_logs; ++i) pass_logs[i] = (uint32_t)sp.logs[i];
    if (flags) {
        bool tuned;
        {
            std::lock_guard<std::mutex> g(g_tune_lock);
            tuned = g_tuned[use][log_len].n != 0;
        }
        *flags = (sp.alt_tiles ? kFormAltTiles : 0u) | (sp.inv16 ? kFormInv16 : 0u) |
                 (sp.per_prime_pack ? kFormPerPrimePack : 0u) | (sp.rows8 ? kFormRows8 : 0u) |
                 (sp.strided8 ? kFormStrided8 : 0u) | (sp.paired ? kFormPaired : 0u) |
                 (sp.waves ? kFormWaves : 0u) |
                 (tuned ? kFormTuned : 0u);
    }
    return sp.n;
}

int ntt_profile_read(char* names, size_t names_len, double* ms, uint64_t* counts, int cap) {
    prof_drain();
    std::lock_guard<std::mutex> g(g_prof_lock);
    int n = 0;
    size_t off = 0;
    for (auto& t : g_prof_totals) {
        if (n >= cap) break;
        const size_t len = t.first.size() + 1;
        if (names && off + len <= names_len) {
            memcpy(names + off, t.first.c_str(), len);
            off += len;
        }
        if (ms) ms[n] = t.second.first;
        if (counts) counts[n] = t.second.second;
        ++n;
    }
    return n;
}


// Diagnostics: intermediate stages of the pipeline, for tests.

int ntt_selftest_host(uint64_t seed) {
    const Catalog& c = cat();
    Rng rng{seed};
    int fails = 0;
    for (int y = 0; y < kMaxPrimes; ++y) {
        const uint64_t p = c.p[y];
        const double pf = c.field.p[y], pinv = c.field.pinv[y];
        const int64_t P = (int64_t)p;
        for (int i = 0; i < 4000; ++i) {
            // mulmod: |a| < 8p against a centered factor, and both operands below 2p.
            const int64_t a = i & 1 ? rng.range(-8 * P + 1, 8 * P) : rng.range(-2 * P + 1, 2 * P);
            const int64_t b = i & 1 ? rng.range(-P / 2, P / 2 + 1) : rng.range(-2 * P + 1, 2 * P);
            const i128 prod = (i128)a * b;
            const i128 lim = (i128)4 * P * P;
            if (prod >= lim || -prod >= lim) continue;
            const double r = mulmod((double)a, (double)b, pf, pinv);
            const bool small = prod < (i128)2 * P * P && -prod < (i128)2 * P * P;
            const double bound = small ? 15.0 * pf / 16.0 : 11.0 * pf / 8.0;
            if (!check_mod(r, prod, p) || fabs(r) >= bound) ++fails;
            const int64_t x = rng.range(-8 * P + 1, 8 * P);
            const double q = reduce((double)x, pf, pinv);
            if (!check_mod(q, x, p) || fabs(q) >= 9.0 * pf / 16.0) ++fails;
        }
        // Packing of every chunk width shape against exact digits.
        static const uint32_t widths[] = {22, 49, 50, 64, 69, 98, 99, 139, 143, 147, 191, 192};
        uint64_t limbs[24];
        int interior = 0;
        for (uint32_t width : widths) {
            for (int t = 0; t < 3; ++t) {
                for (auto& l : limbs) l = t == 2 ? ~0ull : rng.next();
                if (t == 1) limbs[23] = 0;
                Source s;
                s.limbs = limbs;
                s.limb_len = 24;
                s.digits = (24 * 64 + width - 1) / width;
                s.width = width;
                s.value_bits = width;
                s.is_signed = 0;
                s.pad = 0;
                for (uint64_t n = 0; n < s.digits + 2; ++n) {
                    // Exact digit n mod p from the bits.
                    unsigned __int128 r = 0;
                    for (int bit = (int)width - 1; bit >= 0; --bit) {
                        const uint64_t pos = n * width + bit;
                        const uint64_t v = pos / 64 < 24 ? (limbs[pos / 64] >> (pos % 64)) & 1 : 0;
                        r = ((r << 1) | v) % p;
                    }
                    if (n >= s.digits) r = 0;
                    const double got = pack_coef(s, n, c.powers[y], pf, pinv);
                    if (!check_mod(got, (i128)r, p) || fabs(got) >= pf) ++fails;
                    // The unclamped form, wherever the tile test admits it.
                    const int kind = pack_kind(s);
                    double fast = got;
                    bool admitted = false;
                    switch (kind) {
#define NTT_FAST_CASE(K)                                                  \
    case K:                                                               \
        if ((admitted = pack_interior<K>(s, n + 1)))                      \
            fast = pack_as<K, false>(s, n, c.powers[y], pf, pinv);        \
        break;
                        NTT_FAST_CASE(kPackW64)
                        NTT_FAST_CASE(kPack1)
                        NTT_FAST_CASE(kPack2)
                        NTT_FAST_CASE(kPack3)
                        NTT_FAST_CASE(kPack4)
#undef NTT_FAST_CASE
                        default: break;
                    }
                    if (fast != got) ++fails;
                    if (admitted) ++interior;
                    // Parts extracted once and combined per prime (the shared
                    // packing pass), clamped and, where admitted, unclamped.
                    {
                        double dd[kMaxParts] = {0, 0, 0, 0};
                        double shared = 0.0, shared_fast = 0.0;
                        bool fast_ok = false;
                        switch (kind) {
#define NTT_SHARED_CASE(K)                                                        \
    case K:                                                                       \
        extract_as<K, true>(s, n, dd);                                            \
        shared = combine_as<K>(s, dd, c.powers[y], pf, pinv);                     \
        if ((fast_ok = pack_interior<K>(s, n + 1))) {                             \
            extract_as<K, false>(s, n, dd);                                       \
            shared_fast = combine_as<K>(s, dd, c.powers[y], pf, pinv);            \
        }                                                                         \
        break;
                            NTT_SHARED_CASE(kPackW64)
                            NTT_SHARED_CASE(kPack1)
                            NTT_SHARED_CASE(kPack2)
                            NTT_SHARED_CASE(kPack3)
#undef NTT_SHARED_CASE
                            default: shared = got; break;
                        }
                        if (!check_mod(shared, (i128)r, p) || fabs(shared) >= pf) ++fails;
                        if (fast_ok && shared_fast != shared) ++fails;
                    }
                    // Parts extracted once and combined per prime (stacked
                    // packing tiles) give the same residue.
                    if (kind >= kPack1) {
                        double d[kMaxParts];
                        double split = 0.0;
                        switch (kind) {
#define NTT_PARTS_CASE(P)                                                  \
    case kPack1 + P - 1:                                                   \
        extract_parts<P>(s, n, d);                                         \
        split = combine_parts<P>(d, c.powers[y], pf, pinv,                 \
                                 P > 1 && short_high_part((int)s.value_bits - kPartBits * (P - 1))); \
        break;
                            NTT_PARTS_CASE(1)
                            NTT_PARTS_CASE(2)
                            NTT_PARTS_CASE(3)
                            NTT_PARTS_CASE(4)
#undef NTT_PARTS_CASE
                        }
                        if (split != got) ++fails;
                    }
                }
            }
        }
        if (interior == 0) ++fails;  // the unclamped form must have been exercised
        // Signed whole-limb coefficients of 1..64 bits with their sign.
        for (uint32_t vb : {1u, 2u, 16u, 33u, 49u, 50u, 63u, 64u}) {
            for (int t = 0; t < 24; ++t) {
                const int64_t lim = vb >= 64 ? INT64_MAX : (int64_t)1 << (vb - 1);
                int64_t v = (int64_t)rng.next();
                if (vb < 64) v = rng.range(-lim, lim);
                limbs[t] = (uint64_t)v;
            }
            Source s;
            s.limbs = limbs;
            s.limb_len = 24;
            s.digits = 24;
            s.width = 64;
            s.value_bits = vb;
            s.is_signed = 1;
            s.pad = 0;
            for (uint64_t n = 0; n < 24; ++n) {
                const double got = pack_coef(s, n, c.powers[y], pf, pinv);
                if (!check_mod(got, (i128)(int64_t)limbs[n], p) || fabs(got) >= pf) ++fails;
                if (pack_interior<kPackW64Signed>(s, n + 1) &&
                    pack_as<kPackW64Signed, false>(s, n, c.powers[y], pf, pinv) != got)
                    ++fails;
                double dd[2];
                extract_as<kPackW64Signed, true>(s, n, dd);
                const double shared = combine_as<kPackW64Signed>(s, dd, c.powers[y], pf, pinv);
                if (!check_mod(shared, (i128)(int64_t)limbs[n], p) || fabs(shared) >= pf) ++fails;
            }
        }
    }
    fails += garner_extreme<1>(c) + garner_extreme<2>(c) + garner_extreme<3>(c) + garner_extreme<4>(c) +
             garner_extreme<5>(c) + garner_extreme<6>(c) + garner_extreme<7>(c) + garner_extreme<8>(c);
    for (int t = 0; t < 200; ++t) {
        fails += garner_case<1>(c, rng, false) + garner_case<2>(c, rng, false) + garner_case<3>(c, rng, false) +
                 garner_case<4>(c, rng, false) + garner_case<5>(c, rng, false) + garner_case<6>(c, rng, false) +
                 garner_case<7>(c, rng, false) + garner_case<8>(c, rng, false);
        fails += garner_case<1>(c, rng, true) + garner_case<2>(c, rng, true) + garner_case<3>(c, rng, true) +
                 garner_case<4>(c, rng, true);
    }
    // Carry functions: compose(f1, f2) is f2 after f1 for every small carry.
    for (int t = 0; t < 20000; ++t) {
        CarryFn f1{(uint32_t)(rng.next() % 5), (uint32_t)(rng.next() % 7 + 1)};
        CarryFn f2{(uint32_t)(rng.next() % 5), (uint32_t)(rng.next() % 7 + 1)};
        if (rng.next() % 4 == 0) f1.thresh = kInf;
        if (rng.next() % 4 == 0) f2.thresh = kInf;
        const CarryFn g = compose(f1, f2);
        for (uint32_t cin = 0; cin < 8; ++cin) {
            const uint32_t mid = f1.base + (cin >= f1.thresh ? 1u : 0u);
            const uint32_t out = f2.base + (mid >= f2.thresh ? 1u : 0u);
            if (out != g.base + (cin >= g.thresh ? 1u : 0u)) ++fails;
        }
    }
    // Runs of limbs (sum v mod 2^64 and wrap count w each, many of them
    // all-ones): the functions c -> w + [c >= limb_thresh(v)] composed in
    // order, and composed as contiguous slices whose results are composed
    // in order (the carry kernels' two forms), equal adding the carry limb
    // by limb.
    for (int t = 0; t < 20000; ++t) {
        uint64_t v[12];
        uint32_t w[12];
        const int len = 1 + (int)(rng.next() % 12);
        for (int i = 0; i < len; ++i) {
            const uint64_t r = rng.next() % 5;
            v[i] = r == 0 ? rng.next() : r == 1 ? 0 : ~0ull - (rng.next() % 3);
            w[i] = (uint32_t)(rng.next() % 3);
        }
        CarryFn whole{w[0], limb_thresh(v[0])};
        for (int i = 1; i < len; ++i) whole = compose(whole, CarryFn{w[i], limb_thresh(v[i])});
        const int per = 1 + (int)(rng.next() % 4);
        CarryFn sliced{0u, kInf};
        for (int s0 = 0; s0 < len; s0 += per) {
            CarryFn g{w[s0], limb_thresh(v[s0])};
            for (int i = s0 + 1; i < len && i < s0 + per; ++i) g = compose(g, CarryFn{w[i], limb_thresh(v[i])});
            sliced = s0 == 0 ? g : compose(sliced, g);
        }
        for (uint32_t cin = 0; cin < 4; ++cin) {
            uint32_t carry = cin;
            for (int i = 0; i < len; ++i) {
                const uint64_t sum = v[i] + carry;
                carry = w[i] + (sum < v[i] ? 1u : 0u);
            }
            if (whole.base + (cin >= whole.thresh ? 1u : 0u) != carry) ++fails;
            if (sliced.base + (cin >= sliced.thresh ? 1u : 0u) != carry) ++fails;
        }
    }
    return fails;
}

bool ntt_debug_copy(void* d_dst, const void* d_src, uint64_t bytes) {
    if (bytes % 16) return false;
    if (!launch_copy(d_src, d_dst, bytes, nullptr)) return false;
    return hipDeviceSynchronize() == hipSuccess;
}

bool ntt_debug_tables(uint64_t len, double* fwd, double* inv) {
    Executor* ex = plan_executor();
    if (!ex) return false;
    std::lock_guard<std::mutex> guard(ex->lock);
    if (!ex->init()) return false;
    const double *f = nullptr, *i = nullptr;
    uint64_t stride = 0;
    if (!ensure_tables(len, ex->stream, &f, &i, &stride)) return false;
    for (int y = 0; y < kMaxPrimes; ++y) {
        if (hipMemcpy(fwd + (uint64_t)y * len, f + (uint64_t)y * stride, len * 8, hipMemcpyDeviceToHost) != hipSuccess)
            return false;
        if (hipMemcpy(inv + (uint64_t)y * len, i + (uint64_t)y * stride, len * 8, hipMemcpyDeviceToHost) != hipSuccess)
            return false;
    }
    return true;
}

bool ntt_debug_residues(const uint64_t* a, size_t la, const uint64_t* b, size_t lb, int square, double* out,
                        size_t out_len, uint32_t* geometry) {
    Geometry g;
    if (!select_product(la, lb, &g)) return false;
    if (g.blocks) {
        set_error("ntt_debug_residues: the shape runs as a split product (set NTTAMD_SPLIT=0)");
        return false;
    }
    g.school = false;  // the residues of the transform path, whatever a product of this shape runs
    const uint64_t T = row_len(g);
    if (geometry) {
        geometry[0] = g.primes;
        geometry[1] = g.width;
        geometry[2] = (uint32_t)g.log_t;
        geometry[3] = g.tail;
    }
    if (!out) return true;
    if (out_len < g.primes * T) return false;
    Executor* ex = plan_executor();
    if (!ex) return false;
    std::lock_guard<std::mutex> guard(ex->lock);
    if (!ex->init() || !ex->in_a.ensure(la * 8) || !ex->in_b.ensure(lb * 8)) return false;
    if (hipMemcpy(ex->in_a.ptr, a, la * 8, hipMemcpyHostToDevice) != hipSuccess) return false;
    if (hipMemcpy(ex->in_b.ptr, b, lb * 8, hipMemcpyHostToDevice) != hipSuccess) return false;
    OpBuffers buf;
    if (!exec_buffers(*ex, g, square != 0, 0, &buf)) return false;
    double* rows = nullptr;
    if (!residues(ex->stream, g, ex->in_a.as<uint64_t>(), la, ex->in_b.as<uint64_t>(), lb, square != 0, buf, &rows))
        return false;
    if (hipStreamSynchronize(ex->stream) != hipSuccess) return false;
    return hipMemcpy(out, rows, g.primes * T * 8, hipMemcpyDeviceToHost) == hipSuccess;
}

bool ntt_debug_crt(size_t la, size_t lb, const double* rows, uint64_t* out) {
    Geometry g;
    if (!select_product(la, lb, &g) || g.blocks) return false;
    g.school = false;
    const uint64_t T = row_len(g);
    Executor* ex = plan_executor();
    if (!ex) return false;
    std::lock_guard<std::mutex> guard(ex->lock);
    const uint64_t out_len = la + lb;
    if (!ex->init() || !ex->spec_b.ensure(g.primes * T * 8) || !ex->out.ensure(out_len * 8)) return false;
    if (hipMemcpy(ex->spec_b.ptr, rows, g.primes * T * 8, hipMemcpyHostToDevice) != hipSuccess) return false;
    if (!ex->carry.ensure(op_carry_bytes(out_len))) return false;
    if (!crt_product_from(ex->stream, g, ex->spec_b.as<double>(), out_len, ex->carry.as<uint32_t>(),
                          ex->out.as<uint64_t>()))
        return false;
    if (hipStreamSynchronize(ex->stream) != hipSuccess) return false;
    return hipMemcpy(out, ex->out.ptr, out_len * 8, hipMemcpyDeviceToHost) == hipSuccess;
}

bool ntt_row_shape(uint64_t conv_len, uint64_t longest, uint64_t* row, uint64_t* cyclic, uint64_t* table) {
    if (conv_len == 0 || conv_len > ((uint64_t)1 << 40)) return false;
    const int K = std::max(kMinLog, log2_ceil(conv_len));
    Geometry g;
    g.log_t = K;
    if (const uint32_t t = composite_tail(conv_len, longest, K)) {
        g.log_t = K - 1;
        g.tail = t;
    }
    if (row) *row = row_len(g);
    if (cyclic) *cyclic = (uint64_t)1 << g.log_t;
    if (table) *table = table_need(g);
    return true;
}

bool ntt_mul_geometry(size_t len_a, size_t len_b, uint32_t* primes, uint32_t* chunk_bits, uint32_t* log_len,
                      uint32_t* row_tail) {
    Geometry g;
    if (!select_product(len_a, len_b, &g)) return false;
    if (primes) *primes = g.primes;
    if (chunk_bits) *chunk_bits = g.width;
    if (log_len) *log_len = (uint32_t)g.log_t;
    if (row_tail) *row_tail = g.tail;
    return true;
}

bool ntt_mul_split(size_t len_a, size_t len_b, uint32_t* blocks, uint64_t* block_digits) {
    Geometry g;
    if (!select_product(len_a, len_b, &g)) return false;
    if (blocks) *blocks = g.blocks;
    if (block_digits) *block_digits = g.block_digits;
    return true;
}

int ntt_mul_schoolbook(size_t len_a, size_t len_b) {
    Geometry g;
    if (len_a == 0 || len_b == 0 || !select_product(len_a, len_b, &g)) return 0;
    return g.school ? 1 : 0;
}

}  // extern "C"
