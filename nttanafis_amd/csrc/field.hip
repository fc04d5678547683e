// Prime fields 2^49 < p < 2^50 and two-prime Chinese remaindering, the public
// scalar API of the reference (src/field/prime_field.rs:91-442), for host
// callers of the library. Host code only: every operation is exact 64/128-bit
// integer arithmetic on canonical residues in [0, p), so its values are the
// reference's whatever method computes them.
#include <stddef.h>
#include <stdint.h>

#include "../../include/nttanafis.h"

namespace {

using u128 = unsigned __int128;
using i128 = __int128;

constexpr uint64_t kMinModulus = 1ull << 49;  // exclusive bounds (src/field/catalog.rs:19-21)
constexpr uint64_t kMaxModulus = 1ull << 50;
constexpr uint64_t kWitnesses[9] = {2, 3, 5, 7, 11, 13, 17, 19, 23};  // src/field/catalog.rs:25
constexpr uint64_t kCatalog[8] = {1108307720798209ULL, 1086317488242689ULL, 910395627798529ULL,
                                  659706976665601ULL,  699289395265537ULL,  1022545813831681ULL,
                                  1013749720809473ULL, 868614185943041ULL};

uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)a * b % m); }

uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    b %= m;
    while (e) {
        if (e & 1) r = mulmod(r, b, m);
        b = mulmod(b, b, m);
        e >>= 1;
    }
    return r;
}

// Strong probable prime to the nine first prime bases: exact below 3.8e18
// (Jiang and Deng 2014), so exact for every candidate below 2^50.
bool is_prime(uint64_t n) {
    if (n < 2) return false;
    for (uint64_t w : kWitnesses)
        if (n % w == 0) return n == w;
    uint64_t d = n - 1;
    int s = 0;
    while (!(d & 1)) {
        d >>= 1;
        ++s;
    }
    for (uint64_t w : kWitnesses) {
        uint64_t x = powmod(w, d, n);
        if (x == 1 || x == n - 1) continue;
        bool composite = true;
        for (int i = 1; i < s && composite; ++i) {
            x = mulmod(x, x, n);
            if (x == n - 1) composite = false;
        }
        if (composite) return false;
    }
    return true;
}

ntt_prime_field field_of(uint64_t m) {
    ntt_prime_field f;
    f.modulus = m;
    f.reciprocal = 1.0 / (double)m;
    return f;
}

}  // namespace

extern "C" {

bool ntt_field_new(uint64_t modulus, ntt_prime_field* out) {
    if (!out || modulus <= kMinModulus || modulus >= kMaxModulus || !is_prime(modulus)) return false;
    *out = field_of(modulus);
    return true;
}

bool ntt_field_below(uint64_t bound, ntt_prime_field* out) {
    if (!out) return false;
    const uint64_t top = bound < kMaxModulus ? bound : kMaxModulus;
    if (top < 2) return false;
    for (uint64_t c = (top - 2) | 1; c > kMinModulus; c -= 2)
        if (is_prime(c)) {
            *out = field_of(c);
            return true;
        }
    return false;
}

size_t ntt_field_catalog(ntt_prime_field* out, size_t cap) {
    for (size_t i = 0; i < 8 && out && i < cap; ++i) out[i] = field_of(kCatalog[i]);
    return 8;
}

bool ntt_field_residue(const ntt_prime_field* f, uint64_t value, uint64_t* out) {
    if (!f || !out || value >= f->modulus) return false;
    *out = value;
    return true;
}

uint64_t ntt_field_reduce(const ntt_prime_field* f, int64_t value) {
    const int64_t m = (int64_t)f->modulus;
    int64_t r = value % m;
    return (uint64_t)(r < 0 ? r + m : r);
}

uint64_t ntt_field_reduce_u64(const ntt_prime_field* f, uint64_t value) { return value % f->modulus; }

int64_t ntt_field_signed(const ntt_prime_field* f, uint64_t r) {
    return r > (f->modulus >> 1) ? (int64_t)r - (int64_t)f->modulus : (int64_t)r;
}

uint64_t ntt_field_add(const ntt_prime_field* f, uint64_t a, uint64_t b) {
    const uint64_t s = a + b;
    return s >= f->modulus ? s - f->modulus : s;
}

uint64_t ntt_field_sub(const ntt_prime_field* f, uint64_t a, uint64_t b) {
    return a >= b ? a - b : a + f->modulus - b;
}

uint64_t ntt_field_neg(const ntt_prime_field* f, uint64_t a) { return a ? f->modulus - a : 0; }

uint64_t ntt_field_mul(const ntt_prime_field* f, uint64_t a, uint64_t b) { return mulmod(a, b, f->modulus); }

uint64_t ntt_field_pow(const ntt_prime_field* f, uint64_t base, uint64_t exponent) {
    return powmod(base, exponent, f->modulus);
}

// Extended Euclid; zero has no inverse.
bool ntt_field_inv(const ntt_prime_field* f, uint64_t a, uint64_t* out) {
    if (!f || !out || a == 0) return false;
    int64_t t0 = 0, t1 = 1;
    uint64_t r0 = f->modulus, r1 = a;
    while (r1) {
        const uint64_t q = r0 / r1;
        const uint64_t r2 = r0 - q * r1;
        const int64_t t2 = t0 - (int64_t)q * t1;
        r0 = r1;
        r1 = r2;
        t0 = t1;
        t1 = t2;
    }
    if (r0 != 1) return false;
    *out = ntt_field_reduce(f, t0);
    return true;
}

void ntt_field_mul_sub(const ntt_prime_field* f, uint64_t* target, const uint64_t* source, size_t n,
                       uint64_t factor) {
    for (size_t i = 0; i < n; ++i) target[i] = ntt_field_sub(f, target[i], mulmod(source[i], factor, f->modulus));
}

void ntt_field_scale(const ntt_prime_field* f, uint64_t* target, size_t n, uint64_t factor) {
    for (size_t i = 0; i < n; ++i) target[i] = mulmod(target[i], factor, f->modulus);
}

uint64_t ntt_field_evaluate(const ntt_prime_field* f, const uint64_t* coefficients, size_t n, uint64_t point) {
    uint64_t v = 0;
    for (size_t i = n; i-- > 0;) v = ntt_field_add(f, mulmod(v, point, f->modulus), coefficients[i]);
    return v;
}

bool ntt_crt_pair_new(const ntt_prime_field* first, const ntt_prime_field* second, ntt_crt_pair* out) {
    if (!first || !second || !out) return false;
    uint64_t inv = 0;
    if (!ntt_field_inv(second, first->modulus % second->modulus, &inv)) return false;
    out->first = *first;
    out->second = *second;
    out->inverse = inv;
    return true;
}

void ntt_crt_pair_modulus(const ntt_crt_pair* pair, uint64_t* lo, uint64_t* hi) {
    const u128 m = (u128)pair->first.modulus * pair->second.modulus;
    *lo = (uint64_t)m;
    *hi = (uint64_t)(m >> 64);
}

// Garner: x = a + p1 * ((b - a) * p1^-1 mod p2) in [0, M), then shifted to
// the symmetric range -(M-1)/2..(M-1)/2.
void ntt_crt_pair_combine(const ntt_crt_pair* pair, uint64_t first, uint64_t second, uint64_t* lo, int64_t* hi) {
    const ntt_prime_field* f2 = &pair->second;
    const uint64_t digit = mulmod(ntt_field_sub(f2, second, first % f2->modulus), pair->inverse, f2->modulus);
    const u128 m = (u128)pair->first.modulus * f2->modulus;
    const u128 x = (u128)first + (u128)pair->first.modulus * digit;
    const i128 v = x > (m >> 1) ? (i128)x - (i128)m : (i128)x;
    *lo = (uint64_t)v;
    *hi = (int64_t)(v >> 64);
}

}  // extern "C"
