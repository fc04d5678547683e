// Row-pass launches of products (two operands).
#include "fused.h"

namespace nttamd {

bool launch_fused_products(int logl, const PassArgs& a, int primes, bool pack, int form, hipStream_t s) {
    constexpr bool SQ = false;
    switch (logl) {
        case 5: return launch_fused_f<5, SQ>(a, primes, pack, form, s);
        case 6: return launch_fused_f<6, SQ>(a, primes, pack, form, s);
        case 7: return launch_fused_f<7, SQ>(a, primes, pack, form, s);
        case 8: return launch_fused_f<8, SQ>(a, primes, pack, form, s);
        case 9: return launch_fused_f<9, SQ>(a, primes, pack, form, s);
        case 10: return launch_fused_f<10, SQ>(a, primes, pack, form, s);
        case 11: return launch_fused_f<11, SQ>(a, primes, pack, form, s);
        case 12: return launch_fused_f<12, SQ>(a, primes, pack, form, s);
        default: return false;
    }
}

}  // namespace nttamd
