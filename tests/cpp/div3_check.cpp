// div3_check.cpp -- div3_rn(x) (honei_amd/csrc/exact_div.hh, the division
// the matrix-free Q1 weights use on the device) against the division x / 3.0
// on the host, bit for bit:
//   * random normal doubles of every sign and exponent in [-1000, 1000];
//   * runs of consecutive significands at the bottom, the middle and the top
//     of a binade and just below / above 3 * 2^k (quotients next to a power
//     of two);
//   * the operands of the Q1 weights: lognormal cell coefficients, sums of
//     two and of four, doubled sums.
// Prints the first mismatch, if any; exit status 0 when all agree.
//
//   div3_check [random-count]

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "exact_div.hh"

using honei::hipx::div3_rn;

namespace {

uint64_t g_state = 0x243F6A8885A308D3ull;
uint64_t next_u64() {   // SplitMix64
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double u01() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }

double from_bits(uint64_t b) {
    double d;
    std::memcpy(&d, &b, 8);
    return d;
}
uint64_t bits(double d) {
    uint64_t b;
    std::memcpy(&b, &d, 8);
    return b;
}

long g_checked = 0;
bool check(double x) {
    ++g_checked;
    const volatile double three = 3.0;
    const double want = x / three;
    const double got = div3_rn(x);
    if (bits(want) != bits(got)) {
        std::printf("MISMATCH x = %a: x / 3.0 = %a, div3_rn = %a\n", x, want, got);
        return false;
    }
    return true;
}

}  // namespace

int main(int argc, char **argv) {
    const long nrand = argc > 1 ? std::atol(argv[1]) : 20000000L;
    for (long i = 0; i < nrand; ++i) {
        const uint64_t m = next_u64() & 0x000FFFFFFFFFFFFFull;
        const int64_t e = 1023 + (int64_t)(next_u64() % 2001) - 1000;
        const uint64_t s = next_u64() & 1;
        if (!check(from_bits((s << 63) | ((uint64_t)e << 52) | m))) return 1;
    }
    // consecutive significands
    const double bases[] = {1.0, 1.5, 2.0 - 0x1p-30, 3.0 - 0x1p-40, 3.0, 3.0 + 0x1p-40, 0x1.5555555555555p+0,
                            0x1.aaaaaaaaaaaabp+0, 6.0 - 0x1p-35, 0x1p-900, 0x1p+900};
    for (double b : bases)
        for (int sgn = 0; sgn < 2; ++sgn) {
            uint64_t u = bits(sgn ? -b : b);
            for (long k = -(1L << 20); k < (1L << 20); ++k)
                if (!check(from_bits(u + (uint64_t)k))) return 1;
        }
    // operands of the Q1 weights: kappa = exp(sigma z), sums of 2 and 4, doubled sum
    for (long i = 0; i < nrand / 4; ++i) {
        double k[4];
        for (double &v : k) {
            const double u1 = 1.0 - u01(), u2 = u01();
            const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
            v = std::exp(1.0 * z);
        }
        if (!check(k[0]) || !check(k[0] + k[1]) || !check(2.0 * (((k[0] + k[1]) + k[2]) + k[3])) ||
            !check(-(k[2] + k[3])))
            return 1;
    }
    std::printf("div3_rn agrees with x / 3.0 on %ld operands\n", g_checked);
    return 0;
}
