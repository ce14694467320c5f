// Checks hs_div_fast (hindemith_amd/csrc/hs_div.h) against the host's IEEE
// fp32 division: whenever its guard accepts, the quotient must be bitwise
// equal to a / b.  Operand sets:
//   1. every significand of b in [1, 2) at several binades, against a fixed set
//      of numerators (the reciprocal's rounding goes through every case);
//   2. random pairs, log-uniform over and beyond the accepted ranges, mixed
//      with zeros, subnormals, infinities and NaNs;
//   3. near-midpoint quotients: a = RN(m * b) for m halfway between two
//      adjacent floats, the hardest case for a corrected quotient;
//   4. magnitudes of the Horn-Schunck update (|a| up to 1e4, b = alpha^2 +
//      Ix^2 + Iy^2 in [1e-2, 1e6]), which must all be accepted;
//   5. the hardest quotients there are: a/b = (midpoint) + d / (2^24 B) with
//      d = +-1, +-3, built by solving A * 2^24 - M * B = d for significands
//      A, B and an odd midpoint numerator M (the cases only the final
//      correction rounds right).
// Prints one line of counts; exit status 1 on any mismatch.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../hindemith_amd/csrc/hs_div.h"

using hmx::hs_div_fast;
using hmx::hs_recip;

static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static inline uint64_t rnd() {  // xorshift64*
    g_state ^= g_state >> 12;
    g_state ^= g_state << 25;
    g_state ^= g_state >> 27;
    return g_state * 0x2545F4914F6CDD1Dull;
}
static inline float bits_to_f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
static inline uint32_t f_to_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
// random float with exponent in [lo, hi] (unbiased), random significand and sign
static inline float rnd_float(int lo, int hi, bool sign) {
    const uint64_t r = rnd();
    const int e = lo + (int)(r % (uint64_t)(hi - lo + 1));
    const uint32_t mant = (uint32_t)(r >> 32) & 0x7fffff;
    const uint32_t s = sign ? (uint32_t)((r >> 63) & 1) << 31 : 0;
    return bits_to_f(s | (uint32_t)(e + 127) << 23 | mant);
}

static long long n_acc = 0, n_rej = 0, n_bad = 0;

static inline void check(float a, float b) {
    const float y = hs_recip(b);
    float q;
    if (!hs_div_fast(a, b, y, q)) {
        ++n_rej;
        return;
    }
    ++n_acc;
    const float ref = a / b;
    if (f_to_bits(q) != f_to_bits(ref)) {
        if (++n_bad <= 10)
            std::printf("MISMATCH a=%a b=%a fast=%a ieee=%a\n", (double)a, (double)b, (double)q,
                        (double)ref);
    }
}

int main(int argc, char **argv) {
    const long long scale = argc > 1 ? std::atoll(argv[1]) : 1;  // work multiplier

    // 1. every significand of b at several binades
    const int b_exps[] = {-24, -7, 0, 23};
    float nums[24];
    int nn = 0;
    nums[nn++] = 1.0f;
    nums[nn++] = bits_to_f(0x3fffffff);  // 1.99999988
    nums[nn++] = 1.5f;
    nums[nn++] = -3.0f;
    while (nn < 24) nums[nn++] = rnd_float(-50, 50, true);
    const int n_nums = (int)(scale > 1 ? 24 : 8);
    for (int be : b_exps)
        for (uint32_t m = 0; m < (1u << 23); ++m) {
            const float b = bits_to_f((uint32_t)(be + 127) << 23 | m);
            for (int i = 0; i < n_nums; ++i) check(nums[i], b);
        }
    const long long acc1 = n_acc;

    // 2. random pairs over and beyond the accepted ranges, with special values
    const float specials[] = {0.0f, -0.0f, INFINITY, -INFINITY, NAN,
                              bits_to_f(1), bits_to_f(0x807fffff), 0x1p-126f, 0x1p127f};
    for (long long i = 0; i < 40000000LL * scale; ++i) {
        float a = rnd_float(-126, 127, true);
        float b = rnd_float(-40, 40, true);
        const uint64_t r = rnd();
        if ((r & 63) == 0) a = specials[(r >> 8) % 9];
        if (((r >> 16) & 63) == 0) b = specials[(r >> 24) % 9];
        check(a, b);
    }

    // 3. near-midpoint quotients
    for (long long i = 0; i < 20000000LL * scale; ++i) {
        const float b = rnd_float(-24, 23, true);
        const float q = rnd_float(-60, 60, true);
        const double mid = 0.5 * ((double)q + (double)std::nextafter(q, 2.0f * q));
        const float a = (float)(mid * (double)b);  // mid*b is exact in double
        check(a, b);
        check(std::nextafter(a, 0.0f), b);
        check(std::nextafter(a, 2.0f * a), b);
    }

    // 4. Horn-Schunck magnitudes: all must be accepted
    const long long rej_before = n_rej;
    for (long long i = 0; i < 20000000LL * scale; ++i) {
        const uint64_t r = rnd();
        const double u1 = (double)(r & 0xffffffff) / 4294967296.0;
        const double u2 = (double)(r >> 32) / 4294967296.0;
        const float a = (float)((u1 - 0.5) * 2e4 * std::pow(10.0, -8.0 * u2));  // |a| < 1e4
        const float b = (float)std::pow(10.0, -2.0 + 8.0 * u1 * u2 + 1e-3);
        if (a == 0.0f) continue;
        check(a, b);
    }
    const long long hs_rej = n_rej - rej_before;

    // 5. quotients within d * 2^-24 ulp of a midpoint
    long long n_hard = 0;
    const long long acc_before5 = n_acc;
    for (long long i = 0; n_hard < 4000000LL * scale; ++i) {
        const uint32_t B = (uint32_t)(1u << 23) | ((uint32_t)rnd() & 0x7fffff) | 1u;  // odd
        uint32_t inv = B;  // B^-1 mod 2^32 by Newton's iteration
        for (int t = 0; t < 5; ++t) inv *= 2u - B * inv;
        const int ds[] = {1, -1, 3, -3};
        const int d = ds[i & 3];
        // M = -d * B^-1 (mod 2^24), lifted into [2^24, 2^25); odd since d is
        const uint64_t M = ((uint64_t)(uint32_t)(-(int64_t)d * (int64_t)inv) & 0xffffff) |
                           (1ull << 24);
        const int64_t num = (int64_t)(M * B) + d;  // = A * 2^24
        if (num & 0xffffff) continue;              // not a solution (cannot happen)
        const int64_t A = num >> 24;
        if (A < (1 << 23) || A >= (1 << 24)) continue;
        // a = A * 2^(ea-23), b = B * 2^(eb-23); the quotient's binade moves with ea - eb
        const uint64_t r = rnd();
        const int eb = -24 + (int)(r % 48);
        const int ea = eb + (int)((r >> 8) % 100) - 50;
        const float a = std::ldexp((float)A, ea - 23), b = std::ldexp((float)B, eb - 23);
        check(a, b);
        check(-a, b);
        ++n_hard;
    }
    const long long hard_acc = n_acc - acc_before5;

    std::printf("accepted %lld (exhaustive-b %lld, near-midpoint %lld) rejected %lld "
                "mismatches %lld hs-range-rejected %lld\n",
                n_acc, acc1, hard_acc, n_rej, n_bad, hs_rej);
    return (n_bad == 0 && hs_rej == 0) ? 0 : 1;
}
