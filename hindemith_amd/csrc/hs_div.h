// Correctly rounded fp32 quotient a / b from a reciprocal computed once.
//
// The Jacobi step divides a new numerator by the same per-pixel denominator
// every iteration.  With y = RN(1/b) computed once (an IEEE division), each
// step takes the quotient from three corrections instead of a full division:
//
//   q0 = RN(a*y)                  within 1.5 ulp of a/b
//   q1 = RN(q0 + RN(a - b*q0)*y)  within 1/2 ulp + 2^-22 ulp of a/b: faithful
//   q  = RN(q1 + (a - b*q1)*y)    = RN(a/b)
//
// The last line is Markstein's theorem (Markstein, "Computation of elementary
// functions on the IBM RISC System/6000 processor", IBM J. Res. Dev. 34(1),
// 1990; restated as Theorem 4.10 of Muller et al., Handbook of
// Floating-Point Arithmetic): if y is within 1/2 ulp of 1/b, q within one ulp
// of a/b and r = a - b*q is computed exactly (an fma, exact because q is
// faithful), then RN(q + r*y) is the correctly rounded quotient.  Its
// no-overflow / no-underflow conditions hold when |b| lies in [2^-24, 2^24]
// and |a/b| in [2^-76, 2^76]: every residual then lies on a grid no finer
// than 2^-147.  hs_recip gives NaN for any other b, so the quotient is NaN
// and hs_div_fast reports failure; a failed lane (also a zero, tiny, huge or
// non-finite quotient) takes the IEEE division instead.
//
// y itself: on the GPU, y = RN(y0 + RN(1 - b*y0)*y0) from the hardware
// reciprocal y0 = v_rcp_f32(b) (within one ulp), three instructions instead
// of the ten of the division sequence.  One such Newton step from a faithful
// reciprocal is correctly rounded except possibly when b's significand is all
// ones (Markstein's reciprocal theorem, Theorem 4.9 of the same Handbook); for
// this hardware reciprocal it equals the IEEE 1/b for every b in the accepted
// range, which hmx_hs_recip_check verifies exhaustively on the device (every
// significand, every exponent, both signs).  The host computes 1/b.
//
// Plain C++ as well as HIP, so the CPU test can check it against the host's
// IEEE division on hundreds of millions of operand pairs.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define HS_DIV_HD __host__ __device__
#else
#define HS_DIV_HD
#endif

namespace hmx {

HS_DIV_HD inline float hs_recip(float b) {
    const float ab = std::fabs(b);
#if defined(__HIP_DEVICE_COMPILE__)
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float y = std::fma(std::fma(-b, y0, 1.0f), y0, y0);
#else
    const float y = 1.0f / b;
#endif
    return (ab >= 0x1p-24f && ab <= 0x1p24f) ? y : __builtin_nanf("");
}

// q = RN(a / b) when the result is true; otherwise q is unspecified and the
// caller must divide.
HS_DIV_HD inline bool hs_div_fast(float a, float b, float y, float &q) {
    const float q0 = a * y;
    const float r0 = std::fma(-q0, b, a);
    const float q1 = std::fma(r0, y, q0);
    const float r1 = std::fma(-q1, b, a);
    q = std::fma(r1, y, q1);
    const float aq = std::fabs(q);
    return aq >= 0x1p-76f && aq <= 0x1p76f;
}

}  // namespace hmx
