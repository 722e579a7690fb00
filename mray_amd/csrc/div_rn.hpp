// Division by a divisor whose reciprocal is formed once.
//
// For y = RN(1/b): q = RN(a y), r = a - b q (exact with a fused
// multiply-add), q' = RN(q + r y). Markstein's correction step makes q' the
// correctly rounded quotient RN(a/b), i.e. the IEEE division's result, as
// long as no intermediate leaves the normal range. The operation is
// invariant under scaling a or b by powers of two within that range, so
// tests/test_division.py checks it for every dividend significand against
// a set of divisor significands (the adversarial ones next to 1 and 2 and
// random ones), compiled for the host from this header.
//
// The range kept: 2^-30 <= |b| <= 2^30 (host_reciprocal returns NaN
// outside it), 2^-60 <= |a| and |q'| <= 2^60; div_rn_fast_ok() says whether
// a result may be used, else the caller divides.
#pragma once

#if defined(__HIPCC__)
#define MRAY_HD __host__ __device__
#else
#define MRAY_HD
#endif

namespace mray
{

// RN(1 / b) where 2^-30 <= |b| <= 2^30, else NaN (which fails div_rn_fast_ok)
MRAY_HD inline float host_reciprocal(float b)
{
    const float ab = b < 0.0f ? -b : b;
    if (!(ab >= 0x1p-30f && ab <= 0x1p30f))
        return __builtin_nanf("");
    return 1.0f / b;
}

MRAY_HD inline float div_rn_core(float a, float b, float y)
{
    const float q = a * y;
    const float r = __builtin_fmaf(-q, b, a);
    return __builtin_fmaf(r, y, q);
}

MRAY_HD inline bool div_rn_fast_ok(float a, float q)
{
    return __builtin_fabsf(a) >= 0x1p-60f && __builtin_fabsf(q) <= 0x1p60f;
}

} // namespace mray
