// sinf / cosf with the results of the host C library the reference runs on.
//
// The reference's FP32 MESE and prediction-coding code calls sinf / cosf
// (src/mese/MESE_dynamic.h:41, src/moment_prediction_coding.h:94-100), which
// on x86-64 glibc >= 2.28 is the sinf/cosf of Arm's optimized-routines
// (published under MIT / Apache-2.0, adopted by glibc 2.28): the argument is
// widened to double, reduced by the nearest multiple of pi/2 for |x| < 120,
// and evaluated with degree-7 sine and degree-8 cosine polynomials in double.
// On CPUs with FMA, glibc runs a build of it in which every a + b*c is one
// fused multiply-add; that is the variant restated here, with explicit fma()
// so that any compiler gives the same bits on the host and on the GPU.
// The outputs are not correctly rounded (about 0.56 ulp), which is why the
// device library's sinf / cosf cannot stand in for them where the reference's
// FP32 results must be reproduced bit for bit.
//
// Bit-exact against the host library for |x| < 120 (the phases here are in
// [-pi, pi]); larger arguments fall back to the device library.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define MRAY_HD __host__ __device__
#else
#define MRAY_HD
#endif

namespace mray
{
namespace glibc_sincosf
{
// polynomial coefficients: cos(x) ~ 1 + c1 x^2 + ... + c4 x^8,
// sin(x) ~ x + s1 x^3 + s2 x^5 + s3 x^7 on [-pi/4, pi/4]
constexpr double kHpiInv = 0x1.45F306DC9C883p+23; // 2/pi * 2^24
constexpr double kHpi = 0x1.921FB54442D18p0;      // pi/2
constexpr double kC1 = -0x1.ffffffd0c621cp-2;
constexpr double kC2 = 0x1.55553e1068f19p-5;
constexpr double kC3 = -0x1.6c087e89a359dp-10;
constexpr double kC4 = 0x1.99343027bf8c3p-16;
constexpr double kS1 = -0x1.555545995a603p-3;
constexpr double kS2 = 0x1.1107605230bc4p-7;
constexpr double kS3 = -0x1.994eb3774cf24p-13;

MRAY_HD inline uint32_t abstop12(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    return (u >> 20) & 0x7ff;
}

// sin (odd n == 0) or cos (n odd) polynomial of the reduced argument x;
// neg flips the cosine polynomial (quadrants 2 and 3)
MRAY_HD inline float poly(double x, double x2, int n, bool neg)
{
    if ((n & 1) == 0)
    {
        const double x3 = x * x2;
        const double s1 = fma(x2, kS3, kS2);
        const double x7 = x3 * x2;
        const double s = fma(x3, kS1, x);
        return (float)fma(x7, s1, s);
    }
    const double sg = neg ? -1.0 : 1.0;
    const double x4 = x2 * x2;
    const double c2 = fma(x2, sg * kC4, sg * kC3);
    const double c1 = fma(x2, sg * kC1, sg);
    const double x6 = x4 * x2;
    const double c = fma(x4, sg * kC2, c1);
    return (float)fma(x6, c2, c);
}

// x - n pi/2 with n the nearest integer to x 2/pi (|x| < 120)
MRAY_HD inline double reduce(double x, int *np)
{
    const double r = x * kHpiInv;
    const int n = ((int32_t)r + 0x800000) >> 24;
    *np = n;
    return fma(-(double)n, kHpi, x);
}
} // namespace glibc_sincosf

MRAY_HD inline float host_sinf(float y)
{
    using namespace glibc_sincosf;
    const float pio4 = 0x1.921FB6p-1f;
    double x = y;
    if (abstop12(y) < abstop12(pio4))
    {
        if (abstop12(y) < abstop12(0x1p-12f))
            return y;
        return poly(x, x * x, 0, false);
    }
    if (abstop12(y) < abstop12(120.0f))
    {
        int n;
        x = reduce(x, &n);
        const double s = ((n & 3) == 1 || (n & 3) == 2) ? -1.0 : 1.0;
        return poly(x * s, x * x, n, (n & 2) != 0);
    }
    return sinf(y);
}

MRAY_HD inline float host_cosf(float y)
{
    using namespace glibc_sincosf;
    const float pio4 = 0x1.921FB6p-1f;
    double x = y;
    if (abstop12(y) < abstop12(pio4))
    {
        if (abstop12(y) < abstop12(0x1p-12f))
            return 1.0f;
        return poly(x, x * x, 1, false);
    }
    if (abstop12(y) < abstop12(120.0f))
    {
        int n;
        x = reduce(x, &n);
        const double s = ((n & 3) == 1 || (n & 3) == 2) ? -1.0 : 1.0;
        return poly(x * s, x * x, n ^ 1, (n & 2) != 0);
    }
    return cosf(y);
}
} // namespace mray
