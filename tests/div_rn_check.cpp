// Check of mray::div_rn_core (mray_amd/csrc/div_rn.hpp) against IEEE float
// division: for each divisor significand below, every dividend significand
// in [1, 2) and the quotient are compared bit for bit. Built and run by
// tests/test_division.py. Prints "<pairs> <mismatches> <fast-path rejects> <range guards ok>".
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdint.h>

#include "../mray_amd/csrc/div_rn.hpp"

static float from_bits(uint32_t u)
{
    float f;
    memcpy(&f, &u, 4);
    return f;
}

int main(int argc, char **argv)
{
    const int n_random = argc > 1 ? atoi(argv[1]) : 64;
    const int n_edge = 32; // significands next to 1 and next to 2
    long pairs = 0, bad = 0, rejected = 0;
    uint32_t lcg = 12345u;
    for (int ib = 0; ib < 2 * n_edge + n_random; ++ib)
    {
        uint32_t bb;
        if (ib < n_edge)
            bb = 0x3f800000u + (uint32_t)ib;
        else if (ib < 2 * n_edge)
            bb = 0x3fffffffu - (uint32_t)(ib - n_edge);
        else
        {
            lcg = lcg * 1664525u + 1013904223u;
            bb = 0x3f800000u | (lcg >> 9);
        }
        // the divisor at three scales and both signs (a scale of the divisor
        // scales its reciprocal exactly, so these agree when all is well)
        for (int sc = 0; sc < 3; ++sc)
        {
            const float b = from_bits(bb) * (sc == 0 ? 1.0f : sc == 1 ? 0x1p-20f : -0x1p17f);
            const float y = mray::host_reciprocal(b);
            long p = 0, m = 0, rj = 0;
#pragma omp parallel for reduction(+ : p, m, rj) schedule(static, 65536)
            for (uint32_t k = 0; k < (1u << 23); ++k)
            {
                const float a = from_bits(0x3f800000u + k) * (sc == 2 ? 0x1p-40f : 1.0f);
                const float q = mray::div_rn_core(a, b, y);
                const float ref = a / b;
                if (!mray::div_rn_fast_ok(a, q))
                    ++rj;
                else if (memcmp(&q, &ref, 4) != 0)
                    ++m;
                ++p;
            }
            pairs += p;
            bad += m;
            rejected += rj;
        }
    }
    // out of range: no reciprocal, and the fast path refuses the result;
    // zero dividends (whose sign the IEEE division keeps) are refused too
    const float big = mray::host_reciprocal(0x1p31f), tiny = mray::host_reciprocal(0x1p-31f);
    const int guards = (big != big) && (tiny != tiny) && !mray::div_rn_fast_ok(1.0f, mray::div_rn_core(1.0f, 0x1p31f, big)) &&
                       !mray::div_rn_fast_ok(0.0f, 0.0f) && !mray::div_rn_fast_ok(-0.0f, 0.0f);
    printf("%ld %ld %ld %d\n", pairs, bad, rejected, guards);
    return 0;
}
