// Exhaustive check of mray::host_sinf / host_cosf (mray_amd/csrc/host_sincosf.hpp)
// against the host C library's sinf / cosf over every float with |x| < 120.
// Built and run by tests/test_host_sincosf.py. Prints "<n> <sin mismatches> <cos mismatches>".
#include <cstdio>
#include <cstring>

#include "../mray_amd/csrc/host_sincosf.hpp"

int main()
{
    long n = 0, ms = 0, mc = 0;
#pragma omp parallel for reduction(+ : n, ms, mc) schedule(static, 65536)
    for (long u = 0; u < 0x42f00000L; ++u)
        for (int sg = 0; sg < 2; ++sg)
        {
            uint32_t b = (uint32_t)u | (sg ? 0x80000000u : 0u);
            float x;
            memcpy(&x, &b, 4);
            float a = sinf(x), c = cosf(x), ga = mray::host_sinf(x), gc = mray::host_cosf(x);
            ms += memcmp(&a, &ga, 4) != 0;
            mc += memcmp(&c, &gc, 4) != 0;
            ++n;
        }
    printf("%ld %ld %ld\n", n, ms, mc);
    return 0;
}
