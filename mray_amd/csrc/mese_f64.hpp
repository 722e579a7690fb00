// FP64 building blocks of the MIXED preparation and its reverse: eight
// lanes per pixel (eight pixels per wavefront), lane g holding elements
// g, g + 8, ...; per-pixel arrays in LDS, zero past the entries written so
// far, so that inner products read their operands unconditionally.
#pragma once

#include "device_common.hpp"

namespace mray
{

struct dc
{
    double x, y;
};

// Sum over the 8 lanes of a row group, result in every lane.
__device__ __forceinline__ double grp8_sum(double v)
{
    int2 w = *reinterpret_cast<int2 *>(&v);
    int2 o;
    // xor 1 within quads
    o.x = __builtin_amdgcn_update_dpp(0, w.x, 0xB1, 0xF, 0xF, false);
    o.y = __builtin_amdgcn_update_dpp(0, w.y, 0xB1, 0xF, 0xF, false);
    v += *reinterpret_cast<double *>(&o);
    w = *reinterpret_cast<int2 *>(&v);
    // xor 2 within quads
    o.x = __builtin_amdgcn_update_dpp(0, w.x, 0x4E, 0xF, 0xF, false);
    o.y = __builtin_amdgcn_update_dpp(0, w.y, 0x4E, 0xF, 0xF, false);
    v += *reinterpret_cast<double *>(&o);
    w = *reinterpret_cast<int2 *>(&v);
    // lane i <- lane 7-i within each half row (the other quad)
    o.x = __builtin_amdgcn_update_dpp(0, w.x, 0x141, 0xF, 0xF, false);
    o.y = __builtin_amdgcn_update_dpp(0, w.y, 0x141, 0xF, 0xF, false);
    v += *reinterpret_cast<double *>(&o);
    return v;
}

// Orders LDS accesses of the lanes of one wavefront (they execute in lockstep).
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// 1 / d by the hardware reciprocal and two Newton steps (within an ulp of
// the IEEE quotient; d is in (0, 1] here)
__device__ __forceinline__ double recip_nr(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    return fma(r, fma(-d, r, 1.0), r);
}

// One step of the Herglotz recursion for pixel arrays T, E (lane g of the
// pixel's 8): E_l = (2 pi / l) i sum_{j<l} (l - j) T_{l-j} E_j, over the S
// slots that hold j < l (E_j = 0 for j >= l, so surplus terms vanish). A lane
// only reads its own elements of E, so they come from registers (Er, slot s
// = element g + 8s); E in LDS is kept complete for later readers.
template <int S, int C>
__device__ __forceinline__ void herglotz_step(const double *T, dc *E, dc (&Er)[C], int g, int l)
{
    double ax = 0.0, ay = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s)
    {
        const int j = g + 8 * s;
        const int i = (j < l) ? l - j : 0;
        const double w = (double)i * T[i];
        ax = fma(w, Er[s].x, ax);
        ay = fma(w, Er[s].y, ay);
    }
    ax = grp8_sum(ax);
    ay = grp8_sum(ay);
    if (g == (l & 7))
    {
        const double c = 2.0 * kPi / (double)l;
        const dc el = dc{-ay * c, ax * c};
        E[l] = el;
        // slot l / 8: S - 1, or S for l = 8S (the first element of the next octave)
        if ((l >> 3) == S - 1)
            Er[S - 1] = el;
        else if constexpr (S < C)
            Er[S] = el;
    }
    wave_sync();
}

// Steps l = 8(S-1)+1 .. min(8S, nm-1) use S slots; then the next octave.
template <int S, int C>
__device__ __forceinline__ void herglotz_octaves_r(const double *T, dc *E, dc (&Er)[C], int g, int nm)
{
    const int hi = min(8 * S, nm - 1);
    for (int l = 8 * (S - 1) + 1; l <= hi; ++l)
        herglotz_step<S, C>(T, E, Er, g, l);
    if constexpr (S < C)
        herglotz_octaves_r<S + 1, C>(T, E, Er, g, nm);
}

// The Herglotz recursion over the pixel's nm elements; E holds E_0 and zeros
// on entry, all E_l on exit.
template <int S, int C>
__device__ __forceinline__ void herglotz_octaves(const double *T, dc *E, int g, int nm)
{
    static_assert(S == 1, "starts at the first octave");
    dc Er[C];
#pragma unroll
    for (int s = 0; s < C; ++s)
        Er[s] = E[g + 8 * s];
    herglotz_octaves_r<1, C>(T, E, Er, g, nm);
}

// One Levinson step j over the S slots that hold k <= j:
// u = sum_{k<j} A_k E_{j-k}; A_k <- f (A_k - conj(A_{j-k}) u) for k = 0..j
// (A_j = 0 before the step, so k = 0 gives f A_0 and k = j gives -f A_0 u),
// f = 1 / (1 - |u|^2), with the reference's |u|^2 >= 1 guard: from the first
// step where it fires (j_invalid) on, u = 0. A lane's own elements A_k live in
// registers (Ar); LDS holds the array for the partner reads A_{j-k}. With REC
// the step's u is kept in slot S - 1 = j / 8 of lane j % 8 (U).
template <int S, int C, bool REC>
__device__ __forceinline__ void levinson_step(dc *A, dc (&Ar)[C], const dc *E, int g, int j, int &j_invalid,
                                              dc (&U)[C])
{
    dc ar[S];
    double ux = 0.0, uy = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s)
    {
        const int k = g + 8 * s;
        const int r = (k < j) ? j - k : 0;
        ar[s] = A[r];
        const dc e = E[r];
        // A_k = 0 for k >= j, so those terms vanish
        ux = fma(Ar[s].x, e.x, fma(-Ar[s].y, e.y, ux));
        uy = fma(Ar[s].x, e.y, fma(Ar[s].y, e.x, uy));
    }
    ux = grp8_sum(ux);
    uy = grp8_sum(uy);
    double a2 = ux * ux + uy * uy;
    if (j >= j_invalid || a2 >= 1.0)
    {
        ux = uy = 0.0;
        a2 = 0.0;
        j_invalid = min(j_invalid, j);
    }
    if constexpr (REC)
        if (g == (j & 7))
            U[S - 1] = dc{ux, uy};
    const double f = recip_nr(1.0 - a2);
    wave_sync();
#pragma unroll
    for (int s = 0; s < S; ++s)
    {
        const int k = g + 8 * s;
        if (k <= j)
        {
            // conj(A_{j-k}) u (for k = j the partner is A_0, which is real)
            const dc c = ar[s];
            const double px = c.x * ux + c.y * uy;
            const double py = c.x * uy - c.y * ux;
            Ar[s] = dc{f * (Ar[s].x - px), f * (Ar[s].y - py)};
            A[k] = Ar[s];
        }
    }
    wave_sync();
}

// Steps j = max(1, 8(S-1)) .. min(8S-1, nm-1) use S slots; then the next octave.
template <int S, int C, bool REC>
__device__ __forceinline__ void levinson_octaves_r(dc *A, dc (&Ar)[C], const dc *E, int g, int nm, int &j_invalid,
                                                   dc (&U)[C])
{
    const int hi = min(8 * S - 1, nm - 1);
    for (int j = (S == 1) ? 1 : 8 * (S - 1); j <= hi; ++j)
        levinson_step<S, C, REC>(A, Ar, E, g, j, j_invalid, U);
    if constexpr (S < C)
        levinson_octaves_r<S + 1, C, REC>(A, Ar, E, g, nm, j_invalid, U);
}

// The Levinson recursion over the pixel's nm elements; A holds A_0 and zeros
// on entry, the solution on exit.
template <int S, int C, bool REC>
__device__ __forceinline__ void levinson_octaves(dc *A, const dc *E, int g, int nm, int &j_invalid, dc (&U)[C])
{
    static_assert(S == 1, "starts at the first octave");
    dc Ar[C];
#pragma unroll
    for (int s = 0; s < C; ++s)
        Ar[s] = A[g + 8 * s];
    levinson_octaves_r<1, C, REC>(A, Ar, E, g, nm, j_invalid, U);
}

} // namespace mray
