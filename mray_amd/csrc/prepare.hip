// Bounded-MESE preparation: moments -> Lagrange multipliers per pixel
// (MomentImageHost::prepare_moments, src/moment_image.cpp:154-183 ->
// prepare_moments_from_std_coding, src/moment_preparation.h:55-64 ->
// prepareMESELagrange, src/mese/MESE_dynamic.h:194-211).
//
//  PLAIN  the reference FP32 arithmetic operation for operation, one thread
//         per pixel, the three complex temporaries strided through LDS (as
//         prepareMomentsKernel). sinf / cosf of the zeroth exponential
//         moment reproduce the host C library's results (host_sincosf.hpp),
//         so the multipliers are bit-identical to the reference CPU path's.
//
//  MIXED  the Herglotz transform (trigonometric -> exponential moments) and
//         the Levinson recursion in FP64, where conditioning at 50+ moments
//         needs it; the autocorrelation and correlation stay FP64 too since
//         FP64 FMA is full rate on gfx950. Eight lanes work on one pixel
//         (eight pixels per wavefront): lane g owns elements g, g+8, ...;
//         the O(m) inner products of the sequential recursions are reduced
//         across the 8 lanes with DPP butterflies (quad_perm xor 1, xor 2,
//         row_half_mirror), never leaving the 8-lane row group. Per-pixel
//         arrays live in LDS (320 bytes per padded moment per wavefront).
#include "launch.hpp"
#include "mese_f32.hpp"

namespace mray
{

struct dc
{
    double x, y;
};


// ---------------------------------------------------------------------------
// PLAIN (FP32, one thread per pixel)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void prepare_plain_kernel(const float *__restrict__ mom,
                                                           const uint32_t *__restrict__ index, int npix, int m,
                                                           float bias, float *__restrict__ out)
{
    extern __shared__ cf lds_cf[]; // 3 arrays of [m][64]
    const int tid = threadIdx.x;
    const int pix = blockIdx.x * 64 + tid;
    if (pix >= npix)
        return;
    const int nm = index ? min((int)(index[pix + 1] - index[pix]), m) : m;
    const size_t off = index ? (size_t)index[pix] : (size_t)pix * m;
    if (nm <= 0)
        return;
    const Col E{lds_cf + tid, 64}, P{lds_cf + 64 * m + tid, 64}, R{lds_cf + 2 * 64 * m + tid, 64};
    const float *t = mom + off;
    // prepare_moments_from_std_coding (src/moment_preparation.h:55-64):
    // bias lerp toward the uniform density, then prepareMESELagrange
    // (src/mese/MESE_dynamic.h:194-211)
    trig_to_exp(nm, E, [&](int i) { return i == 0 ? lerp_bias(t[0], 0.5f, bias) : lerp_bias(t[i], 0.0f, bias); });
    levinson(nm, P, E);
    lagrange_from_levinson(nm, out + off, E, P, R);
}

// ---------------------------------------------------------------------------
// MIXED (FP64, 8 lanes per pixel)
// ---------------------------------------------------------------------------

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

__device__ __forceinline__ dc dmul(dc a, dc b) { return dc{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

// Orders LDS accesses of the lanes of one wavefront (they execute in lockstep).
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

template <int C>
__global__ __launch_bounds__(64) void prepare_mixed_kernel(const float *__restrict__ mom,
                                                           const uint32_t *__restrict__ index, int npix, int m,
                                                           float bias, float *__restrict__ out)
{
    constexpr int MP = 8 * C; // padded per-pixel stride
    extern __shared__ double lds_d[];
    const int lane = threadIdx.x;
    const int g = lane & 7;
    const int q = lane >> 3;
    const int pix = blockIdx.x * 8 + q;
    const bool valid = pix < npix;
    const int nm = valid ? (index ? min((int)(index[pix + 1] - index[pix]), m) : m) : 0;
    const size_t off = valid ? (index ? (size_t)index[pix] : (size_t)pix * m) : 0;

    double *T = lds_d + q * MP;
    dc *E = reinterpret_cast<dc *>(lds_d + 8 * MP) + q * MP;
    dc *A = reinterpret_cast<dc *>(lds_d + 8 * MP) + 8 * MP + q * MP;

    // biased moments, lerp in FP64 (near the boundary of the moment space the
    // recursion below amplifies a one-ulp change of the input)
    const double b64 = (double)bias;
#pragma unroll
    for (int s = 0; s < C; ++s)
    {
        const int e = g + 8 * s;
        if (e < nm)
        {
            const double v = (double)mom[off + e];
            T[e] = (1.0 - b64) * v + (e == 0 ? b64 * 0.5 : 0.0);
        }
    }
    wave_sync();

    // Herglotz transform: trigonometric -> exponential moments
    // (trigonometricToExponentialMoments, src/mese/MESE_dynamic.h:35-54)
    if (g == 0 && nm > 0)
    {
        double sn, cs;
        sincos(kPi * T[0] - 0.5 * kPi, &sn, &cs);
        E[0] = dc{cs * (0.25 / kPi), sn * (0.25 / kPi)};
    }
    wave_sync();
    for (int l = 1; l < m; ++l)
    {
        if (l < nm)
        {
            double ax = 0.0, ay = 0.0;
#pragma unroll
            for (int s = 0; s < C; ++s)
            {
                const int j = g + 8 * s;
                if (j < l)
                {
                    const double w = (double)(l - j) * T[l - j];
                    const dc ej = E[j];
                    ax = fma(w, ej.x, ax);
                    ay = fma(w, ej.y, ay);
                }
            }
            ax = grp8_sum(ax);
            ay = grp8_sum(ay);
            if (g == (l & 7))
            {
                const double c = 2.0 * kPi / (double)l;
                E[l] = dc{-ay * c, ax * c};
            }
        }
        wave_sync();
    }
    if (g == 0 && nm > 0)
        E[0] = dc{2.0 * E[0].x, 2.0 * E[0].y};
    wave_sync();

    // Levinson recursion on the Hermitian Toeplitz matrix of E, with the
    // reference's |u|^2 >= 1 guard (levinsonsAlgorithm, src/mese/MESE_dynamic.h:100-135)
    if (g == 0 && nm > 0)
        A[0] = dc{1.0 / E[0].x, 0.0};
    wave_sync();
    bool invalid = false;
    for (int j = 1; j < m; ++j)
    {
        if (j < nm)
        {
            double ux = 0.0, uy = 0.0;
            if (!invalid)
            {
#pragma unroll
                for (int s = 0; s < C; ++s)
                {
                    const int k = g + 8 * s;
                    if (k < j)
                    {
                        const dc a = A[k], e = E[j - k];
                        ux = fma(a.x, e.x, fma(-a.y, e.y, ux));
                        uy = fma(a.x, e.y, fma(a.y, e.x, uy));
                    }
                }
            }
            ux = grp8_sum(ux);
            uy = grp8_sum(uy);
            double a2 = ux * ux + uy * uy;
            if (a2 >= 1.0)
            {
                ux = uy = 0.0;
                a2 = 0.0;
                invalid = true;
            }
            const double f = 1.0 / (1.0 - a2);
            dc oldk[C], oldf[C];
            const dc a0 = A[0];
#pragma unroll
            for (int s = 0; s < C; ++s)
            {
                const int k = g + 8 * s;
                if (k >= 1 && k < j)
                {
                    oldk[s] = A[k];
                    oldf[s] = A[j - k];
                }
            }
            wave_sync();
#pragma unroll
            for (int s = 0; s < C; ++s)
            {
                const int k = g + 8 * s;
                if (k >= 1 && k < j)
                {
                    // A_k <- f (A_k - conj(A_{j-k}) u)
                    const dc c = oldf[s];
                    const double px = c.x * ux + c.y * uy;
                    const double py = c.x * uy - c.y * ux;
                    A[k] = dc{f * (oldk[s].x - px), f * (oldk[s].y - py)};
                }
            }
            if (g == (j & 7))
                A[j] = dc{f * (-a0.x * ux), f * (-a0.x * uy)};
            if (g == 0)
                A[0] = dc{f * a0.x, 0.0};
        }
        wave_sync();
    }

    // autocorrelation of P = 2 pi A (computeAutocorrelation), written over A
    const double a0x = (nm > 0) ? A[0].x : 1.0;
    dc R[C];
#pragma unroll
    for (int s = 0; s < C; ++s)
    {
        const int k = g + 8 * s;
        double rx = 0.0, ry = 0.0;
        if (k < nm)
            for (int j = 0; j < nm - k; ++j)
            {
                const dc a = A[j], b = A[j + k];
                // a * conj(b)
                rx = fma(a.x, b.x, fma(a.y, b.y, rx));
                ry = fma(a.y, b.x, fma(-a.x, b.y, ry));
            }
        const double sc = 4.0 * kPi * kPi;
        R[s] = dc{sc * rx, sc * ry};
    }
    wave_sync();
#pragma unroll
    for (int s = 0; s < C; ++s)
    {
        const int k = g + 8 * s;
        if (k < nm)
            A[k] = R[s];
    }
    if (g == 0 && nm > 0)
        E[0] = dc{0.5 * E[0].x, 0.5 * E[0].y};
    wave_sync();

    // imaginary correlation (computeImaginaryCorrelation) and normalisation
    const double nf = 1.0 / (kPi * 2.0 * kPi * a0x);
#pragma unroll
    for (int s = 0; s < C; ++s)
    {
        const int k = g + 8 * s;
        if (k < nm)
        {
            double acc = 0.0;
            for (int j = 0; j < nm - k; ++j)
            {
                const dc r = A[k + j], e = E[j];
                acc = fma(r.x, e.y, fma(r.y, e.x, acc));
            }
            out[off + k] = (float)(nf * acc);
        }
    }
}

size_t g_lds_limit = 65536;

void set_lds_limit(size_t bytes) { g_lds_limit = bytes; }

int launch_prepare(const float *moments, const uint32_t *index, int W, int H, int m, float bias, int form,
                   float *lagrange, hipStream_t stream)
{
    const int npix = W * H;
    if (form == MRAY_FORM_PLAIN)
    {
        size_t lds = (size_t)3 * 64 * m * sizeof(cf);
        if (lds > g_lds_limit)
            return MRAY_ERR_UNSUPPORTED;
        hipLaunchKernelGGL(prepare_plain_kernel, dim3((npix + 63) / 64), dim3(64), lds, stream, moments, index, npix,
                           m, bias, lagrange);
        return 0;
    }
    const int C = (m + 7) / 8;
    const size_t lds = (size_t)8 * 8 * C * (sizeof(double) + 2 * sizeof(dc));
    if (lds > g_lds_limit)
        return MRAY_ERR_UNSUPPORTED;
    dim3 grid((npix + 7) / 8);
#define MRAY_PREP_CASE(CC)                                                                                            \
    case CC:                                                                                                          \
        hipLaunchKernelGGL(prepare_mixed_kernel<CC>, grid, dim3(64), lds, stream, moments, index, npix, m, bias,      \
                           lagrange);                                                                                 \
        return 0;
    switch (C)
    {
        MRAY_PREP_CASE(1)
        MRAY_PREP_CASE(2)
        MRAY_PREP_CASE(3)
        MRAY_PREP_CASE(4)
        MRAY_PREP_CASE(5)
        MRAY_PREP_CASE(6)
        MRAY_PREP_CASE(7)
        MRAY_PREP_CASE(8)
        MRAY_PREP_CASE(9)
        MRAY_PREP_CASE(10)
        MRAY_PREP_CASE(11)
        MRAY_PREP_CASE(12)
        MRAY_PREP_CASE(13)
        MRAY_PREP_CASE(14)
        MRAY_PREP_CASE(15)
        MRAY_PREP_CASE(16)
    default:
        return MRAY_ERR_UNSUPPORTED;
    }
#undef MRAY_PREP_CASE
}

} // namespace mray
