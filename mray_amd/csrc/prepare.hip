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
#include "mese_f64.hpp"

namespace mray
{


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

// per-pixel LDS stride of the mixed kernels, in complex doubles: MP + 4, so
// the eight pixels' arrays start 64 bytes (16 banks) apart (a stride of MP
// elements puts them all in the same banks), except where the padding's 1 KiB
// per workgroup costs a workgroup per CU that is worth more than the bank
// conflicts it saves. Measured at 1024 x 1024 pixels, FP64 multipliers (ms,
// padded / unpadded): C = 10 (80 moments) 12.0 / 11.3, C = 11 (88) 16.1 /
// 14.5, C = 13 (100) 23.9 / 20.7, C = 20 (160) 85.3 / 66.9; where dropping
// the padding also adds a workgroup per CU but loses: C = 5 (40) 2.75 /
// 2.95, C = 7 (50) 4.38 / 4.59, C = 8 (64) 6.81 / 7.17, C = 16 (128) 44.0 /
// 44.8; elsewhere the workgroups per CU are the same and the padding gains
// 1-3 %. At C = 32 the padding would take the two arrays of eight pixels past
// 64 KiB.
constexpr int prepare_pixel_stride(int C)
{
    return (C >= 32 || C == 10 || C == 11 || C == 13 || C == 20) ? 8 * C : 8 * C + 4;
}

// OutT: float multipliers, or FP64 ones for mray_hip_prepare_f64
template <int C, typename OutT = float>
__global__ __launch_bounds__(64) void prepare_mixed_kernel(const float *__restrict__ mom,
                                                           const uint32_t *__restrict__ index, int npix, int m,
                                                           float bias, OutT *__restrict__ out)
{
    // Per pixel (8 per wavefront) two LDS arrays of MP complex doubles: E
    // (exponential moments) and A (Levinson coefficients, later the
    // autocorrelation); the real moments T share A's space until the
    // recursion starts. Every array is zero past the entries written so far,
    // so the inner products read all their operands unconditionally (the
    // indices are clamped into range and the surplus terms are exact zeros):
    // the loads of a step issue together instead of one round trip per term.
    constexpr int MP = 8 * C; // elements per array and pixel
    constexpr int PS = prepare_pixel_stride(C);
    extern __shared__ dc lds_c[];
    const int lane = threadIdx.x;
    const int g = lane & 7;
    const int q = lane >> 3;
    const int pix = blockIdx.x * 8 + q;
    const bool valid = pix < npix;
    const int nm = valid ? (index ? min((int)(index[pix + 1] - index[pix]), m) : m) : 0;
    const size_t off = valid ? (index ? (size_t)index[pix] : (size_t)pix * m) : 0;

    dc *E = lds_c + q * PS;
    dc *A = lds_c + 8 * PS + q * PS;
    double *T = reinterpret_cast<double *>(A);

    // biased moments, lerp in FP64 (near the boundary of the moment space the
    // recursion below amplifies a one-ulp change of the input); zero padding
    const double b64 = (double)bias;
#pragma unroll
    for (int s = 0; s < C; ++s)
    {
        const int e = g + 8 * s;
        const double v = (e < nm) ? (double)mom[off + e] : 0.0;
        T[e] = (e < nm) ? (1.0 - b64) * v + (e == 0 ? b64 * 0.5 : 0.0) : 0.0;
        E[e] = dc{0.0, 0.0};
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
    herglotz_octaves<1, C>(T, E, g, nm);
    if (g == 0 && nm > 0)
        E[0] = dc{2.0 * E[0].x, 2.0 * E[0].y};
    // A: zero, then A_0 = 1 / E_0
#pragma unroll
    for (int s = 0; s < C; ++s)
        A[g + 8 * s] = dc{0.0, 0.0};
    wave_sync();
    if (g == 0 && nm > 0)
        A[0] = dc{1.0 / E[0].x, 0.0};
    wave_sync();

    // Levinson recursion on the Hermitian Toeplitz matrix of E, with the
    // reference's |u|^2 >= 1 guard (levinsonsAlgorithm, src/mese/MESE_dynamic.h:100-135)
    int j_invalid = nm;
    dc u_unused[C];
    levinson_octaves<1, C, false>(A, E, g, nm, j_invalid, u_unused);

    // autocorrelation of P = 2 pi A (computeAutocorrelation), written over A
    // (one pass over j for all slots: A_j, the same for the pixel's 8 lanes,
    // is read once per j instead of once per slot and j)
    const double a0x = (nm > 0) ? A[0].x : 1.0;
    dc R[C];
#pragma unroll
    for (int s = 0; s < C; ++s)
        R[s] = dc{0.0, 0.0};
    for (int j = 0; j < nm; ++j)
    {
        const dc a = A[j];
#pragma unroll
        for (int s = 0; s < C; ++s)
        {
            const int k = g + 8 * s;
            if (j + k < nm)
            {
                const dc b = A[j + k];
                // a * conj(b)
                R[s].x = fma(a.x, b.x, fma(a.y, b.y, R[s].x));
                R[s].y = fma(a.y, b.x, fma(-a.x, b.y, R[s].y));
            }
        }
    }
#pragma unroll
    for (int s = 0; s < C; ++s)
    {
        const double sc = 4.0 * kPi * kPi;
        R[s] = dc{sc * R[s].x, sc * R[s].y};
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
    // (one pass over j for all slots, E_j read once per j)
    const double nf = 1.0 / (kPi * 2.0 * kPi * a0x);
    double acc[C];
#pragma unroll
    for (int s = 0; s < C; ++s)
        acc[s] = 0.0;
    for (int j = 0; j < nm; ++j)
    {
        const dc e = E[j];
#pragma unroll
        for (int s = 0; s < C; ++s)
        {
            const int k = g + 8 * s;
            if (j + k < nm)
            {
                const dc r = A[k + j];
                acc[s] = fma(r.x, e.y, fma(r.y, e.x, acc[s]));
            }
        }
    }
#pragma unroll
    for (int s = 0; s < C; ++s)
    {
        const int k = g + 8 * s;
        if (k < nm)
            out[off + k] = (OutT)(nf * acc[s]);
    }
}

size_t g_lds_limit = 65536;

void set_lds_limit(size_t bytes) { g_lds_limit = bytes; }

namespace
{
// the MIXED preparation into float or FP64 multipliers
template <typename OutT>
int launch_prepare_mixed(const float *moments, const uint32_t *index, int npix, int m, float bias, OutT *lagrange,
                         hipStream_t stream)
{
    int C = (m + 7) / 8;
    if (C > 16)
        C = (C + 3) / 4 * 4; // 20, 24, 28, 32
    const size_t lds = (size_t)8 * prepare_pixel_stride(C) * 2 * sizeof(dc);
    if (lds > g_lds_limit)
        return MRAY_ERR_UNSUPPORTED;
    dim3 grid((npix + 7) / 8);
#define MRAY_PREP_CASE(CC)                                                                                            \
    case CC:                                                                                                          \
        hipLaunchKernelGGL((prepare_mixed_kernel<CC, OutT>), grid, dim3(64), lds, stream, moments, index, npix, m,    \
                           bias, lagrange);                                                                           \
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
        MRAY_PREP_CASE(20)
        MRAY_PREP_CASE(24)
        MRAY_PREP_CASE(28)
        MRAY_PREP_CASE(32)
    default:
        return MRAY_ERR_UNSUPPORTED;
    }
#undef MRAY_PREP_CASE
}
} // namespace

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
    return launch_prepare_mixed<float>(moments, index, npix, m, bias, lagrange, stream);
}

int launch_prepare_f64(const float *moments, const uint32_t *index, int W, int H, int m, float bias,
                       double *lagrange, hipStream_t stream)
{
    return launch_prepare_mixed<double>(moments, index, W * H, m, bias, lagrange, stream);
}

} // namespace mray
