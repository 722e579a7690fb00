// Reconstruct-and-composite march (reconstruct_ray,
// src/raymarch_reconstruct.h:137-200) over prepared Lagrange multipliers.
//
// Both forms replay the reference's phase sequence exactly: the phase is
// *incremented* in float (p += phase_step) while t <= t_far, clamped to
// [-pi, 0] per sample, and compositing stops when alpha >= the early ray
// termination threshold.
//
//  PLAIN  evaluateMESELagrange as written: cosf/sinf of the phase (the host
//         C library's results, host_sincosf.hpp), powers of the circle point
//         by repeated complex multiplication, transmittance by exp in double
//         (the reference's unqualified exp of a float), multipliers strided
//         through LDS (as reconstruct_moments_kernel). Same FP32 operations
//         as the reference CPU path.
//
//  MIXED  multipliers in registers (template M >= m); the series
//         lambda_0 + 2 sum_l lambda_l cos(l theta) by Clenshaw's recurrence
//         in FP64 (2 VALU ops per term, cos theta from an FP64 sine of the
//         half angle per sample): at 50 moments the multipliers reach
//         |lambda| ~ 1e3 and an FP32 series loses ~1e-4 of density to
//         cancellation, while FP64 FMAs cost the same issue slots as FP32
//         ones on gfx950. The 256-entry transfer function is staged in LDS
//         once per workgroup.
#include "device_common.hpp"
#include "launch.hpp"

namespace mray
{

struct RecArgs
{
    int W, H, m;
    float step;
    float emission;
    float ert;
    int fourier;
    Vec3 dmin, dmax;
};

__device__ __forceinline__ void zero_ray(const float4 *tf, float *o)
{
    // reconstruct_zero_ray (src/raymarch_reconstruct.h:127-131)
    float4 c = tf_sample(tf, 0.0f, 0.0f);
    o[0] = c.x;
    o[1] = c.y;
    o[2] = c.z;
    o[3] = 0.0f;
}

__device__ __forceinline__ void store_rgba(float *rgba, size_t pix, float r, float g, float b, float a)
{
    *reinterpret_cast<float4 *>(rgba + pix * 4) = make_float4(r, g, b, a);
}

// ---------------------------------------------------------------------------
// PLAIN
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void reconstruct_plain_kernel(DevCamera cam, RecArgs a,
                                                                const float *__restrict__ lag,
                                                                const uint32_t *__restrict__ index,
                                                                const uint32_t *__restrict__ bounds,
                                                                const float4 *__restrict__ tf_g,
                                                                float *__restrict__ rgba, int32_t *__restrict__ steps)
{
    __shared__ float4 tf[256];
    extern __shared__ float lds_lag[]; // [m][blockDim]
    const int tid = threadIdx.x;
    for (int i = tid; i < 256; i += blockDim.x)
        tf[i] = tf_g[i];
    __syncthreads();

    const int lane = tid & 63;
    const uint32_t ntiles = ((a.W + 7) / 8) * ((a.H + 7) / 8);
    const uint32_t tile = xcd_tile_block(a.W, blockDim.x / 64) * (blockDim.x / 64) + (tid >> 6);
    if (tile >= ntiles)
        return;
    int x, y;
    tile_pixel(tile, a.W, lane, x, y);
    if (x >= a.W || y >= a.H)
        return;
    const size_t pix = (size_t)y * a.W + x;
    const int nm = index ? min((int)(index[pix + 1] - index[pix]), a.m) : a.m;
    const size_t off = index ? (size_t)index[pix] : pix * (size_t)a.m;
    if (nm == 0)
    {
        float o[4];
        zero_ray(tf, o);
        store_rgba(rgba, pix, o[0], o[1], o[2], o[3]);
        if (steps)
            steps[pix] = 0;
        return;
    }
    float *L = lds_lag + tid;
    const int stride = blockDim.x;
    for (int l = 0; l < nm; ++l)
        L[l * stride] = lag[off + l];
    float blo = 0.f, bhi = 1.f;
    if (bounds)
    {
        uint32_t b = bounds[pix];
        blo = (b & 0xFFFFu) / 65535.f;
        bhi = (b >> 16) / 65535.f;
    }
    auto recon = [&](float p) {
        p = rclamp(p, (float)(-kPi), 0.0f);
        float cx = host_cosf(-p), cy = host_sinf(-p);
        // evaluateFourierSeries (src/mese/MESE_dynamic.h:213-226)
        float result = 0.0f;
        float pwx = 1.0f, pwy = 0.0f;
        for (int l = 1; l < nm; ++l)
        {
            float nx = pwx * cx - pwy * cy;
            float ny = pwx * cy + pwy * cx;
            pwx = nx;
            pwy = ny;
            result += L[l * stride] * pwx;
        }
        float sres = 2.0f * result + L[0];
        float rho = a.fourier ? sres : fast_atan(sres) * 0.318309886f + 0.5f;
        return rho * (bhi - blo) + blo;
    };

    Ray r = pixel_ray(cam, a.W, a.H, x, y);
    float t_near, t_far;
    bool hit = intersect_aabb(r, a.dmin, a.dmax, t_near, t_far);
    t_near = rmax(t_near, cam.znear);
    int ns = 0;
    float Lr = 0.f, Lg = 0.f, Lb = 0.f, alpha = 0.f;
    if (hit && (t_far > 0))
    {
        float t = t_near;
        float p = (float)((double)((t - t_near) / (t_far - t_near)) * kPi - kPi);
        const float dphi = (float)(kPi * (double)(a.step / (t_far - t_near)));
        float T = 1.0f, tau = 0.0f;
        float fs = recon(p);
        const int cap = march_cap(t_near, t_far, a.step) + 1;
        for (int it = 0; t <= t_far && it < cap; ++it)
        {
            t += a.step;
            p += dphi;
            ++ns;
            float bs = recon(p);
            float4 c = tf_sample(tf, fs, bs);
            Lr += (a.emission * c.x) * T;
            Lg += (a.emission * c.y) * T;
            Lb += (a.emission * c.z) * T;
            tau += c.w;
            T = (float)exp((double)(-tau));
            alpha = 1.0f - T;
            fs = bs;
            if (alpha >= a.ert)
            {
                alpha = 1.f;
                break;
            }
        }
    }
    store_rgba(rgba, pix, Lr, Lg, Lb, alpha);
    if (steps)
        steps[pix] = ns;
}

// ---------------------------------------------------------------------------
// MIXED
// ---------------------------------------------------------------------------
// sin x in FP64 for x in [0, pi/4]: Taylor series to x^17 (truncation below
// 1e-19); no range reduction needed on this interval
__device__ __forceinline__ double sin_quarter(double x)
{
    const double x2 = x * x;
    double c = 1.0 / 355687428096000.0; // 1/17!
    c = fma(c, -x2, 1.0 / 1307674368000.0);
    c = fma(c, -x2, 1.0 / 6227020800.0);
    c = fma(c, -x2, 1.0 / 39916800.0);
    c = fma(c, -x2, 1.0 / 362880.0);
    c = fma(c, -x2, 1.0 / 5040.0);
    c = fma(c, -x2, 1.0 / 120.0);
    c = fma(c, -x2, 1.0 / 6.0);
    return fma(-x * x2, c, x);
}

// x = cos(theta), theta = -p in [0, pi], from the half angle h = theta / 2
// (exact in float): 1 - 2 sin^2(h), or for h > pi/4, 2 cos^2(h) - 1 with
// cos(h) = sin(pi/2 - h), so that 1 - x (resp. 1 + x) keeps its relative
// accuracy next to theta = 0 (resp. pi), where the high orders need it; the
// sine in FP64 (a float sine puts ~1e-7 of phase error on every order, which
// the large multipliers of 50 moments turn into ~1e-4 of density)
__device__ __forceinline__ double cos_theta_of(float p)
{
    const double h = -0.5 * (double)p;
    const bool reflect = h > 0.78539816339744831;
    const double q = sin_quarter(reflect ? 1.57079632679489662 - h : h);
    const double q2 = 2.0 * q * q;
    return reflect ? q2 - 1.0 : 1.0 - q2;
}

template <int M, typename L>
__device__ __forceinline__ double series_clenshaw(const L (&lam)[M], float p, double *cos_theta = nullptr)
{
    // lambda_0 + 2 sum_{l=1}^{M-1} lambda_l cos(l theta), theta = -p in [0, pi]
    // (entries l >= m hold 0), by Clenshaw's recurrence in FP64 (2 operations
    // per term; gfx950 issues FP64 FMAs at the non-packed FP32 rate):
    //     b_l = lambda_l + 2x b_{l+1} - b_{l+2},  sum_{l>=1} = x b_1 - b_2.
    const double x = cos_theta_of(p);
    if (cos_theta)
        *cos_theta = x;
    const double x2 = x + x;
    double b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int l = M - 1; l >= 1; --l)
    {
        const double t = fma(x2, b1, (double)lam[l] - b2);
        b2 = b1;
        b1 = t;
    }
    return fma(2.0, fma(x, b1, -b2), (double)lam[0]);
}

// series_clenshaw at two phases in one pass, each recurrence with
// series_clenshaw's operations, so both values are series_clenshaw's bit for
// bit (the compiler runs the two recurrences one after the other; forced to
// interleave them term by term, the march took as long)
template <int M, typename L>
__device__ __forceinline__ void series_clenshaw2(const L (&lam)[M], float pa, float pb, double &sa, double &sb)
{
    const double xa = cos_theta_of(pa), xb = cos_theta_of(pb);
    const double x2a = xa + xa, x2b = xb + xb;
    double a1 = 0.0, a2 = 0.0, b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int l = M - 1; l >= 1; --l)
    {
        const double ta = fma(x2a, a1, (double)lam[l] - a2);
        const double tb = fma(x2b, b1, (double)lam[l] - b2);
        a2 = a1;
        a1 = ta;
        b2 = b1;
        b1 = tb;
    }
    sa = fma(2.0, fma(xa, a1, -a2), (double)lam[0]);
    sb = fma(2.0, fma(xb, b1, -b2), (double)lam[0]);
}

// fast_atan with the reciprocal from the hardware approximation (about one
// ulp) instead of the IEEE division and fused multiply-adds in the
// polynomial; the MIXED form only
__device__ __forceinline__ float fast_atan_rcp(float tan)
{
    const bool negative = (tan < 0.0f);
    const float abs_tan = negative ? (-tan) : tan;
    const bool greater_one = (abs_tan > 1.0f);
    float x = __builtin_amdgcn_rcpf(greater_one ? abs_tan : 1.0f);
    float y = greater_one ? 1.0f : abs_tan;
    x = x * y;
    y = x * x;
    float z = __builtin_fmaf(y, 0.020835f, -0.085133f);
    z = __builtin_fmaf(y, z, 0.180141f);
    z = __builtin_fmaf(y, z, -0.330299f);
    y = __builtin_fmaf(y, z, 0.999866f);
    z = __builtin_fmaf(y * x, -2.0f, 1.570796f);
    z = greater_one ? z : 0.0f;
    x = __builtin_fmaf(x, y, z);
    return negative ? (-x) : x;
}

// LagT: float multipliers, or FP64 ones kept from mray_hip_prepare_f64
template <int M, typename LagT = float>
__global__ __launch_bounds__(256) void reconstruct_mixed_kernel(DevCamera cam, RecArgs a,
                                                                const LagT *__restrict__ lag,
                                                                const uint32_t *__restrict__ index,
                                                                const uint32_t *__restrict__ bounds,
                                                                const float4 *__restrict__ tf_g,
                                                                float *__restrict__ rgba, int32_t *__restrict__ steps)
{
    __shared__ float4 tf[256];
    const int tid = threadIdx.x;
    tf[tid] = tf_g[tid];
    __syncthreads();

    const int lane = tid & 63;
    const uint32_t ntiles = ((a.W + 7) / 8) * ((a.H + 7) / 8);
    const uint32_t tile = xcd_tile_block(a.W, 4) * 4 + (tid >> 6);
    if (tile >= ntiles)
        return;
    int x, y;
    tile_pixel(tile, a.W, lane, x, y);
    if (x >= a.W || y >= a.H)
        return;
    const size_t pix = (size_t)y * a.W + x;
    const int nm = index ? min((int)(index[pix + 1] - index[pix]), a.m) : a.m;
    const size_t off = index ? (size_t)index[pix] : pix * (size_t)a.m;
    if (nm == 0)
    {
        float o[4];
        zero_ray(tf, o);
        store_rgba(rgba, pix, o[0], o[1], o[2], o[3]);
        if (steps)
            steps[pix] = 0;
        return;
    }
    // the M loads are unconditional (index clamped into the pixel's nm >= 1
    // entries, the surplus zeroed after), so that they are in flight together:
    // one conditional load and wait after another let L2 evict the rows'
    // cache lines between the loads (3.4x the multipliers' bytes from HBM)
    double lam[M];
#pragma unroll
    for (int l = 0; l < M; ++l)
    {
        const LagT v = lag[off + min(l, nm - 1)];
        lam[l] = (l < nm) ? (double)v : 0.0;
    }
    float blo = 0.f, bhi = 1.f;
    if (bounds)
    {
        uint32_t b = bounds[pix];
        blo = (b & 0xFFFFu) / 65535.f;
        bhi = (b >> 16) / 65535.f;
    }
    const float brange = bhi - blo;
    const bool fourier = a.fourier != 0;
    auto density_of = [&](double series) {
        const float s = (float)series;
        const float rho = fourier ? s : __builtin_fmaf(fast_atan_rcp(s), 0.318309886f, 0.5f);
        return __builtin_fmaf(rho, brange, blo);
    };
    auto recon = [&](float p) { return density_of(series_clenshaw<M>(lam, rclamp(p, (float)(-kPi), 0.0f))); };
    // the densities at p + dphi and p + 2 dphi (the march's float phase
    // steps), from one interleaved pass
    auto recon2 = [&](float p, float dphi, float &d1, float &d2) {
        const float p1 = p + dphi, p2 = p1 + dphi;
        double s1, s2;
        series_clenshaw2<M>(lam, rclamp(p1, (float)(-kPi), 0.0f), rclamp(p2, (float)(-kPi), 0.0f), s1, s2);
        d1 = density_of(s1);
        d2 = density_of(s2);
    };

    Ray r = pixel_ray(cam, a.W, a.H, x, y);
    float t_near, t_far;
    bool hit = intersect_aabb(r, a.dmin, a.dmax, t_near, t_far);
    t_near = rmax(t_near, cam.znear);
    int ns = 0;
    float Lr = 0.f, Lg = 0.f, Lb = 0.f, alpha = 0.f;
    if (hit && (t_far > 0))
    {
        float t = t_near;
        float p = (float)((double)((t - t_near) / (t_far - t_near)) * kPi - kPi);
        const float dphi = (float)(kPi * (double)(a.step / (t_far - t_near)));
        float T = 1.0f, tau = 0.0f;
        float fs = recon(p);
        const int cap = march_cap(t_near, t_far, a.step) + 1;
        // one step of the reference's loop on the back sample bs
        auto step = [&](float bs) {
            t += a.step;
            p += dphi;
            ++ns;
            const float4 c = tf_sample_mixed(tf, fs, bs);
            Lr = __builtin_fmaf(a.emission * c.x, T, Lr);
            Lg = __builtin_fmaf(a.emission * c.y, T, Lg);
            Lb = __builtin_fmaf(a.emission * c.z, T, Lb);
            tau += c.w;
            T = __expf(-tau);
            alpha = 1.0f - T;
            fs = bs;
            const bool stop = alpha >= a.ert;
            if (stop)
                alpha = 1.f;
            return stop;
        };
        // the loop (t <= t_far, at most cap steps, early termination) two
        // steps at a time: both back samples from one pass (the second unused
        // when the first step ends the march), half the loop's tests and
        // branches; above 96 moments one at a time (the second sample's
        // registers would cost occupancy)
        int it = 0;
        if constexpr (M <= 96)
        {
            while (t <= t_far && it < cap)
            {
                float b1, b2;
                recon2(p, dphi, b1, b2);
                if (step(b1))
                    break;
                ++it;
                if (!(t <= t_far && it < cap))
                    break;
                if (step(b2))
                    break;
                ++it;
            }
        }
        else
        {
            for (; t <= t_far && it < cap; ++it)
                if (step(recon(p + dphi)))
                    break;
        }
    }
    store_rgba(rgba, pix, Lr, Lg, Lb, alpha);
    if (steps)
        steps[pix] = ns;
}

// ---------------------------------------------------------------------------
// MIXED with the multipliers split over K lanes per ray (more moments than one
// thread's registers hold)
// ---------------------------------------------------------------------------

// T_64(x), T_63(x), T_128(x), T_127(x), T_192(x), T_191(x) by doubling in
// FP64 from x (T_2a = 2 T_a^2 - 1, T_2a-1 = 2 T_a T_a-1 - x,
// T_a+b = 2 T_a T_b - T_a-b): the Chebyshev polynomials of the same x the
// recurrences use, so that a chunk's Clenshaw sum continues the series
// exactly where the previous chunk stops (a start from an independent cos(j0
// theta) would differ from it by j0 times the error of x).
__device__ __forceinline__ void chunk_starts(double x, double (&tn)[4], double (&tn1)[4])
{
    double a = x, a1 = 1.0; // T_1, T_0
#pragma unroll
    for (int d = 0; d < 6; ++d)
    {
        const double n = fma(2.0 * a, a, -1.0);
        const double n1 = fma(2.0 * a, a1, -x);
        a = n;
        a1 = n1;
    }
    // a = T_64, a1 = T_63
    const double t128 = fma(2.0 * a, a, -1.0), t127 = fma(2.0 * a, a1, -x);
    const double t65 = fma(2.0 * x, a, -a1);
    tn[0] = 1.0;
    tn1[0] = x; // T_-1 = T_1
    tn[1] = a;
    tn1[1] = a1;
    tn[2] = t128;
    tn1[2] = t127;
    tn[3] = fma(2.0 * t128, a, -a);    // T_192 = 2 T_128 T_64 - T_64
    tn1[3] = fma(2.0 * t128, a1, -t65); // T_191 = 2 T_128 T_63 - T_65
}

// The series lambda_0 + 2 sum_l lambda_l cos(l theta), theta = -p, with the
// multipliers split over K adjacent lanes (lane k holds lambda_64k ..
// lambda_64k+63 in lam, zero past the pixel's count): each lane's Clenshaw
// sum over its chunk from T_64k, T_64k-1 of the shared x = cos(theta) (taken
// from the half angle as in series_clenshaw), the K partial sums added by an
// xor butterfly, so every lane of the ray gets the same value bit for bit.
// Also returns x and the lane's T_64k, T_64k-1 (for the reverse's forward
// recurrence over its chunk).
template <int K, typename L>
__device__ __forceinline__ double series_split(const L (&lam)[64], int k, float p, double &x_out, double &t_j0,
                                               double &t_j0m1)
{
    const double xx = cos_theta_of(p);
    double tn[4], tn1[4];
    chunk_starts(xx, tn, tn1);
    double s0 = tn[0], s1 = tn1[0];
#pragma unroll
    for (int c = 1; c < K; ++c)
    {
        s0 = (k == c) ? tn[c] : s0;
        s1 = (k == c) ? tn1[c] : s1;
    }
    const double x2 = xx + xx;
    double b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int i = 63; i >= 0; --i)
    {
        const double t = fma(x2, b1, (double)lam[i] - b2);
        b2 = b1;
        b1 = t;
    }
    // lane 0 takes lambda_0 / 2 off its sum, so that twice the total is
    // lambda_0 + 2 sum_{l>=1} lambda_l T_l
    const double half0 = (k == 0) ? 0.5 * (double)lam[0] : 0.0;
    double part = fma(b1, s0, fma(-b2, s1, -half0));
#pragma unroll
    for (int o = 1; o < K; o <<= 1)
        part += __shfl_xor(part, o, 64);
    x_out = xx;
    t_j0 = s0;
    t_j0m1 = s1;
    return 2.0 * part;
}

// K lanes per ray (K = 2 or 4, adjacent lanes), 64 / K rays per wavefront in
// 8 x (8 / K) pixel tiles; lane k holds multipliers 64k .. 64k + 63. Each lane
// sums its chunk by Clenshaw's recurrence from T_64k, T_64k-1; the K partial
// sums are added by an xor butterfly, which gives every lane of the ray the
// same series value bit for bit, so all K lanes take the same steps and only
// lane 0 writes.
template <int K, typename LagT = float>
__global__ __launch_bounds__(256) void reconstruct_mixed_split_kernel(DevCamera cam, RecArgs a,
                                                                      const LagT *__restrict__ lag,
                                                                      const uint32_t *__restrict__ index,
                                                                      const uint32_t *__restrict__ bounds,
                                                                      const float4 *__restrict__ tf_g,
                                                                      float *__restrict__ rgba,
                                                                      int32_t *__restrict__ steps)
{
    constexpr int M = 64;
    constexpr int RPT = 8 / K; // pixel rows per tile
    __shared__ float4 tf[256];
    const int tid = threadIdx.x;
    tf[tid] = tf_g[tid];
    __syncthreads();

    const int lane = tid & 63;
    const int k = lane % K, ray = lane / K;
    const int tiles_x = (a.W + 7) / 8;
    const uint32_t ntiles = (uint32_t)tiles_x * (uint32_t)((a.H + RPT - 1) / RPT);
    const uint32_t tile = xcd_tile_block(a.W, 4) * 4 + (tid >> 6);
    if (tile >= ntiles)
        return;
    const int x = (int)(tile % tiles_x) * 8 + (ray & 7);
    const int y = (int)(tile / tiles_x) * RPT + (ray >> 3);
    if (x >= a.W || y >= a.H)
        return; // all K lanes of the ray together
    const size_t pix = (size_t)y * a.W + x;
    const int nm = index ? min((int)(index[pix + 1] - index[pix]), a.m) : a.m;
    const size_t off = index ? (size_t)index[pix] : pix * (size_t)a.m;
    if (nm == 0)
    {
        if (k == 0)
        {
            float o[4];
            zero_ray(tf, o);
            store_rgba(rgba, pix, o[0], o[1], o[2], o[3]);
            if (steps)
                steps[pix] = 0;
        }
        return;
    }
    const int j0 = k * M;
    double lam[M];
#pragma unroll
    for (int i = 0; i < M; ++i)
    {
        const LagT v = lag[off + min(j0 + i, nm - 1)];
        lam[i] = (j0 + i < nm) ? (double)v : 0.0;
    }
    float blo = 0.f, bhi = 1.f;
    if (bounds)
    {
        uint32_t b = bounds[pix];
        blo = (b & 0xFFFFu) / 65535.f;
        bhi = (b >> 16) / 65535.f;
    }
    const float brange = bhi - blo;
    const bool fourier = a.fourier != 0;
    auto recon = [&](float p) {
        p = rclamp(p, (float)(-kPi), 0.0f);
        double xx, tj0, tj0m1;
        const float sv = (float)series_split<K>(lam, k, p, xx, tj0, tj0m1);
        float rho = fourier ? sv : __builtin_fmaf(fast_atan_rcp(sv), 0.318309886f, 0.5f);
        return __builtin_fmaf(rho, brange, blo);
    };

    Ray r = pixel_ray(cam, a.W, a.H, x, y);
    float t_near, t_far;
    bool hit = intersect_aabb(r, a.dmin, a.dmax, t_near, t_far);
    t_near = rmax(t_near, cam.znear);
    int ns = 0;
    float Lr = 0.f, Lg = 0.f, Lb = 0.f, alpha = 0.f;
    if (hit && (t_far > 0))
    {
        float t = t_near;
        float p = (float)((double)((t - t_near) / (t_far - t_near)) * kPi - kPi);
        const float dphi = (float)(kPi * (double)(a.step / (t_far - t_near)));
        float T = 1.0f, tau = 0.0f;
        float fs = recon(p);
        const int cap = march_cap(t_near, t_far, a.step) + 1;
        for (int it = 0; t <= t_far && it < cap; ++it)
        {
            t += a.step;
            p += dphi;
            ++ns;
            const float bs = recon(p);
            const float4 c = tf_sample_mixed(tf, fs, bs);
            Lr = __builtin_fmaf(a.emission * c.x, T, Lr);
            Lg = __builtin_fmaf(a.emission * c.y, T, Lg);
            Lb = __builtin_fmaf(a.emission * c.z, T, Lb);
            tau += c.w;
            T = __expf(-tau);
            alpha = 1.0f - T;
            fs = bs;
            if (alpha >= a.ert)
            {
                alpha = 1.f;
                break;
            }
        }
    }
    if (k == 0)
    {
        store_rgba(rgba, pix, Lr, Lg, Lb, alpha);
        if (steps)
            steps[pix] = ns;
    }
}

// ---------------------------------------------------------------------------
// ADJOINT of the MIXED march
// ---------------------------------------------------------------------------

// d fast_atan / dt of the polynomial itself (src/mese/complex_algebra.h:131-152):
// for |t| <= 1 it is x P(x^2), otherwise pi/2 - x P(x^2) with x = 1/|t|.
__device__ __forceinline__ float fast_atan_deriv(float t)
{
    const float at = fabsf(t);
    const bool greater_one = at > 1.0f;
    const float x = greater_one ? 1.0f / at : at;
    const float y = x * x;
    const float P = (((0.020835f * y - 0.085133f) * y + 0.180141f) * y - 0.330299f) * y + 0.999866f;
    const float dP = ((4.0f * 0.020835f * y - 3.0f * 0.085133f) * y + 2.0f * 0.180141f) * y - 0.330299f;
    const float dxp = P + 2.0f * y * dP; // d (x P(x^2)) / dx
    return greater_one ? dxp * y : dxp;
}

// Reverse mode of reconstruct_mixed_kernel for one ray, with the march
// recomputed (per-ray checkpointing: nothing per sample is stored). With the
// forward RGBA L known, the radiance that sample k still shades is L - P_k
// (P_k the prefix sum), so one pass in forward order gives
//     dL/dc_k     = g_rgb e T_k                       (emission of sample k)
//     dL/dsigma_k = -g_rgb . (L - P_k) + g_a T_end     (extinction of sample k)
// with T_k the transmittance in front of sample k and T_end = 1 - alpha
// (0 after early termination). The transfer-function lookup takes its
// entries from fs and its weight from frac(256 bs), so only bs carries a
// derivative: d c / d bs = 256 (T[x1] - T[x0]). Then
// d rho / d S = fast_atan'(S) / pi (hi - lo) and dS / d lambda_l = c_l cos(l theta)
// (c_0 = 1, c_l = 2), cos(l theta) by the Chebyshev recurrence in FP64.
// Transfer-function gradients are summed per workgroup in LDS and added to
// grad_tf with one atomic per entry and workgroup. The LDS sums are 64-bit
// fixed point (integer atomics: float atomics on LDS cost several times as
// much here, above all when the rays of a wavefront add to the same entry);
// the workgroup's scale 2^kfix comes from a bound on what each ray can add to
// one entry, so that no partial sum reaches 2^61.


// K > 1: K adjacent lanes per ray (M = 64 multipliers each, as
// reconstruct_mixed_split_kernel): every lane recomputes the ray's march with
// the shared series value, accumulates the gradient of its own multipliers,
// and only lane 0 adds to the table gradient.
template <int M, int K = 1>
__global__ __launch_bounds__(256) void reconstruct_adjoint_kernel(
    DevCamera cam, RecArgs a, const float *__restrict__ lag, const uint32_t *__restrict__ index,
    const uint32_t *__restrict__ bounds, const float4 *__restrict__ tf_g, const float4 *__restrict__ rgba_fwd,
    const float4 *__restrict__ grad_rgba, float *__restrict__ grad_lag, float *__restrict__ grad_tf)
{
    __shared__ float4 tf[256];
    // table gradient per workgroup, channel-major (ch * 256 + entry), fixed point
    __shared__ unsigned long long gtf[4 * 256];
    __shared__ uint32_t cmax_bits, bound_bits;
    const int tid = threadIdx.x;
    const float4 tv = tf_g[tid];
    tf[tid] = tv;
    if (grad_tf)
        for (int i = tid; i < 4 * 256; i += blockDim.x)
            gtf[i] = 0ull;
    if (tid == 0)
        cmax_bits = bound_bits = 0u;
    __syncthreads();
    // the largest colour of the table (non-negative floats order as their bits)
    atomicMax(&cmax_bits, __float_as_uint(fmaxf(fabsf(tv.x), fmaxf(fabsf(tv.y), fabsf(tv.z)))));

    static_assert(K == 1 || M == 64, "lane-split reverse holds 64 multipliers per lane");
    const int lane = tid & 63;
    const int k = lane % K; // chunk of the lane (0 for K = 1)
    const int j0 = k * M;
    constexpr int RPT = 8 / K; // pixel rows per tile
    const int tiles_x = (a.W + 7) / 8;
    const uint32_t ntiles = (uint32_t)tiles_x * (uint32_t)((a.H + RPT - 1) / RPT);
    const uint32_t tile = xcd_tile_block(a.W, 4) * 4 + (tid >> 6);
    int x = 0, y = 0;
    bool active = tile < ntiles;
    if (active)
    {
        const int ray = lane / K;
        x = (int)(tile % tiles_x) * 8 + (ray & 7);
        y = (int)(tile / tiles_x) * RPT + (ray >> 3);
        active = x < a.W && y < a.H;
    }
    const size_t pix = active ? (size_t)y * a.W + x : 0;
    const int nm = !active ? 0 : index ? min((int)(index[pix + 1] - index[pix]), a.m) : a.m;
    const float4 g = active ? grad_rgba[pix] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 L = active ? rgba_fwd[pix] : make_float4(0.f, 0.f, 0.f, 0.f);
    Ray r = pixel_ray(cam, a.W, a.H, x, y);
    float t_near, t_far;
    const bool hit = intersect_aabb(r, a.dmin, a.dmax, t_near, t_far);
    t_near = rmax(t_near, cam.znear);
    const int cap = march_cap(t_near, t_far, a.step) + 1;
    __syncthreads();
    // bound on what this ray adds to one table entry: per sample at most
    // |dL/dc| = |g_rgb e| T and |dL/dsigma| <= sum |g_ch| (|L_ch| + |P_ch|) + |g_a| T_end
    // (weights w01 + w23 = 1), at most cap samples; twice that for rounding
    float lane_bound = 0.0f;
    if (active && grad_tf && k == 0)
    {
        const float ga = fabsf(g.x) + fabsf(g.y) + fabsf(g.z);
        if (nm == 0)
            lane_bound = fmaxf(fmaxf(fabsf(g.x), fabsf(g.y)), fmaxf(fabsf(g.z), fabsf(g.w)));
        else if (hit && (t_far > 0))
        {
            const float e = fabsf(a.emission), n = (float)cap;
            const float pmax = e * __uint_as_float(cmax_bits) * n;
            const float gc = e * fmaxf(fmaxf(fabsf(g.x), fabsf(g.y)), fabsf(g.z));
            const float gs = fabsf(g.x) * fabsf(L.x) + fabsf(g.y) * fabsf(L.y) + fabsf(g.z) * fabsf(L.z) + ga * pmax +
                             fabsf(g.w) * fabsf(1.0f - L.w);
            lane_bound = 2.0f * n * fmaxf(gc, gs);
        }
    }
    if (grad_tf)
        atomicMax(&bound_bits, __float_as_uint(lane_bound));
    __syncthreads();
    // the same value in every lane: read as a scalar, so the branches on it
    // are uniform
    const float bsum = __uint_as_float(__builtin_amdgcn_readfirstlane(bound_bits)) * (float)blockDim.x;
    const int kfix = fixed_scale(bsum);
    // not finite (a non-finite incoming gradient): the workgroup's table
    // gradient is that value in every entry, rather than summands lost in
    // their conversion
    const bool fixed = fixed_usable(bsum);
    // adds v to channel ch of table entry e
    auto tf_add = [&](int ch, int e, float v) { atomicAdd(&gtf[ch * 256 + e], to_fixed(v, kfix)); };
    if (active)
    {
        const size_t off = index ? (size_t)index[pix] : pix * (size_t)a.m;
        if (nm == 0)
        {
            // reconstruct_zero_ray: the colour is the table at (0, 0) = entry 0
            if (grad_tf && k == 0)
            {
                tf_add(0, 0, g.x);
                tf_add(1, 0, g.y);
                tf_add(2, 0, g.z);
                tf_add(3, 0, g.w);
            }
        }
        else
        {
            float lam[M], glam[M];
#pragma unroll
            for (int l = 0; l < M; ++l)
            {
                const float v = lag[off + min(j0 + l, nm - 1)]; // unconditional, as in the forward
                lam[l] = (j0 + l < nm) ? v : 0.0f;
                glam[l] = 0.0f;
            }
            // the series at phase pc; c1 = cos(theta), tj0 / tj0m1 = T_j0, T_j0-1
            // of this lane's chunk (1 and cos(theta) for K = 1)
            auto series_at = [&](float pc, double &c1, double &tj0, double &tj0m1) {
                if constexpr (K == 1)
                {
                    const double sv = series_clenshaw<M>(lam, pc, &c1);
                    tj0 = 1.0;
                    tj0m1 = c1;
                    return sv;
                }
                else
                    return series_split<K>(lam, k, pc, c1, tj0, tj0m1);
            };
            float blo = 0.f, bhi = 1.f;
            if (bounds)
            {
                const uint32_t b = bounds[pix];
                blo = (b & 0xFFFFu) / 65535.f;
                bhi = (b >> 16) / 65535.f;
            }
            const float brange = bhi - blo;
            const bool fourier = a.fourier != 0;
            const double t_end = 1.0 - (double)L.w;
            const double ge[3] = {(double)g.x * a.emission, (double)g.y * a.emission, (double)g.z * a.emission};

            if (hit && (t_far > 0))
            {
                float t = t_near;
                float p = (float)((double)((t - t_near) / (t_far - t_near)) * kPi - kPi);
                const float dphi = (float)(kPi * (double)(a.step / (t_far - t_near)));
                float T = 1.0f, tau = 0.0f;
                auto rho_of = [&](float s) {
                    const float rho = fourier ? s : __builtin_fmaf(fast_atan_rcp(s), 0.318309886f, 0.5f);
                    return __builtin_fmaf(rho, brange, blo);
                };
                double c1, tj0, tj0m1;
                float fs = rho_of((float)series_at(rclamp(p, (float)(-kPi), 0.0f), c1, tj0, tj0m1));
                double pre[3] = {0.0, 0.0, 0.0};
                // table gradient of the current entry pair (x0, x1), not yet in gtf
                int tf_x0 = -1, tf_x1 = -1;
                float tf_a[4] = {0.f, 0.f, 0.f, 0.f}, tf_b[4] = {0.f, 0.f, 0.f, 0.f};
                auto tf_flush = [&]() {
                    if (tf_x0 < 0)
                        return;
#pragma unroll
                    for (int ch = 0; ch < 4; ++ch)
                    {
                        tf_add(ch, tf_x0, tf_a[ch]);
                        tf_add(ch, tf_x1, tf_b[ch]);
                        tf_a[ch] = tf_b[ch] = 0.0f;
                    }
                };
                for (int it = 0; t <= t_far && it < cap; ++it)
                {
                    t += a.step;
                    p += dphi;
                    const float pc = rclamp(p, (float)(-kPi), 0.0f);
                    const float S = (float)series_at(pc, c1, tj0, tj0m1);
                    const float bs = rho_of(S);
                    const float4 c = tf_sample_mixed(tf, fs, bs);

                    // indices and weights of tf_sample_mixed
                    const float ipx = fs * 256.0f, ipy = bs * 256.0f;
                    const float ay = ipy - floorf(ipy);
                    const int xl = (int)floorf(ipx);
                    const int x0 = iclamp(xl, 0, 255), x1 = iclamp(xl + 1, 0, 255);
                    const float w01 = 1.f - ay, w23 = ay;
                    const float4 lo4 = tf[x0], hi4 = tf[x1];

                    const double Td = (double)T;
                    const double gc[3] = {ge[0] * Td, ge[1] * Td, ge[2] * Td};
                    pre[0] = fma((double)(a.emission * c.x), Td, pre[0]);
                    pre[1] = fma((double)(a.emission * c.y), Td, pre[1]);
                    pre[2] = fma((double)(a.emission * c.z), Td, pre[2]);
                    const double gsig = -((double)g.x * ((double)L.x - pre[0]) + (double)g.y * ((double)L.y - pre[1]) +
                                          (double)g.z * ((double)L.z - pre[2])) +
                                        (double)g.w * t_end;
                    if (grad_tf && k == 0)
                    {
                        // consecutive samples that read the same table pair are
                        // summed in registers; the LDS atomics (contended: the
                        // rays of a wavefront read similar entries) run when
                        // the pair changes
                        if (x0 != tf_x0 || x1 != tf_x1)
                        {
                            tf_flush();
                            tf_x0 = x0;
                            tf_x1 = x1;
                        }
                        const float gv[4] = {(float)gc[0], (float)gc[1], (float)gc[2], (float)gsig};
#pragma unroll
                        for (int ch = 0; ch < 4; ++ch)
                        {
                            tf_a[ch] = __builtin_fmaf(gv[ch], w01, tf_a[ch]);
                            tf_b[ch] = __builtin_fmaf(gv[ch], w23, tf_b[ch]);
                        }
                    }
                    const double drho = 256.0 * (gc[0] * (double)(hi4.x - lo4.x) + gc[1] * (double)(hi4.y - lo4.y) +
                                                 gc[2] * (double)(hi4.z - lo4.z) + gsig * (double)(hi4.w - lo4.w));
                    const float dS = (float)drho * brange * (fourier ? 1.0f : fast_atan_deriv(S) * 0.318309886f);
                    if (dS != 0.0f)
                    {
                        // cos(l theta), theta = -pc, by the Chebyshev recurrence
                        // from the series' own cos(theta), over the lane's
                        // chunk l = j0 .. j0 + M - 1
                        double cm = tj0m1, cl = tj0;
#pragma unroll
                        for (int l = 0; l < M; ++l)
                        {
                            glam[l] = __builtin_fmaf(dS, (float)cl, glam[l]);
                            const double cn = fma(2.0 * c1, cl, -cm);
                            cm = cl;
                            cl = cn;
                        }
                    }
                    tau += c.w;
                    T = __expf(-tau);
                    fs = bs;
                    if (1.0f - T >= a.ert)
                        break;
                }
                if (grad_tf && k == 0)
                    tf_flush();
            }
            if (grad_lag)
            {
                // static indices keep glam in registers
#pragma unroll
                for (int l = 0; l < M; ++l)
                    if (j0 + l < nm)
                        grad_lag[off + j0 + l] = (j0 + l == 0) ? glam[l] : 2.0f * glam[l];
            }
        }
    }
    __syncthreads();
    if (grad_tf)
        for (int i = tid; i < 1024; i += blockDim.x)
        {
            // channel i / 256 of entry i % 256
            const float v = fixed ? from_fixed(gtf[i], kfix) : bsum;
            if (v != 0.0f)
                atomicAdd(&grad_tf[(i & 255) * 4 + (i >> 8)], v);
        }
}

int launch_reconstruct_adjoint(const mray_camera &cam, const float *dmin, const float *dmax, const float *lagrange,
                               const uint32_t *index, const uint32_t *bounds, int W, int H, int m, const float *tf,
                               const mray_rec_params &p, const float *rgba, const float *grad_rgba,
                               float *grad_lagrange, float *grad_tf, hipStream_t stream)
{
    RecArgs a;
    a.W = W;
    a.H = H;
    a.m = m;
    a.step = p.step_size;
    a.emission = p.medium_emission;
    a.ert = p.early_ray_termination;
    a.fourier = p.use_truncated_fourier;
    a.dmin = {dmin[0], dmin[1], dmin[2]};
    a.dmax = {dmax[0], dmax[1], dmax[2]};
    DevCamera c = make_dev_camera(cam);
    const uint32_t ntiles = (uint32_t)(((W + 7) / 8) * ((H + 7) / 8));
    const uint32_t blocks = (ntiles + 3) / 4;
    const float4 *tf4 = reinterpret_cast<const float4 *>(tf);
    const float4 *fw = reinterpret_cast<const float4 *>(rgba);
    const float4 *gr = reinterpret_cast<const float4 *>(grad_rgba);
#define MRAY_ADJ_CASE(MM)                                                                                             \
    if (m <= MM)                                                                                                      \
    {                                                                                                                 \
        hipLaunchKernelGGL(reconstruct_adjoint_kernel<MM>, dim3(blocks), dim3(256), 0, stream, c, a, lagrange, index, \
                           bounds, tf4, fw, gr, grad_lagrange, grad_tf);                                              \
        return 0;                                                                                                     \
    }
    MRAY_ADJ_CASE(8)
    MRAY_ADJ_CASE(16)
    MRAY_ADJ_CASE(24)
    MRAY_ADJ_CASE(32)
    MRAY_ADJ_CASE(40)
    MRAY_ADJ_CASE(48)
    MRAY_ADJ_CASE(56)
    MRAY_ADJ_CASE(64)
#undef MRAY_ADJ_CASE
    // 65..256: two or four lanes per ray, 64 multipliers each
    if (m <= 256)
    {
        if (m <= 128)
        {
            const uint32_t tiles2 = (uint32_t)(((W + 7) / 8) * ((H + 3) / 4));
            hipLaunchKernelGGL((reconstruct_adjoint_kernel<64, 2>), dim3((tiles2 + 3) / 4), dim3(256), 0, stream, c,
                               a, lagrange, index, bounds, tf4, fw, gr, grad_lagrange, grad_tf);
        }
        else
        {
            const uint32_t tiles4 = (uint32_t)(((W + 7) / 8) * ((H + 1) / 2));
            hipLaunchKernelGGL((reconstruct_adjoint_kernel<64, 4>), dim3((tiles4 + 3) / 4), dim3(256), 0, stream, c,
                               a, lagrange, index, bounds, tf4, fw, gr, grad_lagrange, grad_tf);
        }
        return 0;
    }
    return MRAY_ERR_UNSUPPORTED;
}

namespace
{
RecArgs rec_args(const float *dmin, const float *dmax, int W, int H, int m, const mray_rec_params &p)
{
    RecArgs a;
    a.W = W;
    a.H = H;
    a.m = m;
    a.step = p.step_size;
    a.emission = p.medium_emission;
    a.ert = p.early_ray_termination;
    a.fourier = p.use_truncated_fourier;
    a.dmin = {dmin[0], dmin[1], dmin[2]};
    a.dmax = {dmax[0], dmax[1], dmax[2]};
    return a;
}

// the MIXED march on float or FP64 multipliers
template <typename LagT>
int launch_mixed_march(const DevCamera &c, const RecArgs &a, const LagT *lagrange, const uint32_t *index,
                       const uint32_t *bounds, const float4 *tf4, float *rgba, int32_t *steps, hipStream_t stream)
{
    const int W = a.W, H = a.H, m = a.m;
    const uint32_t ntiles = (uint32_t)(((W + 7) / 8) * ((H + 7) / 8));
    const uint32_t blocks = (ntiles + 3) / 4;
#define MRAY_REC_CASE(MM)                                                                                             \
    if (m <= MM)                                                                                                      \
    {                                                                                                                 \
        hipLaunchKernelGGL((reconstruct_mixed_kernel<MM, LagT>), dim3(blocks), dim3(256), 0, stream, c, a, lagrange,  \
                           index, bounds, tf4, rgba, steps);                                                          \
        return 0;                                                                                                     \
    }
    MRAY_REC_CASE(4)
    MRAY_REC_CASE(8)
    MRAY_REC_CASE(12)
    MRAY_REC_CASE(16)
    MRAY_REC_CASE(20)
    MRAY_REC_CASE(24)
    MRAY_REC_CASE(28)
    MRAY_REC_CASE(30)
    MRAY_REC_CASE(32)
    MRAY_REC_CASE(36)
    MRAY_REC_CASE(40)
    MRAY_REC_CASE(44)
    MRAY_REC_CASE(48)
    MRAY_REC_CASE(50)
    MRAY_REC_CASE(52)
    MRAY_REC_CASE(56)
    MRAY_REC_CASE(60)
    MRAY_REC_CASE(64)
    MRAY_REC_CASE(72)
    MRAY_REC_CASE(80)
    MRAY_REC_CASE(88)
    MRAY_REC_CASE(96)
    MRAY_REC_CASE(104)
    MRAY_REC_CASE(112)
    MRAY_REC_CASE(120)
    MRAY_REC_CASE(128)
#undef MRAY_REC_CASE
    if (m <= 256)
    {
        // four lanes per ray, 16 rays per wavefront in 8 x 2 tiles
        const uint32_t tiles4 = (uint32_t)(((W + 7) / 8) * ((H + 1) / 2));
        hipLaunchKernelGGL((reconstruct_mixed_split_kernel<4, LagT>), dim3((tiles4 + 3) / 4), dim3(256), 0, stream, c,
                           a, lagrange, index, bounds, tf4, rgba, steps);
        return 0;
    }
    return MRAY_ERR_UNSUPPORTED;
}
} // namespace

int launch_reconstruct(const mray_camera &cam, const float *dmin, const float *dmax, const float *lagrange,
                       const uint32_t *index, const uint32_t *bounds, int W, int H, int m, const float *tf,
                       const mray_rec_params &p, int form, float *rgba, int32_t *steps, hipStream_t stream)
{
    const RecArgs a = rec_args(dmin, dmax, W, H, m, p);
    const DevCamera c = make_dev_camera(cam);
    const float4 *tf4 = reinterpret_cast<const float4 *>(tf);
    if (form == MRAY_FORM_PLAIN)
    {
        const uint32_t ntiles = (uint32_t)(((W + 7) / 8) * ((H + 7) / 8));
        int bs = (256 * m * (int)sizeof(float) <= 60000) ? 256 : 64;
        hipLaunchKernelGGL(reconstruct_plain_kernel, dim3((ntiles + bs / 64 - 1) / (bs / 64)), dim3(bs),
                           (size_t)bs * m * sizeof(float), stream, c, a, lagrange, index, bounds, tf4, rgba, steps);
        return 0;
    }
    return launch_mixed_march<float>(c, a, lagrange, index, bounds, tf4, rgba, steps, stream);
}

int launch_reconstruct_f64(const mray_camera &cam, const float *dmin, const float *dmax, const double *lagrange,
                           const uint32_t *index, const uint32_t *bounds, int W, int H, int m, const float *tf,
                           const mray_rec_params &p, float *rgba, int32_t *steps, hipStream_t stream)
{
    return launch_mixed_march<double>(make_dev_camera(cam), rec_args(dmin, dmax, W, H, m, p), lagrange, index, bounds,
                                      reinterpret_cast<const float4 *>(tf), rgba, steps, stream);
}

} // namespace mray
