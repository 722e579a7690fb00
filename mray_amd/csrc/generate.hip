// Moment generation: ray march accumulating bounded trigonometric moments of
// the density along each ray (src/raymarch_generate.h:35-148).
//
// Two forms share the sample loop (pixel ray, slab test, t loop, trilinear
// fetch, map_density, min/max bound) so both see exactly the reference's
// samples and produce its 16-bit bounds bit for bit:
//
//  PLAIN  the reference's per-segment update_moments: 2(m-1) sin/cos per
//         sample, one thread per ray, accumulators in LDS (what a direct port
//         of generate_moments_kernel does).
//
//  MIXED  the same integral written per node instead of per segment. For
//         contiguous segments with slopes g_s the i/j terms telescope and
//             trig_j = (1/j^2) sum_k (g_k - g_{k+1}) cos(j phi_k)
//                      - (f_0 / j) sin(j phi_0).
//         The reference's segments are [p_k - dphi, p_k] with p_k recomputed
//         from t, so neighbours overlap by a few ulp (delta), and its lines
//         use the slope (f_k - f_{k-1}) / dphi, so they miss f_k by a few ulp
//         (eps); both enter to first order, exactly enough to reproduce the
//         reference's arithmetic-free result to ~1e-13:
//             + sum_k [ -delta_k f_k cos(j p_k) + h_k sin(j p_k) / j ].
//         A node costs cos and sin of its (float) phase and two Chebyshev
//         recurrences over j (cos and sin of j phi): five operations per
//         moment (accumulate_node) instead of 2(m-1) sin/cos. gfx950 issues
//         FP64 FMAs at the rate of unpacked FP32 ones, so the cosine
//         recurrence and sums (the moments' bulk) and trig_0 (the exact
//         integral of the reference's lines) are FP64; the sine terms carry
//         only the first-order ulp corrections and stay FP32. The moments are
//         rounded to float once. The node's phasor is the previous one turned
//         by the exact FP64 difference of the two float phases (the
//         reference's phases jitter by ulps around the uniform grid, so a
//         fixed step would drift off cos(p_k)). Moments live in registers
//         (template M >= m); more than 60 are split into chunks over
//         workgroups (CHUNK).
#include "device_common.hpp"
#include "launch.hpp"
#include "sph_common.hpp"
#include <type_traits>

namespace mray
{

struct GenArgs
{
    int W, H, m;
    float step;
    float data_min, data_max;
    float data_rinv; // host_reciprocal(data_max - data_min)
    int transform_to_bounds;
    int compact_image;
    float bound_eps;
};

// Sample at t along r (origin + t * dir in float, as the reference).
template <typename T>
__device__ __forceinline__ float sample_at(const VolumeView<T> &vol, const Ray &r, float t, const GenArgs &a)
{
    float px = r.o.x + t * r.d.x, py = r.o.y + t * r.d.y, pz = r.o.z + t * r.d.z;
    return map_density(sample_volume(vol, px, py, pz), a.data_min, a.data_max);
}

// ---------------------------------------------------------------------------
// PLAIN
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void generate_plain_kernel(VolumeView<T> vol, DevCamera cam, GenArgs a,
                                                             float *__restrict__ out, uint32_t *__restrict__ bounds,
                                                             int32_t *__restrict__ samples,
                                                             float2 *__restrict__ ray_range)
{
    extern __shared__ float lds_trig[]; // [m][blockDim.x], strided per thread
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int stride = blockDim.x;
    const uint32_t ntiles = ((a.W + 7) / 8) * ((a.H + 7) / 8);
    const uint32_t tile = xcd_tile_block(a.W, blockDim.x / 64) * (blockDim.x / 64) + (tid >> 6);
    if (tile >= ntiles)
        return;
    int x, y;
    tile_pixel(tile, a.W, lane, x, y);
    if (x >= a.W || y >= a.H)
        return;
    const int m = a.m;
    float *trig = lds_trig + tid;
#define TRIG(j) trig[(j)*stride]
    for (int j = 0; j < m; ++j)
        TRIG(j) = 0.0f;

    Ray r = pixel_ray(cam, a.W, a.H, x, y);
    float t_near, t_far;
    bool hit = intersect_aabb(r, vol.dmin, vol.dmax, t_near, t_far);
    t_near = rmax(t_near, cam.znear);
    size_t pix = (size_t)y * a.W + x;
    float *mo = out + pix * (size_t)m;
    if (!(hit && (t_far > 0)))
    {
        if (!(a.transform_to_bounds && a.compact_image))
            for (int j = 0; j < m; ++j)
                mo[j] = 0.f;
        if (bounds)
            bounds[pix] = 0;
        if (ray_range)
            ray_range[pix] = make_float2(0.0f, 1.0f);
        if (samples)
            samples[pix] = 0;
        return;
    }
    const float s = a.step;
    float lo = 1.f, hi = 0.f;
    float t = t_near;
    float pi_trange_inv = (float)(kPi / (double)(t_far - t_near));
    float p = (float)(-kPi);
    float dphi = (float)(kPi * (double)(s / (t_far - t_near)));
    float fs = sample_at(vol, r, t, a);
    int ns = 1;
    lo = rmin(lo, fs);
    hi = rmax(hi, fs);

    auto update = [&](float p_next, float step_size, float f0, float f1) {
        // update_moments (src/raymarch_generate.h:35-58)
        float p_prev = p_next - step_size;
        float gradient = (f1 - f0) / step_size;
        float yint = f0 - gradient * p_prev;
        for (int j = 1; j < m; ++j)
        {
            float jf = (float)j;
            float jSq = jf * jf;
            float csx = gradient / jSq, csy = yint / jf;
            float sn, cn, sp, cp;
            sincosf(-jf * p_next, &sn, &cn);
            sincosf(-jf * p_prev, &sp, &cp);
            float ay = csy + gradient * jf * p_next / jSq;
            float by = csy + gradient * jf * p_prev / jSq;
            float acc = TRIG(j);
            acc += (csx + 0.0f) * cn - ay * sn;
            acc -= (csx + 0.0f) * cp - by * sp;
            TRIG(j) = acc;
        }
        float t0 = TRIG(0);
        t0 = (float)((double)t0 + (0.5 * gradient * (double)(p_next * p_next) + (double)(yint * p_next)));
        t0 = (float)((double)t0 - (0.5 * gradient * (double)(p_prev * p_prev) + (double)(yint * p_prev)));
        TRIG(0) = t0;
    };

    float bs;
    const int cap = march_cap(t_near, t_far, s);
    for (int it = 0; t + s < t_far && it < cap; ++it)
    {
        t += s;
        p = (float)((double)((t - t_near) * pi_trange_inv) - kPi);
        bs = sample_at(vol, r, t, a);
        ++ns;
        lo = rmin(lo, bs);
        hi = rmax(hi, bs);
        update(p, dphi, fs, bs);
        fs = bs;
    }
    float last_step = 0.0f - p;
    bs = sample_at(vol, r, t_far, a);
    ++ns;
    lo = rmin(lo, bs);
    hi = rmax(hi, bs);
    update(0.0f, last_step, fs, bs);

    // normalisation (src/raymarch_generate.h:116-135)
    if (a.transform_to_bounds)
    {
        lo = rmax(lo - a.bound_eps, 0.0f);
        hi = rmin(hi + a.bound_eps, 1.0f);
        double inv = 1.0 / (double)(hi - lo);
        mo[0] = (float)((double)(TRIG(0) * kInvPiF - lo) * inv);
        for (int j = 1; j < m; ++j)
            mo[j] = (float)((double)(TRIG(j) * kInvPiF) * inv);
    }
    else
        for (int j = 0; j < m; ++j)
            mo[j] = TRIG(j) * kInvPiF;
#undef TRIG
    if (bounds)
        bounds[pix] = pack_bound(lo, hi);
    if (ray_range)
        ray_range[pix] = make_float2(lo, hi);
    if (samples)
        samples[pix] = ns;
}

// ---------------------------------------------------------------------------
// MIXED
// ---------------------------------------------------------------------------

// One node's terms for moments j = 1..M-1:
//     sc[j] += (d + j^2 e) cos(j phi)   (FP64)
//     ss[j] += h sin(j phi)             (FP32)
// the moment being (sc[j] + j ss[j]) / j^2. cos(j phi) by the Chebyshev
// recurrence v_j = 2x v_{j-1} - v_{j-2} in FP64 from x = cos(phi) (gfx950
// issues FP64 FMAs at the rate of unpacked FP32 ones); the sine terms carry
// only the first-order corrections of the reference's ulp offsets (h is a few
// ulps of the slopes), so their recurrence from y = sin(phi) and their sums
// are FP32. The weight d + j^2 e is built up by the odd numbers,
// w_j = w_{j-1} + (2j - 1) e, one FMA whose small odd factor is an
// instruction literal (an FMA by j^2 into a fresh register would need a copy
// of d first: gfx950's three-operand forms take no literal): five operations
// and three registers per moment.
template <int M>
__device__ __forceinline__ void accumulate_node(double (&sc)[M], float (&ss)[M], double d, double e, float h,
                                                double x, float y)
{
    const double x2 = x + x;
    const float x2f = (float)x2;
    double cm2 = 1.0, cm1 = x;
    float sm2 = 0.0f, sm1 = y;
    double w = d + e;
    sc[1] = fma(w, cm1, sc[1]);
    ss[1] = __builtin_fmaf(h, sm1, ss[1]);
#pragma unroll
    for (int j = 2; j < M; ++j)
    {
        const double c = fma(x2, cm1, -cm2);
        const float sn = __builtin_fmaf(x2f, sm1, -sm2);
        w = fma((double)(2 * j - 1), e, w);
        sc[j] = fma(w, c, sc[j]);
        ss[j] = __builtin_fmaf(h, sn, ss[j]);
        cm2 = cm1;
        cm1 = c;
        sm2 = sm1;
        sm1 = sn;
    }
}

// accumulate_node for the moment chunk j = j0 + i, i = 0..M-1, of a march
// whose moments are split into chunks (more than fit in one thread's
// registers): (c0, s0) and (c1, s1) are cos and sin of j0 phi and (j0 + 1) phi,
// the recurrences run from there. The weight d + j^2 e = D + i E2 + i^2 e
// (D = d + j0^2 e, E2 = 2 j0 e) is built up as w_i = w_{i-1} + (2i - 1) e
// + E2: one operation per moment more than accumulate_node's. Slot i = 0 of
// the chunk j0 = 0 (moment 0) gets d, which nothing reads.
template <int M>
__device__ __forceinline__ void accumulate_node_chunk(double (&sc)[M], float (&ss)[M], double d, double e, float h,
                                                      double x, double c0, float s0, double c1, float s1, double j0)
{
    const double x2 = x + x;
    const float x2f = (float)x2;
    const double D = fma(j0 * j0, e, d);
    const double E2 = (j0 + j0) * e;
    double cm2 = c0, cm1 = c1;
    float sm2 = s0, sm1 = s1;
    double w = D;
    sc[0] = fma(w, c0, sc[0]);
    ss[0] = __builtin_fmaf(h, s0, ss[0]);
#pragma unroll
    for (int i = 1; i < M; ++i)
    {
        double c;
        float sn;
        if (i == 1)
        {
            c = c1;
            sn = s1;
        }
        else
        {
            c = fma(x2, cm1, -cm2);
            sn = __builtin_fmaf(x2f, sm1, -sm2);
            cm2 = cm1;
            cm1 = c;
            sm2 = sm1;
            sm1 = sn;
        }
        w = fma((double)(2 * i - 1), e, w) + E2;
        sc[i] = fma(w, c, sc[i]);
        ss[i] = __builtin_fmaf(h, sn, ss[i]);
    }
}

// accumulate_node_chunk for a chunk start J0 known at compile time: the
// weight d + j^2 e continues the odd-number recurrence of accumulate_node
// from D = d + J0^2 e, w_j = w_{j-1} + (2j - 1) e with the odd factor an
// instruction literal, five operations per moment as in the first chunk.
template <int M, int J0>
__device__ __forceinline__ void accumulate_node_chunk_c(double (&sc)[M], float (&ss)[M], double d, double e, float h,
                                                        double x, double c0, float s0, double c1, float s1)
{
    const double x2 = x + x;
    const float x2f = (float)x2;
    double cm2 = c0, cm1 = c1;
    float sm2 = s0, sm1 = s1;
    double w = fma((double)(J0 * J0), e, d);
    sc[0] = fma(w, c0, sc[0]);
    ss[0] = __builtin_fmaf(h, s0, ss[0]);
#pragma unroll
    for (int i = 1; i < M; ++i)
    {
        double c;
        float sn;
        if (i == 1)
        {
            c = c1;
            sn = s1;
        }
        else
        {
            c = fma(x2, cm1, -cm2);
            sn = __builtin_fmaf(x2f, sm1, -sm2);
            cm2 = cm1;
            cm1 = c;
            sm2 = sm1;
            sm1 = sn;
        }
        w = fma((double)(2 * (J0 + i) - 1), e, w);
        sc[i] = fma(w, c, sc[i]);
        ss[i] = __builtin_fmaf(h, sn, ss[i]);
    }
}

// CHUNK: the moments are split into chunks of M (more than one thread's
// registers hold); each chunk's workgroups march the same rays and
// accumulate moments j0 .. j0 + M - 1. The first chunk is the unchunked
// kernel (M moments of an image of m > M: it writes moments 0 .. M - 1, the
// bounds, the ray range and the sample counts, and needs no per-node phasor
// of j0 phi); the CHUNK kernel runs the later chunks, j0 = (blockIdx.y + 1) M,
// or j0 = J0C when that is nonzero (the second of two chunks: its weights
// take literal odd factors, and cos and sin of j0 phi follow the node's
// phasor by a rotation instead of a sincos per node).
template <typename T, int M, bool OFF32, bool POW2, bool CHUNK = false, int J0C = 0>
__global__ __launch_bounds__(256) void generate_mixed_kernel(VolumeView<T> vol, DevCamera cam, GenArgs a,
                                                             float *__restrict__ out, uint32_t *__restrict__ bounds,
                                                             int32_t *__restrict__ samples,
                                                             float2 *__restrict__ ray_range)
{
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const uint32_t ntiles = ((a.W + 7) / 8) * ((a.H + 7) / 8);
    const uint32_t tile = xcd_tile_block(a.W, 4) * 4 + (tid >> 6);
    if (tile >= ntiles)
        return;
    int x, y;
    tile_pixel(tile, a.W, lane, x, y);
    if (x >= a.W || y >= a.H)
        return;
    const int m = a.m;
    const int j0 = CHUNK ? (J0C ? J0C : (int)(blockIdx.y + 1) * M) : 0;
    const bool first_chunk = !CHUNK;
    const int jend = CHUNK ? min(m, j0 + M) : m;
    Ray r = pixel_ray(cam, a.W, a.H, x, y);
    float t_near, t_far;
    bool hit = intersect_aabb(r, vol.dmin, vol.dmax, t_near, t_far);
    t_near = rmax(t_near, cam.znear);
    size_t pix = (size_t)y * a.W + x;
    float *mo = out + pix * (size_t)m;
    if (!(hit && (t_far > 0)))
    {
        if (!(a.transform_to_bounds && a.compact_image))
            for (int j = j0; j < jend; ++j)
                mo[j] = 0.f;
        if (first_chunk)
        {
            if (bounds)
                bounds[pix] = 0;
            if (ray_range)
                ray_range[pix] = make_float2(0.0f, 1.0f);
            if (samples)
                samples[pix] = 0;
        }
        return;
    }

    const float s = a.step;
    const float pi_trange_inv = (float)(kPi / (double)(t_far - t_near));
    const float dphi = (float)(kPi * (double)(s / (t_far - t_near)));

    float t = t_near;
    float f_prev = sample_at(vol, r, t, a);
    const float f_first = f_prev;
    float lo = rmin(1.f, f_prev), hi = rmax(0.f, f_prev);
    int ns = 1;

    // phase of node 0: start of the first segment, p_1 - dphi (or -pi when the
    // ray has no interior sample), as update_moments forms p_prev.
    float node_phase;
    if (t + s < t_far)
    {
        float t1 = t + s;
        float p1 = (float)((double)((t1 - t_near) * pi_trange_inv) - kPi);
        node_phase = p1 - dphi;
    }
    else
        node_phase = (float)(-kPi);
    // cos and sin of the node phase (FP64, of the float phase)
    double zs, zc;
    sincos_small((double)node_phase, zs, zc);
    const double z0s = zs, z0c = zc;
    // CHUNK: cos and sin of j0 phi and (j0 + 1) phi at the current node (the
    // product j0 phi is exact in FP64)
    const double j0d = (double)j0;
    double kc = 1.0, ks = 0.0, kc1 = zc, ks1 = zs;
    auto chunk_start = [&](float ph) {
        if constexpr (CHUNK)
        {
            sincos_small(j0d * (double)ph, ks, kc);
            kc1 = fma(kc, zc, -ks * zs);
            ks1 = fma(ks, zc, kc * zs);
        }
    };
    chunk_start(node_phase);
    const double q0s = ks, q0s1 = ks1;

    // sums of the cosine (FP64) and sine (FP32) terms
    double sc[M];
    float ss[M];
#pragma unroll
    for (int j = 0; j < M; ++j)
    {
        sc[j] = 0.0;
        ss[j] = 0.0f;
    }
    auto accumulate = [&](double d, double e, double h) {
        if constexpr (CHUNK && J0C > 0)
            accumulate_node_chunk_c<M, J0C>(sc, ss, d, e, (float)h, zc, kc, (float)ks, kc1, (float)ks1);
        else if constexpr (CHUNK)
            accumulate_node_chunk<M>(sc, ss, d, e, (float)h, zc, kc, (float)ks, kc1, (float)ks1, j0d);
        else
            accumulate_node<M>(sc, ss, d, e, (float)h, zc, (float)zs);
    };
    double trap = 0.0;
    float g_left = 0.0f;    // slope of the segment ending at the current node
    double eps_left = 0.0;  // its width minus dphi
    int node = 0;

    const int cap = march_cap(t_near, t_far, s);
    const float rdphi = host_reciprocal(dphi); // for the segment slopes
    // the next sample's loads are issued one step ahead (at t + s while the
    // march continues, else at t_far for the last segment)
    auto issue_at = [&](PendingSample<T, OFF32, POW2> &ps, float tt) {
        ps.issue(vol, r.o.x + tt * r.d.x, r.o.y + tt * r.d.y, r.o.z + tt * r.d.z);
    };
    PendingSample<T, OFF32, POW2> pending;
    auto density = [&](float sv) { return map_density_r<POW2>(sv, a.data_min, a.data_max, a.data_rinv); };
    // cos and sin of the next node's phase p: those of the current node turned
    // by dp = p - node_phase, exact in FP64 (the difference of two floats
    // whose magnitudes are within 2^29 of each other), with sin dp and cos dp
    // from their Taylor series to dp^5 and dp^6 (|dp| <= 2^-6: truncation
    // below 5e-17 of sin and 1e-19 of cos); the rotation's rounding adds
    // ~1e-16 per node. Larger steps (rays of a few dozen samples) and the
    // march of 60 moments (whose registers are full) take the direct sincos.
    //
    // A chunk from J0C (compile time) turns cos and sin of j0 phi by j0 dp
    // likewise (j0 dp rounded once), with sincos_quarter (|j0 dp| <= 50 / 64
    // < pi / 4); otherwise, or past that step, chunk_next takes them from a
    // sincos of j0 p.
    constexpr bool kRotate = M <= 56;
    bool chunk_turned = false;
    auto advance_phase = [&](float p) {
        const double dp = (double)p - (double)node_phase;
        chunk_turned = false;
        if (kRotate && fabs(dp) <= 0.015625)
        {
            const double d2 = dp * dp;
            const double sd = fma(dp * d2, fma(d2, 1.0 / 120.0, -1.0 / 6.0), dp);
            const double cd = fma(fma(fma(d2, -1.0 / 720.0, 1.0 / 24.0), d2, -0.5), d2, 1.0);
            const double nc = fma(zc, cd, -zs * sd);
            zs = fma(zs, cd, zc * sd);
            zc = nc;
            if constexpr (CHUNK && J0C > 0)
            {
                static_assert(J0C <= 50, "the turn by J0C dp must stay within pi / 4");
                double sj, cj;
                sincos_quarter((double)J0C * dp, sj, cj);
                const double nkc = fma(kc, cj, -ks * sj);
                ks = fma(ks, cj, kc * sj);
                kc = nkc;
                chunk_turned = true;
            }
        }
        else
            sincos_small((double)p, zs, zc);
    };
    // cos and sin of j0 p and (j0 + 1) p at the node just reached
    auto chunk_next = [&](float ph) {
        if constexpr (CHUNK)
        {
            if (!chunk_turned)
                sincos_small(j0d * (double)ph, ks, kc);
            kc1 = fma(kc, zc, -ks * zs);
            ks1 = fma(ks, zc, kc * zs);
        }
    };
    issue_at(pending, (t + s < t_far) ? t + s : t_far);
    for (int it = 0; t + s < t_far && it < cap; ++it)
    {
        t += s;
        const float p = (float)((double)((t - t_near) * pi_trange_inv) - kPi);
        const float f = density(pending.value());
        // the loaded values die here: keeping the next loads below this point
        // lets them land in the same registers (no copies that would wait
        // for them at the end of the iteration)
        __builtin_amdgcn_sched_barrier(0);
        issue_at(pending, (t + s < t_far && it + 1 < cap) ? t + s : t_far);
        ++ns;
        lo = rmin(lo, f);
        hi = rmax(hi, f);
        // the reference's segment [u, p]: u = p - dphi in float, slope (f - f_prev) / dphi
        const float u = p - dphi;
        const float g = div_rn(f - f_prev, dphi, rdphi);
        const double width = (double)p - (double)u;
        trap += width * ((double)f_prev + 0.5 * (double)g * width);
        // the offset of the segment start from the node (exact: u and the
        // node phase are floats a few ulps apart) and the terms of the node
        const double delta = (node > 0) ? (double)u - (double)node_phase : 0.0;
        const double h = fma((double)g, delta, (double)g_left * eps_left);
        // a node with d = e = h = 0 (inside a stretch of zero density:
        // both slopes and the start value vanish) adds exact zeros; where
        // every lane of the wavefront has one, the update is skipped
        {
            const double dn = (double)g_left - (double)g, en = -delta * (double)f_prev;
            if (dn != 0.0 || en != 0.0 || h != 0.0)
                accumulate(dn, en, h);
        }
        g_left = g;
        eps_left = width - (double)dphi;
        f_prev = f;
        advance_phase(p);
        node_phase = p;
        ++node;
        chunk_next(p);
    }
    // last segment [p_K, 0] (src/raymarch_generate.h:107-114): it starts
    // exactly at p_K and its line ends exactly at f_last
    const float f_last = density(pending.value());
    ++ns;
    lo = rmin(lo, f_last);
    hi = rmax(hi, f_last);
    const float dl = 0.0f - node_phase;
    const float g_last = (f_last - f_prev) / dl;
    trap += (double)dl * ((double)f_prev + 0.5 * (double)g_last * (double)dl);
    accumulate((double)g_left - (double)g_last, 0.0, (double)g_left * eps_left);
    // node at phase 0: cos(j*0) = 1, sin(j*0) = 0
#pragma unroll
    for (int j = CHUNK ? 0 : 1; j < M; ++j)
        sc[j] += (double)g_last;

    // trig_j = (sc_j + j ss_j) / j^2 - f_0 sin(j phi_0) / j in FP64, sin by
    // Chebyshev from phi_0 (from j0 phi_0 and (j0 + 1) phi_0 for a chunk),
    // then moment_j = trig_j / pi, normalised and rounded to float once
    double trig[M];
    {
        const double x2 = z0c + z0c;
        double sm2 = CHUNK ? q0s : 0.0, sm1 = CHUNK ? q0s1 : z0s;
#pragma unroll
        for (int i = CHUNK ? 0 : 1; i < M; ++i)
        {
            double sj;
            if (CHUNK && i == 0)
                sj = q0s;
            else if (i == 1)
                sj = CHUNK ? q0s1 : z0s;
            else
            {
                sj = fma(x2, sm1, -sm2);
                sm2 = sm1;
                sm1 = sj;
            }
            const double jd = j0d + (double)i;
            const double num = fma(jd, (double)ss[i], sc[i]);
            trig[i] = (jd == 0.0) ? 0.0 : num / (jd * jd) - (double)f_first * sj / jd;
        }
    }
    if (first_chunk)
        trig[0] = trap;

    uint32_t b;
    constexpr double kInvPi = 0.31830988618379067154;
    if (a.transform_to_bounds)
    {
        lo = rmax(lo - a.bound_eps, 0.0f);
        hi = rmin(hi + a.bound_eps, 1.0f);
        const double inv = 1.0 / (double)(hi - lo);
        if (first_chunk)
            mo[0] = (float)((trig[0] * kInvPi - (double)lo) * inv);
#pragma unroll
        for (int i = 0; i < M; ++i)
        {
            const int j = j0 + i;
            if (j >= 1 && j < jend)
                mo[j] = (float)(trig[i] * kInvPi * inv);
        }
    }
    else
    {
#pragma unroll
        for (int i = 0; i < M; ++i)
        {
            const int j = j0 + i;
            if (j < jend)
                mo[j] = (float)(trig[i] * kInvPi);
        }
    }
    if (first_chunk)
    {
        b = pack_bound(lo, hi);
        if (bounds)
            bounds[pix] = b;
        if (ray_range)
            ray_range[pix] = make_float2(lo, hi);
        if (samples)
            samples[pix] = ns;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// one launch of generate_mixed_kernel over `chunks` rows of blockIdx.y
template <typename T, int M, bool CHUNK, int J0C>
static void launch_mixed_kernel(const VolumeView<T> &v, const DevCamera &c, const GenArgs &a, float *o, uint32_t *b,
                                int32_t *s, float2 *rr, hipStream_t st, uint32_t blocks, int chunks)
{
    // byte offsets of every voxel fit in 32 bits up to 4 GiB of volume
    const bool off32 = (uint64_t)v.nx * v.ny * v.nz * sizeof(T) <= (uint64_t)0xFFFFFFFFu + 1u;
    // power-of-two domain extents and data range: divisions become products
    const bool pow2 = exact_pow2(v.dmax.x - v.dmin.x) && exact_pow2(v.dmax.y - v.dmin.y) &&
                      exact_pow2(v.dmax.z - v.dmin.z) && exact_pow2(a.data_max - a.data_min);
    const dim3 grid(blocks, chunks);
    if constexpr (std::is_same<T, SphParticle>::value)
    {
        // particle sets: the gather ignores OFF32 and POW2 (the data range
        // division takes the exact path)
        hipLaunchKernelGGL((generate_mixed_kernel<T, M, true, false, CHUNK, J0C>), grid, dim3(256), 0, st, v, c, a,
                           o, b, s, rr);
    }
    else
    {
        if (off32 && pow2)
            hipLaunchKernelGGL((generate_mixed_kernel<T, M, true, true, CHUNK, J0C>), grid, dim3(256), 0, st, v, c, a,
                               o, b, s, rr);
        else if (off32)
            hipLaunchKernelGGL((generate_mixed_kernel<T, M, true, false, CHUNK, J0C>), grid, dim3(256), 0, st, v, c,
                               a, o, b, s, rr);
        else
            hipLaunchKernelGGL((generate_mixed_kernel<T, M, false, false, CHUNK, J0C>), grid, dim3(256), 0, st, v, c,
                               a, o, b, s, rr);
    }
}

// CHUNK marches the moments in chunks of M: the unchunked kernel for the
// first, then the second from j0 = M as a compile-time start, or the later
// ones over blockIdx.y
template <typename T, int M, bool CHUNK = false>
static void launch_mixed(const VolumeView<T> &v, const DevCamera &c, const GenArgs &a, float *o, uint32_t *b,
                         int32_t *s, float2 *rr, hipStream_t st, uint32_t blocks, int chunks = 1)
{
    launch_mixed_kernel<T, M, false, 0>(v, c, a, o, b, s, rr, st, blocks, 1);
    if constexpr (CHUNK)
    {
        if (chunks == 2)
            launch_mixed_kernel<T, M, true, M>(v, c, a, o, b, s, rr, st, blocks, 1);
        else if (chunks > 2)
            launch_mixed_kernel<T, M, true, 0>(v, c, a, o, b, s, rr, st, blocks, chunks - 1);
    }
}

template <typename T>
static int dispatch_mixed(const VolumeView<T> &v, const DevCamera &c, const GenArgs &a, float *o, uint32_t *b,
                          int32_t *s, float2 *rr, hipStream_t st, uint32_t blocks)
{
    const int m = a.m;
#define MRAY_GEN_CASE(MM)                                                                                             \
    if (m <= MM)                                                                                                      \
    {                                                                                                                 \
        launch_mixed<T, MM>(v, c, a, o, b, s, rr, st, blocks);                                                            \
        return 0;                                                                                                     \
    }
    MRAY_GEN_CASE(4)
    MRAY_GEN_CASE(8)
    MRAY_GEN_CASE(12)
    MRAY_GEN_CASE(16)
    MRAY_GEN_CASE(20)
    MRAY_GEN_CASE(24)
    MRAY_GEN_CASE(28)
    MRAY_GEN_CASE(30)
    MRAY_GEN_CASE(32)
    MRAY_GEN_CASE(36)
    MRAY_GEN_CASE(40)
    MRAY_GEN_CASE(44)
    MRAY_GEN_CASE(48)
    MRAY_GEN_CASE(50)
    MRAY_GEN_CASE(52)
    MRAY_GEN_CASE(56)
    MRAY_GEN_CASE(60)
#undef MRAY_GEN_CASE
    // 61..255 moments: K chunks of at most 50 (an unchunked M = 64 loses half
    // its occupancy, and so does a chunk of 56), each M >= ceil(m / K) wide
    if (m <= 255)
    {
        const int chunks = (m + 49) / 50;
        const int per = (m + chunks - 1) / chunks;
#define MRAY_GEN_CHUNK_CASE(MM)                                                                                       \
    if (per <= MM)                                                                                                    \
    {                                                                                                                 \
        launch_mixed<T, MM, true>(v, c, a, o, b, s, rr, st, blocks, chunks);                                          \
        return 0;                                                                                                     \
    }
        MRAY_GEN_CHUNK_CASE(40)
        MRAY_GEN_CHUNK_CASE(50)
#undef MRAY_GEN_CHUNK_CASE
    }
    return MRAY_ERR_UNSUPPORTED;
}

// PLAIN keeps m floats per thread in LDS: 256-thread blocks while that fits
// in 64 KiB, else one wave per block.
static int plain_block(int m) { return (256 * m * (int)sizeof(float) <= 65536) ? 256 : 64; }

// ---------------------------------------------------------------------------
// ADJOINT (F32 volumes): dL/dvolume from dL/dmoments
// ---------------------------------------------------------------------------

// A trilinear sample of the march that keeps what its reverse needs: the
// cell, the weights and the voxel values (same arithmetic as sample_volume;
// loads issued in issue(), used in value()).
template <bool POW2> struct GradSample
{
    float ax, ay, az;
    int cx, cy, cz; // clamped (x0, y0, z0)
    uint32_t fl;    // bit 0/1/2: x1 > x0, y1 > y0, z1 > z0 (0 at a clamped edge)
    float v[8];

    __device__ __forceinline__ void issue(const VolumeView<float> &vol, float px, float py, float pz)
    {
        const float ipx = div_c<POW2>(px - vol.dmin.x, vol.dmax.x - vol.dmin.x, vol.rext.x) * (float)vol.nx;
        const float ipy = div_c<POW2>(py - vol.dmin.y, vol.dmax.y - vol.dmin.y, vol.rext.y) * (float)vol.ny;
        const float ipz = div_c<POW2>(pz - vol.dmin.z, vol.dmax.z - vol.dmin.z, vol.rext.z) * (float)vol.nz;
        const float flx = floorf(ipx), fly = floorf(ipy), flz = floorf(ipz);
        ax = ipx - flx;
        ay = ipy - fly;
        az = ipz - flz;
        const int xl = (int)flx, yl = (int)fly, zl = (int)flz;
        cx = iclamp(xl, 0, vol.nx - 1);
        cy = iclamp(yl, 0, vol.ny - 1);
        cz = iclamp(zl, 0, vol.nz - 1);
        const int x1 = iclamp(xl + 1, 0, vol.nx - 1), y1 = iclamp(yl + 1, 0, vol.ny - 1),
                  z1 = iclamp(zl + 1, 0, vol.nz - 1);
        fl = (uint32_t)(x1 > cx) | ((uint32_t)(y1 > cy) << 1) | ((uint32_t)(z1 > cz) << 2);
        const uint32_t sxy = (uint32_t)(vol.ny * vol.nx);
        const uint32_t dx = fl & 1u, dy = (fl & 2u) ? (uint32_t)vol.nx : 0u;
        const float *p0 = vol.data + ((size_t)cz * sxy + (uint32_t)(cy * vol.nx + cx));
        const float *p1 = p0 + ((fl & 4u) ? sxy : 0u);
        v[0] = p0[0];
        v[1] = p0[dx];
        v[2] = p0[dy];
        v[3] = p0[dy + dx];
        v[4] = p1[0];
        v[5] = p1[dx];
        v[6] = p1[dy];
        v[7] = p1[dy + dx];
    }

    __device__ __forceinline__ float value() const
    {
        return (1.f - ax) * (1.f - ay) * (1.f - az) * v[0] + ax * (1.f - ay) * (1.f - az) * v[1] +
               (1.f - ax) * ay * (1.f - az) * v[2] + ax * ay * (1.f - az) * v[3] + (1.f - ax) * (1.f - ay) * az * v[4] +
               ax * (1.f - ay) * az * v[5] + (1.f - ax) * ay * az * v[6] + ax * ay * az * v[7];
    }
};

// Where the gradient of a cell's eight voxels goes: a workgroup's LDS window
// over a box of voxels (summed there, added to the volume once per voxel and
// chunk of the march) or, for corners
// outside the window, straight to the volume with float atomics. The window's
// 64 KiB hold 8192 voxels in 64-bit fixed point (integer LDS atomics, the
// fast kind) or 16384 in float (LDS float atomics cost several times more,
// but less than the global atomics a box too large for 8192 would take).
enum WinMode : int
{
    kWinGlobal = 0, // no window: float atomics to the volume
    kWinFixed = 1,  // 64-bit fixed point at scale 2^kfix
    kWinFloat = 2,  // float
};
struct GradTarget
{
    float *grad;
    unsigned long long *win; // LDS window, x fastest (as 8192 x u64 or 16384 x f32)
    int kfix;
    int ox, oy, oz, ex, ey, ez;
    int mode; // WinMode, the same in the whole workgroup
    int nx;
    uint32_t sxy;

    // adds w[i] to corner i of the cell (corner i: bit 0 / 1 / 2 of i select
    // x1 / y1 / z1)
    __device__ __forceinline__ void add(int cx, int cy, int cz, uint32_t fl, const float (&w)[8]) const
    {
        const int fx = fl & 1, fy = (fl >> 1) & 1, fz = (fl >> 2) & 1;
        const int lx = cx - ox, ly = cy - oy, lz = cz - oz;
        if (mode != kWinGlobal && lx >= 0 && ly >= 0 && lz >= 0 && lx + fx < ex && ly + fy < ey && lz + fz < ez)
        {
            const int i0 = lx + ex * (ly + ey * lz);
            const int dy = fy * ex, dz = fz * ex * ey;
            if (mode == kWinFixed)
            {
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    atomicAdd(win + i0 + ((i & 1) ? fx : 0) + ((i & 2) ? dy : 0) + ((i & 4) ? dz : 0),
                              to_fixed(w[i], kfix));
            }
            else
            {
                float *wf = reinterpret_cast<float *>(win);
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    atomicAdd(wf + i0 + ((i & 1) ? fx : 0) + ((i & 2) ? dy : 0) + ((i & 4) ? dz : 0), w[i]);
            }
            return;
        }
        float *p0 = grad + ((size_t)cz * sxy + (uint32_t)(cy * nx + cx));
        const uint32_t dy = fy ? (uint32_t)nx : 0u, dz = fz ? sxy : 0u;
#pragma unroll
        for (int i = 0; i < 8; ++i)
            unsafeAtomicAdd(p0 + ((i & 1) ? (uint32_t)fx : 0u) + ((i & 2) ? dy : 0u) + ((i & 4) ? (size_t)dz : 0u), w[i]);
    }
};

// Accumulates gs times the trilinear weights of a sample. Consecutive
// samples of a ray in the same cell are merged in registers; the eight
// corners go to the target when the ray leaves the cell (one code path for
// every lane: keeping a shared face in registers saved atomics but split the
// wavefront over six cases). Clamp-to-edge duplicates receive both weights.
struct CellScatter
{
    int cx = -1, cy = 0, cz = 0;
    uint32_t fl = 0;
    float w[8] = {0, 0, 0, 0, 0, 0, 0, 0};

    __device__ __forceinline__ void flush(const GradTarget &tg)
    {
        if (cx < 0)
            return;
        tg.add(cx, cy, cz, fl, w);
        cx = -1;
    }

    template <bool POW2> __device__ __forceinline__ void add(const GradTarget &tg, const GradSample<POW2> &sm, float gs)
    {
        // the same corner with other neighbour flags is another cell (a
        // sample just outside the domain clamps onto the edge voxel)
        if (!(sm.cx == cx && sm.cy == cy && sm.cz == cz && sm.fl == fl))
        {
            flush(tg);
#pragma unroll
            for (int i = 0; i < 8; ++i)
                w[i] = 0.0f;
            cx = sm.cx;
            cy = sm.cy;
            cz = sm.cz;
            fl = sm.fl;
        }
        const float ax = sm.ax, ay = sm.ay, az = sm.az;
        w[0] += gs * ((1.f - ax) * (1.f - ay) * (1.f - az));
        w[1] += gs * (ax * (1.f - ay) * (1.f - az));
        w[2] += gs * ((1.f - ax) * ay * (1.f - az));
        w[3] += gs * (ax * ay * (1.f - az));
        w[4] += gs * ((1.f - ax) * (1.f - ay) * az);
        w[5] += gs * (ax * (1.f - ay) * az);
        w[6] += gs * ((1.f - ax) * ay * az);
        w[7] += gs * (ax * ay * az);
    }
};

// Reverse mode of the generate march for dL/dmoments (dense W*H*m). The
// derivative of the exact integral of the reference's piecewise-linear
// segment [pp, pn] (length sg, slope denominator seg, r = sg / seg) is
//   d trig_j / d f1 =  (cos j pn - cos j pp) / (j^2 seg) + r sin(j pn) / j
//   d trig_j / d f0 = -(cos j pn - cos j pp) / (j^2 seg) + (1 - r) sin(j pn) / j - sin(j pp) / j
//   d trig_0 / d f1 = sg^2 / (2 seg),  d trig_0 / d f0 = sg - sg^2 / (2 seg).
// Summed over j with the incoming gradients g_j these need, at each end
// point, C(phi) = sum g_j / j^2 cos(j phi) and S(phi) = sum g_j / j sin(j phi):
// two Clenshaw sums in FP64 per node (their differences between neighbouring
// nodes are small next to their values). A segment starts where the previous
// one ended up to the reference's float rounding, pp = phi_k + delta; that
// offset enters to first order, C(pp) = C_k - delta S_k and S(pp) = S_k +
// delta S'_k, with S' = sum g_j cos(j phi) from a central difference of S
// over the neighbouring nodes (delta is a few ulps, so its accuracy hardly
// matters). Then through map_density (slope 1 / (data_max - data_min) where
// it does not clamp) and the trilinear weights to the voxels. The bounds are
// constants (a first pass recomputes them for the normalisation); cos and
// sin of the node phases come from sincos_small. One thread per ray, nothing
// stored per sample.
// CHUNK: moments j0 .. j0 + M - 1 with j0 = blockIdx.y * M (more moments
// than one thread's registers hold); the reverse is linear in the incoming
// gradient, so the chunks' passes add their shares to the volume gradient.
template <int M, bool POW2, bool CHUNK = false>
__global__ __launch_bounds__(256) void generate_adjoint_kernel(VolumeView<float> vol, DevCamera cam, GenArgs a,
                                                               const float *__restrict__ grad_mom,
                                                               const float2 *__restrict__ ray_range,
                                                               float *__restrict__ grad_vol)
{
    // The workgroup's 256 rays advance in chunks of kChunk nodes. Per chunk
    // they agree on the box of voxels they can touch; gradients go to an LDS
    // window over that box (when it fits), which is then added to the volume
    // with one atomic per voxel instead of one per ray, sample and corner.
    constexpr int kWin = 8192; // voxels of LDS window in 64-bit fixed point (twice as many in float)
    constexpr int kChunk = 32; // nodes per chunk and ray
    __shared__ unsigned long long win[kWin];
    __shared__ int bbox[6];
    __shared__ uint32_t bound_bits;
    const int tid = threadIdx.x;
    for (int i = tid; i < kWin; i += blockDim.x)
        win[i] = 0ull;
    if (tid == 0)
        bound_bits = 0u;
    __syncthreads();

    const int lane = tid & 63;
    const uint32_t ntiles = ((a.W + 7) / 8) * ((a.H + 7) / 8);
    const uint32_t tile = xcd_tile_block(a.W, 4) * 4 + (tid >> 6);
    int x = 0, y = 0;
    bool active = tile < ntiles;
    if (active)
    {
        tile_pixel(tile, a.W, lane, x, y);
        active = x < a.W && y < a.H;
    }
    Ray r = pixel_ray(cam, a.W, a.H, active ? x : 0, active ? y : 0);
    float t_near, t_far;
    const bool hit = intersect_aabb(r, vol.dmin, vol.dmax, t_near, t_far);
    t_near = rmax(t_near, cam.znear);
    active = active && hit && (t_far > 0);
    const int m = a.m;
    const size_t pix = active ? (size_t)y * a.W + x : 0;
    const float s = a.step;
    const float pi_trange_inv = (float)(kPi / (double)(t_far - t_near));
    const float dphi = (float)(kPi * (double)(s / (t_far - t_near)));
    const int cap = active ? march_cap(t_near, t_far, s) : 0;
    auto issue_at = [&](GradSample<POW2> &sm, float tt) {
        sm.issue(vol, r.o.x + tt * r.d.x, r.o.y + tt * r.d.y, r.o.z + tt * r.d.z);
    };
    auto density = [&](const GradSample<POW2> &sm) {
        return map_density_r<POW2>(sm.value(), a.data_min, a.data_max, a.data_rinv);
    };

    // normalisation of the forward: inv = 1 / (hi - lo) with the bounds of the
    // ray, as the forward recorded them or from a pass over the samples
    double inv = 1.0;
    if (active && a.transform_to_bounds && ray_range)
    {
        const float2 b = ray_range[pix];
        inv = 1.0 / (double)(b.y - b.x);
    }
    else if (active && a.transform_to_bounds)
    {
        float lo = 1.f, hi = 0.f;
        GradSample<POW2> sm;
        float t = t_near;
        issue_at(sm, t);
        for (int it = 0;; ++it)
        {
            const bool more = t + s < t_far && it < cap;
            const float f = density(sm);
            lo = rmin(lo, f);
            hi = rmax(hi, f);
            if (!more && t == t_far)
                break;
            t = more ? t + s : t_far;
            __builtin_amdgcn_sched_barrier(0);
            issue_at(sm, t);
            if (!more)
            {
                const float fl = density(sm);
                lo = rmin(lo, fl);
                hi = rmax(hi, fl);
                break;
            }
        }
        lo = rmax(lo - a.bound_eps, 0.0f);
        hi = rmin(hi + a.bound_eps, 1.0f);
        inv = 1.0 / (double)(hi - lo);
    }
    // dL/dtrig_j = dL/dmoment_j * inv / pi
    // Chebyshev coefficients g_j / j^2 of C in FP64 registers (slot i holds
    // moment j0 + i; moment 0 enters through g0, in the first chunk)
    const int j0 = CHUNK ? (int)blockIdx.y * M : 0;
    double cc[M];
    const double scale = inv / kPi;
    const double g0 = (active && j0 == 0) ? (double)grad_mom[pix * m] * scale : 0.0;
#pragma unroll
    for (int i = 0; i < M; ++i)
    {
        const int j = j0 + i;
        cc[i] = (active && j >= 1 && j < m) ? (double)grad_mom[pix * m + j] * scale / (double)(j * j) : 0.0;
    }
    // C and S at a node with cos / sin (co, si) of its phase: Clenshaw's
    // recurrence b_l = c_l + 2x b_{l+1} - b_{l+2} gives C = x b_1 - b_2, and
    // its derivative in x, b'_l = 2 b_{l+1} + 2x b'_{l+1} - b'_{l+2}, gives
    // S = -dC/dphi = sin(phi) dC/dx = sin(phi) (b_1 + x b'_1 - b'_2), since
    // S = sum j c_j sin(j phi) for the same coefficients c_j = g_j / j^2.
    //
    // CHUNK: the chunk's sum is C = b_0 T_j0 - b_1 T_j0-1 (Clenshaw down to
    // slot 0), and with T'_n(cos phi) sin(phi) = n sin(n phi),
    // S = sin(phi) (d_0 T_j0 - d_1 T_j0-1) + b_0 j0 sin(j0 phi)
    //     - b_1 (j0 - 1) sin((j0 - 1) phi);
    // cos and sin of j0 phi by binary powering of the node's phasor (j0 is
    // the same in the whole workgroup), so that they are the Chebyshev values
    // of the x the recurrence uses.
    auto sums = [&](double co, double si, double &C, double &S) {
        const double x2 = co + co;
        double b1 = 0.0, b2 = 0.0, d1 = 0.0, d2 = 0.0;
        if constexpr (!CHUNK)
        {
#pragma unroll
            for (int l = M - 1; l >= 1; --l)
            {
                const double tb = fma(x2, b1, cc[l] - b2);
                const double td = fma(x2, d1, fma(2.0, b1, -d2));
                b2 = b1;
                b1 = tb;
                d2 = d1;
                d1 = td;
            }
            C = fma(co, b1, -b2);
            S = si * (b1 + fma(co, d1, -d2));
        }
        else
        {
#pragma unroll
            for (int l = M - 1; l >= 0; --l)
            {
                const double tb = fma(x2, b1, cc[l] - b2);
                const double td = fma(x2, d1, fma(2.0, b1, -d2));
                b2 = b1;
                b1 = tb;
                d2 = d1;
                d1 = td;
            }
            // b1, b2, d1, d2 now hold b_0, b_1, d_0, d_1
            double pc = 1.0, ps = 0.0, bc = co, bs = si;
            for (int e = j0; e > 0; e >>= 1)
            {
                if (e & 1)
                {
                    const double nc = fma(pc, bc, -ps * bs);
                    ps = fma(pc, bs, ps * bc);
                    pc = nc;
                }
                const double sc = fma(bc, bc, -bs * bs);
                bs = 2.0 * bc * bs;
                bc = sc;
            }
            const double qc = fma(pc, co, ps * si), qs = fma(ps, co, -pc * si); // (j0 - 1) phi
            C = fma(b1, pc, -b2 * qc);
            S = fma(si, fma(d1, pc, -d2 * qc), fma(b1 * (double)j0, ps, -b2 * (double)(j0 - 1) * qs));
        }
    };
    const float dscale = 1.0f / (a.data_max - a.data_min);
    // fixed-point scale of the window: a node adds at most |fbar| |dscale| to
    // a voxel (trilinear weights sum to 1), |fbar| <= |A| + |B| <= 6.2 sum_j j |c_j|
    // + 2 pi |g0| (|C'| and |S| <= sum_j j |c_j|); 8 (...) with margin, times the
    // kChunk + 1 nodes a ray adds per chunk, over the workgroup
    {
        double sj = 0.0;
#pragma unroll
        for (int i = 0; i < M; ++i)
            sj = fma((double)(j0 + i), fabs(cc[i]), sj);
        const float lb = active ? (float)(8.0 * (sj + fabs(g0)) * (double)fabsf(dscale) * (kChunk + 1)) : 0.0f;
        atomicMax(&bound_bits, __float_as_uint(lb));
    }
    __syncthreads();
    // the same value in every lane: read as a scalar, so the branches on it
    // are uniform
    const float bsum = __uint_as_float(__builtin_amdgcn_readfirstlane(bound_bits)) * (float)blockDim.x;
    const int kfix = fixed_scale(bsum);
    // not finite: no window, float atomics straight to the volume
    const bool fixed = fixed_usable(bsum);
    // d f / d s of map_density
    auto dmap = [&](float sv) {
        const float u = div_c<POW2>(sv - a.data_min, a.data_max - a.data_min, a.data_rinv);
        return (u > 0.0f && u < 1.0f) ? dscale : 0.0f;
    };
    // clamped cell of the sample at t
    auto cell_at = [&](float tt, int &qx, int &qy, int &qz) {
        const float px = r.o.x + tt * r.d.x, py = r.o.y + tt * r.d.y, pz = r.o.z + tt * r.d.z;
        qx = iclamp((int)floorf(div_c<POW2>(px - vol.dmin.x, vol.dmax.x - vol.dmin.x, vol.rext.x) * (float)vol.nx), 0,
                    vol.nx - 1);
        qy = iclamp((int)floorf(div_c<POW2>(py - vol.dmin.y, vol.dmax.y - vol.dmin.y, vol.rext.y) * (float)vol.ny), 0,
                    vol.ny - 1);
        qz = iclamp((int)floorf(div_c<POW2>(pz - vol.dmin.z, vol.dmax.z - vol.dmin.z, vol.rext.z) * (float)vol.nz), 0,
                    vol.nz - 1);
    };

    const double rdphi = 1.0 / (double)dphi;
    // node 0: the ray's first sample, at the start of the first segment
    // (p_1 - dphi in float, or -pi when the ray has no interior sample)
    float t = t_near;
    float phi_prev;
    if (t + s < t_far && cap > 0)
        phi_prev = (float)((double)((t + s - t_near) * pi_trange_inv) - kPi) - dphi;
    else
        phi_prev = (float)(-kPi);
    // cur: the sample of the node behind the current segment; its loads are
    // issued one node ahead and used after the next node's sums
    GradSample<POW2> cur;
    double co = 1.0, si = 0.0, C_prev = 0.0, S_prev = 0.0;
    if (active)
    {
        issue_at(cur, t);
        sincos_small((double)phi_prev, si, co);
        sums(co, si, C_prev, S_prev);
    }
    double S_prev2 = S_prev, phi_prev2 = (double)phi_prev; // node k-2 for S'
    double fbar = 0.0; // pending gradient of the node behind the current segment
    CellScatter cells;
    GradTarget tg;
    tg.grad = grad_vol;
    tg.win = win;
    tg.kfix = kfix;
    tg.nx = vol.nx;
    tg.sxy = (uint32_t)(vol.nx * vol.ny);
    int it = 0;
    bool done = !active;
    while (__syncthreads_or(!done))
    {
        if (tid == 0)
        {
            bbox[0] = bbox[1] = bbox[2] = 0x7fffffff;
            bbox[3] = bbox[4] = bbox[5] = -1;
        }
        __syncthreads();
        if (!done)
        {
            // this chunk reads samples in [t, t + (kChunk + 1) s], or up to t_far
            // when it ends the ray; one cell of margin on each side
            float tb = t + (float)(kChunk + 1) * s;
            if (tb > t_far || it + kChunk >= cap)
                tb = t_far;
            int ax0, ay0, az0, ax1, ay1, az1;
            cell_at(t, ax0, ay0, az0);
            cell_at(tb, ax1, ay1, az1);
            atomicMin(&bbox[0], max(min(ax0, ax1) - 1, 0));
            atomicMin(&bbox[1], max(min(ay0, ay1) - 1, 0));
            atomicMin(&bbox[2], max(min(az0, az1) - 1, 0));
            atomicMax(&bbox[3], min(max(ax0, ax1) + 2, vol.nx - 1));
            atomicMax(&bbox[4], min(max(ay0, ay1) + 2, vol.ny - 1));
            atomicMax(&bbox[5], min(max(az0, az1) + 2, vol.nz - 1));
        }
        __syncthreads();
        tg.ox = __builtin_amdgcn_readfirstlane(bbox[0]);
        tg.oy = __builtin_amdgcn_readfirstlane(bbox[1]);
        tg.oz = __builtin_amdgcn_readfirstlane(bbox[2]);
        tg.ex = __builtin_amdgcn_readfirstlane(bbox[3]) - tg.ox + 1;
        tg.ey = __builtin_amdgcn_readfirstlane(bbox[4]) - tg.oy + 1;
        tg.ez = __builtin_amdgcn_readfirstlane(bbox[5]) - tg.oz + 1;
        {
            const long long nbox = (tg.ex > 0 && tg.ey > 0 && tg.ez > 0) ? (long long)tg.ex * tg.ey * tg.ez : -1;
            tg.mode = (nbox < 0) ? kWinGlobal
                      : (fixed && nbox <= kWin) ? kWinFixed
                      : (nbox <= 2 * kWin) ? kWinFloat
                                           : kWinGlobal;
        }
        for (int k = 0; k < kChunk && !done; ++k, ++it)
        {
            const bool last = !(t + s < t_far) || it >= cap;
            float p_next, seg;
            if (!last)
            {
                t += s;
                p_next = (float)((double)((t - t_near) * pi_trange_inv) - kPi);
                seg = dphi;
            }
            else
            {
                p_next = 0.0f;
                seg = 0.0f - phi_prev;
            }
            // the segment [pp, pn]: pp = p_next - seg in float (= phi_prev + delta)
            const double pn = (double)p_next;
            const double pp = (double)(p_next - seg);
            const double sg = pn - pp;
            const double delta = pp - (double)phi_prev;
            // cos / sin at the new node: rotated from the previous node by the
            // exact phase difference (Taylor terms to d^8 while it is below
            // 0.05; the rounding drift over a ray stays far below what the
            // sums resolve), else evaluated afresh
            {
                const double d = pn - (double)phi_prev;
                if (fabs(d) < 0.05)
                {
                    const double d2 = d * d;
                    const double cd = fma(d2, fma(d2, fma(d2, fma(d2, 1.0 / 40320, -1.0 / 720), 1.0 / 24), -0.5), 1.0);
                    const double sd = d * fma(d2, fma(d2, fma(d2, -1.0 / 5040, 1.0 / 120), -1.0 / 6), 1.0);
                    const double nc = fma(co, cd, -si * sd);
                    si = fma(si, cd, co * sd);
                    co = nc;
                }
                else
                    sincos_small(pn, si, co);
            }
            double Cn, Sn;
            sums(co, si, Cn, Sn);
            // S' at the previous node by a central difference (one-sided at the
            // first); it only multiplies delta, a few ulps, so float suffices
            const float dS = __fdividef((float)((it == 0) ? Sn - S_prev : Sn - S_prev2),
                                        (float)((it == 0) ? pn - (double)phi_prev : pn - phi_prev2));
            const double Cpp = C_prev - delta * S_prev;
            const double Spp = S_prev + delta * (double)dS;
            const double X = Cn - Cpp, V = Sn - Spp;
            const double rseg = last ? 1.0 / (double)seg : rdphi;
            const double xs = X * rseg;
            const double rr = sg * rseg;
            const double half = 0.5 * sg * rr;
            const double A = -xs + V - rr * Sn + g0 * (sg - half);
            const double B = xs + rr * Sn + g0 * half;
            fbar += A;
            const float gs0 = (float)fbar * dmap(cur.value());
            if (gs0 != 0.0f)
                cells.add(tg, cur, gs0);
            fbar = B;
            S_prev2 = S_prev;
            phi_prev2 = (double)phi_prev;
            C_prev = Cn;
            S_prev = Sn;
            phi_prev = p_next;
            __builtin_amdgcn_sched_barrier(0);
            issue_at(cur, last ? t_far : t);
            if (last)
            {
                const float gl = (float)fbar * dmap(cur.value());
                if (gl != 0.0f)
                    cells.add(tg, cur, gl);
                done = true;
            }
        }
        cells.flush(tg);
        __syncthreads();
        if (tg.mode != kWinGlobal)
        {
            // the window back to all-zero bits (a float sum may end as -0)
            // as its sums go to the volume
            // the window's voxel i at (lx, ly, lz), i = lx + ex (ly + ey lz):
            // each thread steps through i by blockDim.x, carrying lx and ly
            // over (no integer divisions in the loop)
            const int n = tg.ex * tg.ey * tg.ez;
            const int step = (int)blockDim.x;
            const int sx = step % tg.ex, sy = (step / tg.ex) % tg.ey, sz = step / (tg.ex * tg.ey);
            int lx = tid % tg.ex, ly = (tid / tg.ex) % tg.ey, lz = tid / (tg.ex * tg.ey);
            uint32_t *win32 = reinterpret_cast<uint32_t *>(win);
            for (int i = tid; i < n; i += step)
            {
                float v = 0.0f;
                bool any = false;
                if (tg.mode == kWinFixed)
                {
                    const unsigned long long q = win[i];
                    any = q != 0ull;
                    if (any)
                    {
                        win[i] = 0ull;
                        v = from_fixed(q, kfix);
                    }
                }
                else
                {
                    const uint32_t b = win32[i];
                    any = b != 0u;
                    if (any)
                    {
                        win32[i] = 0u;
                        v = __uint_as_float(b);
                    }
                }
                if (any)
                    unsafeAtomicAdd(grad_vol + ((size_t)(tg.oz + lz) * tg.sxy +
                                                (uint32_t)((tg.oy + ly) * vol.nx + tg.ox + lx)),
                                    v);
                lx += sx;
                ly += sy;
                lz += sz;
                if (lx >= tg.ex)
                {
                    lx -= tg.ex;
                    ++ly;
                }
                if (ly >= tg.ey)
                {
                    ly -= tg.ey;
                    ++lz;
                }
            }
        }
    }
}

int launch_generate_adjoint(const float *vox, int nx, int ny, int nz, const float *dmin, const float *dmax,
                            const mray_camera &cam, const mray_gen_params &p, int W, int H, int m,
                            const float *grad_moments, const float *ray_range, float *grad_volume,
                            hipStream_t stream)
{
    const float2 *rr = reinterpret_cast<const float2 *>(ray_range);
    GenArgs a;
    a.W = W;
    a.H = H;
    a.m = m;
    a.step = p.step_size;
    a.data_min = p.data_min;
    a.data_max = p.data_max;
    a.data_rinv = host_reciprocal(p.data_max - p.data_min);
    a.transform_to_bounds = p.transform_to_bounds;
    a.compact_image = p.compact_image;
    a.bound_eps = p.transform_bound_eps;
    DevCamera c = make_dev_camera(cam);
    const uint32_t ntiles = (uint32_t)(((W + 7) / 8) * ((H + 7) / 8));
    const uint32_t blocks = (ntiles + 3) / 4;
    VolumeView<float> v = make_volume_view(vox, nx, ny, nz, dmin, dmax);
    const bool pow2 = exact_pow2(v.dmax.x - v.dmin.x) && exact_pow2(v.dmax.y - v.dmin.y) &&
                      exact_pow2(v.dmax.z - v.dmin.z) && exact_pow2(a.data_max - a.data_min);
#define MRAY_GADJ_CASE(MM)                                                                                            \
    if (m <= MM)                                                                                                      \
    {                                                                                                                 \
        if (pow2)                                                                                                     \
            hipLaunchKernelGGL((generate_adjoint_kernel<MM, true>), dim3(blocks), dim3(256), 0, stream, v, c, a,      \
                               grad_moments, rr, grad_volume);                                                        \
        else                                                                                                          \
            hipLaunchKernelGGL((generate_adjoint_kernel<MM, false>), dim3(blocks), dim3(256), 0, stream, v, c, a,     \
                               grad_moments, rr, grad_volume);                                                        \
        return 0;                                                                                                     \
    }
    // 65..255 moments: chunks of at most 56 over blockIdx.y
#define MRAY_GADJ_CHUNK_CASE(MM)                                                                                      \
    if (m > 64 && m <= 255 && per <= MM)                                                                              \
    {                                                                                                                 \
        const dim3 grid(blocks, chunks);                                                                              \
        if (pow2)                                                                                                     \
            hipLaunchKernelGGL((generate_adjoint_kernel<MM, true, true>), grid, dim3(256), 0, stream, v, c, a,        \
                               grad_moments, rr, grad_volume);                                                        \
        else                                                                                                          \
            hipLaunchKernelGGL((generate_adjoint_kernel<MM, false, true>), grid, dim3(256), 0, stream, v, c, a,       \
                               grad_moments, rr, grad_volume);                                                        \
        return 0;                                                                                                     \
    }
    const int chunks = (m + 55) / 56;
    const int per = (m + chunks - 1) / chunks;
    MRAY_GADJ_CHUNK_CASE(40)
    MRAY_GADJ_CHUNK_CASE(48)
    MRAY_GADJ_CHUNK_CASE(56)
#undef MRAY_GADJ_CHUNK_CASE
    // the Clenshaw sums run over all M slots: widths close to the common
    // moment counts
    MRAY_GADJ_CASE(8)
    MRAY_GADJ_CASE(16)
    MRAY_GADJ_CASE(20)
    MRAY_GADJ_CASE(24)
    MRAY_GADJ_CASE(28)
    MRAY_GADJ_CASE(32)
    MRAY_GADJ_CASE(36)
    MRAY_GADJ_CASE(40)
    MRAY_GADJ_CASE(44)
    MRAY_GADJ_CASE(48)
    MRAY_GADJ_CASE(50)
    MRAY_GADJ_CASE(52)
    MRAY_GADJ_CASE(56)
    MRAY_GADJ_CASE(60)
    MRAY_GADJ_CASE(64)
#undef MRAY_GADJ_CASE
    return MRAY_ERR_UNSUPPORTED;
}

int launch_generate(const void *vox, int dtype, int nx, int ny, int nz, const float *dmin, const float *dmax,
                    const mray_camera &cam, const mray_gen_params &p, int form, int W, int H, int m, float *moments,
                    uint32_t *bounds, int32_t *samples, float *ray_range, hipStream_t stream)
{
    float2 *rr = reinterpret_cast<float2 *>(ray_range);
    GenArgs a;
    a.W = W;
    a.H = H;
    a.m = m;
    a.step = p.step_size;
    a.data_min = p.data_min;
    a.data_max = p.data_max;
    a.data_rinv = host_reciprocal(p.data_max - p.data_min);
    a.transform_to_bounds = p.transform_to_bounds;
    a.compact_image = p.compact_image;
    a.bound_eps = p.transform_bound_eps;
    DevCamera c = make_dev_camera(cam);
    uint32_t ntiles = (uint32_t)(((W + 7) / 8) * ((H + 7) / 8));
    uint32_t blocks = (ntiles + 3) / 4;
    if (dtype == MRAY_DTYPE_PARTICLES)
    {
        // a particle set: vox is its SphBuffers (host struct of device buffers)
        const SphBuffers *sb = static_cast<const SphBuffers *>(vox);
        VolumeView<SphParticle> v;
        v.s = sb->view();
        v.nx = v.ny = v.nz = 1;
        v.dmin = {dmin[0], dmin[1], dmin[2]};
        v.dmax = {dmax[0], dmax[1], dmax[2]};
        v.rext = {host_reciprocal(dmax[0] - dmin[0]), host_reciprocal(dmax[1] - dmin[1]),
                  host_reciprocal(dmax[2] - dmin[2])};
        if (form == MRAY_FORM_PLAIN)
        {
            int bs = plain_block(m);
            hipLaunchKernelGGL(generate_plain_kernel<SphParticle>, dim3((ntiles + bs / 64 - 1) / (bs / 64)), dim3(bs),
                               bs * m * sizeof(float), stream, v, c, a, moments, bounds, samples, rr);
            return 0;
        }
        return dispatch_mixed<SphParticle>(v, c, a, moments, bounds, samples, rr, stream, blocks);
    }
    if (dtype == MRAY_DTYPE_F32)
    {
        VolumeView<float> v = make_volume_view((const float *)vox, nx, ny, nz, dmin, dmax);
        if (form == MRAY_FORM_PLAIN)
        {
            int bs = plain_block(m);
            hipLaunchKernelGGL(generate_plain_kernel<float>, dim3((ntiles + bs / 64 - 1) / (bs / 64)), dim3(bs),
                               bs * m * sizeof(float), stream, v, c, a, moments, bounds, samples, rr);
            return 0;
        }
        return dispatch_mixed<float>(v, c, a, moments, bounds, samples, rr, stream, blocks);
    }
    VolumeView<uint8_t> v = make_volume_view((const uint8_t *)vox, nx, ny, nz, dmin, dmax);
    if (form == MRAY_FORM_PLAIN)
    {
        int bs = plain_block(m);
        hipLaunchKernelGGL(generate_plain_kernel<uint8_t>, dim3((ntiles + bs / 64 - 1) / (bs / 64)), dim3(bs),
                           bs * m * sizeof(float), stream, v, c, a, moments, bounds, samples, rr);
        return 0;
    }
    return dispatch_mixed<uint8_t>(v, c, a, moments, bounds, samples, rr, stream, blocks);
}

} // namespace mray
