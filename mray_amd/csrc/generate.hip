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
//         A node costs one sincos of its (float) phase and two Chebyshev
//         recurrences over j (cos and sin of j phi): 6 VALU ops per moment
//         instead of 2(m-1) sin/cos. The phasor is not carried from node to
//         node by rotation: the reference's node phases jitter by ulps around
//         the uniform grid, and a rotated phasor drifts off cos(p_k) by more
//         than the high orders tolerate.
//         Accumulation is blocked: kFold nodes go into a partial sum that is
//         then added with Kahan compensation; trig_0 (the exact integral of the
//         reference's lines) is accumulated in FP64. Moments live in
//         registers (template M >= m).
#include "device_common.hpp"
#include "launch.hpp"

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
                                                             int32_t *__restrict__ samples)
{
    extern __shared__ float lds_trig[]; // [m][blockDim.x], strided per thread
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int stride = blockDim.x;
    const uint32_t ntiles = ((a.W + 7) / 8) * ((a.H + 7) / 8);
    const uint32_t tile = xcd_swizzle(blockIdx.x, gridDim.x) * (blockDim.x / 64) + (tid >> 6);
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
    if (samples)
        samples[pix] = ns;
}

// ---------------------------------------------------------------------------
// MIXED
// ---------------------------------------------------------------------------
constexpr int kFold = 8;    // nodes per partial sum before the Kahan add

template <int M>
__device__ __forceinline__ void accumulate_node(float (&part)[M], float d, float e, float h, float x, float y)
{
    // part[j] += (d + j^2 e) cos(j phi) + j h sin(j phi) for j = 1..M-1 (the
    // partial sums are scaled by j^2, so the h sin / j term enters as j h sin);
    // cos and sin of j phi by the Chebyshev recurrence v_j = 2x v_{j-1} - v_{j-2}
    // from x = cos(phi), y = sin(phi).
    const float x2 = x + x;
    float cm2 = 1.0f, cm1 = x;
    float sm2 = 0.0f, sm1 = y;
    part[1] = __builtin_fmaf(h, sm1, __builtin_fmaf(d + e, cm1, part[1]));
#pragma unroll
    for (int j = 2; j < M; ++j)
    {
        const float c = __builtin_fmaf(x2, cm1, -cm2);
        const float sn = __builtin_fmaf(x2, sm1, -sm2);
        part[j] = __builtin_fmaf((float)j * h, sn, __builtin_fmaf(__builtin_fmaf((float)(j * j), e, d), c, part[j]));
        cm2 = cm1;
        cm1 = c;
        sm2 = sm1;
        sm1 = sn;
    }
}

template <int M> __device__ __forceinline__ void kahan_fold(float (&part)[M], float (&sum)[M], float (&comp)[M])
{
#pragma unroll
    for (int j = 1; j < M; ++j)
    {
        float yv = part[j] - comp[j];
        float tv = sum[j] + yv;
        comp[j] = (tv - sum[j]) - yv;
        sum[j] = tv;
        part[j] = 0.0f;
    }
}

template <typename T, int M>
__global__ __launch_bounds__(256) void generate_mixed_kernel(VolumeView<T> vol, DevCamera cam, GenArgs a,
                                                             float *__restrict__ out, uint32_t *__restrict__ bounds,
                                                             int32_t *__restrict__ samples)
{
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const uint32_t ntiles = ((a.W + 7) / 8) * ((a.H + 7) / 8);
    const uint32_t tile = xcd_swizzle(blockIdx.x, gridDim.x) * 4 + (tid >> 6);
    if (tile >= ntiles)
        return;
    int x, y;
    tile_pixel(tile, a.W, lane, x, y);
    if (x >= a.W || y >= a.H)
        return;
    const int m = a.m;
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
        if (samples)
            samples[pix] = 0;
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
    float zs, zc;
    sincosf(node_phase, &zs, &zc);
    const float z0s = zs, z0c = zc;

    float part[M], sum[M], comp[M];
#pragma unroll
    for (int j = 0; j < M; ++j)
    {
        part[j] = 0.0f;
        sum[j] = 0.0f;
        comp[j] = 0.0f;
    }
    double trap = 0.0;
    float g_left = 0.0f;   // slope of the segment ending at the current node
    float eps_left = 0.0f; // its width minus dphi
    int node = 0;

    const int cap = march_cap(t_near, t_far, s);
    const float rdphi = host_reciprocal(dphi); // for the segment slopes
    // the next sample's loads are issued one step ahead (at t + s while the
    // march continues, else at t_far for the last segment)
    auto issue_at = [&](PendingSample<T> &ps, float tt) {
        ps.issue(vol, r.o.x + tt * r.d.x, r.o.y + tt * r.d.y, r.o.z + tt * r.d.z);
    };
    PendingSample<T> pending;
    issue_at(pending, (t + s < t_far) ? t + s : t_far);
    for (int it = 0; t + s < t_far && it < cap; ++it)
    {
        t += s;
        const float p = (float)((double)((t - t_near) * pi_trange_inv) - kPi);
        const float f = map_density_r(pending.value(), a.data_min, a.data_max, a.data_rinv);
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
        const float delta = (node > 0) ? (float)((double)u - (double)node_phase) : 0.0f;
        const float h = __builtin_fmaf(g, delta, g_left * eps_left);
        accumulate_node<M>(part, g_left - g, -delta * f_prev, h, zc, zs);
        g_left = g;
        eps_left = (float)(width - (double)dphi);
        f_prev = f;
        node_phase = p;
        ++node;
        sincosf(p, &zs, &zc);
        if (node % kFold == 0)
            kahan_fold<M>(part, sum, comp);
    }
    // last segment [p_K, 0] (src/raymarch_generate.h:107-114): it starts
    // exactly at p_K and its line ends exactly at f_last
    const float f_last = map_density_r(pending.value(), a.data_min, a.data_max, a.data_rinv);
    ++ns;
    lo = rmin(lo, f_last);
    hi = rmax(hi, f_last);
    const float dl = 0.0f - node_phase;
    const float g_last = (f_last - f_prev) / dl;
    trap += (double)dl * ((double)f_prev + 0.5 * (double)g_last * (double)dl);
    accumulate_node<M>(part, g_left - g_last, 0.0f, g_left * eps_left, zc, zs);
    // node at phase 0: cos(j*0) = 1, sin(j*0) = 0
#pragma unroll
    for (int j = 1; j < M; ++j)
        part[j] += g_last;
    kahan_fold<M>(part, sum, comp);

    // trig_j = sum_j / j^2 - f_0 sin(j phi_0) / j, sin by Chebyshev from phi_0.
    float trig[M];
    {
        const float x2 = z0c + z0c;
        float sm2 = 0.0f, sm1 = z0s;
#pragma unroll
        for (int j = 1; j < M; ++j)
        {
            float sj;
            if (j == 1)
                sj = z0s;
            else
            {
                sj = __builtin_fmaf(x2, sm1, -sm2);
                sm2 = sm1;
                sm1 = sj;
            }
            const float jf = (float)j;
            trig[j] = (sum[j] - comp[j]) / (jf * jf) - f_first * sj / jf;
        }
    }
    trig[0] = (float)trap;

    uint32_t b;
    if (a.transform_to_bounds)
    {
        lo = rmax(lo - a.bound_eps, 0.0f);
        hi = rmin(hi + a.bound_eps, 1.0f);
        const double inv = 1.0 / (double)(hi - lo);
        mo[0] = (float)((double)(trig[0] * kInvPiF - lo) * inv);
#pragma unroll
        for (int j = 1; j < M; ++j)
            if (j < m)
                mo[j] = (float)((double)(trig[j] * kInvPiF) * inv);
    }
    else
    {
#pragma unroll
        for (int j = 0; j < M; ++j)
            if (j < m)
                mo[j] = trig[j] * kInvPiF;
    }
    b = pack_bound(lo, hi);
    if (bounds)
        bounds[pix] = b;
    if (samples)
        samples[pix] = ns;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T, int M>
static void launch_mixed(const VolumeView<T> &v, const DevCamera &c, const GenArgs &a, float *o, uint32_t *b,
                         int32_t *s, hipStream_t st, uint32_t blocks)
{
    hipLaunchKernelGGL((generate_mixed_kernel<T, M>), dim3(blocks), dim3(256), 0, st, v, c, a, o, b, s);
}

template <typename T>
static int dispatch_mixed(const VolumeView<T> &v, const DevCamera &c, const GenArgs &a, float *o, uint32_t *b,
                          int32_t *s, hipStream_t st, uint32_t blocks)
{
    const int m = a.m;
#define MRAY_GEN_CASE(MM)                                                                                             \
    if (m <= MM)                                                                                                      \
    {                                                                                                                 \
        launch_mixed<T, MM>(v, c, a, o, b, s, st, blocks);                                                            \
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
    MRAY_GEN_CASE(64)
#undef MRAY_GEN_CASE
    return MRAY_ERR_UNSUPPORTED;
}

// PLAIN keeps m floats per thread in LDS: 256-thread blocks while that fits
// in 64 KiB, else one wave per block.
static int plain_block(int m) { return (256 * m * (int)sizeof(float) <= 65536) ? 256 : 64; }

// ---------------------------------------------------------------------------
// ADJOINT (F32 volumes): dL/dvolume from dL/dmoments
// ---------------------------------------------------------------------------

// Accumulates gs times the trilinear weights of sample_volume at a position
// into the voxels it reads. Consecutive samples of a ray that fall into the
// same cell are merged in registers; a cell's eight sums are added to the
// gradient (float atomics) when the ray leaves it. Clamp-to-edge duplicates
// receive both of their weights.
struct CellScatter
{
    size_t base[4] = {0, 0, 0, 0}; // row offsets r00, r01, r10, r11
    int x0 = -1, x1 = -1;
    float w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool live = false;

    __device__ __forceinline__ void flush(float *grad)
    {
        if (!live)
            return;
        unsafeAtomicAdd(&grad[base[0] + x0], w[0]);
        unsafeAtomicAdd(&grad[base[0] + x1], w[1]);
        unsafeAtomicAdd(&grad[base[1] + x0], w[2]);
        unsafeAtomicAdd(&grad[base[1] + x1], w[3]);
        unsafeAtomicAdd(&grad[base[2] + x0], w[4]);
        unsafeAtomicAdd(&grad[base[2] + x1], w[5]);
        unsafeAtomicAdd(&grad[base[3] + x0], w[6]);
        unsafeAtomicAdd(&grad[base[3] + x1], w[7]);
        live = false;
    }

    __device__ __forceinline__ void add(float *grad, const VolumeView<float> &v, float px, float py, float pz, float gs)
    {
        float ipx = ((px - v.dmin.x) / (v.dmax.x - v.dmin.x)) * (float)v.nx;
        float ipy = ((py - v.dmin.y) / (v.dmax.y - v.dmin.y)) * (float)v.ny;
        float ipz = ((pz - v.dmin.z) / (v.dmax.z - v.dmin.z)) * (float)v.nz;
        float flx = floorf(ipx), fly = floorf(ipy), flz = floorf(ipz);
        float ax = ipx - flx, ay = ipy - fly, az = ipz - flz;
        int xl = (int)flx, yl = (int)fly, zl = (int)flz;
        int nx0 = iclamp(xl, 0, v.nx - 1), nx1 = iclamp(xl + 1, 0, v.nx - 1);
        int y0 = iclamp(yl, 0, v.ny - 1), y1 = iclamp(yl + 1, 0, v.ny - 1);
        int z0 = iclamp(zl, 0, v.nz - 1), z1 = iclamp(zl + 1, 0, v.nz - 1);
        size_t sxy = (size_t)(v.ny * v.nx);
        size_t r00 = (size_t)z0 * sxy + (size_t)(y0 * v.nx);
        size_t r01 = (size_t)z0 * sxy + (size_t)(y1 * v.nx);
        size_t r10 = (size_t)z1 * sxy + (size_t)(y0 * v.nx);
        size_t r11 = (size_t)z1 * sxy + (size_t)(y1 * v.nx);
        if (!(live && nx0 == x0 && nx1 == x1 && r00 == base[0] && r01 == base[1] && r10 == base[2] && r11 == base[3]))
        {
            flush(grad);
            x0 = nx0;
            x1 = nx1;
            base[0] = r00;
            base[1] = r01;
            base[2] = r10;
            base[3] = r11;
            for (int i = 0; i < 8; ++i)
                w[i] = 0.0f;
            live = true;
        }
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

// cos and sin of (a + d) from those of a, for the small differences between
// consecutive node phases (FP64 Taylor terms up to d^9; sincos beyond 0.1)
__device__ __forceinline__ void rotate_small(double &c, double &s, double a_plus_d, double d)
{
    if (fabs(d) > 0.1)
    {
        sincos(a_plus_d, &s, &c);
        return;
    }
    const double d2 = d * d;
    const double cd = 1.0 - d2 * (0.5 - d2 * (1.0 / 24 - d2 * (1.0 / 720 - d2 * (1.0 / 40320))));
    const double sd = d * (1.0 - d2 * (1.0 / 6 - d2 * (1.0 / 120 - d2 * (1.0 / 5040 - d2 * (1.0 / 362880)))));
    const double nc = c * cd - s * sd;
    const double ns = s * cd + c * sd;
    c = nc;
    s = ns;
}

// Reverse mode of the generate march for dL/dmoments (dense W*H*m): the
// derivative of the exact integral of the reference's piecewise-linear
// segments (its float end points, slope (f1 - f0) / dphi and last segment),
// per segment [pp, pn] of length sg with denominator seg, r = sg / seg:
//   d trig_j / d f1 =  (cos j pn - cos j pp) / (j^2 seg) + r sin(j pn) / j
//   d trig_j / d f0 = -(cos j pn - cos j pp) / (j^2 seg) + (1 - r) sin(j pn) / j - sin(j pp) / j
//   d trig_0 / d f1 = sg^2 / (2 seg),  d trig_0 / d f0 = sg - sg^2 / (2 seg)
// then through map_density (slope 1 / (data_max - data_min) where it does not
// clamp) and the trilinear weights to the voxels. The bounds are constants
// (a first pass recomputes them for the normalisation). cos / sin of j x by
// Chebyshev recurrences in FP64 from cos / sin of each segment's end points,
// which are carried from node to node by exact phase differences. One
// thread per ray, nothing stored per sample.
template <int M>
__global__ __launch_bounds__(256) void generate_adjoint_kernel(VolumeView<float> vol, DevCamera cam, GenArgs a,
                                                               const float *__restrict__ grad_mom,
                                                               float *__restrict__ grad_vol)
{
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const uint32_t ntiles = ((a.W + 7) / 8) * ((a.H + 7) / 8);
    const uint32_t tile = xcd_swizzle(blockIdx.x, gridDim.x) * 4 + (tid >> 6);
    if (tile >= ntiles)
        return;
    int x, y;
    tile_pixel(tile, a.W, lane, x, y);
    if (x >= a.W || y >= a.H)
        return;
    const int m = a.m;
    Ray r = pixel_ray(cam, a.W, a.H, x, y);
    float t_near, t_far;
    const bool hit = intersect_aabb(r, vol.dmin, vol.dmax, t_near, t_far);
    t_near = rmax(t_near, cam.znear);
    if (!(hit && (t_far > 0)))
        return;
    const size_t pix = (size_t)y * a.W + x;
    const float s = a.step;
    const float pi_trange_inv = (float)(kPi / (double)(t_far - t_near));
    const float dphi = (float)(kPi * (double)(s / (t_far - t_near)));
    const int cap = march_cap(t_near, t_far, s);
    auto raw = [&](float t, float &px, float &py, float &pz) {
        px = r.o.x + t * r.d.x;
        py = r.o.y + t * r.d.y;
        pz = r.o.z + t * r.d.z;
        return sample_volume(vol, px, py, pz);
    };

    // normalisation of the forward: inv = 1 / (hi - lo) with the bounds of the ray
    double inv = 1.0;
    if (a.transform_to_bounds)
    {
        float lo = 1.f, hi = 0.f, px, py, pz;
        float t = t_near;
        float f = map_density(raw(t, px, py, pz), a.data_min, a.data_max);
        lo = rmin(lo, f);
        hi = rmax(hi, f);
        for (int it = 0; t + s < t_far && it < cap; ++it)
        {
            t += s;
            f = map_density(raw(t, px, py, pz), a.data_min, a.data_max);
            lo = rmin(lo, f);
            hi = rmax(hi, f);
        }
        f = map_density(raw(t_far, px, py, pz), a.data_min, a.data_max);
        lo = rmin(lo, f);
        hi = rmax(hi, f);
        lo = rmax(lo - a.bound_eps, 0.0f);
        hi = rmin(hi + a.bound_eps, 1.0f);
        inv = 1.0 / ((double)hi - (double)lo);
    }
    // dL/dtrig_j = dL/dmoment_j * inv / pi (the factors 1/j and 1/j^2 are
    // float literals of the unrolled loop)
    float gj[M];
    const double scale = inv / kPi;
    const double g0 = (double)grad_mom[pix * m] * scale;
#pragma unroll
    for (int j = 1; j < M; ++j)
        gj[j] = j < m ? (float)((double)grad_mom[pix * m + j] * scale) : 0.0f;
    const float dscale = 1.0f / (a.data_max - a.data_min);
    // d f / d s of map_density
    auto dmap = [&](float sv) {
        const float u = (sv - a.data_min) / (a.data_max - a.data_min);
        return (u > 0.0f && u < 1.0f) ? dscale : 0.0f;
    };

    float t = t_near, p = (float)(-kPi);
    float px0, py0, pz0;
    float s0 = raw(t, px0, py0, pz0);
    double fbar0 = 0.0; // pending gradient of the node behind the current segment
    CellScatter cells;
    // cos / sin of the previous node's phase, carried forward by exact
    // differences (consecutive phases differ by ~dphi, or by ulps)
    double ph_prev = (double)p, cprev, sprev;
    sincos(ph_prev, &sprev, &cprev);
    for (int it = 0;; ++it)
    {
        const bool last = !(t + s < t_far) || it >= cap;
        float p_next, seg, px1, py1, pz1, s1;
        if (!last)
        {
            t += s;
            p_next = (float)((double)((t - t_near) * pi_trange_inv) - kPi);
            seg = dphi;
            s1 = raw(t, px1, py1, pz1);
        }
        else
        {
            p_next = 0.0f;
            seg = 0.0f - p;
            s1 = raw(t_far, px1, py1, pz1);
        }
        const double pn = p_next, pp = (double)(p_next - seg), sg = pn - pp;
        double cp1 = cprev, sp1 = sprev;
        rotate_small(cp1, sp1, pp, pp - ph_prev);
        double cn1 = cp1, sn1 = sp1;
        rotate_small(cn1, sn1, pn, sg);
        ph_prev = pn;
        cprev = cn1;
        sprev = sn1;
        // X = sum g_j / j^2 (cos j pn - cos j pp), V = sum g_j / j (sin j pn -
        // sin j pp), Yn = sum g_j / j sin j pn: the differences are formed in FP64 (they are
        // small next to their terms), the weighted sums in FP32
        float X = 0.0f, V = 0.0f, Yn = 0.0f;
        double cn = cn1, sn = sn1, cp = cp1, sp = sp1;      // j = 1
        double cnm = 1.0, snm = 0.0, cpm = 1.0, spm = 0.0;  // j = 0
#pragma unroll
        for (int j = 1; j < M; ++j)
        {
            const float g1 = gj[j] * (1.0f / (float)j);
            X = __builtin_fmaf(g1 * (1.0f / (float)j), (float)(cn - cp), X);
            V = __builtin_fmaf(g1, (float)(sn - sp), V);
            Yn = __builtin_fmaf(g1, (float)sn, Yn);
            const double cnn = fma(2.0 * cn1, cn, -cnm), snn = fma(2.0 * cn1, sn, -snm);
            const double cpn = fma(2.0 * cp1, cp, -cpm), spn = fma(2.0 * cp1, sp, -spm);
            cnm = cn;
            snm = sn;
            cpm = cp;
            spm = sp;
            cn = cnn;
            sn = snn;
            cp = cpn;
            sp = spn;
        }
        const double segd = (double)seg;
        const double rr = sg / segd;
        const double half = 0.5 * sg * sg / segd;
        // A = -X / seg + (1 - r) Y - Z with Y - Z = V, Y = Yn
        const double A = -(double)X / segd + (double)V - rr * (double)Yn + g0 * (sg - half);
        const double B = (double)X / segd + rr * (double)Yn + g0 * half;
        fbar0 += A;
        const float gs0 = (float)fbar0 * dmap(s0);
        if (gs0 != 0.0f)
            cells.add(grad_vol, vol, px0, py0, pz0, gs0);
        fbar0 = B;
        s0 = s1;
        px0 = px1;
        py0 = py1;
        pz0 = pz1;
        p = p_next;
        if (last)
            break;
    }
    const float gs0 = (float)fbar0 * dmap(s0);
    if (gs0 != 0.0f)
        cells.add(grad_vol, vol, px0, py0, pz0, gs0);
    cells.flush(grad_vol);
}

int launch_generate_adjoint(const float *vox, int nx, int ny, int nz, const float *dmin, const float *dmax,
                            const mray_camera &cam, const mray_gen_params &p, int W, int H, int m,
                            const float *grad_moments, float *grad_volume, hipStream_t stream)
{
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
#define MRAY_GADJ_CASE(MM)                                                                                            \
    if (m <= MM)                                                                                                      \
    {                                                                                                                 \
        hipLaunchKernelGGL(generate_adjoint_kernel<MM>, dim3(blocks), dim3(256), 0, stream, v, c, a, grad_moments,    \
                           grad_volume);                                                                              \
        return 0;                                                                                                     \
    }
    MRAY_GADJ_CASE(8)
    MRAY_GADJ_CASE(16)
    MRAY_GADJ_CASE(24)
    MRAY_GADJ_CASE(32)
    MRAY_GADJ_CASE(40)
    MRAY_GADJ_CASE(48)
    MRAY_GADJ_CASE(56)
    MRAY_GADJ_CASE(64)
#undef MRAY_GADJ_CASE
    return MRAY_ERR_UNSUPPORTED;
}

int launch_generate(const void *vox, int dtype, int nx, int ny, int nz, const float *dmin, const float *dmax,
                    const mray_camera &cam, const mray_gen_params &p, int form, int W, int H, int m, float *moments,
                    uint32_t *bounds, int32_t *samples, hipStream_t stream)
{
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
    if (dtype == MRAY_DTYPE_F32)
    {
        VolumeView<float> v = make_volume_view((const float *)vox, nx, ny, nz, dmin, dmax);
        if (form == MRAY_FORM_PLAIN)
        {
            int bs = plain_block(m);
            hipLaunchKernelGGL(generate_plain_kernel<float>, dim3((ntiles + bs / 64 - 1) / (bs / 64)), dim3(bs),
                               bs * m * sizeof(float), stream, v, c, a, moments, bounds, samples);
            return 0;
        }
        return dispatch_mixed<float>(v, c, a, moments, bounds, samples, stream, blocks);
    }
    VolumeView<uint8_t> v = make_volume_view((const uint8_t *)vox, nx, ny, nz, dmin, dmax);
    if (form == MRAY_FORM_PLAIN)
    {
        int bs = plain_block(m);
        hipLaunchKernelGGL(generate_plain_kernel<uint8_t>, dim3((ntiles + bs / 64 - 1) / (bs / 64)), dim3(bs),
                           bs * m * sizeof(float), stream, v, c, a, moments, bounds, samples);
        return 0;
    }
    return dispatch_mixed<uint8_t>(v, c, a, moments, bounds, samples, stream, blocks);
}

} // namespace mray
