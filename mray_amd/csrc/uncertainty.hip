// Error bounds and uncertainty rendering of moment images
// (src/raymarch_reconstruct.h:203-367, src/error_measures.h,
// src/moment_raymarcher.cu:693-950), in the reference's FP32 arithmetic (the
// PLAIN form): multipliers strided through LDS, the series by powers of the
// circle point with the host C library's cosf / sinf, fast_atan, unfused
// float operations in the reference's order. Results are bit-identical to
// the reference CPU path.
//
//  error_bounds_kernel   per pixel, the error of the reconstruction against
//                        the mapped volume sample at t = t_near, t_near + s,
//                        ... <= t_far (the phase recomputed from t at every
//                        sample): maximum, RMSE, or the P^2 estimate of a
//                        percentile of |pred - ref|; stored as the uint16
//                        e * 65535 (truncated)
//  uncertainty_kernel    MODE 0: the preintegrated march over
//                        (1 - s) lower + s upper, lower / upper = the
//                        reconstruction -/+ the pixel's error bound clamped
//                        to its density bounds; MODE 1: one bound sample per
//                        step, (lower, upper) or (mean, error) classified by
//                        a 256 x 256 uncertainty table (box- or
//                        gaussian-filtered transfer function)
#include "device_common.hpp"
#include "launch.hpp"
#include "sph_common.hpp"

namespace mray
{

namespace
{

struct UArgs
{
    int W, H, m;
    float step;
    float emission;
    float ert;
    int fourier;
    Vec3 dmin, dmax;
    // error bounds
    int error_type;
    float q; // percentile / 100
    float data_min, data_max;
    // uncertainty
    float interp;
    int mean_std;
};

// MESEReconstruction / TruncatedFourierReconstruction::reconstruct
// (src/reconstruction_techniques.h:43-94) on multipliers at L[l * stride]
__device__ __forceinline__ float plain_density(const float *L, int stride, int nm, float p, float blo, float bhi,
                                               bool fourier)
{
    p = rclamp(p, (float)(-kPi), 0.0f);
    const float cx = host_cosf(-p), cy = host_sinf(-p);
    // evaluateFourierSeries (src/mese/MESE_dynamic.h:213-226)
    float result = 0.0f;
    float pwx = 1.0f, pwy = 0.0f;
    for (int l = 1; l < nm; ++l)
    {
        const float nx = pwx * cx - pwy * cy;
        const float ny = pwx * cy + pwy * cx;
        pwx = nx;
        pwy = ny;
        result += L[l * stride] * pwx;
    }
    const float sres = 2.0f * result + L[0];
    const float rho = fourier ? sres : fast_atan(sres) * 0.318309886f + 0.5f;
    return rho * (bhi - blo) + blo;
}

// ErrorMax / ErrorRMSE / ErrorPercentile (src/error_measures.h:29-231)
struct ErrAcc
{
    int type;
    float q;
    float maxe, squares;
    int count, init;
    float h[5];
    int pos[5];
    float des[5];

    __device__ void start(int t, float qq)
    {
        type = t;
        q = qq;
        maxe = squares = 0.0f;
        count = init = 0;
#pragma unroll
        for (int i = 0; i < 5; ++i)
        {
            h[i] = 0.0f;
            pos[i] = 0;
            des[i] = 0.0f;
        }
    }

    // the P^2 marker updates index the five markers with loop counters;
    // fully unrolled they stay in registers
    __device__ float parabolic(int i, int d) const
    {
        const float t1 = d / (float)(pos[i + 1] - pos[i - 1]);
        const float t2 = (pos[i] - pos[i - 1] + d) * (h[i + 1] - h[i]) / (float)(pos[i + 1] - pos[i]);
        const float t3 = (pos[i + 1] - pos[i] - d) * (h[i] - h[i - 1]) / (float)(pos[i] - pos[i - 1]);
        return h[i] + t1 * (t2 + t3);
    }

    __device__ float linear(int i, int d) const
    {
        // d = +-1: h[i + d] and pos[i + d] spelled out for register indexing
        const float hn = (d > 0) ? h[i + 1] : h[i - 1];
        const int pn = (d > 0) ? pos[i + 1] : pos[i - 1];
        return h[i] + d * (hn - h[i]) / (float)(pn - pos[i]);
    }

    __device__ void update(float ref, float pred)
    {
        if (type == MRAY_ERROR_MAX)
        {
            maxe = rmax(maxe, fabsf(pred - ref));
            return;
        }
        if (type == MRAY_ERROR_RMSE)
        {
            count += 1;
            const float d = pred - ref;
            squares += d * d;
            return;
        }
        const float value = fabsf(pred - ref);
        if (init < 5)
        {
#pragma unroll
            for (int i = 0; i < 5; ++i)
                if (i == init)
                    h[i] = value;
            ++init;
            if (init == 5)
            {
                // sort_five_numbers (src/error_measures.h:58-98): ascending
#pragma unroll
                for (int i = 1; i < 5; ++i)
#pragma unroll
                    for (int j = i; j > 0; --j)
                        if (h[j] < h[j - 1])
                        {
                            const float t = h[j];
                            h[j] = h[j - 1];
                            h[j - 1] = t;
                        }
#pragma unroll
                for (int i = 0; i < 5; ++i)
                    pos[i] = 1 + i;
                des[0] = 1.0f;
                des[1] = 1.0f + 2.0f * q;
                des[2] = 1.0f + 4.0f * q;
                des[3] = 3.0f + 2.0f * q;
                des[4] = 5.0f;
            }
            return;
        }
        int move = 0;
        if (value < h[0])
        {
            h[0] = value;
            move = 1;
        }
        else if (value < h[1])
            move = 1;
        else if (value < h[2])
            move = 2;
        else if (value < h[3])
            move = 3;
        else if (value < h[4])
            move = 4;
        else if (value > h[4])
        {
            h[4] = value;
            move = 4;
        }
#pragma unroll
        for (int k = 0; k < 5; ++k)
            if (k >= move)
                pos[k] += 1;
        des[1] += q / 2.0f;
        des[2] += q;
        des[3] += (1.0f + q) / 2.f;
        des[4] += 1.0f;
        // adjust_heights: markers 1..3 (marker 4's desired position and
        // position both grow by exactly 1 per update: it never moves)
#pragma unroll
        for (int i = 1; i < 4; ++i)
        {
            const float di = des[i] - pos[i];
            if ((di >= 1 && (pos[i + 1] - pos[i] > 1)) || (di <= -1 && (pos[i - 1] - pos[i] < -1)))
            {
                const int sdi = (di > 0) ? 1 : -1;
                const float qp = parabolic(i, sdi);
                if ((h[i - 1] < qp) && (qp < h[i + 1]))
                    h[i] = qp;
                else
                    h[i] = linear(i, sdi);
                pos[i] += sdi;
            }
        }
    }

    // fewer than 3 percentile samples read marker heights the reference
    // leaves uninitialised; here they are 0
    __device__ float value() const
    {
        if (type == MRAY_ERROR_MAX)
            return maxe;
        if (type == MRAY_ERROR_RMSE)
            return sqrtf(squares / (float)count);
        return h[2];
    }
};

// MomentImageHost::set_error_bound (src/moment_image.h:102-106): the float to
// uint16 conversion truncates and keeps the low 16 bits of the integer
__device__ __forceinline__ uint16_t quantize_error_bound(float e) { return (uint16_t)(uint32_t)(int32_t)(e * 65535.f); }

__device__ __forceinline__ void density_bounds(const float *L, int stride, int nm, float p, float blo, float bhi,
                                               bool fourier, float e, int mean_std, float &b0, float &b1)
{
    const float rho = plain_density(L, stride, nm, p, blo, bhi, fourier);
    if (mean_std)
    {
        b0 = rho;
        b1 = e;
    }
    else
    {
        b0 = rclamp(rho - e, blo, bhi);
        b1 = rclamp(rho + e, blo, bhi);
    }
}

// TransferFunction::sample on a full 256 x 256 table (src/transfer_function.h:44-59):
// the (x, y + 1) corner is weighted by a.x (1 - a.y), (x + 1, y) by
// (1 - a.x) a.y, as the reference writes it
__device__ __forceinline__ float4 tf2d_sample(const float4 *__restrict__ tf, float fs, float bs)
{
    const float ipx = fs * 256.0f, ipy = bs * 256.0f;
    const float ax = ipx - floorf(ipx), ay = ipy - floorf(ipy);
    const int xl = (int)floorf(ipx), yl = (int)floorf(ipy);
    const int x0 = iclamp(xl, 0, 255), x1 = iclamp(xl + 1, 0, 255);
    const int y0 = iclamp(yl, 0, 255), y1 = iclamp(yl + 1, 0, 255);
    const float4 ll = tf[y0 * 256 + x0], lr = tf[y1 * 256 + x0], rr = tf[y1 * 256 + x1], rl = tf[y0 * 256 + x1];
    const float w0 = (1.f - ax) * (1.f - ay), w1 = ax * (1.f - ay), w2 = (1.f - ax) * ay, w3 = ax * ay;
    float4 o;
    o.x = w0 * ll.x + w1 * lr.x + w2 * rl.x + w3 * rr.x;
    o.y = w0 * ll.y + w1 * lr.y + w2 * rl.y + w3 * rr.y;
    o.z = w0 * ll.z + w1 * lr.z + w2 * rl.z + w3 * rr.z;
    o.w = w0 * ll.w + w1 * lr.w + w2 * rl.w + w3 * rr.w;
    return o;
}

// the pixel of this thread (8 x 8 tiles per wavefront, as the reconstruct
// kernels); false past the image
__device__ __forceinline__ bool pixel_of(const UArgs &a, int &x, int &y)
{
    const int tid = threadIdx.x, lane = tid & 63;
    const uint32_t ntiles = ((a.W + 7) / 8) * ((a.H + 7) / 8);
    const uint32_t tile = xcd_tile_block(a.W, blockDim.x / 64) * (blockDim.x / 64) + (tid >> 6);
    if (tile >= ntiles)
        return false;
    tile_pixel(tile, a.W, lane, x, y);
    return x < a.W && y < a.H;
}

// reconstruct_error (src/raymarch_reconstruct.h:323-367) per pixel, as
// reconstruct_error_kernel / reconstruct_error_host (src/moment_raymarcher.cu:693-714,
// src/moment_raymarcher.cpp:516-539) run it; a pixel without moments
// reconstructs with an empty series (the reference reads the first entry of
// an empty multiplier array)
template <typename T>
__global__ __launch_bounds__(256) void error_bounds_kernel(VolumeView<T> vol, DevCamera cam, UArgs a,
                                                           const float *__restrict__ lag,
                                                           const uint32_t *__restrict__ index,
                                                           const uint32_t *__restrict__ bounds,
                                                           uint16_t *__restrict__ out)
{
    extern __shared__ float lds_lag[]; // [m][blockDim]
    int x, y;
    if (!pixel_of(a, x, y))
        return;
    const size_t pix = (size_t)y * a.W + x;
    const int nm = index ? min((int)(index[pix + 1] - index[pix]), a.m) : a.m;
    const size_t off = index ? (size_t)index[pix] : pix * (size_t)a.m;
    float *L = lds_lag + threadIdx.x;
    const int stride = blockDim.x;
    for (int l = 0; l < nm; ++l)
        L[l * stride] = lag[off + l];
    if (nm == 0)
        L[0] = 0.0f;
    const int ne = max(nm, 1);
    float blo = 0.f, bhi = 1.f;
    if (bounds)
    {
        const uint32_t b = bounds[pix];
        blo = (b & 0xFFFFu) / 65535.f;
        bhi = (b >> 16) / 65535.f;
    }
    const Ray r = pixel_ray(cam, a.W, a.H, x, y);
    float t_near, t_far;
    const bool hit = intersect_aabb(r, vol.dmin, vol.dmax, t_near, t_far);
    t_near = rmax(t_near, cam.znear);
    float e = 0.0f;
    if (hit && (t_far > 0))
    {
        ErrAcc acc;
        acc.start(a.error_type, a.q);
        float t = t_near;
        const int cap = march_cap(t_near, t_far, a.step) + 1;
        for (int it = 0; t <= t_far && it < cap; ++it)
        {
            const float p = (float)((double)((t - t_near) / (t_far - t_near)) * kPi - kPi);
            const float fs = plain_density(L, stride, ne, p, blo, bhi, false);
            const float px = r.o.x + t * r.d.x, py = r.o.y + t * r.d.y, pz = r.o.z + t * r.d.z;
            const float ref = map_density(sample_volume(vol, px, py, pz), a.data_min, a.data_max);
            acc.update(ref, fs);
            t += a.step;
        }
        e = acc.value();
    }
    out[pix] = quantize_error_bound(e);
}

// MODE 0: reconstruct_ray with ErrorBoundInterpolator
// (src/raymarch_reconstruct.h:137-200, src/reconstruction_techniques.h:140-156);
// MODE 1: reconstruct_ray_uncertainty_tf (src/raymarch_reconstruct.h:203-257)
template <int MODE>
__global__ __launch_bounds__(256) void uncertainty_kernel(DevCamera cam, UArgs a, const float *__restrict__ lag,
                                                          const uint32_t *__restrict__ index,
                                                          const uint32_t *__restrict__ bounds,
                                                          const uint16_t *__restrict__ error_bounds,
                                                          const float4 *__restrict__ tf_g, float *__restrict__ rgba)
{
    __shared__ float4 tf1[MODE == 0 ? 256 : 1];
    extern __shared__ float lds_lag[];
    if (MODE == 0)
    {
        for (int i = threadIdx.x; i < 256; i += blockDim.x)
            tf1[i] = tf_g[i];
        __syncthreads();
    }
    int x, y;
    if (!pixel_of(a, x, y))
        return;
    const size_t pix = (size_t)y * a.W + x;
    const int nm = index ? min((int)(index[pix + 1] - index[pix]), a.m) : a.m;
    const size_t off = index ? (size_t)index[pix] : pix * (size_t)a.m;
    float4 *o4 = reinterpret_cast<float4 *>(rgba) + pix;
    if (nm == 0)
    {
        // reconstruct_zero_ray (src/raymarch_reconstruct.h:127-131)
        const float4 c = (MODE == 0) ? tf_sample(tf1, 0.0f, 0.0f) : tf2d_sample(tf_g, 0.0f, 0.0f);
        *o4 = make_float4(c.x, c.y, c.z, 0.0f);
        return;
    }
    float *L = lds_lag + threadIdx.x;
    const int stride = blockDim.x;
    for (int l = 0; l < nm; ++l)
        L[l * stride] = lag[off + l];
    float blo = 0.f, bhi = 1.f;
    if (bounds)
    {
        const uint32_t b = bounds[pix];
        blo = (b & 0xFFFFu) / 65535.f;
        bhi = (b >> 16) / 65535.f;
    }
    const float e = (float)error_bounds[pix] / 65535.f;
    const bool fourier = a.fourier != 0;
    const Ray r = pixel_ray(cam, a.W, a.H, x, y);
    float t_near, t_far;
    const bool hit = intersect_aabb(r, a.dmin, a.dmax, t_near, t_far);
    t_near = rmax(t_near, cam.znear);
    float Lr = 0.f, Lg = 0.f, Lb = 0.f, alpha = 0.f;
    if (hit && (t_far > 0))
    {
        float t = t_near;
        float p = (float)((double)((t - t_near) / (t_far - t_near)) * kPi - kPi);
        const float dphi = (float)(kPi * (double)(a.step / (t_far - t_near)));
        float T = 1.0f, tau = 0.0f;
        const int cap = march_cap(t_near, t_far, a.step) + 1;
        float b0, b1, fs = 0.0f;
        if (MODE == 0)
        {
            density_bounds(L, stride, nm, p, blo, bhi, fourier, e, 0, b0, b1);
            fs = (1 - a.interp) * b0 + a.interp * b1;
        }
        for (int it = 0; t <= t_far && it < cap; ++it)
        {
            t += a.step;
            p += dphi;
            float4 c;
            if (MODE == 0)
            {
                density_bounds(L, stride, nm, p, blo, bhi, fourier, e, 0, b0, b1);
                const float bs = (1 - a.interp) * b0 + a.interp * b1;
                c = tf_sample(tf1, fs, bs);
                fs = bs;
            }
            else
            {
                density_bounds(L, stride, nm, p, blo, bhi, fourier, e, a.mean_std, b0, b1);
                c = tf2d_sample(tf_g, b0, b1);
            }
            Lr += (a.emission * c.x) * T;
            Lg += (a.emission * c.y) * T;
            Lb += (a.emission * c.z) * T;
            tau += c.w;
            T = (float)exp((double)(-tau));
            alpha = 1.0f - T;
            if (alpha >= a.ert)
            {
                alpha = 1.f;
                break;
            }
        }
    }
    *o4 = make_float4(Lr, Lg, Lb, alpha);
}

int block_for(int m) { return (256 * m * (int)sizeof(float) <= 60000) ? 256 : 64; }

} // namespace

int launch_error_bounds(const void *vox, int dtype, int nx, int ny, int nz, const float *dmin, const float *dmax,
                        const mray_camera &cam, const float *lagrange, const uint32_t *index, const uint32_t *bounds,
                        int W, int H, int m, const mray_error_params &p, uint16_t *out, hipStream_t stream)
{
    if (p.error_type < MRAY_ERROR_MAX || p.error_type > MRAY_ERROR_PERCENTILE)
        return MRAY_ERR_ARG;
    UArgs a{};
    a.W = W;
    a.H = H;
    a.m = m;
    a.step = p.step_size;
    a.error_type = p.error_type;
    a.q = (float)p.error_percentile / 100.f;
    a.data_min = p.data_min;
    a.data_max = p.data_max;
    const DevCamera c = make_dev_camera(cam);
    const uint32_t ntiles = (uint32_t)(((W + 7) / 8) * ((H + 7) / 8));
    const int bs = block_for(m);
    const dim3 grid((ntiles + bs / 64 - 1) / (bs / 64));
    const size_t lds = (size_t)bs * m * sizeof(float);
    if (dtype == MRAY_DTYPE_PARTICLES)
    {
        const SphBuffers *sb = static_cast<const SphBuffers *>(vox);
        VolumeView<SphParticle> v;
        v.s = sb->view();
        v.nx = v.ny = v.nz = 1;
        v.dmin = {dmin[0], dmin[1], dmin[2]};
        v.dmax = {dmax[0], dmax[1], dmax[2]};
        v.rext = {host_reciprocal(dmax[0] - dmin[0]), host_reciprocal(dmax[1] - dmin[1]),
                  host_reciprocal(dmax[2] - dmin[2])};
        hipLaunchKernelGGL(error_bounds_kernel<SphParticle>, grid, dim3(bs), lds, stream, v, c, a, lagrange, index,
                           bounds, out);
    }
    else if (dtype == MRAY_DTYPE_F32)
        hipLaunchKernelGGL(error_bounds_kernel<float>, grid, dim3(bs), lds, stream,
                           make_volume_view((const float *)vox, nx, ny, nz, dmin, dmax), c, a, lagrange, index, bounds,
                           out);
    else
        hipLaunchKernelGGL(error_bounds_kernel<uint8_t>, grid, dim3(bs), lds, stream,
                           make_volume_view((const uint8_t *)vox, nx, ny, nz, dmin, dmax), c, a, lagrange, index,
                           bounds, out);
    return hipGetLastError() == hipSuccess ? 0 : MRAY_ERR_HIP;
}

int launch_uncertainty(const mray_camera &cam, const float *dmin, const float *dmax, const float *lagrange,
                       const uint32_t *index, const uint32_t *bounds, const uint16_t *error_bounds, int W, int H, int m,
                       const float *tf, const mray_rec_params &p, int mode, int error_type, float bound_interpolation,
                       float *rgba, hipStream_t stream)
{
    if ((mode != 0 && mode != 1) || error_type < MRAY_ERROR_MAX || error_type > MRAY_ERROR_PERCENTILE)
        return MRAY_ERR_ARG;
    UArgs a{};
    a.W = W;
    a.H = H;
    a.m = m;
    a.step = p.step_size;
    a.emission = p.medium_emission;
    a.ert = p.early_ray_termination;
    a.fourier = p.use_truncated_fourier;
    a.dmin = {dmin[0], dmin[1], dmin[2]};
    a.dmax = {dmax[0], dmax[1], dmax[2]};
    a.interp = bound_interpolation;
    // reconstruct_uncertainty_convolution_kernel (src/moment_raymarcher.cu:882-924):
    // (lower, upper) for the maximum error, (mean, error) otherwise
    a.mean_std = error_type != MRAY_ERROR_MAX;
    const DevCamera c = make_dev_camera(cam);
    const uint32_t ntiles = (uint32_t)(((W + 7) / 8) * ((H + 7) / 8));
    const int bs = block_for(m);
    const dim3 grid((ntiles + bs / 64 - 1) / (bs / 64));
    const size_t lds = (size_t)bs * m * sizeof(float);
    const float4 *tf4 = reinterpret_cast<const float4 *>(tf);
    if (mode == 0)
        hipLaunchKernelGGL(uncertainty_kernel<0>, grid, dim3(bs), lds, stream, c, a, lagrange, index, bounds,
                           error_bounds, tf4, rgba);
    else
        hipLaunchKernelGGL(uncertainty_kernel<1>, grid, dim3(bs), lds, stream, c, a, lagrange, index, bounds,
                           error_bounds, tf4, rgba);
    return hipGetLastError() == hipSuccess ? 0 : MRAY_ERR_HIP;
}

} // namespace mray
