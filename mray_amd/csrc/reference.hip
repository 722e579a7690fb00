// Reference direct volume rendering: the march that renders the volume itself
// through the transfer function (no moments), the ground truth the
// moment-based images are compared against. reference_ray
// (src/raymarch_reference.h:93-162) as called by reference_host
// (src/moment_raymarcher.cpp:47-61) / generate_reference_device, without
// single scattering (its default).
//
// One thread per ray, 8x8 pixel tiles per wavefront as the other marches.
// Every value is the reference CPU path's operation for operation: the
// trilinear sample (divisions by the domain extent and data range by div_rn,
// the IEEE quotient), map_density, the table lookup with its weight quirk
// (tf_sample), float compositing without contraction and T = exp(-tau) in
// double, so the RGBA is bit-identical to the CPU oracle's. The next
// sample's eight loads are issued one step ahead (PendingSample).
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "launch.hpp"

namespace mray
{

namespace
{

struct RefArgs
{
    int W, H;
    float step, emission, scattering, ert;
    float data_min, data_max, data_rinv;
};

template <typename T, bool OFF32>
__global__ __launch_bounds__(256) void reference_kernel(VolumeView<T> vol, DevCamera cam, RefArgs a,
                                                        const float4 *__restrict__ tf_g, float4 *__restrict__ rgba,
                                                        int32_t *__restrict__ steps)
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
    Ray r = pixel_ray(cam, a.W, a.H, x, y);
    float t_near, t_far;
    const bool hit = intersect_aabb(r, vol.dmin, vol.dmax, t_near, t_far);
    t_near = rmax(t_near, cam.znear);
    float L0 = 0.0f, L1 = 0.0f, L2 = 0.0f, alpha = 0.0f;
    int ns = 0;
    if (hit && (t_far > 0))
    {
        auto issue_at = [&](PendingSample<T, OFF32> &ps, float tt) {
            ps.issue(vol, r.o.x + tt * r.d.x, r.o.y + tt * r.d.y, r.o.z + tt * r.d.z);
        };
        auto density = [&](float sv) { return map_density_r(sv, a.data_min, a.data_max, a.data_rinv); };
        float t = t_near;
        PendingSample<T, OFF32> ps;
        issue_at(ps, t);
        float fs = density(ps.value());
        issue_at(ps, t + a.step);
        float T_ = 1.0f, tau = 0.0f;
        // the reference loops while t < t_far; the cap only ends a march whose
        // t + step stops advancing
        const int cap = march_cap(t_near, t_far, a.step) + 1;
        for (int it = 0; t < t_far && it < cap; ++it)
        {
            t += a.step;
            ++ns;
            const float s = ps.value();
            // the loaded values die here; the next sample's loads go after
            __builtin_amdgcn_sched_barrier(0);
            issue_at(ps, t + a.step);
            const float bs = density(s);
            float4 c = tf_sample(tf, fs, bs);
            if (s >= 1e10f) // the reference's colour of an invalid sample
                c = make_float4(0.0f, 1.0f, 0.0f, 0.05f);
            // L += (Le + Ls sigma_s) T with Ls = 0 (no single scattering)
            L0 = L0 + (a.emission * c.x + 0.0f * a.scattering) * T_;
            L1 = L1 + (a.emission * c.y + 0.0f * a.scattering) * T_;
            L2 = L2 + (a.emission * c.z + 0.0f * a.scattering) * T_;
            tau += c.w;
            T_ = (float)exp((double)(-tau));
            alpha = 1.0f - T_;
            fs = bs;
            if (alpha >= a.ert)
            {
                alpha = 1.0f;
                break;
            }
        }
    }
    rgba[pix] = make_float4(L0, L1, L2, alpha);
    if (steps)
        steps[pix] = ns;
}

template <typename T>
int launch_typed(const T *vox, int nx, int ny, int nz, const float *dmin, const float *dmax, const DevCamera &c,
                 const RefArgs &a, const float4 *tf, float4 *rgba, int32_t *steps, hipStream_t stream)
{
    const VolumeView<T> v = make_volume_view(vox, nx, ny, nz, dmin, dmax);
    const uint32_t ntiles = (uint32_t)(((a.W + 7) / 8) * ((a.H + 7) / 8));
    const dim3 grid((ntiles + 3) / 4);
    // byte offsets of every voxel fit in 32 bits up to 4 GiB of volume
    const bool off32 = (uint64_t)nx * ny * nz * sizeof(T) <= (uint64_t)0xFFFFFFFFu + 1u;
    if (off32)
        hipLaunchKernelGGL((reference_kernel<T, true>), grid, dim3(256), 0, stream, v, c, a, tf, rgba, steps);
    else
        hipLaunchKernelGGL((reference_kernel<T, false>), grid, dim3(256), 0, stream, v, c, a, tf, rgba, steps);
    return 0;
}

} // namespace

int launch_reference(const void *vox, int dtype, int nx, int ny, int nz, const float *dmin, const float *dmax,
                     const mray_camera &cam, const mray_ref_params &p, int W, int H, const float *tf, float *rgba,
                     int32_t *steps, hipStream_t stream)
{
    // reference_ray's single-scattering term runs when ss.enable && sigma_s > 0
    if (p.single_scattering && p.medium_scattering > 0.0f)
        return MRAY_ERR_UNSUPPORTED;
    RefArgs a;
    a.W = W;
    a.H = H;
    a.step = p.step_size;
    a.emission = p.medium_emission;
    a.scattering = p.medium_scattering;
    a.ert = p.early_ray_termination;
    a.data_min = p.data_min;
    a.data_max = p.data_max;
    a.data_rinv = host_reciprocal(p.data_max - p.data_min);
    const DevCamera c = make_dev_camera(cam);
    const float4 *tf4 = reinterpret_cast<const float4 *>(tf);
    float4 *out = reinterpret_cast<float4 *>(rgba);
    if (dtype == MRAY_DTYPE_F32)
        return launch_typed((const float *)vox, nx, ny, nz, dmin, dmax, c, a, tf4, out, steps, stream);
    if (dtype == MRAY_DTYPE_U8)
        return launch_typed((const uint8_t *)vox, nx, ny, nz, dmin, dmax, c, a, tf4, out, steps, stream);
    return MRAY_ERR_ARG;
}

} // namespace mray
