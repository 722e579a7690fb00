// C ABI of libmray_hip.so (declared in include/mray_hip.h): argument
// validation, device/volume handles, status codes and the thread-local error
// message. Kernels live in generate.hip, prepare.hip, reconstruct.hip,
// synth.hip.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

#include "../../include/mray_hip.h"
#include "launch.hpp"
#include "sph_common.hpp"

namespace mray
{
void set_lds_limit(size_t bytes);
}

struct mray_volume_s
{
    const void *data;
    bool owned;
    int dtype;
    int nx, ny, nz;
    float dmin[3], dmax[3];
    int device;
};

namespace
{
thread_local char g_err[512] = "";
bool g_init = false;

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(MRAY_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    return MRAY_OK;
}

int check_status(int rc, const char *what)
{
    if (rc == MRAY_ERR_UNSUPPORTED)
        return fail(rc, "%s: configuration not supported by this build", what);
    if (rc == MRAY_ERR_OOM)
        return fail(rc, "%s: device allocation failed", what);
    if (rc != 0)
        return fail(rc, "%s: failed (%d)", what, rc);
    return check_launch(what);
}

bool bad_image(int W, int H, int m) { return W <= 0 || H <= 0 || m <= 0 || m > 255 || (long long)W * H > (1LL << 31); }
} // namespace

extern "C" {

const char *mray_hip_last_error(void) { return g_err; }

int mray_hip_init(int device)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(MRAY_ERR_HIP, "mray_hip_init: no HIP device available");
    if (device < 0 || device >= count)
        return fail(MRAY_ERR_ARG, "mray_hip_init: device %d out of range [0, %d)", device, count);
    if (hipSetDevice(device) != hipSuccess)
        return fail(MRAY_ERR_HIP, "mray_hip_init: hipSetDevice(%d) failed", device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess)
        return fail(MRAY_ERR_HIP, "mray_hip_init: hipGetDeviceProperties failed");
    mray::set_lds_limit(prop.sharedMemPerBlock);
    // the library's scratch comes from the device's default pool
    // (hipMallocAsync); with the default release threshold of zero the pool
    // returns freed memory at every synchronisation, and the next call maps
    // it again. The pool keeps what it holds instead.
    hipMemPool_t pool;
    uint64_t keep = UINT64_MAX;
    if (hipDeviceGetDefaultMemPool(&pool, device) != hipSuccess ||
        hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess)
        return fail(MRAY_ERR_HIP, "mray_hip_init: setting the memory pool's release threshold failed");
    // the device address of the allocation-free entry points' scratch array
    if (mray::noalloc_scratch_init(device, nullptr) != 0)
        return fail(MRAY_ERR_HIP, "mray_hip_init: hipGetSymbolAddress of the scratch array failed");
    g_init = true;
    return MRAY_OK;
}

int mray_hip_volume_create(const void *host_voxels, int dtype, int nx, int ny, int nz, const float dom_min[3],
                           const float dom_max[3], mray_volume *out)
{
    if (!out || !host_voxels || !dom_min || !dom_max)
        return fail(MRAY_ERR_ARG, "mray_hip_volume_create: null argument");
    if ((dtype != MRAY_DTYPE_F32 && dtype != MRAY_DTYPE_U8) || nx <= 0 || ny <= 0 || nz <= 0)
        return fail(MRAY_ERR_ARG, "mray_hip_volume_create: bad dtype or size");
    size_t bytes = (size_t)nx * ny * nz * (dtype == MRAY_DTYPE_F32 ? 4 : 1);
    void *d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess)
        return fail(MRAY_ERR_OOM, "mray_hip_volume_create: hipMalloc(%zu) failed", bytes);
    if (hipMemcpy(d, host_voxels, bytes, hipMemcpyHostToDevice) != hipSuccess)
    {
        (void)hipFree(d);
        return fail(MRAY_ERR_HIP, "mray_hip_volume_create: upload failed");
    }
    int rc = mray_hip_volume_wrap(d, dtype, nx, ny, nz, dom_min, dom_max, out);
    if (rc != MRAY_OK)
    {
        (void)hipFree(d);
        return rc;
    }
    (*out)->owned = true;
    return MRAY_OK;
}

int mray_hip_volume_wrap(const void *d_voxels, int dtype, int nx, int ny, int nz, const float dom_min[3],
                         const float dom_max[3], mray_volume *out)
{
    if (!out || !d_voxels || !dom_min || !dom_max)
        return fail(MRAY_ERR_ARG, "mray_hip_volume_wrap: null argument");
    if ((dtype != MRAY_DTYPE_F32 && dtype != MRAY_DTYPE_U8) || nx <= 0 || ny <= 0 || nz <= 0)
        return fail(MRAY_ERR_ARG, "mray_hip_volume_wrap: bad dtype or size");
    if ((long long)nx * ny > (1LL << 31))
        return fail(MRAY_ERR_ARG, "mray_hip_volume_wrap: nx*ny must fit in int");
    for (int i = 0; i < 3; ++i)
        if (!(dom_max[i] > dom_min[i]))
            return fail(MRAY_ERR_ARG, "mray_hip_volume_wrap: empty domain on axis %d", i);
    auto *v = new (std::nothrow) mray_volume_s;
    if (!v)
        return fail(MRAY_ERR_OOM, "mray_hip_volume_wrap: out of host memory");
    v->data = d_voxels;
    v->owned = false;
    v->dtype = dtype;
    v->nx = nx;
    v->ny = ny;
    v->nz = nz;
    std::memcpy(v->dmin, dom_min, sizeof(v->dmin));
    std::memcpy(v->dmax, dom_max, sizeof(v->dmax));
    (void)hipGetDevice(&v->device);
    *out = v;
    return MRAY_OK;
}

int mray_hip_volume_from_particles(const float *d_x, const float *d_y, const float *d_z, const float *d_v, uint32_t n,
                                   float h, const float dom_min[3], const float dom_max[3], mray_volume *out)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_volume_from_particles: call mray_hip_init first");
    if (!out || !d_x || !d_y || !d_z || !d_v || !dom_min || !dom_max || n == 0)
        return fail(MRAY_ERR_ARG, "mray_hip_volume_from_particles: null argument or no particles");
    if (!(h > 0.0f))
        return fail(MRAY_ERR_ARG, "mray_hip_volume_from_particles: smoothing length must be > 0");
    for (int i = 0; i < 3; ++i)
        if (!(dom_max[i] > dom_min[i]))
            return fail(MRAY_ERR_ARG, "mray_hip_volume_from_particles: empty domain on axis %d", i);
    auto *sb = new (std::nothrow) mray::SphBuffers;
    auto *v = new (std::nothrow) mray_volume_s;
    if (!sb || !v)
    {
        delete sb;
        delete v;
        return fail(MRAY_ERR_OOM, "mray_hip_volume_from_particles: out of host memory");
    }
    const int rc = mray::sph_build(d_x, d_y, d_z, d_v, n, h, true, *sb, nullptr);
    if (rc != 0 || hipStreamSynchronize(nullptr) != hipSuccess)
    {
        delete sb;
        delete v;
        return fail(rc ? rc : MRAY_ERR_HIP, "mray_hip_volume_from_particles: building the cell grid failed (%d)", rc);
    }
    v->data = sb;
    v->owned = true;
    v->dtype = MRAY_DTYPE_PARTICLES;
    v->nx = v->ny = v->nz = 1;
    std::memcpy(v->dmin, dom_min, sizeof(v->dmin));
    std::memcpy(v->dmax, dom_max, sizeof(v->dmax));
    (void)hipGetDevice(&v->device);
    *out = v;
    return MRAY_OK;
}

int mray_hip_volume_destroy(mray_volume vol)
{
    if (!vol)
        return MRAY_OK;
    int rc = MRAY_OK;
    if (vol->dtype == MRAY_DTYPE_PARTICLES)
    {
        auto *sb = static_cast<mray::SphBuffers *>(const_cast<void *>(vol->data));
        mray::sph_free(*sb, true, nullptr);
        delete sb;
    }
    else if (vol->owned && hipFree(const_cast<void *>(vol->data)) != hipSuccess)
        rc = fail(MRAY_ERR_HIP, "mray_hip_volume_destroy: hipFree failed");
    delete vol;
    return rc;
}

const void *mray_hip_volume_data(mray_volume vol)
{
    if (vol && vol->dtype == MRAY_DTYPE_PARTICLES)
        return static_cast<const mray::SphBuffers *>(vol->data)->parts;
    return vol ? vol->data : nullptr;
}

int mray_hip_generate(mray_volume vol, const mray_camera *cam, const mray_gen_params *p, int form, int W, int H,
                      int m, float *d_moments, uint32_t *d_bounds, int32_t *d_samples, float *d_ray_range,
                      mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_generate: call mray_hip_init first");
    if (!vol || !cam || !p || !d_moments)
        return fail(MRAY_ERR_ARG, "mray_hip_generate: null argument");
    if (bad_image(W, H, m))
        return fail(MRAY_ERR_ARG, "mray_hip_generate: bad image size %dx%d or moment count %d", W, H, m);
    if (form != MRAY_FORM_PLAIN && form != MRAY_FORM_MIXED)
        return fail(MRAY_ERR_ARG, "mray_hip_generate: unknown form %d", form);
    if (p->transform_to_bounds && !d_bounds)
        return fail(MRAY_ERR_ARG, "mray_hip_generate: transform_to_bounds needs d_bounds");
    if (!(p->step_size > 0.0f) || !(p->data_max != p->data_min))
        return fail(MRAY_ERR_ARG, "mray_hip_generate: step_size must be > 0 and data_max != data_min");
    int rc = mray::launch_generate(vol->data, vol->dtype, vol->nx, vol->ny, vol->nz, vol->dmin, vol->dmax, *cam, *p,
                                   form, W, H, m, d_moments, d_bounds, d_samples, d_ray_range, (hipStream_t)stream);
    return check_status(rc, "mray_hip_generate");
}

int mray_hip_reference(mray_volume vol, const mray_camera *cam, const float *d_tf, const mray_ref_params *p, int W,
                       int H, float *d_rgba, int32_t *d_steps, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_reference: call mray_hip_init first");
    if (!vol || !cam || !d_tf || !p || !d_rgba)
        return fail(MRAY_ERR_ARG, "mray_hip_reference: null argument");
    if (bad_image(W, H, 1))
        return fail(MRAY_ERR_ARG, "mray_hip_reference: bad image size %dx%d", W, H);
    if (!(p->step_size > 0.0f) || !(p->data_max != p->data_min))
        return fail(MRAY_ERR_ARG, "mray_hip_reference: step_size must be > 0 and data_max != data_min");
    if (vol->dtype == MRAY_DTYPE_PARTICLES)
        return fail(MRAY_ERR_UNSUPPORTED, "mray_hip_reference: regular grids only");
    int rc = mray::launch_reference(vol->data, vol->dtype, vol->nx, vol->ny, vol->nz, vol->dmin, vol->dmax, *cam, *p,
                                    W, H, d_tf, d_rgba, d_steps, (hipStream_t)stream);
    if (rc == MRAY_ERR_UNSUPPORTED)
        return fail(rc, "mray_hip_reference: single scattering is not supported by this build");
    return check_status(rc, "mray_hip_reference");
}

int mray_hip_prepare(const float *d_moments, const uint32_t *d_index, int W, int H, int m, float bias, int form,
                     float *d_lagrange, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_prepare: call mray_hip_init first");
    if (!d_moments || !d_lagrange)
        return fail(MRAY_ERR_ARG, "mray_hip_prepare: null argument");
    if (bad_image(W, H, m))
        return fail(MRAY_ERR_ARG, "mray_hip_prepare: bad image size %dx%d or moment count %d", W, H, m);
    if (form != MRAY_FORM_PLAIN && form != MRAY_FORM_MIXED)
        return fail(MRAY_ERR_ARG, "mray_hip_prepare: unknown form %d", form);
    int rc = mray::launch_prepare(d_moments, d_index, W, H, m, bias, form, d_lagrange, (hipStream_t)stream);
    return check_status(rc, "mray_hip_prepare");
}

int mray_hip_reconstruct(const mray_camera *cam, const float dom_min[3], const float dom_max[3],
                         const float *d_lagrange, const uint32_t *d_index, const uint32_t *d_bounds, int W, int H,
                         int m, const float *d_tf, const mray_rec_params *p, int form, float *d_rgba,
                         int32_t *d_steps, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_reconstruct: call mray_hip_init first");
    if (!cam || !dom_min || !dom_max || !d_lagrange || !d_tf || !p || !d_rgba)
        return fail(MRAY_ERR_ARG, "mray_hip_reconstruct: null argument");
    if (bad_image(W, H, m))
        return fail(MRAY_ERR_ARG, "mray_hip_reconstruct: bad image size %dx%d or moment count %d", W, H, m);
    if (form != MRAY_FORM_PLAIN && form != MRAY_FORM_MIXED)
        return fail(MRAY_ERR_ARG, "mray_hip_reconstruct: unknown form %d", form);
    if (!(p->step_size > 0.0f))
        return fail(MRAY_ERR_ARG, "mray_hip_reconstruct: step_size must be > 0");
    int rc = mray::launch_reconstruct(*cam, dom_min, dom_max, d_lagrange, d_index, d_bounds, W, H, m, d_tf, *p, form,
                                      d_rgba, d_steps, (hipStream_t)stream);
    return check_status(rc, "mray_hip_reconstruct");
}

int mray_hip_error_bounds(mray_volume vol, const mray_camera *cam, const float *d_lagrange, const uint32_t *d_index,
                          const uint32_t *d_bounds, int W, int H, int m, const mray_error_params *p,
                          uint16_t *d_error_bounds, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_error_bounds: call mray_hip_init first");
    if (!vol || !cam || !d_lagrange || !p || !d_error_bounds)
        return fail(MRAY_ERR_ARG, "mray_hip_error_bounds: null argument");
    if (bad_image(W, H, m))
        return fail(MRAY_ERR_ARG, "mray_hip_error_bounds: bad image size %dx%d or moment count %d", W, H, m);
    if (!(p->step_size > 0.0f) || !(p->data_max != p->data_min))
        return fail(MRAY_ERR_ARG, "mray_hip_error_bounds: step_size must be > 0 and data_max != data_min");
    if (p->error_type < MRAY_ERROR_MAX || p->error_type > MRAY_ERROR_PERCENTILE)
        return fail(MRAY_ERR_ARG, "mray_hip_error_bounds: unknown error_type %d", p->error_type);
    int rc = mray::launch_error_bounds(vol->data, vol->dtype, vol->nx, vol->ny, vol->nz, vol->dmin, vol->dmax, *cam,
                                       d_lagrange, d_index, d_bounds, W, H, m, *p, d_error_bounds,
                                       (hipStream_t)stream);
    return check_status(rc, "mray_hip_error_bounds");
}

int mray_hip_reconstruct_uncertainty(const mray_camera *cam, const float dom_min[3], const float dom_max[3],
                                     const float *d_lagrange, const uint32_t *d_index, const uint32_t *d_bounds,
                                     const uint16_t *d_error_bounds, int W, int H, int m, const float *d_tf,
                                     const mray_rec_params *p, int mode, int error_type, float bound_interpolation,
                                     float *d_rgba, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_reconstruct_uncertainty: call mray_hip_init first");
    if (!cam || !dom_min || !dom_max || !d_lagrange || !d_error_bounds || !d_tf || !p || !d_rgba)
        return fail(MRAY_ERR_ARG, "mray_hip_reconstruct_uncertainty: null argument");
    if (bad_image(W, H, m))
        return fail(MRAY_ERR_ARG, "mray_hip_reconstruct_uncertainty: bad image size %dx%d or moment count %d", W, H,
                    m);
    if (mode != 0 && mode != 1)
        return fail(MRAY_ERR_ARG, "mray_hip_reconstruct_uncertainty: mode must be 0 (interpolation) or 1 "
                                  "(convolution)");
    if (error_type < MRAY_ERROR_MAX || error_type > MRAY_ERROR_PERCENTILE)
        return fail(MRAY_ERR_ARG, "mray_hip_reconstruct_uncertainty: unknown error_type %d", error_type);
    if (!(p->step_size > 0.0f))
        return fail(MRAY_ERR_ARG, "mray_hip_reconstruct_uncertainty: step_size must be > 0");
    int rc = mray::launch_uncertainty(*cam, dom_min, dom_max, d_lagrange, d_index, d_bounds, d_error_bounds, W, H, m,
                                      d_tf, *p, mode, error_type, bound_interpolation, d_rgba, (hipStream_t)stream);
    return check_status(rc, "mray_hip_reconstruct_uncertainty");
}

int mray_hip_prepare_f64(const float *d_moments, const uint32_t *d_index, int W, int H, int m, float bias,
                         double *d_lagrange, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_prepare_f64: call mray_hip_init first");
    if (!d_moments || !d_lagrange)
        return fail(MRAY_ERR_ARG, "mray_hip_prepare_f64: null argument");
    if (bad_image(W, H, m))
        return fail(MRAY_ERR_ARG, "mray_hip_prepare_f64: bad image size %dx%d or moment count %d", W, H, m);
    int rc = mray::launch_prepare_f64(d_moments, d_index, W, H, m, bias, d_lagrange, (hipStream_t)stream);
    return check_status(rc, "mray_hip_prepare_f64");
}

int mray_hip_reconstruct_f64(const mray_camera *cam, const float dom_min[3], const float dom_max[3],
                             const double *d_lagrange, const uint32_t *d_index, const uint32_t *d_bounds, int W,
                             int H, int m, const float *d_tf, const mray_rec_params *p, float *d_rgba,
                             int32_t *d_steps, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_reconstruct_f64: call mray_hip_init first");
    if (!cam || !dom_min || !dom_max || !d_lagrange || !d_tf || !p || !d_rgba)
        return fail(MRAY_ERR_ARG, "mray_hip_reconstruct_f64: null argument");
    if (bad_image(W, H, m))
        return fail(MRAY_ERR_ARG, "mray_hip_reconstruct_f64: bad image size %dx%d or moment count %d", W, H, m);
    if (!(p->step_size > 0.0f))
        return fail(MRAY_ERR_ARG, "mray_hip_reconstruct_f64: step_size must be > 0");
    int rc = mray::launch_reconstruct_f64(*cam, dom_min, dom_max, d_lagrange, d_index, d_bounds, W, H, m, d_tf, *p,
                                          d_rgba, d_steps, (hipStream_t)stream);
    return check_status(rc, "mray_hip_reconstruct_f64");
}

int mray_hip_reconstruct_backward(const mray_camera *cam, const float dom_min[3], const float dom_max[3],
                                  const float *d_lagrange, const uint32_t *d_index, const uint32_t *d_bounds, int W,
                                  int H, int m, const float *d_tf, const mray_rec_params *p, const float *d_rgba,
                                  const float *d_grad_rgba, float *d_grad_lagrange, float *d_grad_tf,
                                  mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_reconstruct_backward: call mray_hip_init first");
    if (!cam || !dom_min || !dom_max || !d_lagrange || !d_tf || !p || !d_rgba || !d_grad_rgba)
        return fail(MRAY_ERR_ARG, "mray_hip_reconstruct_backward: null argument");
    if (!d_grad_lagrange && !d_grad_tf)
        return fail(MRAY_ERR_ARG, "mray_hip_reconstruct_backward: no gradient requested");
    if (bad_image(W, H, m))
        return fail(MRAY_ERR_ARG, "mray_hip_reconstruct_backward: bad image size %dx%d or moment count %d", W, H, m);
    if (!(p->step_size > 0.0f))
        return fail(MRAY_ERR_ARG, "mray_hip_reconstruct_backward: step_size must be > 0");
    int rc = mray::launch_reconstruct_adjoint(*cam, dom_min, dom_max, d_lagrange, d_index, d_bounds, W, H, m, d_tf, *p,
                                              d_rgba, d_grad_rgba, d_grad_lagrange, d_grad_tf, (hipStream_t)stream);
    if (rc == MRAY_ERR_UNSUPPORTED)
        return fail(rc, "mray_hip_reconstruct_backward: configuration not supported");
    return check_status(rc, "mray_hip_reconstruct_backward");
}

int mray_hip_generate_backward(mray_volume vol, const mray_camera *cam, const mray_gen_params *p, int W, int H, int m,
                               const float *d_grad_moments, const float *d_ray_range, float *d_grad_volume,
                               mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_generate_backward: call mray_hip_init first");
    if (!vol || !cam || !p || !d_grad_moments || !d_grad_volume)
        return fail(MRAY_ERR_ARG, "mray_hip_generate_backward: null argument");
    if (bad_image(W, H, m))
        return fail(MRAY_ERR_ARG, "mray_hip_generate_backward: bad image size %dx%d or moment count %d", W, H, m);
    if (vol->dtype != MRAY_DTYPE_F32)
        return fail(MRAY_ERR_UNSUPPORTED, "mray_hip_generate_backward: gradients need an F32 volume");
    if (!(p->step_size > 0.0f) || !(p->data_max > p->data_min))
        return fail(MRAY_ERR_ARG, "mray_hip_generate_backward: step_size must be > 0 and data_max > data_min");
    int rc = mray::launch_generate_adjoint((const float *)vol->data, vol->nx, vol->ny, vol->nz, vol->dmin, vol->dmax,
                                           *cam, *p, W, H, m, d_grad_moments, d_ray_range, d_grad_volume,
                                           (hipStream_t)stream);
    if (rc == MRAY_ERR_UNSUPPORTED)
        return fail(rc, "mray_hip_generate_backward: configuration not supported");
    return check_status(rc, "mray_hip_generate_backward");
}

int mray_hip_prepare_backward(const float *d_moments, const uint32_t *d_index, int W, int H, int m, float bias,
                              const float *d_grad_lagrange, float *d_grad_moments, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_prepare_backward: call mray_hip_init first");
    if (!d_moments || !d_grad_lagrange || !d_grad_moments)
        return fail(MRAY_ERR_ARG, "mray_hip_prepare_backward: null argument");
    if (bad_image(W, H, m))
        return fail(MRAY_ERR_ARG, "mray_hip_prepare_backward: bad image size %dx%d or moment count %d", W, H, m);
    int rc = mray::launch_prepare_adjoint(d_moments, d_index, W, H, m, bias, d_grad_lagrange, d_grad_moments,
                                          (hipStream_t)stream);
    if (rc == MRAY_ERR_OOM)
        return fail(rc, "mray_hip_prepare_backward: scratch allocation failed");
    return check_status(rc, "mray_hip_prepare_backward");
}

int mray_hip_compute_num_moments(const float *d_moments, const uint32_t *d_bounds, int W, int H, int m,
                                 float error_threshold, uint32_t *d_counts, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_compute_num_moments: call mray_hip_init first");
    if (!d_moments || !d_counts)
        return fail(MRAY_ERR_ARG, "mray_hip_compute_num_moments: null argument");
    if (bad_image(W, H, m))
        return fail(MRAY_ERR_ARG, "mray_hip_compute_num_moments: bad image size %dx%d or moment count %d", W, H, m);
    int rc = mray::launch_num_moments(d_moments, d_bounds, W, H, m, error_threshold, d_counts, (hipStream_t)stream);
    if (rc == MRAY_ERR_UNSUPPORTED)
        return fail(rc, "mray_hip_compute_num_moments: %d moments do not fit the LDS", m);
    return check_status(rc, "mray_hip_compute_num_moments");
}

int mray_hip_moment_index(const uint32_t *d_counts, int W, int H, uint32_t *d_index, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_moment_index: call mray_hip_init first");
    if (!d_counts || !d_index)
        return fail(MRAY_ERR_ARG, "mray_hip_moment_index: null argument");
    if (bad_image(W, H, 1))
        return fail(MRAY_ERR_ARG, "mray_hip_moment_index: bad image size %dx%d", W, H);
    int rc = mray::launch_moment_index(d_counts, W * H, d_index, (hipStream_t)stream);
    if (rc == MRAY_ERR_OOM)
        return fail(rc, "mray_hip_moment_index: scratch allocation failed");
    return check_status(rc, "mray_hip_moment_index");
}

int mray_hip_clamp_num_moments(uint32_t *d_counts, int W, int H, int min_moments, int max_moments,
                               mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_clamp_num_moments: call mray_hip_init first");
    if (!d_counts)
        return fail(MRAY_ERR_ARG, "mray_hip_clamp_num_moments: null argument");
    if (bad_image(W, H, 1) || min_moments < 1 || max_moments < min_moments || max_moments > 255)
        return fail(MRAY_ERR_ARG, "mray_hip_clamp_num_moments: bad image size %dx%d or range [%d, %d]", W, H,
                    min_moments, max_moments);
    int rc = mray::launch_clamp_counts(d_counts, W * H, min_moments, max_moments, (hipStream_t)stream);
    return check_status(rc, "mray_hip_clamp_num_moments");
}

int mray_hip_moment_index_noalloc(const uint32_t *d_counts, int W, int H, uint32_t *d_index, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_moment_index_noalloc: call mray_hip_init first");
    if (!d_counts || !d_index)
        return fail(MRAY_ERR_ARG, "mray_hip_moment_index_noalloc: null argument");
    if (bad_image(W, H, 1))
        return fail(MRAY_ERR_ARG, "mray_hip_moment_index_noalloc: bad image size %dx%d", W, H);
    if ((size_t)W * H > mray::noalloc_max_pixels())
        return fail(MRAY_ERR_ARG, "mray_hip_moment_index_noalloc: %dx%d pixels exceed its scratch limit of %zu pixels",
                    W, H, mray::noalloc_max_pixels());
    int rc = mray::launch_moment_index_noalloc(d_counts, W * H, d_index, (hipStream_t)stream);
    if (rc == MRAY_ERR_HIP)
        return fail(rc, "mray_hip_moment_index_noalloc: resolving the scratch array's device address failed");
    return check_status(rc, "mray_hip_moment_index_noalloc");
}

int mray_hip_compact(const float *d_dense, const uint32_t *d_index, int W, int H, int m, float *d_compact,
                     mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_compact: call mray_hip_init first");
    if (!d_dense || !d_index || !d_compact)
        return fail(MRAY_ERR_ARG, "mray_hip_compact: null argument");
    if (bad_image(W, H, m) || (long long)W * H * m > 0xFFFFFFFFLL)
        return fail(MRAY_ERR_ARG, "mray_hip_compact: bad image size %dx%d or moment count %d", W, H, m);
    int rc = mray::launch_compact(d_dense, d_index, W, H, m, d_compact, (hipStream_t)stream);
    return check_status(rc, "mray_hip_compact");
}

int mray_hip_prediction_encode(float *d_data, const uint32_t *d_index, int W, int H, int m, int coding_warp,
                               float bias, uint16_t *d_coding_params, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_prediction_encode: call mray_hip_init first");
    if (!d_data || !d_index)
        return fail(MRAY_ERR_ARG, "mray_hip_prediction_encode: null argument");
    if (bad_image(W, H, m))
        return fail(MRAY_ERR_ARG, "mray_hip_prediction_encode: bad image size %dx%d or moment count %d", W, H, m);
    if (coding_warp != MRAY_CODING_WARP_DEFAULT_TRANSFORMED && coding_warp != MRAY_CODING_WARP_NONE)
        return fail(MRAY_ERR_ARG, "mray_hip_prediction_encode: unknown coding warp %d", coding_warp);
    if (coding_warp == MRAY_CODING_WARP_DEFAULT_TRANSFORMED && m > 1 && !d_coding_params)
        return fail(MRAY_ERR_ARG, "mray_hip_prediction_encode: the default warp needs d_coding_params");
    int rc = mray::launch_prediction_encode(d_data, d_index, W, H, m, coding_warp, bias, d_coding_params,
                                            (hipStream_t)stream);
    if (rc == MRAY_ERR_UNSUPPORTED)
        return fail(rc, "mray_hip_prediction_encode: %d moments do not fit the LDS", m);
    if (rc == MRAY_ERR_OOM)
        return fail(rc, "mray_hip_prediction_encode: scratch allocation failed");
    return check_status(rc, "mray_hip_prediction_encode");
}

int mray_hip_prepare_coded(const float *d_moments, const uint32_t *d_index, int W, int H, int m, int coding_warp,
                           const uint16_t *d_coding_params, float *d_lagrange, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_prepare_coded: call mray_hip_init first");
    if (!d_moments || !d_lagrange)
        return fail(MRAY_ERR_ARG, "mray_hip_prepare_coded: null argument");
    if (bad_image(W, H, m))
        return fail(MRAY_ERR_ARG, "mray_hip_prepare_coded: bad image size %dx%d or moment count %d", W, H, m);
    if (coding_warp != MRAY_CODING_WARP_DEFAULT_TRANSFORMED && coding_warp != MRAY_CODING_WARP_NONE)
        return fail(MRAY_ERR_ARG, "mray_hip_prepare_coded: unknown coding warp %d", coding_warp);
    if (coding_warp == MRAY_CODING_WARP_DEFAULT_TRANSFORMED && m > 1 && !d_coding_params)
        return fail(MRAY_ERR_ARG, "mray_hip_prepare_coded: the default warp needs d_coding_params");
    int rc = mray::launch_prepare_coded(d_moments, d_index, W, H, m, coding_warp, d_coding_params, d_lagrange,
                                        (hipStream_t)stream);
    if (rc == MRAY_ERR_UNSUPPORTED)
        return fail(rc, "mray_hip_prepare_coded: %d moments do not fit the LDS", m);
    return check_status(rc, "mray_hip_prepare_coded");
}

int mray_hip_quantize(const float *d_data, const uint32_t *d_index, int W, int H, int prediction_coded,
                      const uint8_t *bits, int n_bits, uint8_t *d_stream, size_t capacity, size_t *stream_bytes,
                      mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_quantize: call mray_hip_init first");
    if (!d_index || !bits || (d_stream && !d_data))
        return fail(MRAY_ERR_ARG, "mray_hip_quantize: null argument");
    if (bad_image(W, H, 1) || n_bits < 1 || n_bits > 256)
        return fail(MRAY_ERR_ARG, "mray_hip_quantize: bad image size %dx%d or table length %d", W, H, n_bits);
    size_t nbytes = 0;
    int rc = mray::launch_quantize(d_data, d_index, W, H, prediction_coded, bits, n_bits, d_stream, capacity, &nbytes,
                                   (hipStream_t)stream);
    if (stream_bytes)
        *stream_bytes = nbytes;
    if (rc == MRAY_ERR_ARG)
        return fail(rc, "mray_hip_quantize: widths must be 1..32, every pixel at most %d moments, and the stream "
                        "(%zu bytes) within the capacity (%zu)", n_bits, nbytes, capacity);
    if (rc == MRAY_ERR_OOM)
        return fail(rc, "mray_hip_quantize: scratch allocation failed");
    return check_status(rc, "mray_hip_quantize");
}

int mray_hip_quantization_table(const float *d_data, const uint32_t *d_index, int W, int H, int m, int coding_warp,
                                const uint16_t *d_coding_params, int start_bits, uint8_t *table, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_quantization_table: call mray_hip_init first");
    if (!d_data || !d_index || !table)
        return fail(MRAY_ERR_ARG, "mray_hip_quantization_table: null argument");
    if (bad_image(W, H, m) || start_bits < 2 || start_bits > 31)
        return fail(MRAY_ERR_ARG, "mray_hip_quantization_table: bad image size %dx%d, moment count %d or start bits %d",
                    W, H, m, start_bits);
    if (coding_warp != MRAY_CODING_WARP_DEFAULT_TRANSFORMED && coding_warp != MRAY_CODING_WARP_NONE)
        return fail(MRAY_ERR_ARG, "mray_hip_quantization_table: unknown coding warp %d", coding_warp);
    if (coding_warp == MRAY_CODING_WARP_DEFAULT_TRANSFORMED && m > 1 && !d_coding_params)
        return fail(MRAY_ERR_ARG, "mray_hip_quantization_table: the default warp needs d_coding_params");
    int rc = mray::launch_quantization_table(d_data, d_index, W, H, m, coding_warp, d_coding_params, start_bits, table,
                                             (hipStream_t)stream);
    if (rc == MRAY_ERR_UNSUPPORTED)
        return fail(rc, "mray_hip_quantization_table: %d moments do not fit the LDS", m);
    if (rc == MRAY_ERR_ARG)
        return fail(rc, "mray_hip_quantization_table: a pixel has more than %d moments", m);
    if (rc == MRAY_ERR_OOM)
        return fail(rc, "mray_hip_quantization_table: scratch allocation failed");
    return check_status(rc, "mray_hip_quantization_table");
}

int mray_hip_dequantize(const uint8_t *d_stream, size_t stream_bytes, const uint32_t *d_index, int W, int H,
                        int prediction_coded, const uint8_t *bits, int n_bits, float *d_data, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_dequantize: call mray_hip_init first");
    if (!d_index || !bits || !d_data || (!d_stream && stream_bytes))
        return fail(MRAY_ERR_ARG, "mray_hip_dequantize: null argument");
    if (bad_image(W, H, 1) || n_bits < 1 || n_bits > 256)
        return fail(MRAY_ERR_ARG, "mray_hip_dequantize: bad image size %dx%d or table length %d", W, H, n_bits);
    int rc = mray::launch_dequantize(d_stream, stream_bytes, d_index, W, H, prediction_coded, bits, n_bits, d_data,
                                     (hipStream_t)stream);
    if (rc == MRAY_ERR_ARG)
        return fail(rc, "mray_hip_dequantize: widths must be 1..32, every pixel at most %d moments, and the "
                        "stream (%zu bytes) must hold every whole byte of the words", n_bits, stream_bytes);
    if (rc == MRAY_ERR_OOM)
        return fail(rc, "mray_hip_dequantize: scratch allocation failed");
    return check_status(rc, "mray_hip_dequantize");
}

int mray_hip_quantize_dense(const float *d_data, size_t n, int m, int bits, uint8_t *d_out, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_quantize_dense: call mray_hip_init first");
    if ((n && (!d_data || !d_out)) || m < 1 || m > 255 || bits < 1 || bits > 32)
        return fail(MRAY_ERR_ARG, "mray_hip_quantize_dense: null argument, moment count %d or width %d", m, bits);
    return check_status(mray::launch_quantize_dense(d_data, n, m, bits, d_out, (hipStream_t)stream),
                        "mray_hip_quantize_dense");
}

int mray_hip_dequantize_dense(const uint8_t *d_in, size_t n, int m, int bits, float *d_data, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_dequantize_dense: call mray_hip_init first");
    if ((n && (!d_in || !d_data)) || m < 1 || m > 255 || bits < 1 || bits > 32)
        return fail(MRAY_ERR_ARG, "mray_hip_dequantize_dense: null argument, moment count %d or width %d", m, bits);
    return check_status(mray::launch_dequantize_dense(d_in, n, m, bits, d_data, (hipStream_t)stream),
                        "mray_hip_dequantize_dense");
}

int mray_hip_entropy_encode(const float *d_data, const uint32_t *d_index, int W, int H, int m, int prediction_coded,
                            const uint8_t *bits, uint8_t *h_out, size_t capacity, size_t *out_bytes,
                            mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_entropy_encode: call mray_hip_init first");
    if (!d_index || !bits || !d_data)
        return fail(MRAY_ERR_ARG, "mray_hip_entropy_encode: null argument");
    if (bad_image(W, H, 1) || m < 1 || m > 255)
        return fail(MRAY_ERR_ARG, "mray_hip_entropy_encode: bad image size %dx%d or moment count %d", W, H, m);
    size_t n = 0;
    int rc = mray::launch_entropy_encode(d_data, d_index, W, H, m, prediction_coded, bits, h_out, capacity, &n,
                                         (hipStream_t)stream);
    if (out_bytes)
        *out_bytes = n;
    if (rc == MRAY_ERR_ARG)
        return fail(rc, "mray_hip_entropy_encode: widths must be 1..32 for moment 0 and 1..14 above, every pixel at "
                        "most %d moments, and the section (%zu bytes) within the capacity (%zu)", m, n, capacity);
    if (rc == MRAY_ERR_UNSUPPORTED)
        return fail(rc, "mray_hip_entropy_encode: the device cannot hold the model of the widest alphabet in LDS");
    if (rc == MRAY_ERR_OOM)
        return fail(rc, "mray_hip_entropy_encode: scratch allocation failed");
    return check_status(rc, "mray_hip_entropy_encode");
}

int mray_hip_entropy_decode(const uint8_t *h_in, size_t n_in, const uint32_t *d_index, int W, int H, int m,
                            int prediction_coded, const uint8_t *bits, float *d_data, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_entropy_decode: call mray_hip_init first");
    if (!h_in || !d_index || !bits || !d_data)
        return fail(MRAY_ERR_ARG, "mray_hip_entropy_decode: null argument");
    if (bad_image(W, H, 1) || m < 1 || m > 255)
        return fail(MRAY_ERR_ARG, "mray_hip_entropy_decode: bad image size %dx%d or moment count %d", W, H, m);
    int rc = mray::launch_entropy_decode(h_in, n_in, d_index, W, H, m, prediction_coded, bits, d_data,
                                         (hipStream_t)stream);
    if (rc == MRAY_ERR_ARG)
        return fail(rc, "mray_hip_entropy_decode: widths must be 1..32 for moment 0 and 1..14 above, and the "
                        "section (%zu bytes) must hold every length its headers give", n_in);
    if (rc == MRAY_ERR_UNSUPPORTED)
        return fail(rc, "mray_hip_entropy_decode: the device cannot hold the model of the widest alphabet in LDS");
    if (rc == MRAY_ERR_OOM)
        return fail(rc, "mray_hip_entropy_decode: scratch allocation failed");
    return check_status(rc, "mray_hip_entropy_decode");
}

int mray_hip_particles_to_grid(const float *d_x, const float *d_y, const float *d_z, const float *d_v, uint32_t n,
                               float h, const float dom_min[3], const float dom_max[3], int nx, int ny, int nz,
                               float *d_grid, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_particles_to_grid: call mray_hip_init first");
    if (!d_x || !d_y || !d_z || !d_v || !dom_min || !dom_max || !d_grid)
        return fail(MRAY_ERR_ARG, "mray_hip_particles_to_grid: null argument");
    if (n == 0 || n > 0x7fffffffu || !(h > 0.0f) || nx <= 0 || ny <= 0 || nz <= 0)
        return fail(MRAY_ERR_ARG, "mray_hip_particles_to_grid: bad particle count, smoothing length or grid size");
    int rc = mray::launch_particles_to_grid(d_x, d_y, d_z, d_v, n, h, dom_min, dom_max, nx, ny, nz, d_grid,
                                            (hipStream_t)stream);
    if (rc == MRAY_ERR_OOM)
        return fail(rc, "mray_hip_particles_to_grid: device allocation failed");
    if (rc == MRAY_ERR_UNSUPPORTED)
        return fail(rc, "mray_hip_particles_to_grid: the cell grid of size h exceeds 2^32 cells");
    return check_status(rc, "mray_hip_particles_to_grid");
}

int mray_hip_synth_mixing_layer(float *d_out, int n, const double *amp, const double *theta, const double *phi,
                                uint64_t seed, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_synth_mixing_layer: call mray_hip_init first");
    if (!d_out || !amp || !theta || !phi || n <= 0)
        return fail(MRAY_ERR_ARG, "mray_hip_synth_mixing_layer: bad argument");
    int rc = mray::launch_synth_mixing_layer(d_out, n, amp, theta, phi, seed, (hipStream_t)stream);
    return check_status(rc, "mray_hip_synth_mixing_layer");
}

int mray_hip_synth_noise_blobs(uint8_t *d_out, int nx, int ny, int nz, const float *centres, const float *sigmas,
                               const float *amps, int n_blobs, uint64_t seed, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_synth_noise_blobs: call mray_hip_init first");
    if (!d_out || nx <= 0 || ny <= 0 || nz <= 0 || n_blobs < 0 || n_blobs > 64 ||
        (n_blobs > 0 && (!centres || !sigmas || !amps)))
        return fail(MRAY_ERR_ARG, "mray_hip_synth_noise_blobs: bad argument");
    int rc = mray::launch_synth_noise_blobs(d_out, nx, ny, nz, centres, sigmas, amps, n_blobs, seed,
                                            (hipStream_t)stream);
    return check_status(rc, "mray_hip_synth_noise_blobs");
}

int mray_hip_synth_marschner_lobb(float *d_out, int n, mray_stream stream)
{
    if (!g_init)
        return fail(MRAY_ERR_NOT_INIT, "mray_hip_synth_marschner_lobb: call mray_hip_init first");
    if (!d_out || n <= 0)
        return fail(MRAY_ERR_ARG, "mray_hip_synth_marschner_lobb: bad argument");
    int rc = mray::launch_synth_marschner_lobb(d_out, n, (hipStream_t)stream);
    return check_status(rc, "mray_hip_synth_marschner_lobb");
}

} // extern "C"
