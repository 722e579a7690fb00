// Device buffers and the C ABI plumbing of the host library.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../../include/mray_hip.h"
#include "../include/cut/camera.h"
#include "../include/parameters.h"

namespace mray_host
{

inline void check(int status, const char *what)
{
    if (status != MRAY_OK)
        throw std::runtime_error(std::string(what) + ": " + mray_hip_last_error());
}

inline void check_hip(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// selects device 0 once per process (MRAY_DEVICE overrides)
inline void ensure_init()
{
    static std::once_flag once;
    static int status = MRAY_OK;
    std::call_once(once, [] {
        const char *d = std::getenv("MRAY_DEVICE");
        status = mray_hip_init(d ? std::atoi(d) : 0);
    });
    check(status, "mray_hip_init");
}

// owning device allocation
struct DeviceBuffer
{
    void *ptr = nullptr;
    size_t bytes = 0;

    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) : bytes(n)
    {
        if (n)
            check_hip(hipMalloc(&ptr, n), "hipMalloc");
    }
    DeviceBuffer(const void *host, size_t n) : DeviceBuffer(n)
    {
        if (n)
            check_hip(hipMemcpy(ptr, host, n, hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer()
    {
        if (ptr)
            (void)hipFree(ptr);
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
    void download(void *host, size_t n) const
    {
        if (n)
            check_hip(hipMemcpy(host, ptr, n, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }
};

// a device copy of v that is never a null allocation (empty compact images)
template <class T> DeviceBuffer *upload_nonempty(const std::vector<T> &v)
{
    auto *b = new DeviceBuffer(std::max<size_t>(v.size(), 1) * sizeof(T));
    if (!v.empty())
        check_hip(hipMemcpy(b->ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    return b;
}

inline mray_camera to_c(const cut::Camera &c)
{
    mray_camera o{};
    o.orthographic = c.orthographic ? 1 : 0;
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i)
            o.transform[j * 4 + i] = c.transform[j][i];
    const Vec3f *v[4] = {&c.position, &c.right, &c.up, &c.dir};
    float *d[4] = {o.position, o.right, o.up, o.dir};
    for (int k = 0; k < 4; ++k)
    {
        d[k][0] = v[k]->x;
        d[k][1] = v[k]->y;
        d[k][2] = v[k]->z;
    }
    o.scale = c.scale;
    o.aspect = c.aspect;
    o.znear = c.znear;
    return o;
}

inline mray_gen_params to_c(const GenerationParameters &p)
{
    mray_gen_params g{};
    g.step_size = p.step_size;
    g.data_min = p.data_min;
    g.data_max = p.data_max;
    g.error_threshold = p.error_threshold;
    g.bias = p.bias;
    g.compact_image = p.compact_image;
    g.prediction_coding = p.prediction_coding;
    g.coding_warp = p.coding_warp;
    g.transform_to_bounds = p.transform_to_bounds;
    g.transform_bound_eps = p.transform_bound_eps;
    return g;
}

inline mray_rec_params to_c(const ReconstructionParameters &p)
{
    mray_rec_params r{};
    r.bias = p.bias;
    r.medium_emission = p.medium_emission;
    r.medium_scattering = p.medium_scattering;
    r.early_ray_termination = p.early_ray_termination;
    r.data_min = p.data_min;
    r.data_max = p.data_max;
    r.step_size = p.step_size;
    r.use_truncated_fourier = p.use_truncated_fourier;
    return r;
}

inline mray_ref_params to_c(const ReferenceParameters &p)
{
    mray_ref_params r{};
    r.step_size = p.step_size;
    r.medium_emission = p.medium_emission;
    r.medium_scattering = p.medium_scattering;
    r.early_ray_termination = p.early_ray_termination;
    r.data_min = p.data_min;
    r.data_max = p.data_max;
    r.single_scattering = 0;
    return r;
}

} // namespace mray_host
