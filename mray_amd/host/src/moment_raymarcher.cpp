// generate / reconstruct (src/moment_raymarcher.cpp:121-183, 283-353) over
// the C ABI: host images are uploaded, the kernels run on the current HIP
// device, results are downloaded into the caller's MomentImageHost /
// ColorImageHost.
#include <algorithm>
#include <cmath>

#include "../include/moment_raymarcher.h"
#include "device.hpp"

using mray_host::check;
using mray_host::DeviceBuffer;

namespace
{
void require_form(int form)
{
    if (form != MRAY_FORM_PLAIN && form != MRAY_FORM_MIXED)
        throw invalid_args("kernel_form must be MRAY_FORM_PLAIN (0) or MRAY_FORM_MIXED (1)");
}

struct VolumeHandle
{
    mray_volume v = nullptr;
    ~VolumeHandle()
    {
        if (v)
            (void)mray_hip_volume_destroy(v);
    }
};

void require_size(const MomentImageHost &img)
{
    if (img.width <= 0 || img.height <= 0 || img.num_moments <= 0 || img.num_moments > 255)
        throw invalid_args("moment image needs width, height > 0 and 1..255 moments");
}

// a device copy of the grid (freed with the handle)
void upload_volume(const RegularGrid &volume, VolumeHandle &vh)
{
    if (volume.data.size() * sizeof(float) < volume.num_voxels() * volume.get_byte_size())
        throw invalid_data("volume data smaller than its size");
    const float dmin[3] = {volume.domain.min.x, volume.domain.min.y, volume.domain.min.z};
    const float dmax[3] = {volume.domain.max.x, volume.domain.max.y, volume.domain.max.z};
    check(mray_hip_volume_create(volume.voxels(),
                                 volume.type == RegularGrid::DataType::U8 ? MRAY_DTYPE_U8 : MRAY_DTYPE_F32,
                                 volume.size.x, volume.size.y, volume.size.z, dmin, dmax, &vh.v),
          "volume_create");
}

// the particle set sorted into its cell grid on the device: every ray sample
// is Particles::sample
void upload_particles(const Particles &volume, VolumeHandle &vh)
{
    const uint32_t n = volume.size();
    if (n == 0 || volume.y.size() != n || volume.z.size() != n || volume.values.size() != n)
        throw invalid_data("particle arrays must be non-empty and of equal length");
    const Vec3f lo = volume.domain.min, hi = volume.domain.max;
    DeviceBuffer dx(volume.x.data(), n * 4), dy(volume.y.data(), n * 4), dz(volume.z.data(), n * 4),
        dv(volume.values.data(), n * 4);
    const float dmin[3] = {lo.x, lo.y, lo.z};
    const float dmax[3] = {hi.x, hi.y, hi.z};
    check(mray_hip_volume_from_particles(dx.as<float>(), dy.as<float>(), dz.as<float>(), dv.as<float>(), n,
                                         volume.smoothing_length, dmin, dmax, &vh.v),
          "volume_from_particles");
}

// clamped and truncated RGBA8 (ColorImageHost::set_color)
void to_rgba8(const vector<Vec4f> &rgba, int W, int H, ColorImageHost &cimg)
{
    if (cimg.width != W || cimg.height != H)
        throw invalid_args("colour image and moment image differ in size");
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            cimg.set_color(x, y, rgba[(size_t)y * W + x]);
}

// the generate kernel on a device-resident volume, then the compaction and
// prediction coding that generate_host applies
void generate_on(mray_volume vol, const cut::Camera &cam, MomentImageHost &img, const GenerationParameters &params)
{
    require_size(img);
    require_form(params.kernel_form);
    if (params.prediction_coding && !params.compact_image)
        throw invalid_args("prediction coding needs a compact image (compact_image=1)");
    const int W = img.width, H = img.height, m = img.num_moments;
    const size_t npix = (size_t)W * H;
    mray_camera c = mray_host::to_c(cam);
    mray_gen_params g = mray_host::to_c(params);
    DeviceBuffer d_mom(npix * m * sizeof(float));
    const bool want_bounds = img.has_bounds() || params.transform_to_bounds;
    DeviceBuffer d_bounds(want_bounds ? npix * 4 : 0);
    check(mray_hip_generate(vol, &c, &g, params.kernel_form, W, H, m, d_mom.as<float>(), d_bounds.as<uint32_t>(), nullptr, nullptr,
                            nullptr),
          "generate");
    if (img.has_bounds())
        d_bounds.download(img.bounds.data(), npix * 4);

    if (!params.compact_image)
    {
        img.is_compact = false;
        img.data.resize(npix * m);
        d_mom.download(img.data.data(), npix * m * sizeof(float));
        return;
    }
    DeviceBuffer d_index((npix + 1) * 4);
    check(mray_hip_compute_num_moments(d_mom.as<float>(), img.has_bounds() ? d_bounds.as<uint32_t>() : nullptr, W,
                                       H, m, params.error_threshold, d_index.as<uint32_t>(), nullptr),
          "compute_num_moments");
    check(mray_hip_moment_index(d_index.as<uint32_t>(), W, H, d_index.as<uint32_t>(), nullptr), "moment_index");
    img.index.resize(npix + 1);
    d_index.download(img.index.data(), (npix + 1) * 4);
    const size_t total = img.index[npix];
    DeviceBuffer d_comp(std::max<size_t>(total, 1) * sizeof(float));
    check(mray_hip_compact(d_mom.as<float>(), d_index.as<uint32_t>(), W, H, m, d_comp.as<float>(), nullptr),
          "compact");
    img.is_compact = true;
    if (params.prediction_coding)
    {
        const bool with_params = requires_coding_parameters(params.coding_warp) && m > 1;
        DeviceBuffer d_q(with_params ? (size_t)(m - 1) * sizeof(CodingParamType) : 0);
        check(mray_hip_prediction_encode(d_comp.as<float>(), d_index.as<uint32_t>(), W, H, m, params.coding_warp,
                                         params.bias, d_q.as<uint16_t>(), nullptr),
              "prediction_encode");
        img.prediction_code = true;
        img.coding_warp = params.coding_warp;
        img.coding_params.clear();
        if (with_params)
        {
            img.coding_params.resize(m - 1);
            d_q.download(img.coding_params.data(), (size_t)(m - 1) * sizeof(CodingParamType));
        }
    }
    img.data.resize(total);
    d_comp.download(img.data.data(), total * sizeof(float));
}

// multipliers (or trigonometric moments for the truncated Fourier
// reconstruction) of every pixel into d_lag, layout of mi.data
void prepare_on_device(const MomentImageHost &mi, const DeviceBuffer &d_data, const DeviceBuffer &d_index,
                       float bias, bool truncated_fourier, int form, DeviceBuffer &d_lag)
{
    const int W = mi.width, H = mi.height, m = mi.num_moments;
    const uint32_t *idx = mi.is_compact ? d_index.as<uint32_t>() : nullptr;
    if (truncated_fourier)
    {
        if (mi.prediction_code)
            throw invalid_args("the truncated Fourier reconstruction of a prediction-coded image is not supported");
        if (!mi.data.empty())
            mray_host::check_hip(hipMemcpy(d_lag.ptr, d_data.ptr, mi.data.size() * sizeof(float),
                                           hipMemcpyDeviceToDevice),
                                 "copy");
        return;
    }
    if (mi.prediction_code)
    {
        const bool with_params = requires_coding_parameters(mi.coding_warp) && m > 1;
        if (with_params && mi.coding_params.size() != (size_t)(m - 1))
            throw invalid_data("prediction-coded image without its coding parameters");
        DeviceBuffer d_q(mi.coding_params.data(), with_params ? (size_t)(m - 1) * sizeof(CodingParamType) : 0);
        check(mray_hip_prepare_coded(d_data.as<float>(), idx, W, H, m, mi.coding_warp, d_q.as<uint16_t>(),
                                     d_lag.as<float>(), nullptr),
              "prepare_coded");
        return;
    }
    check(mray_hip_prepare(d_data.as<float>(), idx, W, H, m, bias, form, d_lag.as<float>(), nullptr), "prepare");
}

void reconstruct_float(const cut::Camera &cam, const MomentImageHost &mi, const TransferFunction &tf,
                       vector<Vec4f> &rgba, const ReconstructionParameters &params)
{
    mray_host::ensure_init();
    require_size(mi);
    require_form(params.kernel_form);
    const int W = mi.width, H = mi.height, m = mi.num_moments;
    const size_t npix = (size_t)W * H;
    if (mi.is_compact && mi.index.size() != npix + 1)
        throw invalid_data("compact moment image without index");
    // data may be empty for a compact image without moments: keep a valid pointer
    std::unique_ptr<DeviceBuffer> d_data_p(mray_host::upload_nonempty(mi.data));
    const DeviceBuffer &d_data = *d_data_p;
    DeviceBuffer d_index(mi.is_compact ? mi.index.data() : nullptr, mi.is_compact ? (npix + 1) * 4 : 0);
    DeviceBuffer d_bounds(mi.has_bounds() ? mi.bounds.data() : nullptr, mi.has_bounds() ? npix * 4 : 0);
    DeviceBuffer d_lag(std::max<size_t>(mi.data.size(), 1) * sizeof(float));
    prepare_on_device(mi, d_data, d_index, params.bias, params.use_truncated_fourier, params.kernel_form, d_lag);
    DeviceBuffer d_tf(tf.data.data(), sizeof(tf.data));
    DeviceBuffer d_rgba(npix * sizeof(Vec4f));
    mray_camera c = mray_host::to_c(cam);
    mray_rec_params r = mray_host::to_c(params);
    const float dmin[3] = {mi.domain.min.x, mi.domain.min.y, mi.domain.min.z};
    const float dmax[3] = {mi.domain.max.x, mi.domain.max.y, mi.domain.max.z};
    check(mray_hip_reconstruct(&c, dmin, dmax, d_lag.as<float>(), mi.is_compact ? d_index.as<uint32_t>() : nullptr,
                               mi.has_bounds() ? d_bounds.as<uint32_t>() : nullptr, W, H, m, d_tf.as<float>(), &r,
                               params.kernel_form, d_rgba.as<float>(), nullptr, nullptr),
          "reconstruct");
    rgba.resize(npix);
    d_rgba.download(rgba.data(), npix * sizeof(Vec4f));
}
} // namespace

void generate(cut::Camera cam, const RegularGrid &volume, MomentImageHost &img, const GenerationParameters &params)
{
    mray_host::ensure_init();
    img.domain = volume.domain;
    VolumeHandle vh;
    upload_volume(volume, vh);
    generate_on(vh.v, cam, img, params);
}

void generate(cut::Camera cam, const Particles &volume, MomentImageHost &img, const GenerationParameters &params)
{
    mray_host::ensure_init();
    img.domain = volume.domain;
    VolumeHandle vh;
    upload_particles(volume, vh);
    generate_on(vh.v, cam, img, params);
}

void reconstruct(cut::Camera cam, const MomentImageHost &mi, const TransferFunction &tf, vector<Vec4f> &rgba,
                 const ReconstructionParameters &params)
{
    reconstruct_float(cam, mi, tf, rgba, params);
}

void reconstruct(cut::Camera cam, const MomentImageHost &mi, const TransferFunction &tf, ColorImageHost &cimg,
                 const ReconstructionParameters &params)
{
    if (cimg.width != mi.width || cimg.height != mi.height)
        throw invalid_args("colour image and moment image differ in size");
    vector<Vec4f> rgba;
    reconstruct_float(cam, mi, tf, rgba, params);
    to_rgba8(rgba, mi.width, mi.height, cimg);
}

void generate_reference(cut::Camera cam, const TransferFunction &tf, const RegularGrid &volume, int W, int H,
                        vector<Vec4f> &rgba, const ReferenceParameters &params)
{
    mray_host::ensure_init();
    if (W <= 0 || H <= 0)
        throw invalid_args("image size must be positive");
    VolumeHandle vh;
    upload_volume(volume, vh);
    const size_t npix = (size_t)W * H;
    DeviceBuffer d_tf(tf.data.data(), sizeof(tf.data));
    DeviceBuffer d_rgba(npix * sizeof(Vec4f));
    mray_camera c = mray_host::to_c(cam);
    mray_ref_params r = mray_host::to_c(params);
    check(mray_hip_reference(vh.v, &c, d_tf.as<float>(), &r, W, H, d_rgba.as<float>(), nullptr, nullptr),
          "reference");
    rgba.resize(npix);
    d_rgba.download(rgba.data(), npix * sizeof(Vec4f));
}

void generate_reference(cut::Camera cam, const TransferFunction &tf, const RegularGrid &volume, ColorImageHost &img,
                        const ReferenceParameters &params)
{
    vector<Vec4f> rgba;
    generate_reference(cam, tf, volume, img.width, img.height, rgba, params);
    for (int y = 0; y < img.height; ++y)
        for (int x = 0; x < img.width; ++x)
            img.set_color(x, y, rgba[(size_t)y * img.width + x]);
}

namespace
{
void error_bounds_on(mray_volume vol, const cut::Camera &cam, MomentImageHost &mi,
                     const ErrorReconstructionParameters &params)
{
    mray_host::ensure_init();
    require_size(mi);
    const int W = mi.width, H = mi.height, m = mi.num_moments;
    const size_t npix = (size_t)W * H;
    if (mi.is_compact && mi.index.size() != npix + 1)
        throw invalid_data("compact moment image without index");
    std::unique_ptr<DeviceBuffer> d_data_p(mray_host::upload_nonempty(mi.data));
    DeviceBuffer d_index(mi.is_compact ? mi.index.data() : nullptr, mi.is_compact ? (npix + 1) * 4 : 0);
    DeviceBuffer d_bounds(mi.has_bounds() ? mi.bounds.data() : nullptr, mi.has_bounds() ? npix * 4 : 0);
    DeviceBuffer d_lag(std::max<size_t>(mi.data.size(), 1) * sizeof(float));
    prepare_on_device(mi, *d_data_p, d_index, params.bias, false, MRAY_FORM_PLAIN, d_lag);
    DeviceBuffer d_err(npix * sizeof(uint16_t));
    mray_camera c = mray_host::to_c(cam);
    mray_error_params e{};
    e.error_type = params.error_type;
    e.error_percentile = params.error_percentile;
    e.bias = params.bias;
    e.data_min = params.data_min;
    e.data_max = params.data_max;
    e.step_size = params.step_size;
    check(mray_hip_error_bounds(vol, &c, d_lag.as<float>(), mi.is_compact ? d_index.as<uint32_t>() : nullptr,
                                mi.has_bounds() ? d_bounds.as<uint32_t>() : nullptr, W, H, m, &e,
                                d_err.as<uint16_t>(), nullptr),
          "error_bounds");
    mi.error_bounds.resize(npix);
    d_err.download(mi.error_bounds.data(), npix * sizeof(uint16_t));
}
} // namespace

void reconstruct_error(cut::Camera cam, const RegularGrid &volume, MomentImageHost &mi,
                       const ErrorReconstructionParameters &params)
{
    mray_host::ensure_init();
    VolumeHandle h;
    upload_volume(volume, h);
    error_bounds_on(h.v, cam, mi, params);
}

void reconstruct_error(cut::Camera cam, const Particles &volume, MomentImageHost &mi,
                       const ErrorReconstructionParameters &params)
{
    mray_host::ensure_init();
    VolumeHandle h;
    upload_particles(volume, h);
    error_bounds_on(h.v, cam, mi, params);
}

void reconstruct_uncertainty(cut::Camera cam, const MomentImageHost &mi, const TransferFunction &tf,
                             vector<Vec4f> &rgba, const UncertaintyReconstructionParameters &params, int mode)
{
    mray_host::ensure_init();
    require_size(mi);
    if (mode == 1 && tf.table2d.size() != (size_t)TransferFunction::RESOLUTION * TransferFunction::RESOLUTION)
        throw invalid_args("the uncertainty convolution needs an uncertainty transfer function "
                           "(read_from_disk_uncertainty)");
    if (!mi.has_error_bounds())
        throw invalid_args("the uncertainty renderings need a moment image with error bounds");
    const int W = mi.width, H = mi.height, m = mi.num_moments;
    const size_t npix = (size_t)W * H;
    if (mi.is_compact && mi.index.size() != npix + 1)
        throw invalid_data("compact moment image without index");
    std::unique_ptr<DeviceBuffer> d_data_p(mray_host::upload_nonempty(mi.data));
    DeviceBuffer d_index(mi.is_compact ? mi.index.data() : nullptr, mi.is_compact ? (npix + 1) * 4 : 0);
    DeviceBuffer d_bounds(mi.has_bounds() ? mi.bounds.data() : nullptr, mi.has_bounds() ? npix * 4 : 0);
    DeviceBuffer d_err(mi.error_bounds.data(), npix * sizeof(uint16_t));
    DeviceBuffer d_lag(std::max<size_t>(mi.data.size(), 1) * sizeof(float));
    prepare_on_device(mi, *d_data_p, d_index, params.rec.bias, params.rec.use_truncated_fourier, MRAY_FORM_PLAIN,
                      d_lag);
    DeviceBuffer d_tf(mode == 0 ? (const void *)tf.data.data() : (const void *)tf.table2d.data(),
                      mode == 0 ? sizeof(tf.data) : tf.table2d.size() * sizeof(Vec4f));
    DeviceBuffer d_rgba(npix * sizeof(Vec4f));
    mray_camera c = mray_host::to_c(cam);
    mray_rec_params r = mray_host::to_c(params.rec);
    const float dmin[3] = {mi.domain.min.x, mi.domain.min.y, mi.domain.min.z};
    const float dmax[3] = {mi.domain.max.x, mi.domain.max.y, mi.domain.max.z};
    check(mray_hip_reconstruct_uncertainty(&c, dmin, dmax, d_lag.as<float>(),
                                           mi.is_compact ? d_index.as<uint32_t>() : nullptr,
                                           mi.has_bounds() ? d_bounds.as<uint32_t>() : nullptr,
                                           d_err.as<uint16_t>(), W, H, m, d_tf.as<float>(), &r, mode,
                                           params.error_type, params.bound_interpolation, d_rgba.as<float>(),
                                           nullptr),
          "reconstruct_uncertainty");
    rgba.resize(npix);
    d_rgba.download(rgba.data(), npix * sizeof(Vec4f));
}

void reconstruct_uncertainty_interpolation(cut::Camera cam, const MomentImageHost &mi, const TransferFunction &tf,
                                           ColorImageHost &cimg, const UncertaintyReconstructionParameters &params)
{
    vector<Vec4f> rgba;
    reconstruct_uncertainty(cam, mi, tf, rgba, params, 0);
    to_rgba8(rgba, mi.width, mi.height, cimg);
}

void reconstruct_uncertainty_convolution(cut::Camera cam, const MomentImageHost &mi, const TransferFunction &tf,
                                         ColorImageHost &cimg, const UncertaintyReconstructionParameters &params)
{
    vector<Vec4f> rgba;
    reconstruct_uncertainty(cam, mi, tf, rgba, params, 1);
    to_rgba8(rgba, mi.width, mi.height, cimg);
}
