// MomentImageHost operations (src/moment_image.cpp:69-183) on the GPU.
#include <algorithm>

#include "../include/moment_image.h"
#include "../include/moment_raymarcher.h"
#include "device.hpp"

using mray_host::check;
using mray_host::DeviceBuffer;

void MomentImageHost::compact()
{
    mray_host::ensure_init();
    if (is_compact)
        throw invalid_args("moment image is already compact");
    const size_t npix = (size_t)width * height;
    if (index.size() != npix + 1)
        throw invalid_data("compact() needs index[] with the number of moments per pixel");
    DeviceBuffer d_dense(data.data(), data.size() * sizeof(float));
    DeviceBuffer d_index(index.data(), (npix + 1) * 4);
    check(mray_hip_moment_index(d_index.as<uint32_t>(), width, height, d_index.as<uint32_t>(), nullptr),
          "moment_index");
    d_index.download(index.data(), (npix + 1) * 4);
    const size_t total = index[npix];
    DeviceBuffer d_out(std::max<size_t>(total, 1) * sizeof(float));
    check(mray_hip_compact(d_dense.as<float>(), d_index.as<uint32_t>(), width, height, num_moments,
                           d_out.as<float>(), nullptr),
          "compact");
    data.resize(total);
    d_out.download(data.data(), total * sizeof(float));
    is_compact = true;
}

void MomentImageHost::prediction_encode(int warp, float bias)
{
    mray_host::ensure_init();
    if (!is_compact)
        throw invalid_args("prediction_encode needs a compact image");
    const size_t npix = (size_t)width * height;
    std::unique_ptr<DeviceBuffer> d_data_p(mray_host::upload_nonempty(data));
    DeviceBuffer &d_data = *d_data_p;
    DeviceBuffer d_index(index.data(), (npix + 1) * 4);
    const bool with_params = requires_coding_parameters(warp) && num_moments > 1;
    DeviceBuffer d_q(with_params ? (size_t)(num_moments - 1) * sizeof(CodingParamType) : 0);
    check(mray_hip_prediction_encode(d_data.as<float>(), d_index.as<uint32_t>(), width, height, num_moments, warp,
                                     bias, d_q.as<uint16_t>(), nullptr),
          "prediction_encode");
    d_data.download(data.data(), data.size() * sizeof(float));
    coding_params.clear();
    if (with_params)
    {
        coding_params.resize(num_moments - 1);
        d_q.download(coding_params.data(), coding_params.size() * sizeof(CodingParamType));
    }
    prediction_code = true;
    coding_warp = warp;
}

vector<float> MomentImageHost::prepare_all_moments(float bias, int form) const
{
    mray_host::ensure_init();
    const size_t npix = (size_t)width * height;
    std::unique_ptr<DeviceBuffer> d_data_p(mray_host::upload_nonempty(data));
    DeviceBuffer &d_data = *d_data_p;
    DeviceBuffer d_index(is_compact ? index.data() : nullptr, is_compact ? (npix + 1) * 4 : 0);
    DeviceBuffer d_lag(std::max<size_t>(data.size(), 1) * sizeof(float));
    const uint32_t *idx = is_compact ? d_index.as<uint32_t>() : nullptr;
    if (prediction_code)
    {
        const bool with_params = requires_coding_parameters(coding_warp) && num_moments > 1;
        DeviceBuffer d_q(coding_params.data(), with_params ? coding_params.size() * sizeof(CodingParamType) : 0);
        check(mray_hip_prepare_coded(d_data.as<float>(), idx, width, height, num_moments, coding_warp,
                                     d_q.as<uint16_t>(), d_lag.as<float>(), nullptr),
              "prepare_coded");
    }
    else
        check(mray_hip_prepare(d_data.as<float>(), idx, width, height, num_moments, bias, form,
                               d_lag.as<float>(), nullptr),
              "prepare");
    vector<float> out(data.size());
    d_lag.download(out.data(), out.size() * sizeof(float));
    return out;
}

vector<float> MomentImageHost::prepare_moments(int x, int y, float bias, int form) const
{
    auto all = prepare_all_moments(bias, form);
    const size_t idx = get_idx(x, y);
    return vector<float>(all.begin() + idx, all.begin() + idx + get_num_moments(x, y));
}
