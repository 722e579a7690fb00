// C entry points of the host library (include/mray_host.h).
#include <cstring>
#include <exception>
#include <string>

#include "../../../include/mray_host.h"
#include "../include/moment_image_io.h"
#include "../include/transfer_function.h"

namespace
{
thread_local std::string g_error;

template <class F> int guarded(const char *what, F &&f)
{
    try
    {
        f();
        return MRAY_OK;
    }
    catch (const invalid_args &e)
    {
        g_error = std::string(what) + ": " + e.what();
        return MRAY_ERR_ARG;
    }
    catch (const std::exception &e)
    {
        g_error = std::string(what) + ": " + e.what();
        return MRAY_ERR_HIP;
    }
}
} // namespace

extern "C" {

const char *mray_host_last_error(void) { return g_error.c_str(); }

int mray_host_write_moment_image(const char *path, int W, int H, int m, int prediction_coded, int coding_warp,
                                 const float dmin[3], const float dmax[3], const uint32_t *index, const float *data,
                                 size_t n_data, const uint32_t *bounds, const uint16_t *error_bounds,
                                 const uint16_t *coding_params, const uint8_t *table, int compress, int bits,
                                 int entropy_coding)
{
    return guarded("mray_host_write_moment_image", [&] {
        if (!path || !dmin || !dmax || (n_data && !data) || W <= 0 || H <= 0 || m < 1 || m > 255)
            throw invalid_args("null argument or bad image size");
        MomentImageHost mi;
        mi.width = W;
        mi.height = H;
        mi.num_moments = m;
        mi.is_compact = index != nullptr;
        mi.prediction_code = prediction_coded != 0;
        mi.coding_warp = static_cast<CodingWarpType>(coding_warp);
        mi.domain = Range<Vec3f>{Vec3f{dmin[0], dmin[1], dmin[2]}, Vec3f{dmax[0], dmax[1], dmax[2]}};
        const size_t npix = (size_t)W * H;
        if (index)
        {
            mi.index.assign(index, index + npix + 1);
            if (index[npix] != n_data)
                throw invalid_args("index[W*H] must equal the number of data floats");
        }
        else if (n_data != npix * (size_t)m)
            throw invalid_args("a dense image holds W*H*m floats");
        mi.data.assign(data, data + n_data);
        if (bounds)
        {
            mi.bounds.resize(npix);
            std::memcpy(mi.bounds.data(), bounds, npix * sizeof(uint32_t));
        }
        if (error_bounds)
            mi.error_bounds.assign(error_bounds, error_bounds + npix);
        if (coding_params)
        {
            mi.coding_params.resize((size_t)m - 1);
            std::memcpy(mi.coding_params.data(), coding_params, ((size_t)m - 1) * 2 * sizeof(uint16_t));
        }
        if (table)
            set_prediction_coding_quantization_table(vector<Byte>(table, table + m));
        write_moment_image(mi, path, compress != 0, static_cast<Byte>(bits), entropy_coding != 0);
    });
}

int mray_host_load_moment_image(const char *path, int64_t hdr[10], uint32_t *index, float *data, uint32_t *bounds,
                                uint16_t *error_bounds, uint16_t *coding_params)
{
    return guarded("mray_host_load_moment_image", [&] {
        if (!path || !hdr)
            throw invalid_args("null argument");
        const MomentImageHost mi = load_moment_image(path);
        const int64_t h[10] = {mi.width,        mi.height,           mi.num_moments,         (int64_t)mi.coding_warp,
                               mi.is_compact,   mi.prediction_code,  mi.has_bounds(),        mi.has_error_bounds(),
                               mi.has_coding_params(), (int64_t)mi.data.size()};
        std::memcpy(hdr, h, sizeof h);
        if (index && !mi.index.empty())
            std::memcpy(index, mi.index.data(), mi.index.size() * sizeof(uint32_t));
        if (data)
            std::memcpy(data, mi.data.data(), mi.data.size() * sizeof(float));
        if (bounds && mi.has_bounds())
            std::memcpy(bounds, mi.bounds.data(), mi.bounds.size() * sizeof(uint32_t));
        if (error_bounds && mi.has_error_bounds())
            std::memcpy(error_bounds, mi.error_bounds.data(), mi.error_bounds.size() * sizeof(uint16_t));
        if (coding_params && mi.has_coding_params())
            std::memcpy(coding_params, mi.coding_params.data(), mi.coding_params.size() * 2 * sizeof(uint16_t));
    });
}

int mray_host_uncertainty_tf(const uint8_t *rgba8, float step_size, int error_type, float *table)
{
    return guarded("mray_host_uncertainty_tf", [&] {
        if (!rgba8 || !table)
            throw invalid_args("null argument");
        const auto tf1d = tf_from_rgba8(rgba8);
        TransferFunction tf;
        if (error_type == MRAY_ERROR_MAX || error_type == MRAY_ERROR_PERCENTILE)
            tf = box_filter(tf1d, step_size);
        else if (error_type == MRAY_ERROR_RMSE)
            tf = gaussian_filter(tf1d, step_size);
        else
            throw invalid_args("error_type must be 0 (max), 1 (RMSE) or 2 (percentile)");
        std::memcpy(table, tf.table2d.data(), tf.table2d.size() * sizeof(Vec4f));
    });
}

int mray_host_quantization_table(uint8_t table[255])
{
    if (!uses_prediction_coding_quantization_table())
        return 0;
    const auto t = get_prediction_coding_quantization_table();
    if (table)
        std::memcpy(table, t.data(), std::min<size_t>(t.size(), 255));
    return 1;
}
}
