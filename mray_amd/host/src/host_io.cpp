// Camera files, transfer functions and parameter parsing
// (src/cut/camera.cpp:30-75, src/transfer_function.cpp:99-119, src/parameters.h).
#include <cmath>
#include <fstream>
#include <iostream>

#include "../../../include/mray_types.h"
#include "../include/cut/camera.h"
#include "../include/parameters.h"
#include "../include/transfer_function.h"

cut::Camera cut::load_from_file(string file, int width, int height)
{
    std::ifstream fs(file, std::ios::in);
    if (fs.fail())
        throw invalid_args("Could not open camera file " + file);
    cut::Camera cam;
    fs >> cam.orthographic;
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i)
            fs >> cam.transform[j][i];
    if (cam.orthographic)
    {
        fs >> cam.position.x >> cam.position.y >> cam.position.z;
        fs >> cam.dir.x >> cam.dir.y >> cam.dir.z;
        fs >> cam.up.x >> cam.up.y >> cam.up.z;
        fs >> cam.right.x >> cam.right.y >> cam.right.z;
    }
    else
    {
        cam.position = Vec3f(cam.transform[0][3], cam.transform[1][3], cam.transform[2][3]);
        fs >> cam.scale;
        fs >> cam.znear;
        cam.aspect = width / static_cast<float>(height);
    }
    if (fs.fail())
        throw invalid_data("Malformed camera file " + file);
    return cam;
}

std::array<Vec4f, TransferFunction::RESOLUTION> tf_from_rgba8(const unsigned char *data)
{
    std::array<Vec4f, TransferFunction::RESOLUTION> float_tf;
    for (int i = 0; i < TransferFunction::RESOLUTION; ++i)
        float_tf[i] = Vec4f(std::pow(data[i * 4] / 255.f, 2.2f), std::pow(data[i * 4 + 1] / 255.f, 2.2f),
                            std::pow(data[i * 4 + 2] / 255.f, 2.2f), data[i * 4 + 3] / 255.f);
    return float_tf;
}

std::array<Vec4f, TransferFunction::RESOLUTION> read_tf(const char *filepath)
{
    std::ifstream file(filepath, std::ios::in | std::ios::binary);
    if (file.fail())
        throw invalid_args(string("Invalid transfer function file given: ") + filepath);
    unsigned char data[TransferFunction::RESOLUTION * 4];
    file.read(reinterpret_cast<char *>(data), sizeof(data));
    if (file.gcount() != (std::streamsize)sizeof(data))
        throw invalid_data(string("Transfer function file too short: ") + filepath);
    return tf_from_rgba8(data);
}

TransferFunction preintegrate(const std::array<Vec4f, TransferFunction::RESOLUTION> &tf1d, float step_size)
{
    // the entry the reference keeps in every row (src/transfer_function.cpp:82-90)
    TransferFunction tf;
    for (int x = 0; x < TransferFunction::RESOLUTION; ++x)
    {
        Vec4f t = tf1d[x];
        t.w *= step_size;
        t.x = t.x * t.w;
        t.y = t.y * t.w;
        t.z = t.z * t.w;
        tf.data[x] = t;
    }
    return tf;
}

TransferFunction read_from_disk(const char *filepath, float step_size)
{
    return preintegrate(read_tf(filepath), step_size);
}

namespace
{
// sample1D (src/transfer_function.cpp:32-44)
Vec4f sample1d(const std::array<Vec4f, TransferFunction::RESOLUTION> &tf1d, float x)
{
    const float ip = x * static_cast<float>(TransferFunction::RESOLUTION);
    const float a = ip - std::floor(ip);
    const int xl = (int)std::floor(ip);
    const Vec4f &l = tf1d[std::min(std::max(xl, 0), TransferFunction::RESOLUTION - 1)];
    const Vec4f &r = tf1d[std::min(std::max(xl + 1, 0), TransferFunction::RESOLUTION - 1)];
    return Vec4f((1.f - a) * l.x + a * r.x, (1.f - a) * l.y + a * r.y, (1.f - a) * l.z + a * r.z,
                 (1.f - a) * l.w + a * r.w);
}

// normal_cdf (src/cut/gaussian.h:33-59): Abramowitz and Stegun 7.1.26 with
// the reference's float / double promotions and the float exp
float normal_cdf(float x)
{
    const float a1 = 0.254829592f, a2 = -0.284496736f, a3 = 1.421413741f, a4 = -1.453152027f, a5 = 1.061405429f,
                p = 0.3275911f;
    const int sign = (x < 0) ? -1 : 1;
    x = (float)((double)std::fabs(x) / 1.41421356237);
    const float t = (float)(1.0 / (1.0 + (double)(p * x)));
    const float y = (float)(1.0 - (double)((((((a5 * t + a4) * t) + a3) * t + a2) * t + a1) * t * std::exp(-x * x)));
    return (float)(0.5 * (1.0 + (double)((float)sign * y)));
}
} // namespace

// box_filter (src/transfer_function.cpp:152-185)
TransferFunction box_filter(const std::array<Vec4f, TransferFunction::RESOLUTION> &tf1d, float step_size)
{
    const int R = TransferFunction::RESOLUTION;
    TransferFunction tf;
    tf.table2d.assign((size_t)R * R, Vec4f());
    for (int y = 0; y < R; ++y)
        for (int x = 0; x < R; ++x)
        {
            const float sf = x / float(R), sb = y / float(R);
            const int num_steps = std::abs(y - x) + 2;
            const float dw = 1.0f / float(num_steps);
            Vec4f rgba(0.0f);
            for (int i = 0; i < num_steps; ++i)
            {
                const float w = float(i) / float(num_steps - 1);
                const Vec4f t = sample1d(tf1d, w * sb + (1 - w) * sf);
                const float a = dw * t.w * step_size;
                rgba.x += t.x * a;
                rgba.y += t.y * a;
                rgba.z += t.z * a;
                rgba.w += a;
            }
            tf.table2d[(size_t)y * R + x] = rgba;
        }
    return tf;
}

// gaussian_filter (src/transfer_function.cpp:187-231)
TransferFunction gaussian_filter(const std::array<Vec4f, TransferFunction::RESOLUTION> &tf1d, float step_size)
{
    const int R = TransferFunction::RESOLUTION;
    TransferFunction tf;
    tf.table2d.assign((size_t)R * R, Vec4f());
    for (int y = 0; y < R; ++y)
        for (int x = 0; x < R; ++x)
        {
            const float mean = x / float(R), sd = y / float(R);
            Vec4f rgba(0.0f);
            if (sd != 0.0f)
            {
                const float delta = 1.0f / static_cast<float>(R);
                for (int i = 0; i < R; ++i)
                {
                    const float xi = i * delta;
                    const float weight = normal_cdf((xi + delta - mean) / sd) - normal_cdf((xi - mean) / sd);
                    const float a = weight * tf1d[i].w * step_size;
                    rgba.x += tf1d[i].x * a;
                    rgba.y += tf1d[i].y * a;
                    rgba.z += tf1d[i].z * a;
                    rgba.w += a;
                }
            }
            else
            {
                rgba = tf1d[x];
                rgba.w *= step_size;
                rgba.x = rgba.x * rgba.w;
                rgba.y = rgba.y * rgba.w;
                rgba.z = rgba.z * rgba.w;
            }
            tf.table2d[(size_t)y * R + x] = rgba;
        }
    return tf;
}

TransferFunction read_from_disk_uncertainty(const char *filepath, float step_size, int error_type)
{
    const auto tf1d = read_tf(filepath);
    if (error_type == MRAY_ERROR_MAX || error_type == MRAY_ERROR_PERCENTILE)
        return box_filter(tf1d, step_size);
    if (error_type == MRAY_ERROR_RMSE)
        return gaussian_filter(tf1d, step_size);
    throw invalid_args("error_type must be 0 (max), 1 (RMSE) or 2 (percentile)");
}

namespace
{
template <class T> void read_f(std::unordered_map<string, string> &p, const char *k, T &v)
{
    auto it = p.find(k);
    if (it != p.end())
        v = std::stof(it->second);
}
template <class T> void read_i(std::unordered_map<string, string> &p, const char *k, T &v)
{
    auto it = p.find(k);
    if (it != p.end())
        v = static_cast<T>(std::stoi(it->second));
}
void read_b(std::unordered_map<string, string> &p, const char *k, bool &v)
{
    auto it = p.find(k);
    if (it != p.end())
        v = std::stoi(it->second) != 0;
}
} // namespace

GenerationParameters parse_gen_params(std::unordered_map<string, string> &params)
{
    GenerationParameters p;
    read_f(params, "step_size", p.step_size);
    read_f(params, "step_size_write_samples", p.step_size_write_samples);
    read_f(params, "data_min", p.data_min);
    read_f(params, "data_max", p.data_max);
    read_f(params, "error_threshold", p.error_threshold);
    read_f(params, "bias", p.bias);
    read_b(params, "compact_image", p.compact_image);
    read_b(params, "prediction_coding", p.prediction_coding);
    read_b(params, "compress", p.compress);
    read_b(params, "entropy_coding", p.entropy_coding);
    read_i(params, "quantization_bits", p.quantization_bits);
    read_i(params, "coding_warp", p.coding_warp);
    read_i(params, "quantization_curve_sampling", p.quantization_curve_sampling);
    read_b(params, "transform_to_bounds", p.transform_to_bounds);
    read_f(params, "transform_bound_eps", p.transform_bound_eps);
    return p;
}

// parse_ref_params (src/parameters.h:352-377)
ReferenceParameters parse_ref_params(std::unordered_map<string, string> &params)
{
    ReferenceParameters p;
    read_f(params, "step_size", p.step_size);
    read_f(params, "medium_emission", p.medium_emission);
    read_f(params, "medium_scattering", p.medium_scattering);
    read_f(params, "early_ray_termination", p.early_ray_termination);
    read_f(params, "data_min", p.data_min);
    read_f(params, "data_max", p.data_max);
    return p;
}

ErrorReconstructionParameters parse_error_params(std::unordered_map<string, string> &params)
{
    // parse_error_params (src/parameters.h:139-158): the file options are not
    // parsed there, so --error-bounds writes uncompressed float files
    ErrorReconstructionParameters p;
    read_i(params, "error_type", p.error_type);
    read_i(params, "error_percentile", p.error_percentile);
    read_f(params, "bias", p.bias);
    read_f(params, "data_min", p.data_min);
    read_f(params, "data_max", p.data_max);
    read_f(params, "step_size", p.step_size);
    return p;
}

UncertaintyReconstructionParameters parse_uncertainty_rec_params(std::unordered_map<string, string> &params)
{
    UncertaintyReconstructionParameters p;
    p.rec = parse_rec_params(params);
    read_f(params, "bound_interpolation", p.bound_interpolation);
    read_i(params, "error_type", p.error_type);
    return p;
}

ReconstructionParameters parse_rec_params(std::unordered_map<string, string> &params)
{
    ReconstructionParameters p;
    read_f(params, "bias", p.bias);
    read_f(params, "medium_emission", p.medium_emission);
    read_f(params, "medium_scattering", p.medium_scattering);
    read_f(params, "early_ray_termination", p.early_ray_termination);
    read_f(params, "data_min", p.data_min);
    read_f(params, "data_max", p.data_max);
    read_f(params, "step_size", p.step_size);
    read_f(params, "step_size_write_samples", p.step_size_write_samples);
    read_b(params, "use_truncated_fourier", p.use_truncated_fourier);
    return p;
}
