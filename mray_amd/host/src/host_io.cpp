// Camera files, transfer functions and parameter parsing
// (src/cut/camera.cpp:30-75, src/transfer_function.cpp:99-119, src/parameters.h).
#include <cmath>
#include <fstream>
#include <iostream>

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
