// mray command line (src/main.cpp:563-660): the generate and reconstruct
// modes with the reference's "k=v,k=v" parameter strings, on one MI355X.
//
//   mray --generate <moment_image.bin> <volume.raw> "k=v,..."
//       width, height, num_moments, camera, GenerationParameters keys,
//       dims=NX:NY:NZ, dtype=f32|u8, domain=X0:Y0:Z0:X1:Y1:Z1 (default the
//       unit cube); volumes are raw voxel files [z][y][x] (netCDF and
//       .h5part loaders are not part of this build)
//   mray --reconstruct <moment_image.bin> <output.ppm> "k=v,..."
//       tf, camera, preintegration_step_size_factor, background_r/g/b,
//       ReconstructionParameters keys; float_rgba=<file> also writes the
//       un-clamped float RGBA (4 floats per pixel, rows top-down)
//   mray --reference <output.ppm> <volume.raw> "k=v,..."
//       width, height, tf, camera, preintegration_step_size_factor,
//       background_r/g/b, ReferenceParameters keys, dims / dtype / domain as
//       for --generate; float_rgba=<file> as for --reconstruct
//       (src/main.cpp:134-168, 740-752)
//   mray --error-bounds <moment_image.bin> <volume.raw> "k=v,..."
//       camera, ErrorReconstructionParameters keys (error_type,
//       error_percentile, bias, data_min, data_max, step_size), dims / dtype
//       / domain as for --generate; adds the per-pixel error bounds to the
//       image and rewrites it with float moments (src/main.cpp:333-367)
//   mray --uncertainty-interpolate <moment_image.bin> <output.ppm> "k=v,..."
//   mray --uncertainty-convolution <moment_image.bin> <output.ppm> "k=v,..."
//       as --reconstruct, plus bound_interpolation / error_type
//       (src/main.cpp:475-519); the convolution classifies with
//       read_from_disk_uncertainty's table of the tf at step_size *
//       preintegration_step_size_factor; both in the reference's FP32
//       arithmetic
//   generate and reconstruct: form=mixed (default) | plain
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <unordered_map>

#include "../include/moment_image_io.h"
#include "../include/moment_raymarcher.h"
#include "../../../include/mray_types.h"

using std::unordered_map;

namespace
{

unordered_map<string, string> parse_parameters(const string &params)
{
    unordered_map<string, string> out;
    size_t start = 0;
    while (start <= params.size())
    {
        size_t end = params.find(',', start);
        const string pair = params.substr(start, end == string::npos ? string::npos : end - start);
        const size_t eq = pair.find('=');
        if (!pair.empty())
            out[pair.substr(0, eq)] = eq == string::npos ? "" : pair.substr(eq + 1);
        if (end == string::npos)
            break;
        start = end + 1;
    }
    return out;
}

vector<float> split_floats(const string &s)
{
    vector<float> v;
    std::stringstream ss(s);
    string item;
    while (std::getline(ss, item, ':'))
        v.push_back(std::stof(item));
    return v;
}

const string &need(unordered_map<string, string> &p, const char *key)
{
    auto it = p.find(key);
    if (it == p.end())
        throw invalid_args(string("missing parameter ") + key);
    return it->second;
}

// the form=plain|mixed option: the kernel form of this run
int parse_form(unordered_map<string, string> &p)
{
    auto it = p.find("form");
    if (it == p.end() || it->second == "mixed")
        return MRAY_FORM_MIXED;
    if (it->second == "plain")
        return MRAY_FORM_PLAIN;
    throw invalid_args("form must be plain or mixed");
}

RegularGrid load_raw(const string &file, unordered_map<string, string> &p)
{
    auto dims = split_floats(need(p, "dims"));
    if (dims.size() != 3)
        throw invalid_args("dims must be NX:NY:NZ");
    const bool u8 = p.count("dtype") && p["dtype"] == "u8";
    Range<Vec3f> dom(Vec3f(0, 0, 0), Vec3f(1, 1, 1));
    if (p.count("domain"))
    {
        auto d = split_floats(p["domain"]);
        if (d.size() != 6)
            throw invalid_args("domain must be X0:Y0:Z0:X1:Y1:Z1");
        dom = Range<Vec3f>(Vec3f(d[0], d[1], d[2]), Vec3f(d[3], d[4], d[5]));
    }
    RegularGrid g(Vec3i((int)dims[0], (int)dims[1], (int)dims[2]), dom,
                  u8 ? RegularGrid::DataType::U8 : RegularGrid::DataType::F32);
    const size_t bytes = g.num_voxels() * g.get_byte_size();
    g.data.resize((bytes + 3) / 4);
    std::ifstream in(file, std::ios::binary);
    if (!in)
        throw invalid_args("cannot open volume " + file);
    in.read(reinterpret_cast<char *>(g.data.data()), bytes);
    if ((size_t)in.gcount() != bytes)
        throw invalid_data("volume file " + file + " is smaller than dims x dtype");
    return g;
}

void run_generate(const string &mi_file, const string &dataset, unordered_map<string, string> p)
{
    const int width = std::stoi(need(p, "width")), height = std::stoi(need(p, "height"));
    const int num_moments = std::stoi(need(p, "num_moments"));
    if (num_moments > 255)
        throw invalid_args("Number of moments cannot be greater than 255");
    auto gen_params = parse_gen_params(p);
    gen_params.kernel_form = parse_form(p);
    auto cam = cut::load_from_file(need(p, "camera"), width, height);
    MomentImageHost mi(width, height, num_moments, gen_params.compact_image, gen_params.transform_to_bounds);
    auto volume = load_raw(dataset, p);
    mi.domain = volume.domain;
    generate(cam, volume, mi, gen_params);
    // the quantization curve of a quantized prediction-coded image, on a
    // sample of it (src/main.cpp:318-327)
    if (gen_params.prediction_coding && gen_params.quantization_bits < 32)
    {
        const int sw = std::min(mi.width, gen_params.quantization_curve_sampling);
        const int sh = std::min(mi.height, gen_params.quantization_curve_sampling);
        create_quantization_table(sample_moment_image(mi, sw, sh), gen_params.quantization_bits);
    }
    write_moment_image(mi, mi_file, gen_params.compress, gen_params.quantization_bits, gen_params.entropy_coding);
}

void run_reconstruct(const string &mi_file, const string &output, unordered_map<string, string> p)
{
    auto rec_params = parse_rec_params(p);
    rec_params.kernel_form = parse_form(p);
    float factor = 1.f;
    if (p.count("preintegration_step_size_factor"))
        factor = std::stof(p["preintegration_step_size_factor"]);
    auto tf = read_from_disk(need(p, "tf").c_str(), rec_params.step_size * factor);
    auto mi = load_moment_image(mi_file);
    auto cam = cut::load_from_file(need(p, "camera"), mi.width, mi.height);
    vector<Vec4f> rgba;
    reconstruct(cam, mi, tf, rgba, rec_params);
    ColorImageHost img(mi.width, mi.height);
    for (int y = 0; y < mi.height; ++y)
        for (int x = 0; x < mi.width; ++x)
            img.set_color(x, y, rgba[(size_t)y * mi.width + x]);
    Vec3f bg(1.f, 1.f, 1.f);
    if (p.count("background_r"))
        bg.x = std::stof(p["background_r"]);
    if (p.count("background_g"))
        bg.y = std::stof(p["background_g"]);
    if (p.count("background_b"))
        bg.z = std::stof(p["background_b"]);
    write_PPM(output, img, bg);
    if (p.count("float_rgba"))
    {
        std::ofstream f(p["float_rgba"], std::ios::binary);
        f.write(reinterpret_cast<const char *>(rgba.data()), rgba.size() * sizeof(Vec4f));
    }
}

Vec3f background(unordered_map<string, string> &p)
{
    Vec3f bg(1.f, 1.f, 1.f);
    if (p.count("background_r"))
        bg.x = std::stof(p["background_r"]);
    if (p.count("background_g"))
        bg.y = std::stof(p["background_g"]);
    if (p.count("background_b"))
        bg.z = std::stof(p["background_b"]);
    return bg;
}

void write_outputs(const string &output, const vector<Vec4f> &rgba, int W, int H, unordered_map<string, string> &p)
{
    ColorImageHost img(W, H);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            img.set_color(x, y, rgba[(size_t)y * W + x]);
    write_PPM(output, img, background(p));
    if (p.count("float_rgba"))
    {
        std::ofstream f(p["float_rgba"], std::ios::binary);
        f.write(reinterpret_cast<const char *>(rgba.data()), rgba.size() * sizeof(Vec4f));
    }
}

// compute_errors (src/main.cpp:333-367): raw volumes only
void run_error_bounds(const string &mi_file, const string &dataset, unordered_map<string, string> p)
{
    auto error_params = parse_error_params(p);
    auto mi = load_moment_image(mi_file);
    auto cam = cut::load_from_file(need(p, "camera"), mi.width, mi.height);
    auto volume = load_raw(dataset, p);
    mi.domain = volume.domain;
    reconstruct_error(cam, volume, mi, error_params);
    write_moment_image(mi, mi_file, error_params.compress, error_params.quantization_bits,
                       error_params.entropy_coding);
}

// reconstruct_uncertainty_interpolate / _convolution (src/main.cpp:475-519)
void run_uncertainty(const string &mi_file, const string &output, unordered_map<string, string> p, int mode)
{
    auto up = parse_uncertainty_rec_params(p);
    float factor = 1.f;
    if (p.count("preintegration_step_size_factor"))
        factor = std::stof(p["preintegration_step_size_factor"]);
    auto tf = mode == 0 ? read_from_disk(need(p, "tf").c_str(), up.rec.step_size * factor)
                        : read_from_disk_uncertainty(need(p, "tf").c_str(), up.rec.step_size * factor,
                                                     up.error_type);
    auto mi = load_moment_image(mi_file);
    auto cam = cut::load_from_file(need(p, "camera"), mi.width, mi.height);
    vector<Vec4f> rgba;
    reconstruct_uncertainty(cam, mi, tf, rgba, up, mode);
    write_outputs(output, rgba, mi.width, mi.height, p);
}

// the reference renderer (src/main.cpp:134-168): raw volumes only
void run_reference(const string &output, const string &dataset, unordered_map<string, string> p)
{
    auto ref_params = parse_ref_params(p);
    float factor = 1.f;
    if (p.count("preintegration_step_size_factor"))
        factor = std::stof(p["preintegration_step_size_factor"]);
    auto tf = read_from_disk(need(p, "tf").c_str(), ref_params.step_size * factor);
    const int width = std::stoi(need(p, "width")), height = std::stoi(need(p, "height"));
    auto cam = cut::load_from_file(need(p, "camera"), width, height);
    auto volume = load_raw(dataset, p);
    vector<Vec4f> rgba;
    generate_reference(cam, tf, volume, width, height, rgba, ref_params);
    ColorImageHost img(width, height);
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x)
            img.set_color(x, y, rgba[(size_t)y * width + x]);
    Vec3f bg(1.f, 1.f, 1.f);
    if (p.count("background_r"))
        bg.x = std::stof(p["background_r"]);
    if (p.count("background_g"))
        bg.y = std::stof(p["background_g"]);
    if (p.count("background_b"))
        bg.z = std::stof(p["background_b"]);
    write_PPM(output, img, bg);
    if (p.count("float_rgba"))
    {
        std::ofstream f(p["float_rgba"], std::ios::binary);
        f.write(reinterpret_cast<const char *>(rgba.data()), rgba.size() * sizeof(Vec4f));
    }
}

} // namespace

int main(int argc, char *argv[])
{
    if (argc < 5)
    {
        std::cerr << "usage: mray --generate <moment_image> <volume.raw> \"k=v,...\"\n"
                     "       mray --reconstruct <moment_image> <output.ppm> \"k=v,...\"\n"
                     "       mray --reference <output.ppm> <volume.raw> \"k=v,...\"\n"
                     "       mray --error-bounds <moment_image> <volume.raw> \"k=v,...\"\n"
                     "       mray --uncertainty-interpolate <moment_image> <output.ppm> \"k=v,...\"\n"
                     "       mray --uncertainty-convolution <moment_image> <output.ppm> \"k=v,...\"\n";
        return 2;
    }
    const string mode(argv[1]);
    try
    {
        if (mode == "--generate")
            run_generate(argv[2], argv[3], parse_parameters(argv[4]));
        else if (mode == "--reconstruct")
            run_reconstruct(argv[2], argv[3], parse_parameters(argv[4]));
        else if (mode == "--reference")
            run_reference(argv[2], argv[3], parse_parameters(argv[4]));
        else if (mode == "--error-bounds")
            run_error_bounds(argv[2], argv[3], parse_parameters(argv[4]));
        else if (mode == "--uncertainty-interpolate")
            run_uncertainty(argv[2], argv[3], parse_parameters(argv[4]), 0);
        else if (mode == "--uncertainty-convolution")
            run_uncertainty(argv[2], argv[3], parse_parameters(argv[4]), 1);
        else
        {
            std::cerr << "unknown mode " << mode << "\n";
            return 2;
        }
    }
    catch (const std::exception &e)
    {
        std::cerr << "mray: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
