// Moment-image files (src/moment_image_io.cpp): uncompressed, with the moment
// data as floats or, for compact images, as the quantized bit stream of
// moment_quantization (packed and unpacked on the GPU by mray_hip_quantize /
// mray_hip_dequantize). LZ4 compression and entropy coding are not supported.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iterator>

#include "../include/color_image.h"
#include "../include/moment_image_io.h"
#include "device.hpp"

namespace
{
template <class T> void put(std::fstream &out, const T &v) { out.write(reinterpret_cast<const char *>(&v), sizeof(T)); }
template <class T> void get(std::fstream &in, T &v) { in.read(reinterpret_cast<char *>(&v), sizeof(T)); }
void put_bytes(std::fstream &out, const void *p, size_t n) { out.write(static_cast<const char *>(p), n); }
void get_bytes(std::fstream &in, void *p, size_t n, const std::string &file)
{
    in.read(static_cast<char *>(p), n);
    if ((size_t)in.gcount() != n)
        throw invalid_data("truncated moment image " + file);
}

// the word widths of the quantized stream: for prediction-coded images the
// file's table when it carries one, else get_prediction_coding_quantization_table's default (16 bits for
// moment 0, then quantization_bits) for prediction-coded images and
// quantization_bits for every moment otherwise
// (src/moment_quantization.cpp:186-203, 285-360)
bool g_use_table = false; // s_UseQuantizationTable, s_QuantizationTable
vector<Byte> g_table(255, 0);

vector<Byte> quant_widths(const MomentImageHost &mi, int bits, const vector<Byte> &file_table)
{
    vector<Byte> w(std::max(mi.num_moments, 1), static_cast<Byte>(bits));
    if (mi.prediction_code && !file_table.empty())
        std::copy(file_table.begin(), file_table.begin() + std::min(file_table.size(), w.size()), w.begin());
    else if (mi.prediction_code)
        w[0] = 16;
    return w;
}

vector<Byte> quantize_on_device(const MomentImageHost &mi, const vector<Byte> &widths)
{
    using namespace mray_host;
    ensure_init();
    DeviceBuffer d_data(mi.data.data(), std::max<size_t>(mi.data.size(), 1) * sizeof(float));
    DeviceBuffer d_index(mi.index.data(), mi.index.size() * sizeof(uint32_t));
    const size_t cap = mi.data.size() * sizeof(float) + 8; // words of at most 32 bits
    DeviceBuffer d_out(cap);
    size_t n = 0;
    check(mray_hip_quantize(d_data.as<float>(), d_index.as<uint32_t>(), mi.width, mi.height, mi.prediction_code,
                            widths.data(), (int)widths.size(), d_out.as<uint8_t>(), cap, &n, nullptr),
          "mray_hip_quantize");
    vector<Byte> out(n);
    d_out.download(out.data(), n);
    return out;
}

void dequantize_on_device(MomentImageHost &mi, const vector<Byte> &widths, const vector<Byte> &bytes)
{
    using namespace mray_host;
    ensure_init();
    DeviceBuffer d_stream(bytes.data(), std::max<size_t>(bytes.size(), 1));
    DeviceBuffer d_index(mi.index.data(), mi.index.size() * sizeof(uint32_t));
    DeviceBuffer d_data(std::max<size_t>(mi.data.size(), 1) * sizeof(float));
    check(mray_hip_dequantize(d_stream.as<uint8_t>(), bytes.size(), d_index.as<uint32_t>(), mi.width, mi.height,
                              mi.prediction_code, widths.data(), (int)widths.size(), d_data.as<float>(), nullptr),
          "mray_hip_dequantize");
    d_data.download(mi.data.data(), mi.data.size() * sizeof(float));
}
} // namespace

void write_moment_image(const MomentImageHost &mi, const std::string filename, bool compress, Byte quantizationBits,
                        bool entropy_coding)
{
    if (compress || entropy_coding)
        throw invalid_args("moment-image compression and entropy coding are not supported here");
    if (quantizationBits < 1 || quantizationBits > 32)
        throw invalid_args("quantization_bits must be 1..32");
    if (quantizationBits < 32 && !mi.is_compact)
        throw invalid_args("quantised dense moment images (LZ4-compressed in the reference) are not supported here");
    if (mi.width > 65535 || mi.height > 65535 || mi.num_moments > 255)
        throw invalid_args("moment image too large for the file header");
    std::fstream out(filename, std::fstream::binary | std::fstream::out);
    if (!out)
        throw invalid_args("cannot write " + filename);
    put(out, static_cast<uint16_t>(mi.width));
    put(out, static_cast<uint16_t>(mi.height));
    put(out, static_cast<uint8_t>(mi.num_moments));
    put(out, static_cast<uint8_t>(mi.coding_warp));
    uint8_t flags = 0;
    flags |= mi.is_compact << 1;
    flags |= mi.prediction_code << 2;
    flags |= mi.has_bounds() << 4;
    flags |= mi.has_error_bounds() << 5;
    flags |= mi.has_coding_params() << 7;
    put(out, flags);
    put(out, static_cast<uint8_t>(quantizationBits | (g_use_table ? 0x80 : 0))); // no entropy coding
    put(out, mi.domain.min);
    put(out, mi.domain.max);
    if (g_use_table) // write_quantization_table (src/moment_image_io.cpp:340-344)
        put_bytes(out, g_table.data(), mi.num_moments);
    if (mi.is_compact)
    {
        vector<Byte> counts((size_t)mi.width * mi.height);
        for (size_t i = 0; i < counts.size(); ++i)
        {
            const uint32_t n = mi.index[i + 1] - mi.index[i];
            if (n > 255)
                throw invalid_data("more than 255 moments in a pixel");
            counts[i] = static_cast<Byte>(n);
        }
        put_bytes(out, counts.data(), counts.size());
    }
    if (mi.has_bounds())
        put_bytes(out, mi.bounds.data(), mi.bounds.size() * sizeof(DensityBound));
    if (mi.has_error_bounds())
        put_bytes(out, mi.error_bounds.data(), mi.error_bounds.size() * sizeof(ErrorBound));
    if (mi.has_coding_params())
        put_bytes(out, mi.coding_params.data(), mi.coding_params.size() * sizeof(CodingParamType));
    if (quantizationBits < 32)
    {
        // moment_quantization::quantize_prediction_coding / quantize (src/moment_image_io.cpp:397-408)
        const vector<Byte> table =
            g_use_table ? vector<Byte>(g_table.begin(), g_table.begin() + mi.num_moments) : vector<Byte>();
        const vector<Byte> stream = quantize_on_device(mi, quant_widths(mi, quantizationBits, table));
        put_bytes(out, stream.data(), stream.size());
    }
    else
        put_bytes(out, mi.data.data(), mi.data.size() * sizeof(float));
    if (!out)
        throw invalid_data("write error on " + filename);
}

MomentImageHost load_moment_image(const std::string filename)
{
    std::fstream in(filename, std::fstream::binary | std::fstream::in);
    if (!in)
        throw invalid_args("cannot open " + filename);
    MomentImageHost mi;
    uint16_t w, h;
    uint8_t num_moments, warp, flags, qbits;
    get(in, w);
    get(in, h);
    get(in, num_moments);
    get(in, warp);
    get(in, flags);
    get(in, qbits);
    get(in, mi.domain.min);
    get(in, mi.domain.max);
    if (!in)
        throw invalid_data("truncated moment-image header in " + filename);
    const int bits = qbits & 0x3F;
    if ((flags & 1) || (qbits & 0x40))
        throw invalid_args("compressed or entropy-coded moment images are not supported here");
    if (bits < 1 || bits > 32 || (bits < 32 && !(flags & 2)))
        throw invalid_args("quantised moment images are supported compact, with 1..32 bits");
    mi.width = w;
    mi.height = h;
    mi.num_moments = num_moments;
    mi.coding_warp = warp;
    mi.is_compact = flags & 2;
    mi.prediction_code = flags & 4;
    const size_t npix = (size_t)w * h;
    // the quantization table of a prediction-coded image follows the header
    // (read_quantization_table, src/moment_image_io.cpp:110-115)
    vector<Byte> file_table;
    if (qbits & 0x80)
    {
        file_table.resize(num_moments);
        get_bytes(in, file_table.data(), file_table.size(), filename);
    }
    if (mi.is_compact)
    {
        vector<Byte> counts(npix);
        get_bytes(in, counts.data(), npix, filename);
        mi.index.resize(npix + 1);
        uint32_t sum = 0;
        for (size_t i = 0; i < npix; ++i)
        {
            mi.index[i] = sum;
            sum += counts[i];
        }
        mi.index[npix] = sum;
    }
    if (flags & 16)
    {
        mi.bounds.resize(npix);
        get_bytes(in, mi.bounds.data(), npix * sizeof(DensityBound), filename);
    }
    if (flags & 32)
    {
        mi.error_bounds.resize(npix);
        get_bytes(in, mi.error_bounds.data(), npix * sizeof(ErrorBound), filename);
    }
    if (flags & 128)
    {
        mi.coding_params.resize(num_moments > 0 ? num_moments - 1 : 0);
        get_bytes(in, mi.coding_params.data(), mi.coding_params.size() * sizeof(CodingParamType), filename);
    }
    const size_t n = mi.is_compact ? mi.index[npix] : npix * num_moments;
    mi.data.resize(n);
    if (bits < 32)
    {
        // the rest of the file is the stream (read_moment_bytes, src/moment_image_io.cpp:91-107)
        const vector<Byte> stream((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        dequantize_on_device(mi, quant_widths(mi, bits, file_table), stream);
    }
    else
        get_bytes(in, mi.data.data(), n * sizeof(float), filename);
    return mi;
}

void write_PPM(string file, const ColorImageHost &img, Vec3f background_color)
{
    FILE *fp = fopen(file.c_str(), "wb");
    if (!fp)
        throw invalid_args("cannot write " + file);
    fprintf(fp, "P6\n%d %d\n255\n", img.width, img.height);
    for (int j = img.height - 1; j >= 0; --j)
        for (int i = 0; i < img.width; ++i)
        {
            const auto rgba = img.get_color(i, j);
            const Vec4f c(rgba.r / 255.f, rgba.g / 255.f, rgba.b / 255.f, rgba.a / 255.f);
            const float bx = c.x * c.w + background_color.x * (1.f - c.w);
            const float by = c.y * c.w + background_color.y * (1.f - c.w);
            const float bz = c.z * c.w + background_color.z * (1.f - c.w);
            const ColorImageHost::rgba8_t out(
                Vec4f(std::pow(bx, 1.f / 2.2f), std::pow(by, 1.f / 2.2f), std::pow(bz, 1.f / 2.2f), 1.0f));
            fwrite(&out, 1, 3, fp);
        }
    fclose(fp);
}

void set_prediction_coding_quantization_table(const vector<Byte> &table)
{
    g_use_table = true;
    for (size_t i = 0; i < table.size() && i < g_table.size(); ++i)
        g_table[i] = table[i];
}

bool uses_prediction_coding_quantization_table() { return g_use_table; }

vector<Byte> get_prediction_coding_quantization_table() { return g_table; }

void create_quantization_table(const MomentImageHost &img, int start_bits)
{
    using namespace mray_host;
    if (!img.prediction_code || !img.is_compact)
        throw invalid_args("the quantization curve needs a compact prediction-coded image");
    ensure_init();
    DeviceBuffer d_data(img.data.data(), std::max<size_t>(img.data.size(), 1) * sizeof(float));
    DeviceBuffer d_index(img.index.data(), img.index.size() * sizeof(uint32_t));
    DeviceBuffer d_params(img.coding_params.data(),
                          std::max<size_t>(img.coding_params.size(), 1) * sizeof(CodingParamType));
    vector<Byte> table(img.num_moments);
    check(mray_hip_quantization_table(d_data.as<float>(), d_index.as<uint32_t>(), img.width, img.height,
                                      img.num_moments, img.coding_warp,
                                      img.coding_params.empty() ? nullptr : d_params.as<uint16_t>(), start_bits,
                                      table.data(), nullptr),
          "mray_hip_quantization_table");
    set_prediction_coding_quantization_table(table);
}

MomentImageHost sample_moment_image(const MomentImageHost &img, int w, int h)
{
    if (!img.is_compact || w <= 0 || h <= 0)
        throw invalid_args("sample_moment_image needs a compact image and a positive size");
    const int fx = img.width / w, fy = img.height / h;
    // the reference's loops: rows y < w, columns x < h, pixel y * w + x
    if (h > w || (w - 1) * fy >= img.height || (h - 1) * fx >= img.width)
        throw invalid_args("sample_moment_image: the reference's loops leave the images at this size");
    const int m = img.num_moments;
    vector<uint32_t> counts((size_t)w * h, 0);
    vector<float> dense((size_t)w * h * m, 0.0f);
    for (int y = 0; y < w; ++y)
        for (int x = 0; x < h; ++x)
        {
            const size_t src = (size_t)(y * fy) * img.width + x * fx;
            const uint32_t n = img.index[src + 1] - img.index[src];
            const size_t dst = (size_t)y * w + x;
            counts[dst] = n;
            for (uint32_t l = 0; l < n; ++l)
                dense[dst * m + l] = img.data[img.index[src] + l];
        }
    MomentImageHost s;
    s.width = w;
    s.height = h;
    s.num_moments = m;
    s.is_compact = true;
    s.prediction_code = img.prediction_code;
    s.coding_warp = img.coding_warp;
    s.coding_params = img.coding_params;
    s.index.resize((size_t)w * h + 1);
    uint32_t sum = 0;
    for (size_t i = 0; i < counts.size(); ++i)
    {
        s.index[i] = sum;
        sum += counts[i];
    }
    s.index[counts.size()] = sum;
    s.data.resize(sum);
    for (size_t i = 0; i < counts.size(); ++i)
        for (uint32_t l = 0; l < counts[i]; ++l)
            s.data[s.index[i] + l] = dense[i * m + l];
    return s;
}
