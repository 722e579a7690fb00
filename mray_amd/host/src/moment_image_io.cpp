// Moment-image files (src/moment_image_io.cpp), every variant the reference
// writes and reads: the header, the optional prediction-coding quantization
// table, then for compact images the per-pixel counts, bounds, error bounds,
// coding parameters and moments, for dense images the same without the
// counts. Each section may be LZ4-compressed (liblz4, a u32 length prefix,
// stored raw when compression does not shrink it). Compact moments are
// floats, the packed quantized stream (mray_hip_quantize) or the
// entropy-coded section (mray_hip_entropy_encode); dense moments are floats
// or, with compression, LZ4 of byte-aligned quantized values
// (mray_hip_quantize_dense). Quantization and entropy coding run on the GPU;
// LZ4 runs on the host.
#include <lz4.h>

#include <climits>
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
void put_bytes(std::fstream &out, const void *p, size_t n) { out.write(static_cast<const char *>(p), (std::streamsize)n); }
void get_bytes(std::fstream &in, void *p, size_t n, const std::string &file)
{
    in.read(static_cast<char *>(p), (std::streamsize)n);
    if ((size_t)in.gcount() != n)
        throw invalid_data("truncated moment image " + file);
}

vector<Byte> rest_of(std::fstream &in)
{
    return vector<Byte>((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
}

// LZ4_compress_default into a buffer of dst_cap bytes; 0 when it fails
int lz4_compress(const void *src, size_t n, vector<Byte> &dst, size_t dst_cap)
{
    if (n > (size_t)LZ4_MAX_INPUT_SIZE || dst_cap > (size_t)INT_MAX)
        return 0;
    dst.resize(std::max<size_t>(dst_cap, 1));
    return LZ4_compress_default(static_cast<const char *>(src), reinterpret_cast<char *>(dst.data()), (int)n,
                                (int)dst_cap);
}

// write_bytes (src/moment_image_io.cpp:315-338): with compression a u32
// length, then the LZ4 block, or the raw bytes when LZ4 does not shrink them
void write_section(std::fstream &out, const void *p, size_t n, bool compress)
{
    if (!compress)
    {
        put_bytes(out, p, n);
        return;
    }
    if (n > UINT32_MAX)
        throw invalid_args("a compressed moment-image section is limited to 4 GiB");
    vector<Byte> c;
    const int r = lz4_compress(p, n, c, n);
    if (r <= 0 || (size_t)r >= n)
    {
        put(out, static_cast<uint32_t>(n));
        put_bytes(out, p, n);
    }
    else
    {
        put(out, static_cast<uint32_t>(r));
        put_bytes(out, c.data(), (size_t)r);
    }
}

// read_bytes (src/moment_image_io.cpp:32-61): a section of known length n
void read_section(std::fstream &in, void *p, size_t n, bool compressed, const std::string &file)
{
    if (!compressed)
    {
        get_bytes(in, p, n, file);
        return;
    }
    uint32_t cs;
    get(in, cs);
    if (!in || cs > n)
        throw invalid_data("bad compressed section length in " + file);
    if (cs == n)
    {
        get_bytes(in, p, n, file);
        return;
    }
    vector<Byte> c(cs);
    get_bytes(in, c.data(), cs, file);
    if (n > (size_t)INT_MAX || LZ4_decompress_safe(reinterpret_cast<const char *>(c.data()), static_cast<char *>(p),
                                                   (int)cs, (int)n) < 0)
        throw invalid_data("corrupt LZ4 section in " + file);
}

// read_moment_bytes (src/moment_image_io.cpp:63-107): the compact moment
// section. Compressed, its length is at most 4 bytes per moment; a block
// that does not decompress is the raw section written when LZ4 did not
// shrink it. Uncompressed, it is the rest of the file.
vector<Byte> read_moment_section(std::fstream &in, size_t n_moments, bool compressed, const std::string &file)
{
    if (!compressed)
        return rest_of(in);
    uint32_t cs;
    get(in, cs);
    if (!in)
        throw invalid_data("truncated moment image " + file);
    const size_t bound = 4 * n_moments;
    vector<Byte> bytes(bound, 0);
    if (cs == bound || cs == 0)
    {
        in.read(reinterpret_cast<char *>(bytes.data()), (std::streamsize)bound);
        return bytes;
    }
    vector<Byte> c(cs);
    get_bytes(in, c.data(), cs, file);
    const int r = bound > (size_t)INT_MAX ? -1
                                          : LZ4_decompress_safe(reinterpret_cast<const char *>(c.data()),
                                                                reinterpret_cast<char *>(bytes.data()), (int)cs,
                                                                (int)bound);
    if (r < 0)
        return c;
    bytes.resize((size_t)r);
    return bytes;
}

// moment_quantization's table (s_UseQuantizationTable, s_QuantizationTable):
// set by create_quantization_table, set_prediction_coding_quantization_table
// and by loading a file that carries one; once set, every file written
// carries it and prediction-coded images are quantized with it
bool g_use_table = false;
vector<Byte> g_table(255, 0);

// get_prediction_coding_quantization_table(max_moments, numBits)
// (src/moment_quantization.cpp:186-203): the table when set, else 16 bits
// for moment 0 and numBits for the others
vector<Byte> prediction_widths(int m, int bits)
{
    vector<Byte> w(std::max(m, 1), static_cast<Byte>(bits));
    if (g_use_table)
        std::copy(g_table.begin(), g_table.begin() + w.size(), w.begin());
    else
        w[0] = 16;
    return w;
}

// the widths of the entropy coder: the table for prediction-coded images
// (get_prediction_coding_quantization_table(), which needs it set), else
// quantization_bits for every moment (src/moment_image_io.cpp:383-396)
vector<Byte> entropy_widths(const MomentImageHost &mi, int bits)
{
    if (mi.prediction_code)
    {
        if (!g_use_table)
            throw invalid_args("entropy coding a prediction-coded image needs the quantization table");
        return vector<Byte>(g_table.begin(), g_table.begin() + std::max(mi.num_moments, 1));
    }
    return vector<Byte>(std::max(mi.num_moments, 1), static_cast<Byte>(bits));
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

vector<Byte> entropy_encode_on_device(const MomentImageHost &mi, const vector<Byte> &widths)
{
    using namespace mray_host;
    ensure_init();
    DeviceBuffer d_data(mi.data.data(), std::max<size_t>(mi.data.size(), 1) * sizeof(float));
    DeviceBuffer d_index(mi.index.data(), mi.index.size() * sizeof(uint32_t));
    const size_t npix = (size_t)mi.width * mi.height;
    // at most 4 bytes per symbol, 2 per moment 0, and the headers
    const size_t cap = 4 + 2 * npix + 20 * (size_t)mi.num_moments + 4 * mi.data.size() + 64;
    vector<Byte> out(cap);
    size_t n = 0;
    check(mray_hip_entropy_encode(d_data.as<float>(), d_index.as<uint32_t>(), mi.width, mi.height, mi.num_moments,
                                  mi.prediction_code, widths.data(), out.data(), cap, &n, nullptr),
          "mray_hip_entropy_encode");
    out.resize(n);
    return out;
}

void entropy_decode_on_device(MomentImageHost &mi, const vector<Byte> &widths, const vector<Byte> &bytes)
{
    using namespace mray_host;
    ensure_init();
    DeviceBuffer d_index(mi.index.data(), mi.index.size() * sizeof(uint32_t));
    DeviceBuffer d_data(std::max<size_t>(mi.data.size(), 1) * sizeof(float));
    check(mray_hip_entropy_decode(bytes.data(), bytes.size(), d_index.as<uint32_t>(), mi.width, mi.height,
                                  mi.num_moments, mi.prediction_code, widths.data(), d_data.as<float>(), nullptr),
          "mray_hip_entropy_decode");
    d_data.download(mi.data.data(), mi.data.size() * sizeof(float));
}

vector<Byte> quantize_dense_on_device(const vector<float> &data, int m, int bits)
{
    using namespace mray_host;
    ensure_init();
    const size_t nb = (size_t)(bits + 7) / 8;
    DeviceBuffer d_data(data.data(), std::max<size_t>(data.size(), 1) * sizeof(float));
    DeviceBuffer d_out(std::max<size_t>(data.size() * nb, 1));
    check(mray_hip_quantize_dense(d_data.as<float>(), data.size(), m, bits, d_out.as<uint8_t>(), nullptr),
          "mray_hip_quantize_dense");
    check_hip(hipDeviceSynchronize(), "mray_hip_quantize_dense");
    vector<Byte> out(data.size() * nb);
    d_out.download(out.data(), out.size());
    return out;
}

void dequantize_dense_on_device(const vector<Byte> &qs, int m, int bits, vector<float> &data)
{
    using namespace mray_host;
    ensure_init();
    DeviceBuffer d_in(qs.data(), std::max<size_t>(qs.size(), 1));
    DeviceBuffer d_data(std::max<size_t>(data.size(), 1) * sizeof(float));
    check(mray_hip_dequantize_dense(d_in.as<uint8_t>(), data.size(), m, bits, d_data.as<float>(), nullptr),
          "mray_hip_dequantize_dense");
    check_hip(hipDeviceSynchronize(), "mray_hip_dequantize_dense");
    d_data.download(data.data(), data.size() * sizeof(float));
}

// the moment section of a compact image (write_moment_image,
// src/moment_image_io.cpp:376-409)
vector<Byte> compact_moment_bytes(const MomentImageHost &mi, int bits, bool entropy_coding)
{
    if (entropy_coding)
    {
        const vector<Byte> w = entropy_widths(mi, bits);
        for (size_t l = 1; l < w.size(); ++l)
            if (w[l] < 1 || w[l] > 14)
                throw invalid_args("entropy coding needs 1..14 bits per moment above moment 0");
        return entropy_encode_on_device(mi, w);
    }
    if (bits < 32)
        return quantize_on_device(mi, mi.prediction_code ? prediction_widths(mi.num_moments, bits)
                                                         : vector<Byte>(std::max(mi.num_moments, 1), (Byte)bits));
    const Byte *p = reinterpret_cast<const Byte *>(mi.data.data());
    return vector<Byte>(p, p + mi.data.size() * sizeof(float));
}

// the moments of a dense image (write_static_image,
// src/moment_image_io.cpp:411-461): with compression and quantization a byte
// holding quantization_bits, then LZ4 of the byte-aligned quantized values
// (raw when LZ4 fails); with compression alone LZ4 of the floats (when LZ4
// fails the reference writes data.size() bytes, a quarter of the floats, and
// so does this); else the floats
void write_dense_moments(std::fstream &out, const MomentImageHost &mi, bool compress, int bits)
{
    if (compress && bits < 32)
    {
        put(out, static_cast<uint8_t>(bits));
        const vector<Byte> q = quantize_dense_on_device(mi.data, mi.num_moments, bits);
        vector<Byte> c;
        const int r = lz4_compress(q.data(), q.size(), c, q.size() * 4);
        if (r <= 0 || (size_t)r >= q.size() * 4)
            put_bytes(out, q.data(), q.size());
        else
            put_bytes(out, c.data(), (size_t)r);
    }
    else if (compress)
    {
        vector<Byte> c;
        const int r = lz4_compress(mi.data.data(), mi.data.size() * sizeof(float), c, mi.data.size() * 4);
        if (r <= 0 || (size_t)r >= mi.data.size() * 4)
            put_bytes(out, mi.data.data(), mi.data.size());
        else
            put_bytes(out, c.data(), (size_t)r);
    }
    else
        put_bytes(out, mi.data.data(), mi.data.size() * sizeof(float));
}

// read_static_image (src/moment_image_io.cpp:199-254). The compressed
// quantized variant starts with the byte holding quantization_bits that
// write_static_image puts there; it is skipped (the reference's reader hands
// it to LZ4 with the block and cannot read these files).
void read_dense_moments(std::fstream &in, MomentImageHost &mi, bool compressed, int bits, const std::string &file)
{
    const size_t n = (size_t)mi.width * mi.height * mi.num_moments;
    mi.data.resize(n);
    if (!compressed)
    {
        get_bytes(in, mi.data.data(), n * sizeof(float), file);
        return;
    }
    if (bits < 32)
    {
        uint8_t stored;
        get(in, stored);
        if (!in || stored != bits)
            throw invalid_data("bad quantization byte in " + file);
    }
    const vector<Byte> c = rest_of(in);
    const size_t nb = bits < 32 ? (size_t)(bits + 7) / 8 : sizeof(float);
    vector<Byte> raw(n * nb);
    const int r = (c.size() > (size_t)INT_MAX || raw.size() > (size_t)INT_MAX)
                      ? -1
                      : LZ4_decompress_safe(reinterpret_cast<const char *>(c.data()),
                                            reinterpret_cast<char *>(raw.data()), (int)c.size(), (int)raw.size());
    if (r != (int)raw.size())
    {
        // raw values when LZ4 did not shrink them
        if (bits < 32 && c.size() == raw.size())
            raw = c;
        else
            throw invalid_data("corrupt LZ4 moment block in " + file);
    }
    if (bits < 32)
        dequantize_dense_on_device(raw, mi.num_moments, bits, mi.data);
    else
        std::memcpy(mi.data.data(), raw.data(), raw.size());
}
} // namespace

void write_moment_image(const MomentImageHost &mi, const std::string filename, bool compress, Byte quantizationBits,
                        bool entropy_coding)
{
    if (quantizationBits < 1 || quantizationBits > 32)
        throw invalid_args("quantization_bits must be 1..32");
    if (mi.width > 65535 || mi.height > 65535 || mi.num_moments > 255 || mi.num_moments < 1)
        throw invalid_args("moment image too large for the file header");
    // the moment section is built first: nothing is written when it fails
    vector<Byte> moments;
    if (mi.is_compact)
        moments = compact_moment_bytes(mi, quantizationBits, entropy_coding);
    std::fstream out(filename, std::fstream::binary | std::fstream::out);
    if (!out)
        throw invalid_args("cannot write " + filename);
    // write_header (src/moment_image_io.cpp:481-513)
    put(out, static_cast<uint16_t>(mi.width));
    put(out, static_cast<uint16_t>(mi.height));
    put(out, static_cast<uint8_t>(mi.num_moments));
    put(out, static_cast<uint8_t>(mi.coding_warp));
    uint8_t flags = 0;
    flags |= compress;
    flags |= mi.is_compact << 1;
    flags |= mi.prediction_code << 2;
    flags |= mi.has_bounds() << 4;
    flags |= mi.has_error_bounds() << 5;
    flags |= mi.has_coding_params() << 7;
    put(out, flags);
    put(out, static_cast<uint8_t>(quantizationBits | (entropy_coding ? 0x40 : 0) | (g_use_table ? 0x80 : 0)));
    put(out, mi.domain.min);
    put(out, mi.domain.max);
    if (g_use_table) // write_quantization_table (src/moment_image_io.cpp:340-344)
        write_section(out, g_table.data(), (size_t)mi.num_moments, compress);
    if (mi.is_compact)
    {
        // write_index: per-pixel counts as bytes (src/moment_image_io.cpp:346-360)
        vector<Byte> counts((size_t)mi.width * mi.height);
        for (size_t i = 0; i < counts.size(); ++i)
        {
            const uint32_t n = mi.index[i + 1] - mi.index[i];
            if (n > 255)
                throw invalid_data("more than 255 moments in a pixel");
            counts[i] = static_cast<Byte>(n);
        }
        write_section(out, counts.data(), counts.size(), compress);
    }
    if (mi.has_bounds())
        write_section(out, mi.bounds.data(), mi.bounds.size() * sizeof(DensityBound), compress);
    if (mi.has_error_bounds())
        write_section(out, mi.error_bounds.data(), mi.error_bounds.size() * sizeof(ErrorBound), compress);
    if (mi.has_coding_params())
        write_section(out, mi.coding_params.data(), mi.coding_params.size() * sizeof(CodingParamType), compress);
    if (mi.is_compact)
        write_section(out, moments.data(), moments.size(), compress);
    else
        write_dense_moments(out, mi, compress, quantizationBits);
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
    const bool entropy = qbits & 0x40, compressed = flags & 1;
    if (bits < 1 || bits > 32)
        throw invalid_data("quantization_bits out of range in " + filename);
    mi.width = w;
    mi.height = h;
    mi.num_moments = num_moments;
    mi.coding_warp = static_cast<CodingWarpType>(warp);
    mi.is_compact = flags & 2;
    mi.prediction_code = flags & 4;
    const size_t npix = (size_t)w * h;
    // the quantization table is installed as the reference's loader does
    // (read_quantization_table, src/moment_image_io.cpp:110-115)
    if (qbits & 0x80)
    {
        vector<Byte> table(num_moments);
        read_section(in, table.data(), table.size(), compressed, filename);
        set_prediction_coding_quantization_table(table);
    }
    if (mi.is_compact)
    {
        // read_index (src/moment_image_io.cpp:117-133)
        vector<Byte> counts(npix);
        read_section(in, counts.data(), npix, compressed, filename);
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
        read_section(in, mi.bounds.data(), npix * sizeof(DensityBound), compressed, filename);
    }
    if (flags & 32)
    {
        mi.error_bounds.resize(npix);
        read_section(in, mi.error_bounds.data(), npix * sizeof(ErrorBound), compressed, filename);
    }
    if (flags & 128)
    {
        mi.coding_params.resize(num_moments > 0 ? num_moments - 1 : 0);
        read_section(in, mi.coding_params.data(), mi.coding_params.size() * sizeof(CodingParamType), compressed,
                     filename);
    }
    if (!mi.is_compact)
    {
        read_dense_moments(in, mi, compressed, bits, filename);
        return mi;
    }
    // read_moment_image (src/moment_image_io.cpp:153-181)
    const size_t n = mi.index[npix];
    mi.data.resize(n);
    const vector<Byte> bytes = read_moment_section(in, n, compressed, filename);
    if (entropy)
        entropy_decode_on_device(mi, entropy_widths(mi, bits), bytes);
    else if (bits < 32)
        dequantize_on_device(mi, mi.prediction_code ? prediction_widths(num_moments, bits)
                                                    : vector<Byte>(std::max<int>(num_moments, 1), (Byte)bits),
                             bytes);
    else
    {
        if (bytes.size() < n * sizeof(float))
            throw invalid_data("truncated moment data in " + filename);
        std::memcpy(mi.data.data(), bytes.data(), n * sizeof(float));
    }
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
