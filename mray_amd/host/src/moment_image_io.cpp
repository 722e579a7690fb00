// Moment-image files, uncompressed and unquantised (src/moment_image_io.cpp).
#include <cmath>
#include <cstdio>
#include <fstream>

#include "../include/color_image.h"
#include "../include/moment_image_io.h"

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
} // namespace

void write_moment_image(const MomentImageHost &mi, const std::string filename, bool compress, Byte quantizationBits,
                        bool entropy_coding)
{
    if (compress || quantizationBits != 32 || entropy_coding)
        throw invalid_args("moment-image compression, quantisation and entropy coding are not supported here");
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
    put(out, static_cast<uint8_t>(32));
    put(out, mi.domain.min);
    put(out, mi.domain.max);
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
    if ((flags & 1) || (qbits & 0x3F) != 32 || (qbits & 0xC0))
        throw invalid_args("compressed, quantised or entropy-coded moment images are not supported here");
    mi.width = w;
    mi.height = h;
    mi.num_moments = num_moments;
    mi.coding_warp = warp;
    mi.is_compact = flags & 2;
    mi.prediction_code = flags & 4;
    const size_t npix = (size_t)w * h;
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
