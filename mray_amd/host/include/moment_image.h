// Moment image (MomentImageHost, src/moment_image.h:33-147): same fields,
// layouts and accessors as the reference. compact(), prediction_encode() and
// prepare_moments() run on the GPU through libmray_hip.so.
#pragma once

#include "common.h"

using ErrorBound = uint16_t;

struct MomentImageHost
{
    int width, height;
    int num_moments; // maximum number of moments per pixel
    bool is_compact;
    bool prediction_code;
    CodingWarpType coding_warp;
    Range<Vec3f> domain; // AABB of the data
    vector<uint32_t> index;
    vector<DensityBound> bounds;
    vector<ErrorBound> error_bounds;
    vector<CodingParamType> coding_params;
    vector<float> data;

    MomentImageHost() = default;
    MomentImageHost(int w, int h, int num_moments, bool create_index, bool create_bounds = false)
        : width(w), height(h), num_moments(num_moments), is_compact(false), prediction_code(false),
          coding_warp(CODING_WARP_DEFAULT_TRANSFORMED)
    {
        if (create_index)
            index.resize((size_t)width * height + 1);
        if (create_bounds)
            bounds.resize((size_t)width * height);
        data.resize((size_t)get_elements_per_pixel() * width * height);
    }

    uint32_t get_elements_per_pixel() const { return num_moments; }
    uint32_t get_idx_compact(int x, int y) const { return index[(size_t)y * width + x]; }
    size_t get_idx(int x, int y) const
    {
        return is_compact ? get_idx_compact(x, y) : ((size_t)y * width + x) * get_elements_per_pixel();
    }
    int get_num_moments(int x, int y) const
    {
        return is_compact ? (int)(index[(size_t)y * width + x + 1] - index[(size_t)y * width + x]) : num_moments;
    }
    bool has_bounds() const { return !bounds.empty(); }
    void set_bounds(int x, int y, DensityBound bound) { bounds[(size_t)y * width + x] = bound; }
    Vec2f get_bounds(int x, int y) const
    {
        return has_bounds() ? bounds[(size_t)y * width + x].get() : Vec2f(0.f, 1.f);
    }
    void add_error_bounds() { error_bounds.resize((size_t)width * height); }
    bool has_error_bounds() const { return !error_bounds.empty(); }
    bool has_coding_params() const { return !coding_params.empty(); }
    void set_coding_params(const vector<float> &params) { coding_params = CodingParamType::quantize(params); }
    vector<float> get_coding_params() const { return CodingParamType::dequantize(coding_params); }

    // Compacts a dense image whose index[] holds the number of moments per
    // pixel (MomentImageHost::compact, src/moment_image.cpp:69-94).
    void compact();

    // Prediction coding of a compact image (src/moment_image.cpp:96-152).
    void prediction_encode(int warp, float bias);

    // Lagrange multipliers of one pixel (src/moment_image.cpp:154-183);
    // form 1 = mixed (FP64 recursion, the default), 0 = plain (the
    // reference's FP32 arithmetic).
    vector<float> prepare_moments(int x, int y, float bias, int form = 1) const;

    // Lagrange multipliers of all pixels, in the layout of data[].
    vector<float> prepare_all_moments(float bias, int form = 1) const;
};
