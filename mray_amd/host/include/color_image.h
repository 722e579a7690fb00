// RGBA8 image (src/color_image.h:30-70): colours clamped to [0, 1] and
// truncated to bytes.
#pragma once

#include "common.h"

struct ColorImageHost
{
    struct rgba8_t
    {
        unsigned char r, g, b, a;
        rgba8_t() = default;
        explicit rgba8_t(Vec4f c)
        {
            r = static_cast<unsigned char>(c.x * 255.f);
            g = static_cast<unsigned char>(c.y * 255.f);
            b = static_cast<unsigned char>(c.z * 255.f);
            a = static_cast<unsigned char>(c.w * 255.f);
        }
    };

    int width, height;
    vector<rgba8_t> rgba;

    ColorImageHost(int w, int h) : width(w), height(h), rgba((size_t)w * h) {}

    void set_color(int x, int y, Vec4f c)
    {
        auto cl = [](float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); };
        rgba[(size_t)y * width + x] = rgba8_t(Vec4f(cl(c.x), cl(c.y), cl(c.z), cl(c.w)));
    }
    rgba8_t get_color(int x, int y) const { return rgba[(size_t)y * width + x]; }
};

// Binary PPM of the image blended over a background and gamma-encoded with
// 1 / 2.2, rows written bottom-up (write_PPM, src/color_image.cpp:29-57).
void write_PPM(string file, const ColorImageHost &img, Vec3f background_color);
