// Small value types of the moment-image API, named and laid out as the
// reference's (src/common.h:130-150, src/cut/cuda_math.h): vectors, ranges,
// DensityBound, CodingParamType (src/coding_params.h:30-78).
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

using std::string;
using std::vector;

using Byte = unsigned char;

#define MAX_UINT16 65535.f

struct Vec2f
{
    float x = 0.f, y = 0.f;
    Vec2f() = default;
    Vec2f(float x, float y) : x(x), y(y) {}
};

struct Vec3f
{
    float x = 0.f, y = 0.f, z = 0.f;
    Vec3f() = default;
    Vec3f(float x, float y, float z) : x(x), y(y), z(z) {}
};

struct Vec3i
{
    int x = 0, y = 0, z = 0;
    Vec3i() = default;
    Vec3i(int x, int y, int z) : x(x), y(y), z(z) {}
};

struct Vec4f
{
    float x = 0.f, y = 0.f, z = 0.f, w = 0.f;
    Vec4f() = default;
    explicit Vec4f(float v) : x(v), y(v), z(v), w(v) {}
    Vec4f(float x, float y, float z, float w) : x(x), y(y), z(z), w(w) {}
};

// row-major 4x4 matrix, m[j][i] = row j, column i (cut::Mat4f)
struct Mat4f
{
    float m[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    float *operator[](int j) { return m[j]; }
    const float *operator[](int j) const { return m[j]; }
};

template <typename T> struct Range
{
    T min, max;
    Range() = default;
    Range(T min, T max) : min(min), max(max) {}
};

// errors of the reference (src/common.h:69-97)
struct invalid_args : std::invalid_argument
{
    using std::invalid_argument::invalid_argument;
};
struct invalid_data : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

// 16-bit truncated density bounds (src/common.h:135-150)
struct DensityBound
{
    uint16_t lower;
    uint16_t upper;

    DensityBound() = default;
    DensityBound(float l, float u) { set(l, u); }
    void set(float l, float u)
    {
        lower = static_cast<uint16_t>(l * MAX_UINT16);
        upper = static_cast<uint16_t>(u * MAX_UINT16);
    }
    Vec2f get() const { return Vec2f(lower / (MAX_UINT16), upper / (MAX_UINT16)); }
};
static_assert(sizeof(DensityBound) == 4, "DensityBound must be two uint16");

// coding warps (src/coding_params.h:32-39)
enum CodingWarp
{
    CODING_WARP_DEFAULT_TRANSFORMED = 0,
    CODING_WARP_NONE = 1,
};
using CodingWarpType = int;

inline bool requires_coding_parameters(CodingWarpType type) { return type == CODING_WARP_DEFAULT_TRANSFORMED; }

// quantized coding parameters (src/coding_params.h:47-78)
struct CodingParamType
{
    uint16_t param1;
    uint16_t param2;

    static float get_param1(CodingParamType c) { return (c.param1 / MAX_UINT16); }
    static float get_param2(CodingParamType c) { return (c.param2 / MAX_UINT16); }

    static vector<CodingParamType> quantize(const vector<float> &params)
    {
        auto clamp01 = [](float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); };
        vector<CodingParamType> res;
        for (size_t i = 0; i < params.size() / 2; ++i)
        {
            uint16_t p1 = static_cast<uint16_t>(std::ceil(clamp01(params[i * 2]) * MAX_UINT16));
            uint16_t p2 = static_cast<uint16_t>(std::ceil(clamp01(params[i * 2 + 1]) * MAX_UINT16));
            res.push_back(CodingParamType{p1, p2});
        }
        return res;
    }

    static vector<float> dequantize(const vector<CodingParamType> &coding_params)
    {
        vector<float> params;
        for (auto c : coding_params)
        {
            params.push_back(get_param1(c));
            params.push_back(get_param2(c));
        }
        return params;
    }
};
static_assert(sizeof(CodingParamType) == 4, "CodingParamType must be two uint16");
