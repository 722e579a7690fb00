// Volumes (src/volume.h:31-160): a regular grid of F32 or U8 voxels and SPH
// particles.
#pragma once

#include "common.h"

struct RegularGrid
{
    enum class DataType
    {
        F32,
        U8
    };

    Vec3i size;
    Range<Vec3f> domain;
    DataType type;
    vector<float> data; // U8 voxels are stored as bytes inside this vector, as in the reference

    RegularGrid(Vec3i size, Range<Vec3f> spatial_domain, DataType dtype)
        : size(size), domain(spatial_domain), type(dtype)
    {
    }
    RegularGrid(Vec3i size, Range<Vec3f> spatial_domain, const vector<float> &data, DataType dtype)
        : size(size), domain(spatial_domain), type(dtype), data(data)
    {
    }

    size_t num_voxels() const { return (size_t)size.x * size.y * size.z; }
    int get_byte_size() const { return type == DataType::U8 ? 1 : 4; }
    const void *voxels() const { return data.data(); }
};

struct Particles
{
    Range<Vec3f> domain;
    vector<float> x, y, z, values;
    float smoothing_length;

    Particles(Range<Vec3f> domain, const vector<float> &x, const vector<float> &y, const vector<float> &z,
              const vector<float> &values, float smoothing_length)
        : domain(domain), x(x), y(y), z(z), values(values), smoothing_length(smoothing_length)
    {
    }

    uint32_t size() const { return (uint32_t)x.size(); }
    float get_step_size() const { return smoothing_length * 0.5f; }
};
