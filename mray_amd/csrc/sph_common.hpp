// SPH particle sampling shared by the grid splat (sph.hip) and the march over
// particles (generate.hip): the reference's Particles::sample
// (src/volume.cpp:83-94) through its uniform grid of cell size h
// (src/particle_grid.cuh:46-403, src/particle_grid.cpp:38-55) and
// InterpolateParticlesOp (src/particle_interpolate.h:35-58).
//
// Particles are sorted by their 4^3-tiled cell index (stably, so the order
// inside a cell is the input order); a sample at p is
//     sum_q W(|x_q - p| / h) v_q / sum_q W(|x_q - p| / h)
// over the particles of the 27 cells around p with |x_q - p|^2 < h^2, cells
// in z, y, x order and particles in sorted order, W the cubic spline
// (src/cut/kernels.h:55-64); 0 where no particle is that close.
#pragma once

#include "device_common.hpp"

namespace mray
{

constexpr uint32_t kSphEmpty = 0xFFFFFFFFu;

struct CellGrid
{
    uint32_t sx, sy, sz; // padded cell counts
    float minp[3];
    float inv_cell;
    float r2;
};

// getIndexFromCell with 4^3 tiles (src/particle_grid.cuh:64-83)
__host__ __device__ __forceinline__ uint32_t cell_index(const CellGrid &g, int cx, int cy, int cz)
{
    const int tsx = 4, tsy = 4, tsz = (g.sz == 1) ? 1 : 4;
    const int width_tiles = ((int)g.sx + tsx - 1) / tsx;
    const int height_tiles = ((int)g.sy + tsy - 1) / tsy;
    const int tile_x = cx / tsx, tile_y = cy / tsy, tile_z = cz / tsz;
    const int in_x = cx % tsx, in_y = cy % tsy, in_z = cz % tsz;
    return (uint32_t)((tile_z * height_tiles * width_tiles + tile_y * width_tiles + tile_x) * (tsx * tsy * tsz) +
                      in_z * tsx * tsy + in_y * tsx + in_x);
}

__device__ __forceinline__ int cell_coord(const CellGrid &g, float p, int k)
{
    return (int)((p - g.minp[k]) * g.inv_cell);
}

// cubicSpline (src/cut/kernels.h:55-64)
__device__ __forceinline__ float cubic_spline(float x)
{
    const float sigma = 0.5641895835477563f;
    if (x < 1.f)
        return sigma * (1.f - (3.f / 2.f) * x * x * (1.f - x * 0.5f));
    else if (x < 2.f)
        return sigma / 4.f * (2.f - x) * (2.f - x) * (2.f - x);
    return 0.0f;
}

// The cell-sorted particles: records (x, y, z, value) and per-cell [start,
// end) ranges (start kSphEmpty for an empty cell).
struct SphView
{
    const float4 *parts;
    const uint32_t *start, *end;
    CellGrid g;
    float inv_h;
};

// Particles::sample at p. The loads are issued ahead of their use, in the
// same summation order: a row's (up to) three cells' ranges together, and a
// cell's particles four records at a time (one dependent load per particle
// left the gather waiting on memory).
__device__ __forceinline__ float sph_sample(const SphView &s, float px, float py, float pz)
{
    const CellGrid &g = s.g;
    const int c[3] = {cell_coord(g, px, 0), cell_coord(g, py, 1), cell_coord(g, pz, 2)};
    float value = 0.0f, wsum = 0.0f;
    const int z0 = max(c[2] - 1, 0), z1 = min(c[2] + 2, (int)g.sz);
    const int y0 = max(c[1] - 1, 0), y1 = min(c[1] + 2, (int)g.sy);
    const int x0 = max(c[0] - 1, 0), x1 = min(c[0] + 2, (int)g.sx);
    auto add = [&](const float4 &pt) {
        const float dx = px - pt.x, dy = py - pt.y, dz = pz - pt.z;
        const float d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < g.r2)
        {
            const float ex = pt.x - px, ey = pt.y - py, ez = pt.z - pz;
            const float w = cubic_spline(sqrtf(ex * ex + ey * ey + ez * ez) * s.inv_h);
            wsum += w;
            value += w * pt.w;
        }
    };
    for (int cz = z0; cz < z1; ++cz)
        for (int cy = y0; cy < y1; ++cy)
        {
            // the row's cells x0 .. x1 - 1 (at most three)
            uint32_t st[3], en[3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
            {
                const bool in = x0 + i < x1;
                const uint32_t ci = cell_index(g, in ? x0 + i : x0, cy, cz);
                st[i] = in ? s.start[ci] : kSphEmpty;
                en[i] = in ? s.end[ci] : 0u;
            }
#pragma unroll
            for (int i = 0; i < 3; ++i)
            {
                if (st[i] == kSphEmpty)
                    continue;
                const uint32_t e = en[i];
                uint32_t q = st[i];
                for (; q + 4 <= e; q += 4)
                {
                    const float4 p0 = s.parts[q], p1 = s.parts[q + 1], p2 = s.parts[q + 2], p3 = s.parts[q + 3];
                    add(p0);
                    add(p1);
                    add(p2);
                    add(p3);
                }
                for (; q < e; ++q)
                    add(s.parts[q]);
            }
        }
    if (wsum != 0.0f)
        value /= wsum;
    return value;
}

// The device buffers of a cell-sorted particle set (sph.hip builds them).
struct SphBuffers
{
    float4 *parts = nullptr;
    uint32_t *start = nullptr, *end = nullptr;
    CellGrid g{};
    float inv_h = 0.0f;
    uint32_t n = 0;
    size_t ncells = 0;
    SphView view() const { return SphView{parts, start, end, g, inv_h}; }
};

// Sorts n particles into the cell grid of cell size h (synchronises the
// stream once: the grid is sized from the particles' bounding box). With
// persistent the buffers come from hipMalloc (the caller frees them with
// sph_free), else from the stream-ordered pool.
int sph_build(const float *x, const float *y, const float *z, const float *v, uint32_t n, float h, bool persistent,
              SphBuffers &out, hipStream_t stream);
void sph_free(SphBuffers &b, bool persistent, hipStream_t stream);

// Tag type of a particle set marched directly: VolumeView<SphParticle> is
// what the generate kernels take for it.
struct SphParticle
{
};

template <> struct VolumeView<SphParticle>
{
    SphView s;
    int nx, ny, nz; // 1: no voxels
    Vec3 dmin, dmax;
    Vec3 rext;
};

__device__ __forceinline__ float sample_volume(const VolumeView<SphParticle> &v, float px, float py, float pz)
{
    return sph_sample(v.s, px, py, pz);
}

// the gather has no loads to hide: issue() samples, value() returns it
template <bool OFF32, bool POW2> struct PendingSample<SphParticle, OFF32, POW2>
{
    float val;
    __device__ __forceinline__ void issue(const VolumeView<SphParticle> &v, float px, float py, float pz)
    {
        val = sph_sample(v.s, px, py, pz);
    }
    __device__ __forceinline__ float value() const { return val; }
};

} // namespace mray
