// SPH particles -> regular grid (the C4 input of the moment-image path).
//
// The reference samples particles through a uniform grid of cell size h
// (Particles::create_grid / Particles::sample, src/volume.cpp:55-94;
// UniformGrid, src/particle_grid.cpp:38-55, src/particle_grid.cuh:46-403;
// InterpolateParticlesOp, src/particle_interpolate.h:35-58): particles are
// sorted by their 4^3-tiled cell index, and a sample at p is
//     sum_q W(|x_q - p| / h) v_q / sum_q W(|x_q - p| / h)
// over the particles of the 27 cells around p with |x_q - p|^2 < h^2, W the
// cubic spline (src/cut/kernels.h:55-64). Here that sample is evaluated at
// every voxel position of an nx*ny*nz grid (voxel i at dmin + i/n (dmax - dmin),
// the RegularGrid::sample convention), producing the volume that generate
// marches.
//
// Passes: bounding box (float min/max as ordered-integer atomics, exact),
// cell keys, a stable radix sort of (key, particle) (hipCUB; the reference
// uses thrust::sort_by_key), particles gathered into sorted float4 records,
// per-cell [start, end) tables, then one thread per voxel in 8x8x4 voxel
// blocks (voxels of a block share their 27-cell neighbourhood). The voxel's
// sum runs over cells in z, y, x order and particles in sorted order: with a
// stable sort that is the oracle's order, so the grid is bit-identical to it
// (the reference's std::sort leaves the order inside a cell unspecified).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstring>

#include "device_common.hpp"
#include "launch.hpp"

namespace mray
{

namespace
{

constexpr uint32_t kEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ int float_key(float f)
{
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__host__ __device__ inline float key_float(int k)
{
    const int i = k >= 0 ? k : k ^ 0x7fffffff;
    float f;
    std::memcpy(&f, &i, 4);
    return f;
}

struct CellGrid
{
    uint32_t sx, sy, sz; // padded cell counts
    float minp[3];
    float inv_cell;
    float r2;
};

// getIndexFromCell with 4^3 tiles (src/particle_grid.cuh:64-83)
__device__ __forceinline__ uint32_t cell_index(const CellGrid &g, int cx, int cy, int cz)
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

__global__ void bbox_init_kernel(int *__restrict__ keys)
{
    const int i = threadIdx.x;
    if (i < 6)
        keys[i] = i < 3 ? 0x7fffffff : (int)0x80000000; // min slots: +max key, max slots: -max key
}

// min / max of the positions; NaN coordinates are dropped, as by the
// reference's (q < mn) ? q : mn
__global__ __launch_bounds__(256) void bbox_kernel(const float *__restrict__ x, const float *__restrict__ y,
                                                   const float *__restrict__ z, uint32_t n, int *__restrict__ keys)
{
    int lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
    int hi[3] = {(int)0x80000000, (int)0x80000000, (int)0x80000000};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const float q[3] = {x[i], y[i], z[i]};
        for (int k = 0; k < 3; ++k)
            if (q[k] == q[k])
            {
                const int kk = float_key(q[k] + 0.0f);
                lo[k] = min(lo[k], kk);
                hi[k] = max(hi[k], kk);
            }
    }
    for (int k = 0; k < 3; ++k)
    {
        for (int o = 32; o > 0; o >>= 1)
        {
            lo[k] = min(lo[k], __shfl_xor(lo[k], o));
            hi[k] = max(hi[k], __shfl_xor(hi[k], o));
        }
    }
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; ++k)
        {
            atomicMin(&keys[k], lo[k]);
            atomicMax(&keys[3 + k], hi[k]);
        }
}

__global__ __launch_bounds__(256) void cell_key_kernel(const float *__restrict__ x, const float *__restrict__ y,
                                                       const float *__restrict__ z, uint32_t n, CellGrid g,
                                                       uint32_t *__restrict__ keys, uint32_t *__restrict__ ids)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    // The cell of the largest coordinate can round to one past the grid when
    // the padded size is a multiple of 4 ((mx - mn) / h just below an integer
    // while inv_cell rounds up); the reference would index past its cell
    // array there. Clamping keeps the key inside the tables.
    const int cx = min(max(cell_coord(g, x[i], 0), 0), (int)g.sx - 1);
    const int cy = min(max(cell_coord(g, y[i], 1), 0), (int)g.sy - 1);
    const int cz = min(max(cell_coord(g, z[i], 2), 0), (int)g.sz - 1);
    keys[i] = cell_index(g, cx, cy, cz);
    ids[i] = i;
}

__global__ __launch_bounds__(256) void gather_particles_kernel(const float *__restrict__ x, const float *__restrict__ y,
                                                               const float *__restrict__ z, const float *__restrict__ v,
                                                               const uint32_t *__restrict__ ids, uint32_t n,
                                                               float4 *__restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t q = ids[i];
    out[i] = make_float4(x[q], y[q], z[q], v[q]);
}

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t *__restrict__ p, size_t n, uint32_t v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// [start, end) of every occupied cell from the sorted keys
__global__ __launch_bounds__(256) void cell_range_kernel(const uint32_t *__restrict__ keys, uint32_t n,
                                                         uint32_t *__restrict__ start, uint32_t *__restrict__ end)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t k = keys[i];
    if (i == 0 || keys[i - 1] != k)
        start[k] = i;
    if (i == n - 1 || keys[i + 1] != k)
        end[k] = i + 1;
}

// Particles::sample at every voxel position
__global__ __launch_bounds__(256) void sph_gather_kernel(const float4 *__restrict__ parts,
                                                         const uint32_t *__restrict__ start,
                                                         const uint32_t *__restrict__ end, CellGrid g, float inv_h,
                                                         float3 dmin, float3 dmax, int nx, int ny, int nz,
                                                         float *__restrict__ grid)
{
    const int ix = blockIdx.x * 8 + (threadIdx.x & 7);
    const int iy = blockIdx.y * 8 + ((threadIdx.x >> 3) & 7);
    const int iz = blockIdx.z * 4 + (threadIdx.x >> 6);
    if (ix >= nx || iy >= ny || iz >= nz)
        return;
    const float p[3] = {dmin.x + (float)ix / (float)nx * (dmax.x - dmin.x),
                        dmin.y + (float)iy / (float)ny * (dmax.y - dmin.y),
                        dmin.z + (float)iz / (float)nz * (dmax.z - dmin.z)};
    const int c[3] = {cell_coord(g, p[0], 0), cell_coord(g, p[1], 1), cell_coord(g, p[2], 2)};
    float value = 0.0f, wsum = 0.0f;
    const int z0 = max(c[2] - 1, 0), z1 = min(c[2] + 2, (int)g.sz);
    const int y0 = max(c[1] - 1, 0), y1 = min(c[1] + 2, (int)g.sy);
    const int x0 = max(c[0] - 1, 0), x1 = min(c[0] + 2, (int)g.sx);
    for (int cz = z0; cz < z1; ++cz)
        for (int cy = y0; cy < y1; ++cy)
            for (int cx = x0; cx < x1; ++cx)
            {
                const uint32_t ci = cell_index(g, cx, cy, cz);
                const uint32_t s = start[ci];
                if (s == kEmpty)
                    continue;
                const uint32_t e = end[ci];
                for (uint32_t q = s; q < e; ++q)
                {
                    const float4 pt = parts[q];
                    const float dx = p[0] - pt.x, dy = p[1] - pt.y, dz = p[2] - pt.z;
                    const float d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 < g.r2)
                    {
                        const float ex = pt.x - p[0], ey = pt.y - p[1], ez = pt.z - p[2];
                        const float w = cubic_spline(sqrtf(ex * ex + ey * ey + ez * ez) * inv_h);
                        wsum += w;
                        value += w * pt.w;
                    }
                }
            }
    if (wsum != 0.0f)
        value /= wsum;
    grid[((size_t)iz * ny + iy) * nx + ix] = value;
}

uint32_t next_multiple(uint32_t x, int mul) { return ((x - 1) / mul + 1) * mul; }

} // namespace

int launch_particles_to_grid(const float *x, const float *y, const float *z, const float *v, uint32_t n, float h,
                             const float *dmin, const float *dmax, int nx, int ny, int nz, float *grid,
                             hipStream_t stream)
{
    // bounding box (createGridFromDomainCPU, src/particle_grid.cpp:38-55)
    int *d_keys6 = nullptr;
    if (hipMallocAsync((void **)&d_keys6, 6 * sizeof(int), stream) != hipSuccess)
        return MRAY_ERR_OOM;
    hipLaunchKernelGGL(bbox_init_kernel, dim3(1), dim3(64), 0, stream, d_keys6);
    const unsigned bb_blocks = (unsigned)std::min<uint64_t>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(bbox_kernel, dim3(bb_blocks), dim3(256), 0, stream, x, y, z, n, d_keys6);
    int h_keys6[6];
    if (hipMemcpyAsync(h_keys6, d_keys6, sizeof(h_keys6), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return MRAY_ERR_HIP;
    (void)hipFreeAsync(d_keys6, stream);
    float mn[3], mx[3];
    for (int k = 0; k < 3; ++k)
    {
        mn[k] = key_float(h_keys6[k]);
        mx[k] = key_float(h_keys6[3 + k]);
    }

    // grid of cell size h, sizes padded to the 4^3 tiles (createGridFromDomain,
    // src/particle_grid.cuh:389-400); Vec3 / scalar multiplies by the reciprocal
    CellGrid g;
    const float inv_cellf = 1 / h;
    uint32_t gs[3];
    for (int k = 0; k < 3; ++k)
        gs[k] = (uint32_t)((mx[k] - mn[k]) * inv_cellf + 1);
    g.sx = next_multiple(gs[0], 4);
    g.sy = next_multiple(gs[1], 4);
    g.sz = (gs[2] > 1) ? next_multiple(gs[2], 4) : gs[2];
    for (int k = 0; k < 3; ++k)
        g.minp[k] = mn[k];
    g.inv_cell = (float)(1.0 / (double)h);
    g.r2 = h * h;
    const size_t ncells = (size_t)g.sx * g.sy * g.sz;
    if (ncells >= (size_t)kEmpty)
        return MRAY_ERR_UNSUPPORTED;
    int end_bit = 1;
    while (end_bit < 32 && ((uint64_t)1 << end_bit) < ncells)
        ++end_bit;

    // keys, stable sort, sorted records, cell ranges
    uint32_t *keys = nullptr, *ids = nullptr, *keys_s = nullptr, *ids_s = nullptr, *start = nullptr, *end = nullptr;
    float4 *parts = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    int rc = 0;
    auto alloc = [&](void **p, size_t bytes) {
        if (rc == 0 && hipMallocAsync(p, bytes, stream) != hipSuccess)
            rc = MRAY_ERR_OOM;
    };
    alloc((void **)&keys, (size_t)n * 4);
    alloc((void **)&ids, (size_t)n * 4);
    alloc((void **)&keys_s, (size_t)n * 4);
    alloc((void **)&ids_s, (size_t)n * 4);
    alloc((void **)&parts, (size_t)n * sizeof(float4));
    alloc((void **)&start, ncells * 4);
    alloc((void **)&end, ncells * 4);
    if (rc == 0 && hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys_s, ids, ids_s, (int)n, 0, end_bit,
                                                      stream) != hipSuccess)
        rc = MRAY_ERR_HIP;
    alloc(&tmp, tmp_bytes);
    if (rc == 0)
    {
        const unsigned pb = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(cell_key_kernel, dim3(pb), dim3(256), 0, stream, x, y, z, n, g, keys, ids);
        if (hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys_s, ids, ids_s, (int)n, 0, end_bit,
                                               stream) != hipSuccess)
            rc = MRAY_ERR_HIP;
    }
    if (rc == 0)
    {
        const unsigned pb = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(gather_particles_kernel, dim3(pb), dim3(256), 0, stream, x, y, z, v, ids_s, n, parts);
        const unsigned fb = (unsigned)std::min<size_t>((ncells + 255) / 256, 65536);
        hipLaunchKernelGGL(fill_u32_kernel, dim3(fb), dim3(256), 0, stream, start, ncells, kEmpty);
        hipLaunchKernelGGL(cell_range_kernel, dim3(pb), dim3(256), 0, stream, keys_s, n, start, end);
        dim3 blocks((nx + 7) / 8, (ny + 7) / 8, (nz + 3) / 4);
        hipLaunchKernelGGL(sph_gather_kernel, blocks, dim3(256), 0, stream, parts, start, end, g, 1.f / h,
                           make_float3(dmin[0], dmin[1], dmin[2]), make_float3(dmax[0], dmax[1], dmax[2]), nx, ny,
                           nz, grid);
    }
    for (void *p : {(void *)keys, (void *)ids, (void *)keys_s, (void *)ids_s, (void *)parts, (void *)start, (void *)end,
                    tmp})
        if (p)
            (void)hipFreeAsync(p, stream);
    return rc;
}

} // namespace mray
