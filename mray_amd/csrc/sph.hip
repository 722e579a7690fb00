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

#include "launch.hpp"
#include "sph_common.hpp"

namespace mray
{

namespace
{

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
__global__ __launch_bounds__(256) void sph_gather_kernel(SphView sv, float3 dmin, float3 dmax, int nx, int ny, int nz,
                                                         float *__restrict__ grid)
{
    const int ix = blockIdx.x * 8 + (threadIdx.x & 7);
    const int iy = blockIdx.y * 8 + ((threadIdx.x >> 3) & 7);
    const int iz = blockIdx.z * 4 + (threadIdx.x >> 6);
    if (ix >= nx || iy >= ny || iz >= nz)
        return;
    const float px = dmin.x + (float)ix / (float)nx * (dmax.x - dmin.x);
    const float py = dmin.y + (float)iy / (float)ny * (dmax.y - dmin.y);
    const float pz = dmin.z + (float)iz / (float)nz * (dmax.z - dmin.z);
    grid[((size_t)iz * ny + iy) * nx + ix] = sph_sample(sv, px, py, pz);
}

uint32_t next_multiple(uint32_t x, int mul) { return ((x - 1) / mul + 1) * mul; }

} // namespace

int sph_build(const float *x, const float *y, const float *z, const float *v, uint32_t n, float h, bool persistent,
              SphBuffers &out, hipStream_t stream)
{
    auto alloc = [&](void **p, size_t bytes) {
        return persistent ? hipMalloc(p, bytes) : hipMallocAsync(p, bytes, stream);
    };
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
    if (ncells >= (size_t)kSphEmpty)
        return MRAY_ERR_UNSUPPORTED;
    int end_bit = 1;
    while (end_bit < 32 && ((uint64_t)1 << end_bit) < ncells)
        ++end_bit;

    // keys, stable sort, sorted records, cell ranges
    uint32_t *keys = nullptr, *ids = nullptr, *keys_s = nullptr, *ids_s = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    int rc = 0;
    auto scratch = [&](void **p, size_t bytes) {
        if (rc == 0 && hipMallocAsync(p, bytes, stream) != hipSuccess)
            rc = MRAY_ERR_OOM;
    };
    out = SphBuffers{};
    out.g = g;
    out.inv_h = 1.f / h;
    out.n = n;
    out.ncells = ncells;
    scratch((void **)&keys, (size_t)n * 4);
    scratch((void **)&ids, (size_t)n * 4);
    scratch((void **)&keys_s, (size_t)n * 4);
    scratch((void **)&ids_s, (size_t)n * 4);
    if (rc == 0 && (alloc((void **)&out.parts, (size_t)n * sizeof(float4)) != hipSuccess ||
                    alloc((void **)&out.start, ncells * 4) != hipSuccess ||
                    alloc((void **)&out.end, ncells * 4) != hipSuccess))
        rc = MRAY_ERR_OOM;
    if (rc == 0 && hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys_s, ids, ids_s, (int)n, 0, end_bit,
                                                      stream) != hipSuccess)
        rc = MRAY_ERR_HIP;
    scratch(&tmp, tmp_bytes);
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
        hipLaunchKernelGGL(gather_particles_kernel, dim3(pb), dim3(256), 0, stream, x, y, z, v, ids_s, n, out.parts);
        const unsigned fb = (unsigned)std::min<size_t>((ncells + 255) / 256, 65536);
        hipLaunchKernelGGL(fill_u32_kernel, dim3(fb), dim3(256), 0, stream, out.start, ncells, kSphEmpty);
        hipLaunchKernelGGL(cell_range_kernel, dim3(pb), dim3(256), 0, stream, keys_s, n, out.start, out.end);
    }
    for (void *p : {(void *)keys, (void *)ids, (void *)keys_s, (void *)ids_s, tmp})
        if (p)
            (void)hipFreeAsync(p, stream);
    if (rc != 0)
        sph_free(out, persistent, stream);
    return rc;
}

void sph_free(SphBuffers &b, bool persistent, hipStream_t stream)
{
    for (void *p : {(void *)b.parts, (void *)b.start, (void *)b.end})
        if (p)
            (void)(persistent ? hipFree(p) : hipFreeAsync(p, stream));
    b.parts = nullptr;
    b.start = b.end = nullptr;
}

int launch_particles_to_grid(const float *x, const float *y, const float *z, const float *v, uint32_t n, float h,
                             const float *dmin, const float *dmax, int nx, int ny, int nz, float *grid,
                             hipStream_t stream)
{
    SphBuffers b;
    int rc = sph_build(x, y, z, v, n, h, false, b, stream);
    if (rc != 0)
        return rc;
    dim3 blocks((nx + 7) / 8, (ny + 7) / 8, (nz + 3) / 4);
    hipLaunchKernelGGL(sph_gather_kernel, blocks, dim3(256), 0, stream, b.view(),
                       make_float3(dmin[0], dmin[1], dmin[2]), make_float3(dmax[0], dmax[1], dmax[2]), nx, ny, nz,
                       grid);
    sph_free(b, false, stream);
    return 0;
}

} // namespace mray
