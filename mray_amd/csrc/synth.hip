// Device-side generators for the synthetic benchmark volumes. Same formulas
// as mray_amd/synth.py (SplitMix64-hashed value-noise lattice, hashed
// mixing-layer modes and blob parameters passed in from the host);
// deterministic, but not bit identical to the numpy versions (device vs libm
// transcendental functions).
#include "device_common.hpp"
#include "launch.hpp"

namespace mray
{

__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double hash_uniform(uint64_t seed, uint64_t stream, uint64_t counter)
{
    uint64_t key = splitmix64(seed * 0x100000001B3ull + stream);
    uint64_t h = splitmix64(key ^ splitmix64(counter));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ double lattice(uint64_t seed, int octave, int64_t ix, int64_t iy, int64_t iz)
{
    uint64_t key = ((uint64_t)ix * 73856093ull) ^ ((uint64_t)iy * 19349663ull) ^ ((uint64_t)iz * 83492791ull);
    return 2.0 * hash_uniform(seed, 100 + octave, key) - 1.0;
}

// octaves of value noise at lattice base * 2^o over normalised coordinates
// (weights 2^-o, normalised to [-1, 1]), the lattice hashed per octave
__device__ double value_noise(uint64_t seed, int octaves, double base, double x, double y, double z)
{
    double total = 0.0, norm = 0.0;
    for (int o = 0; o < octaves; ++o)
    {
        const double f = base * (double)(1 << o);
        const double px = x * f, py = y * f, pz = z * f;
        const double ix = floor(px), iy = floor(py), iz = floor(pz);
        const double fx = px - ix, fy = py - iy, fz = pz - iz;
        const double sx = fx * fx * (3 - 2 * fx), sy = fy * fy * (3 - 2 * fy), sz = fz * fz * (3 - 2 * fz);
        double acc = 0.0;
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx)
                {
                    const double w = (dx ? sx : 1 - sx) * (dy ? sy : 1 - sy) * (dz ? sz : 1 - sz);
                    acc += w * lattice(seed, o, (int64_t)ix + dx, (int64_t)iy + dy, (int64_t)iz + dz);
                }
        const double amp = 1.0 / (double)(1 << o);
        total += acc * amp;
        norm += amp;
    }
    return total / norm;
}

// the mixing layer's noise: 4 octaves from a lattice of 8
__device__ double value_noise4(uint64_t seed, double x, double y, double z) { return value_noise(seed, 4, 8.0, x, y, z); }

struct Modes
{
    double amp[16], theta[16], phi[16];
};

__global__ void mixing_layer_height(double *h, int n, Modes md)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y;
    if (ix >= n)
        return;
    const double x = (double)ix / n, y = (double)iy / n;
    double acc = 0.0;
    for (int k = 0; k < 16; ++k)
        acc += md.amp[k] *
               sin(2.0 * kPi * (k + 1) * (x * cos(md.theta[k]) + y * sin(md.theta[k])) + md.phi[k]);
    h[(size_t)iy * n + ix] = acc;
}

__global__ void mixing_layer_fill(float *out, const double *h, int n, uint64_t seed)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y;
    const int iz = blockIdx.z;
    if (ix >= n)
        return;
    const double x = (double)ix / n, y = (double)iy / n, z = (double)iz / n;
    const double layer = 0.5 + 0.5 * tanh((z - 0.5 - h[(size_t)iy * n + ix]) / 0.01);
    const double noise = value_noise4(seed, x, y, z);
    out[((size_t)iz * n + iy) * n + ix] = (float)(layer + 0.05 * noise);
}

__global__ void marschner_lobb_fill(float *out, int n)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y;
    const int iz = blockIdx.z;
    if (ix >= n)
        return;
    const double x = -1.0 + 2.0 * ix / n, y = -1.0 + 2.0 * iy / n, z = -1.0 + 2.0 * iz / n;
    const double r = sqrt(x * x + y * y);
    const double fm = 6.0, a = 0.25;
    const double rr = cos(2.0 * kPi * fm * cos(kPi * r / 2.0));
    out[((size_t)iz * n + iy) * n + ix] = (float)((1.0 - sin(kPi * z / 2.0) + a * (1.0 + rr)) / (2.0 * (1.0 + a)));
}

struct Blobs
{
    float cx[64], cy[64], cz[64], inv2s2[64], amp[64];
    int n;
};

// the noise-and-blobs volume (mray_amd/synth.py, blob_params): voxel (x, y,
// z) at p = (x, y, z) / nx; v = clamp(w smoothstep(e0, e1, 0.5 + 0.5 noise) +
// sum_k a_k exp(-|p - c_k|^2 / (2 s_k^2)), 0, 1) stored as floor(255 v + 0.5)
__global__ void noise_blobs_fill(uint8_t *out, int nx, int ny, int nz, Blobs bl, uint64_t seed, int octaves,
                                 double base, double weight, double e0, double e1)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y;
    const int iz = blockIdx.z;
    if (ix >= nx)
        return;
    const double inv = 1.0 / nx;
    const double x = ix * inv, y = iy * inv, z = iz * inv;
    const double nv = 0.5 + 0.5 * value_noise(seed, octaves, base, x, y, z);
    const double t = fmin(fmax((nv - e0) / (e1 - e0), 0.0), 1.0);
    double v = weight * t * t * (3.0 - 2.0 * t);
    const float xf = (float)x, yf = (float)y, zf = (float)z;
    float b = 0.0f;
    for (int k = 0; k < bl.n; ++k)
    {
        const float dx = xf - bl.cx[k], dy = yf - bl.cy[k], dz = zf - bl.cz[k];
        b += bl.amp[k] * expf(-(dx * dx + dy * dy + dz * dz) * bl.inv2s2[k]);
    }
    v = fmin(fmax(v + (double)b, 0.0), 1.0);
    out[((size_t)iz * ny + iy) * nx + ix] = (uint8_t)floor(255.0 * v + 0.5);
}

int launch_synth_noise_blobs(uint8_t *out, int nx, int ny, int nz, const float *centres, const float *sigmas,
                             const float *amps, int n_blobs, uint64_t seed, hipStream_t stream)
{
    if (n_blobs < 0 || n_blobs > 64)
        return MRAY_ERR_ARG;
    Blobs bl;
    bl.n = n_blobs;
    for (int k = 0; k < n_blobs; ++k)
    {
        bl.cx[k] = centres[3 * k];
        bl.cy[k] = centres[3 * k + 1];
        bl.cz[k] = centres[3 * k + 2];
        bl.inv2s2[k] = 1.0f / (2.0f * sigmas[k] * sigmas[k]);
        bl.amp[k] = amps[k];
    }
    hipLaunchKernelGGL(noise_blobs_fill, dim3((nx + 255) / 256, ny, nz), dim3(256), 0, stream, out, nx, ny, nz, bl,
                       seed, 6, 4.0, 0.5, 0.3, 0.7);
    return 0;
}

int launch_synth_mixing_layer(float *out, int n, const double *amp, const double *theta, const double *phi,
                              uint64_t seed, hipStream_t stream)
{
    Modes md;
    for (int k = 0; k < 16; ++k)
    {
        md.amp[k] = amp[k];
        md.theta[k] = theta[k];
        md.phi[k] = phi[k];
    }
    double *h = nullptr;
    if (hipMallocAsync((void **)&h, sizeof(double) * (size_t)n * n, stream) != hipSuccess)
        return MRAY_ERR_OOM;
    dim3 blk(256);
    hipLaunchKernelGGL(mixing_layer_height, dim3((n + 255) / 256, n), blk, 0, stream, h, n, md);
    hipLaunchKernelGGL(mixing_layer_fill, dim3((n + 255) / 256, n, n), blk, 0, stream, out, h, n, seed);
    if (hipFreeAsync(h, stream) != hipSuccess)
        return MRAY_ERR_HIP;
    return 0;
}

int launch_synth_marschner_lobb(float *out, int n, hipStream_t stream)
{
    hipLaunchKernelGGL(marschner_lobb_fill, dim3((n + 255) / 256, n, n), dim3(256), 0, stream, out, n);
    return 0;
}

} // namespace mray
