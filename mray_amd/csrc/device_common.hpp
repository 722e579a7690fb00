// Device-side building blocks shared by the moment-image kernels (gfx950).
//
// Everything that decides *which* samples a ray takes and *what value* each
// sample has (pixel -> ray, slab test, step loop, trilinear fetch,
// map_density, bound quantisation, TF lookup) follows the reference CPU
// semantics operation by operation, so that sample values and 16-bit bounds
// are bit-identical to the CPU oracle. This file is compiled with
// -ffp-contract=off; kernels request FMAs explicitly with __builtin_fmaf.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mray_types.h"
#include "div_rn.hpp"
#include "host_sincosf.hpp"

namespace mray
{

constexpr double kPi = 3.14159265358979323846;
constexpr float kInvPiF = 0.3183098861837907f; // M_INV_PI (src/raymarch_generate.h:28)

struct Vec3
{
    float x, y, z;
};

// cut::min / cut::max: (x < y) ? x : y, NaN handling by operand order
// (src/cut/cuda_math.h:77-85). Never fminf/fmaxf.
__device__ __forceinline__ float rmin(float x, float y) { return (x < y) ? x : y; }
__device__ __forceinline__ float rmax(float x, float y) { return (x > y) ? x : y; }
__device__ __forceinline__ float rclamp(float v, float lo, float hi) { return rmin(rmax(v, lo), hi); }
__device__ __forceinline__ int iclamp(int v, int lo, int hi)
{
    int a = (v > lo) ? v : lo;
    return (a < hi) ? a : hi;
}

// Camera passed by value as a kernel argument (no __constant__ state).
struct DevCamera
{
    int orthographic;
    float m[12]; // rows 0..2 of the camera-to-world matrix
    Vec3 pos, right, up, dir;
    float scale, aspect, znear;
};

inline DevCamera make_dev_camera(const mray_camera &c)
{
    DevCamera d;
    d.orthographic = c.orthographic;
    for (int i = 0; i < 12; ++i)
        d.m[i] = c.transform[i];
    d.pos = {c.position[0], c.position[1], c.position[2]};
    d.right = {c.right[0], c.right[1], c.right[2]};
    d.up = {c.up[0], c.up[1], c.up[2]};
    d.dir = {c.dir[0], c.dir[1], c.dir[2]};
    d.scale = c.scale;
    d.aspect = c.aspect;
    d.znear = c.znear;
    return d;
}

struct Ray
{
    Vec3 o, d;
};

// get_eye_ray (src/cut/raytracing.h:39-65) with the pixel -> NDC mapping of
// src/moment_raymarcher.cpp:133 (no half-pixel offset).
__device__ __forceinline__ Ray pixel_ray(const DevCamera &cam, int W, int H, int x, int y)
{
    float ndcx = x / (float)W * 2.f - 1.f;
    float ndcy = y / (float)H * 2.f - 1.f;
    Ray r;
    if (cam.orthographic)
    {
        r.o.x = (cam.pos.x + ndcx * cam.right.x) + ndcy * cam.up.x;
        r.o.y = (cam.pos.y + ndcx * cam.right.y) + ndcy * cam.up.y;
        r.o.z = (cam.pos.z + ndcx * cam.right.z) + ndcy * cam.up.z;
        r.d = cam.dir;
    }
    else
    {
        r.o = cam.pos;
        float px = ndcx * cam.aspect * cam.scale;
        float py = ndcy * cam.scale;
        float v[3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            v[k] = cam.m[k * 4 + 0] * px + cam.m[k * 4 + 1] * py + cam.m[k * 4 + 2] * -1.0f + cam.m[k * 4 + 3] * 0.0f;
        float inv = 1 / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        r.d = {v[0] * inv, v[1] * inv, v[2] * inv};
    }
    return r;
}

// intersect_AABB (src/cut/raytracing.h:67-86)
__device__ __forceinline__ bool intersect_aabb(const Ray &r, const Vec3 &bmin, const Vec3 &bmax, float &t_near,
                                               float &t_far)
{
    float ix = 1.0f / r.d.x, iy = 1.0f / r.d.y, iz = 1.0f / r.d.z;
    float tbx = ix * (bmin.x - r.o.x), tby = iy * (bmin.y - r.o.y), tbz = iz * (bmin.z - r.o.z);
    float ttx = ix * (bmax.x - r.o.x), tty = iy * (bmax.y - r.o.y), ttz = iz * (bmax.z - r.o.z);
    float tminx = rmin(ttx, tbx), tminy = rmin(tty, tby), tminz = rmin(ttz, tbz);
    float tmaxx = rmax(ttx, tbx), tmaxy = rmax(tty, tby), tmaxz = rmax(ttz, tbz);
    float largest_tmin = rmax(rmax(tminx, tminy), rmax(tminx, tminz));
    float smallest_tmax = rmin(rmin(tmaxx, tmaxy), rmin(tmaxx, tmaxz));
    t_near = largest_tmin;
    t_far = smallest_tmax;
    return smallest_tmax > largest_tmin;
}

// Device view of a regular grid, voxel (x,y,z) at data[z*nx*ny + y*nx + x].
template <typename T> struct VolumeView
{
    const T *__restrict__ data;
    int nx, ny, nz;
    Vec3 dmin, dmax;
    Vec3 rext; // host_reciprocal(dmax - dmin) per axis
};

// a / b with y = host_reciprocal(b) formed once (div_rn.hpp): the IEEE
// quotient, by three operations where the operands are in range
__device__ __forceinline__ float div_rn(float a, float b, float y)
{
    const float q = div_rn_core(a, b, y);
    if (__builtin_expect(!div_rn_fast_ok(a, q), 0))
        return a / b;
    return q;
}

// a / b with y = host_reciprocal(b): for a power-of-two b (POW2, see
// exact_pow2) the exact product a * y, else div_rn
template <bool POW2> __device__ __forceinline__ float div_c(float a, float b, float y)
{
    if constexpr (POW2)
        return a * y;
    else
        return div_rn(a, b, y);
}

// b > 0 is a power of two whose reciprocal host_reciprocal gives (exactly)
inline bool exact_pow2(float b)
{
    int e;
    return b > 0.0f && frexpf(b, &e) == 0.5f && host_reciprocal(b) == host_reciprocal(b);
}

template <typename T>
inline VolumeView<T> make_volume_view(const T *data, int nx, int ny, int nz, const float *dmin, const float *dmax)
{
    VolumeView<T> v;
    v.data = data;
    v.nx = nx;
    v.ny = ny;
    v.nz = nz;
    v.dmin = {dmin[0], dmin[1], dmin[2]};
    v.dmax = {dmax[0], dmax[1], dmax[2]};
    v.rext = {host_reciprocal(dmax[0] - dmin[0]), host_reciprocal(dmax[1] - dmin[1]),
              host_reciprocal(dmax[2] - dmin[2])};
    return v;
}

__device__ __forceinline__ float voxel_value(float v) { return v; }
__device__ __forceinline__ float voxel_value(uint8_t v) { return (float)(v / 255.f); }

// RegularGrid::sample (src/volume.h:97-123): corner-aligned trilinear with
// clamp-to-edge indices and the reference's weight order.
template <typename T> __device__ __forceinline__ float sample_volume(const VolumeView<T> &v, float px, float py, float pz)
{
    float ipx = ((px - v.dmin.x) / (v.dmax.x - v.dmin.x)) * (float)v.nx;
    float ipy = ((py - v.dmin.y) / (v.dmax.y - v.dmin.y)) * (float)v.ny;
    float ipz = ((pz - v.dmin.z) / (v.dmax.z - v.dmin.z)) * (float)v.nz;
    float flx = floorf(ipx), fly = floorf(ipy), flz = floorf(ipz);
    float ax = ipx - flx, ay = ipy - fly, az = ipz - flz;
    int xl = (int)flx, yl = (int)fly, zl = (int)flz;
    int x0 = iclamp(xl, 0, v.nx - 1), x1 = iclamp(xl + 1, 0, v.nx - 1);
    int y0 = iclamp(yl, 0, v.ny - 1), y1 = iclamp(yl + 1, 0, v.ny - 1);
    int z0 = iclamp(zl, 0, v.nz - 1), z1 = iclamp(zl + 1, 0, v.nz - 1);
    size_t sxy = (size_t)(v.ny * v.nx);
    size_t r00 = (size_t)z0 * sxy + (size_t)(y0 * v.nx);
    size_t r01 = (size_t)z0 * sxy + (size_t)(y1 * v.nx);
    size_t r10 = (size_t)z1 * sxy + (size_t)(y0 * v.nx);
    size_t r11 = (size_t)z1 * sxy + (size_t)(y1 * v.nx);
    float lll = voxel_value(v.data[r00 + x0]);
    float rll = voxel_value(v.data[r00 + x1]);
    float lrl = voxel_value(v.data[r01 + x0]);
    float rrl = voxel_value(v.data[r01 + x1]);
    float llr = voxel_value(v.data[r10 + x0]);
    float rlr = voxel_value(v.data[r10 + x1]);
    float lrr = voxel_value(v.data[r11 + x0]);
    float rrr = voxel_value(v.data[r11 + x1]);
    return (1.f - ax) * (1.f - ay) * (1.f - az) * lll + ax * (1.f - ay) * (1.f - az) * rll +
           (1.f - ax) * ay * (1.f - az) * lrl + ax * ay * (1.f - az) * rrl + (1.f - ax) * (1.f - ay) * az * llr +
           ax * (1.f - ay) * az * rlr + (1.f - ax) * ay * az * lrr + ax * ay * az * rrr;
}

// sample_volume split in two: issue() computes the weights and starts the
// eight loads, value() combines them later, so that the loads of the next
// sample are in flight while the current one is used (same arithmetic and
// result as sample_volume). OFF32: the volume spans at most 4 GiB, so the
// corners are addressed by 32-bit byte offsets from the base (one scalar
// base, no 64-bit address arithmetic per load). POW2: every extent of the
// domain is a power of two, so dividing by it is multiplying by its exact
// reciprocal.
template <typename T, bool OFF32 = false, bool POW2 = false> struct PendingSample
{
    float ax, ay, az;
    T lll, rll, lrl, rrl, llr, rlr, lrr, rrr;

    __device__ __forceinline__ void issue(const VolumeView<T> &v, float px, float py, float pz)
    {
        const float ipx = div_c<POW2>(px - v.dmin.x, v.dmax.x - v.dmin.x, v.rext.x) * (float)v.nx;
        const float ipy = div_c<POW2>(py - v.dmin.y, v.dmax.y - v.dmin.y, v.rext.y) * (float)v.ny;
        const float ipz = div_c<POW2>(pz - v.dmin.z, v.dmax.z - v.dmin.z, v.rext.z) * (float)v.nz;
        float flx = floorf(ipx), fly = floorf(ipy), flz = floorf(ipz);
        ax = ipx - flx;
        ay = ipy - fly;
        az = ipz - flz;
        int xl = (int)flx, yl = (int)fly, zl = (int)flz;
        int x0 = iclamp(xl, 0, v.nx - 1), x1 = iclamp(xl + 1, 0, v.nx - 1);
        int y0 = iclamp(yl, 0, v.ny - 1), y1 = iclamp(yl + 1, 0, v.ny - 1);
        int z0 = iclamp(zl, 0, v.nz - 1), z1 = iclamp(zl + 1, 0, v.nz - 1);
        const uint32_t sxy = (uint32_t)(v.ny * v.nx);
        if constexpr (OFF32)
        {
            constexpr uint32_t B = sizeof(T);
            const char *base = reinterpret_cast<const char *>(v.data);
            const uint32_t o0 = ((uint32_t)z0 * sxy + (uint32_t)(y0 * v.nx + x0)) * B;
            const uint32_t dxb = (uint32_t)(x1 - x0) * B;
            const uint32_t dyb = (uint32_t)((y1 - y0) * v.nx) * B;
            const uint32_t o1 = o0 + ((z1 > z0) ? sxy * B : 0u);
            lll = *reinterpret_cast<const T *>(base + o0);
            rll = *reinterpret_cast<const T *>(base + (o0 + dxb));
            lrl = *reinterpret_cast<const T *>(base + (o0 + dyb));
            rrl = *reinterpret_cast<const T *>(base + (o0 + dyb + dxb));
            llr = *reinterpret_cast<const T *>(base + o1);
            rlr = *reinterpret_cast<const T *>(base + (o1 + dxb));
            lrr = *reinterpret_cast<const T *>(base + (o1 + dyb));
            rrr = *reinterpret_cast<const T *>(base + (o1 + dyb + dxb));
            return;
        }
        // one 64-bit corner address, the others by 32-bit offsets (the
        // clamped neighbour is 0 or 1 cell away on each axis)
        const T *p0 = v.data + ((size_t)z0 * sxy + (uint32_t)(y0 * v.nx + x0));
        const uint32_t dx = (uint32_t)(x1 - x0);
        const uint32_t dy = (uint32_t)((y1 - y0) * v.nx);
        const T *p1 = p0 + ((z1 > z0) ? sxy : 0u);
        lll = p0[0];
        rll = p0[dx];
        lrl = p0[dy];
        rrl = p0[dy + dx];
        llr = p1[0];
        rlr = p1[dx];
        lrr = p1[dy];
        rrr = p1[dy + dx];
    }

    __device__ __forceinline__ float value() const
    {
        return (1.f - ax) * (1.f - ay) * (1.f - az) * voxel_value(lll) + ax * (1.f - ay) * (1.f - az) * voxel_value(rll) +
               (1.f - ax) * ay * (1.f - az) * voxel_value(lrl) + ax * ay * (1.f - az) * voxel_value(rrl) +
               (1.f - ax) * (1.f - ay) * az * voxel_value(llr) + ax * (1.f - ay) * az * voxel_value(rlr) +
               (1.f - ax) * ay * az * voxel_value(lrr) + ax * ay * az * voxel_value(rrr);
    }
};

// map_density (src/parameters.h:31-34)
__device__ __forceinline__ float map_density(float s, float dmin, float dmax)
{
    return rclamp((s - dmin) / (dmax - dmin), 0.0f, 1.0f);
}

// map_density with host_reciprocal(dmax - dmin) formed once (same result)
template <bool POW2 = false>
__device__ __forceinline__ float map_density_r(float s, float dmin, float dmax, float rinv)
{
    return rclamp(div_c<POW2>(s - dmin, dmax - dmin, rinv), 0.0f, 1.0f);
}

// DensityBound::set (src/common.h:143-147)
__device__ __forceinline__ uint32_t pack_bound(float l, float u)
{
    uint32_t lo = (uint16_t)(l * 65535.f);
    uint32_t hi = (uint16_t)(u * 65535.f);
    return lo | (hi << 16);
}

// fast_atan (src/mese/complex_algebra.h:131-152)
__device__ __forceinline__ float fast_atan(float tan)
{
    bool negative = (tan < 0.0f);
    float abs_tan = negative ? (-tan) : tan;
    bool greater_one = (abs_tan > 1.0f);
    float x = greater_one ? abs_tan : 1.0f;
    x = 1.0f / x;
    float y = greater_one ? 1.0f : abs_tan;
    x = x * y;
    y = x * x;
    float z = y * 0.020835f - 0.085133f;
    z = y * z + 0.180141f;
    z = y * z - 0.330299f;
    y = y * z + 0.999866f;
    z = y * x;
    z = z * -2.0f + 1.570796f;
    z = greater_one ? z : 0.0f;
    x = x * y + z;
    return negative ? (-x) : x;
}

// TransferFunction::sample on the row-constant 256-entry table
// (src/transfer_function.h:44-59): only frac(256*bs) weights the two entries.
__device__ __forceinline__ float4 tf_sample(const float4 *tf, float fs, float bs)
{
    float ipx = fs * 256.0f, ipy = bs * 256.0f;
    float ax = ipx - floorf(ipx);
    float ay = ipy - floorf(ipy);
    int xl = (int)floorf(ipx);
    int x0 = iclamp(xl, 0, 255), x1 = iclamp(xl + 1, 0, 255);
    float4 l = tf[x0], r = tf[x1];
    float w0 = (1.f - ax) * (1.f - ay), w1 = ax * (1.f - ay), w2 = (1.f - ax) * ay, w3 = ax * ay;
    float4 o;
    o.x = w0 * l.x + w1 * l.x + w2 * r.x + w3 * r.x;
    o.y = w0 * l.y + w1 * l.y + w2 * r.y + w3 * r.y;
    o.z = w0 * l.z + w1 * l.z + w2 * r.z + w3 * r.z;
    o.w = w0 * l.w + w1 * l.w + w2 * r.w + w3 * r.w;
    return o;
}

// tf_sample with the weights of each table entry summed first: w0 + w1 =
// (1 - ax)(1 - ay) + ax (1 - ay) = 1 - ay and w2 + w3 = ay exactly, so the
// front sample's fraction drops out; fused multiply-adds, two operations per
// channel instead of seven. The same lookup up to rounding, closer to its
// exact value (MIXED form only).
__device__ __forceinline__ float4 tf_sample_mixed(const float4 *tf, float fs, float bs)
{
    const float ipx = fs * 256.0f, ipy = bs * 256.0f;
    const float ay = ipy - floorf(ipy);
    const int xl = (int)floorf(ipx);
    const int x0 = iclamp(xl, 0, 255), x1 = iclamp(xl + 1, 0, 255);
    const float4 l = tf[x0], r = tf[x1];
    const float w01 = 1.f - ay, w23 = ay;
    float4 o;
    o.x = __builtin_fmaf(w23, r.x, w01 * l.x);
    o.y = __builtin_fmaf(w23, r.y, w01 * l.y);
    o.z = __builtin_fmaf(w23, r.z, w01 * l.z);
    o.w = __builtin_fmaf(w23, r.w, w01 * l.w);
    return o;
}

// Upper bound on the iterations of a march loop from t_near to t_far with
// step s. Float rounding can shorten each increment by up to half an ulp, so
// the bound has 1% plus 16 iterations of slack; it only ends loops whose
// t + s no longer advances (which would never terminate).
__device__ __forceinline__ int march_cap(float t_near, float t_far, float s)
{
    float n = (t_far - t_near) / s;
    n = rmin(rmax(n, 0.0f), 1.0e8f);
    return (int)(n * 1.01f) + 16;
}

// v 2^k as a 64-bit two's complement integer, the fraction truncated
// (|v| 2^k < 2^53): the summand of the fixed-point LDS accumulators, whose
// integer atomics cost a fraction of float atomics on LDS
__device__ __forceinline__ unsigned long long to_fixed(float v, int k)
{
    const double x = ldexp((double)v, k);
    const double hi = floor(x * 0x1p-32);
    const double lo = fma(-hi, 0x1p32, x); // in [0, 2^32), exact
    return ((unsigned long long)(uint32_t)(int)hi << 32) | (unsigned long long)(uint32_t)lo;
}

// the value of a fixed-point sum q at scale 2^k
__device__ __forceinline__ float from_fixed(unsigned long long q, int k)
{
    return (float)ldexp((double)(long long)q, -k);
}

// whether sums bounded by b can be kept in fixed point (b finite: the
// incoming gradients are; otherwise the callers add floats, so that infinities
// and NaNs propagate as they would through float sums)
__device__ __forceinline__ bool fixed_usable(float b) { return b < 3.0e38f; }

// exponent of a scale for sums bounded by b: 2^k b < 2^61 (0 for b = 0 or
// not finite)
__device__ __forceinline__ int fixed_scale(float b)
{
    int ex = 0;
    if (!(b > 0.0f && b < 3.0e38f))
        return 0;
    frexpf(b, &ex); // b < 2^ex
    return 61 - ex;
}

// sin and cos of r in FP64 for |r| <= pi/4: Taylor series to r^19 and r^20
// (truncation below 1e-21)
__device__ __forceinline__ void sincos_quarter(double r, double &sr, double &cr)
{
    const double r2 = r * r;
    double ps = 1.0 / 121645100408832000.0;   // 1/19!
    ps = fma(ps, -r2, 1.0 / 355687428096000.0); // 1/17!
    ps = fma(ps, -r2, 1.0 / 1307674368000.0);
    ps = fma(ps, -r2, 1.0 / 6227020800.0);
    ps = fma(ps, -r2, 1.0 / 39916800.0);
    ps = fma(ps, -r2, 1.0 / 362880.0);
    ps = fma(ps, -r2, 1.0 / 5040.0);
    ps = fma(ps, -r2, 1.0 / 120.0);
    ps = fma(ps, -r2, 1.0 / 6.0);
    sr = fma(-r * r2, ps, r);
    double pc = 1.0 / 2432902008176640000.0;  // 1/20!
    pc = fma(pc, -r2, 1.0 / 6402373705728000.0); // 1/18!
    pc = fma(pc, -r2, 1.0 / 20922789888000.0);
    pc = fma(pc, -r2, 1.0 / 87178291200.0);
    pc = fma(pc, -r2, 1.0 / 479001600.0);
    pc = fma(pc, -r2, 1.0 / 3628800.0);
    pc = fma(pc, -r2, 1.0 / 40320.0);
    pc = fma(pc, -r2, 1.0 / 720.0);
    pc = fma(pc, -r2, 1.0 / 24.0);
    pc = fma(pc, -r2, 0.5);
    cr = fma(-r2, pc, 1.0);
}

// sin and cos of x in FP64 for |x| < 1e5 (node phases in [-pi, 0], and
// their products with moment indices below 256): the nearest multiple k of
// pi/2 is taken off in two parts (the high part has 33 significant bits, so
// k * hi is exact for |k| < 2^20), then sincos_quarter on |r| <= pi/4.
__device__ __forceinline__ void sincos_small(double x, double &s, double &c)
{
    const double k = rint(x * 0.63661977236758134);
    const double r = fma(-k, 6.07710050650619224932e-11, fma(-k, 1.57079632673412561417e+00, x));
    double sr, cr;
    sincos_quarter(r, sr, cr);
    const int q = (int)k & 3;
    s = (q == 0) ? sr : (q == 1) ? cr : (q == 2) ? -sr : -cr;
    c = (q == 0) ? cr : (q == 1) ? -sr : (q == 2) ? -cr : sr;
}

// Blocks are dealt round-robin to the 8 XCDs (block b runs on XCD b mod 8).
// Remap so that each XCD takes runs of `run` consecutive blocks (a row of
// tiles) in turn: the j-th block on XCD k is block j mod run of run
// 8 (j / run) + k. Within a run neighbouring rays share the XCD's L2, and
// the runs are dealt out so that every XCD gets rows from all over the
// image: a perspective frame's rays are far shorter near its top and bottom
// edges, and a contiguous band of rows per XCD left the busiest XCD with
// 1.29x the mean work (1024 x 1024 rays into a cube at 25 degrees; 1.00x
// dealt out by rows). Blocks past the last whole round of 8 runs keep their
// index.
__device__ __forceinline__ uint32_t xcd_swizzle(uint32_t bid, uint32_t nblocks, uint32_t run)
{
    constexpr uint32_t kXcd = 8;
    const uint32_t span = kXcd * run;
    if (bid >= nblocks / span * span)
        return bid;
    const uint32_t j = bid / kXcd, k = bid % kXcd;
    return ((j / run) * kXcd + k) * run + j % run;
}

// This block's index in the tile order of an image W pixels wide, its blocks
// holding tiles_per_block 8-pixel-wide tiles each, dealt out by tile rows
// (xcd_swizzle)
__device__ __forceinline__ uint32_t xcd_tile_block(int W, uint32_t tiles_per_block)
{
    const uint32_t tiles_x = (uint32_t)(W + 7) / 8;
    const uint32_t run = max(1u, (tiles_x + tiles_per_block - 1) / tiles_per_block);
    return xcd_swizzle(blockIdx.x, gridDim.x, run);
}

// 8x8 pixel tile per 64-lane wavefront; tiles in row-major order.
__device__ __forceinline__ void tile_pixel(uint32_t tile, int W, int lane, int &x, int &y)
{
    int tiles_x = (W + 7) / 8;
    x = (int)(tile % tiles_x) * 8 + (lane & 7);
    y = (int)(tile / tiles_x) * 8 + (lane >> 3);
}

} // namespace mray
