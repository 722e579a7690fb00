// Moment-image compaction and prediction coding, reproducing the reference
// CPU path bit for bit (FP32, the reference's operation order, the host C
// library's sinf/cosf):
//
//   compute_num_moments     per-pixel moment count from the rRMSE of moments
//                           predicted by the MESE extension
//                           (src/moment_compaction.h:35-174)
//   index scan, compaction  exclusive scan of the counts into index[W*H+1]
//                           and the copy into the compact layout
//                           (MomentImageHost::compact, src/moment_image.cpp:69-94;
//                           src/moment_image_device.cu:65-110)
//   prediction_encode       bias, Herglotz transform, encoding into reflection
//                           coefficients, rotation to the real axis, coding
//                           parameters and warp
//                           (src/moment_image.cpp:96-152,
//                           src/moment_prediction_coding.h:47-159,
//                           src/coding_transform.h:37-68, src/coding_params.h:56-78)
//   prepare_coded           decoding of a prediction-coded image into
//                           Lagrange multipliers
//                           (prepare_moments_from_pred_coding, src/moment_preparation.h:32-53)
//
// Per-pixel work is sequential (the reference's sums must be formed in its
// order to round alike), one thread per pixel with its complex arrays in LDS
// columns; the workgroup size adapts so that the columns fit the LDS.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <vector>

#include <algorithm>
#include <type_traits>

#include "launch.hpp"
#include "mese_f32.hpp"

namespace mray
{

// ---------------------------------------------------------------------------
// compaction: number of moments per pixel
// ---------------------------------------------------------------------------

// The search keeps two per-pixel arrays of m complex numbers: the exponential
// moments E (fixed for the pixel; a global scratch column) and a work array X
// in LDS that holds at once the three arrays the reference keeps separately,
// since their live parts never overlap:
//   X[0 .. n_known-1]  the Levinson solution lev (scaled by 2 pi),
//   X[n_known .. m-1]  the extended exponential moments newE, each of which
//                      is overwritten by the bounded moment B[l] computed
//                      from it;
// newE[l] for l < n_known is E[l] (the reference copies it), and B[l] for
// l < n_known is the real moment l. One LDS column per pixel instead of three
// fits three times the pixels per CU. The sums run in the reference's order
// with its operations, so the counts are bit for bit its counts.

// evaluate_rrmse (src/moment_compaction.h:84-111) with evaluate_unknown_moments
// (:35-60: extend the first n_known exponential moments to m with the MESE,
// Peters 2019, Eq. B.3) and exponential_to_bounded_moments (:65-79, including
// its use of the l-th exponential moment inside the j loop).
// The search runs about one wavefront per SIMD (the LDS columns bound the
// pixels per CU), so nothing hides the latency of a load but the loads issued
// with it: the loops read the operands of a run of kRun terms (global E, LDS
// X) before they combine them, in the reference's order. Registers are not
// what bounds the occupancy here, so the runs are long.
constexpr int kRun = 16;

// The saved Levinson states are written once and read once, long after; the
// exponential moments are read thousands of times per search, and the
// resident searches' columns of them (about 4.9 MB per XCD at 100 moments)
// already exceed its 4 MB L2. The states' loads and stores are nontemporal so
// that they do not evict the exponential moments (238 -> 233 ms for the
// counts of a 2048 x 2048 frame of 100 moments).
__device__ __forceinline__ cf load_nt(const cf &a)
{
    const unsigned long long b = __builtin_nontemporal_load(reinterpret_cast<const unsigned long long *>(&a));
    cf r;
    __builtin_memcpy(&r, &b, sizeof(r));
    return r;
}

__device__ __forceinline__ void store_nt(cf &a, cf v)
{
    unsigned long long b;
    __builtin_memcpy(&b, &v, sizeof(b));
    __builtin_nontemporal_store(b, reinterpret_cast<unsigned long long *>(&a));
}

// acc + sum_{j = j0 .. j1} a(j) b(j), added in order of j
template <class FA, class FB> __device__ __forceinline__ cf dot_run(cf acc, int j0, int j1, FA a, FB b)
{
    int j = j0;
    for (; j + kRun - 1 <= j1; j += kRun)
    {
        cf av[kRun], bv[kRun];
#pragma unroll
        for (int i = 0; i < kRun; ++i)
        {
            av[i] = a(j + i);
            bv[i] = b(j + i);
        }
#pragma unroll
        for (int i = 0; i < kRun; ++i)
            acc = cadd(acc, cmul(av[i], bv[i]));
    }
    for (; j <= j1; ++j)
        acc = cadd(acc, cmul(a(j), b(j)));
    return acc;
}

// levinson (mese_f32.hpp) with its sums in runs, and its pair updates
// (P_k, P_{j-k}) four pairs at a time: the pairs of one step are disjoint, so
// reading four before writing them changes nothing.
//
// Resumable: the recursion's state after its steps j < n (P_0 .. P_{n-1} and
// whether the |u|^2 >= 1 guard has fired) is the result for n_known = n, and
// the steps from there on are the same operations whatever came before, so
// levinson(n') for n' > n continues from that state bit for bit. `from`
// holds such a state for n_from (0: none), and the state reached is stored to
// `to` (before the caller scales the solution).
__device__ __forceinline__ void levinson_resume(int m, Col P, Col E, Col from, int n_from, bool inv_from, Col to,
                                                bool &inv_to)
{
    int j0 = 1;
    bool invalid = false;
    if (n_from > 0 && n_from <= m)
    {
        int i = 0;
        for (; i + kRun <= n_from; i += kRun)
        {
            cf v[kRun];
#pragma unroll
            for (int r = 0; r < kRun; ++r)
                v[r] = load_nt(from[i + r]);
#pragma unroll
            for (int r = 0; r < kRun; ++r)
                P[i + r] = v[r];
        }
        for (; i < n_from; ++i)
            P[i] = load_nt(from[i]);
        j0 = n_from;
        invalid = inv_from;
    }
    else
        P[0] = cmk(1.0f / E[0].x, 0.0f);
    for (int j = j0; j < m; ++j)
    {
        cf u = cmk(0.0f, 0.0f);
        if (!invalid)
            u = dot_run(u, 0, j - 1, [&](int k) { return P[k]; }, [&](int k) { return E[j - k]; });
        float a2 = u.x * u.x + u.y * u.y;
        if (a2 >= 1.0f)
        {
            u = cmk(0.0f, 0.0f);
            a2 = 0;
            invalid = true;
        }
        const float factor = 1.0f / (1.0f - a2);
        const float p0 = P[0].x;
        const int kend = j / 2;
        int k = 1;
        for (; k + 3 <= kend; k += 4)
        {
            cf av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
            {
                av[i] = P[k + i];
                bv[i] = P[j - k - i];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
            {
                const int q = j - k - i;
                P[k + i] = crmul(factor, csub(av[i], cmul(cconj(bv[i]), u)));
                if (q != k + i)
                    P[q] = crmul(factor, csub(bv[i], cmul(cconj(av[i]), u)));
            }
        }
        for (; k <= kend; ++k)
        {
            const int q = j - k;
            const cf a = P[k], b = P[q];
            P[k] = crmul(factor, csub(a, cmul(cconj(b), u)));
            if (q != k)
                P[q] = crmul(factor, csub(b, cmul(cconj(a), u)));
        }
        P[0] = cmk(factor * p0, 0.0f);
        P[j] = crmul(factor, crmul(-p0, u));
    }
    {
        int i = 0;
        for (; i + kRun <= m; i += kRun)
        {
            cf v[kRun];
#pragma unroll
            for (int r = 0; r < kRun; ++r)
                v[r] = P[i + r];
#pragma unroll
            for (int r = 0; r < kRun; ++r)
                store_nt(to[i + r], v[r]);
        }
        for (; i < m; ++i)
            store_nt(to[i], P[i]);
    }
    inv_to = invalid;
}

// evaluate_rrmse with the Levinson recursion resumed from the state `from`
// (the result for n_from <= n_known, or none), its own state stored to `to`
__device__ float evaluate_rrmse(int n_known, int m, const float *moments, Col E, Col X, float norm, Col from,
                                int n_from, bool inv_from, Col to, bool &inv_to)
{
    const int n = n_known - 1;
    const int mm = m - 1;
    levinson_resume(n_known, X, E, from, n_from, inv_from, to, inv_to);
    for (int i = 0; i < n + 1; ++i)
        X[i] = crmul(6.28318531f, X[i]);
    const cf inv_lev0 = cdiv(cmk(-1.0f, 0.0f), X[0]);
    for (int k = 0; k < mm - n; ++k)
    {
        // acc = sum_{j=1..n} newE[j + k] lev[n + 1 - j]: newE[j + k] is E up to
        // index n (j <= n - k), then the extension
        const int jsplit = n - k;
        cf acc = dot_run(
            cmk(0.0f, 0.0f), 1, jsplit, [&](int j) { return E[j + k]; }, [&](int j) { return X[n + 1 - j]; });
        acc = dot_run(
            acc, (jsplit < 0 ? 0 : jsplit) + 1, n, [&](int j) { return X[j + k]; },
            [&](int j) { return X[n + 1 - j]; });
        X[n + k + 1] = cmul(acc, inv_lev0);
    }
    // lev is no longer needed: its slots take the known bounded moments, the
    // real moments (as the reference's ext array holds them)
    {
        int q = 0;
        for (; q + kRun <= n_known; q += kRun)
        {
            float mv[kRun];
#pragma unroll
            for (int i = 0; i < kRun; ++i)
                mv[i] = moments[q + i];
#pragma unroll
            for (int i = 0; i < kRun; ++i)
                X[q + i] = cmk(mv[i], 0.0f);
        }
        for (; q < n_known; ++q)
            X[q] = cmk(moments[q], 0.0f);
    }
    // bounded moments B[l], l >= n_known, over the extension
    const cf gamma_0 = cdivr(E[0], 2.f);
    for (int l = n_known; l < m; ++l)
    {
        const cf El = X[l];
        const cf d = crmul((float)l, gamma_0);
        // j = 0 reads B[l] itself, still zero
        cf acc = cmk(0.0f, 0.0f);
        acc = cadd(acc, cmul(cmul(cdiv(cmk((float)l, 0.0f), d), El), acc));
        // c = (l - j) / d by the reference's complex division of the real
        // (l - j, 0): ((l - j) d.x + 0 d.y, -(l - j) d.y + 0 d.x) / |d|^2.
        // With d.x and d.y normal, (l - j) d.x and (l - j) d.y are nonzero
        // and finite, and adding the product by zero (+-0) leaves them as
        // they are: the same bits without those two products and sums.
        const float dinv = 1.0f / cabs2(d);
        auto sum_terms = [&](auto plain) {
            auto quotient = [&](int r) {
                const float fr = (float)r;
                if constexpr (decltype(plain)::value)
                    return cmk((fr * d.x) * dinv, (-fr * d.y) * dinv);
                else
                    return cdiv(cmk(fr, 0.0f), d);
            };
            int j = 1;
            for (; j + kRun - 1 < l; j += kRun)
            {
                cf bv[kRun];
#pragma unroll
                for (int i = 0; i < kRun; ++i)
                    bv[i] = X[l - j - i];
#pragma unroll
                for (int i = 0; i < kRun; ++i)
                    acc = cadd(acc, cmul(cmul(quotient(l - j - i), El), bv[i]));
            }
            for (; j < l; ++j)
                acc = cadd(acc, cmul(cmul(quotient(l - j), El), X[l - j]));
        };
        if (__builtin_isnormal(d.x) && __builtin_isnormal(d.y))
            sum_terms(std::true_type{});
        else
            sum_terms(std::false_type{});
        const cf den = cmul(cmul(cmk(0.0f, 2.0f), cmk((float)kPi, 0.0f)), gamma_0);
        X[l] = csub(cdiv(El, den), acc);
    }
    float sum = 0.0f;
    int k = n_known;
    for (; k + kRun <= m; k += kRun)
    {
        float xv[kRun], mv[kRun];
#pragma unroll
        for (int i = 0; i < kRun; ++i)
        {
            xv[i] = X[k + i].x;
            mv[i] = moments[k + i];
        }
#pragma unroll
        for (int i = 0; i < kRun; ++i)
        {
            const float diff = xv[i] - mv[i];
            sum += diff * diff;
        }
    }
    for (; k < m; ++k)
    {
        const float diff = X[k].x - moments[k];
        sum += diff * diff;
    }
    return sqrtf(sum) * norm;
}

// trig_to_exp (mese_f32.hpp, src/mese/MESE_dynamic.h:35-54) for the count
// search: the recurrence runs on the pixel's LDS column W (free until the
// search starts) with its sums in runs (dot_run: the same operations in the
// same order, so the same bits), and the result is copied to the global
// column E, where the search reads it. Reading E from global memory one term
// at a time, the transform took half the search's time (116 of 233 ms for a
// 2048 x 2048 frame of 100 moments).
template <class T> __device__ __forceinline__ void trig_to_exp_runs(int m, Col E, Col W, T t)
{
    const float ph = 3.14159265f * t(0) - 1.57079633f;
    W[0] = crmul(0.0795774715f, cmk(host_cosf(ph), host_sinf(ph)));
    for (int l = 1; l < m; ++l)
    {
        const cf acc = dot_run(
            cmk(0.0f, 0.0f), 0, l - 1, [&](int j) { return crmul((float)(l - j), W[j]); },
            [&](int j) { return cmk(t(l - j), 0.0f); });
        // (2 pi / l) is evaluated in double by the reference (M_PI)
        W[l] = cmul(acc, cmk(0.0f, (float)((double)(2.0f * kPi) / (double)(float)l)));
    }
    W[0] = crmul(2.0f, W[0]);
    int i = 0;
    for (; i + kRun <= m; i += kRun)
    {
        cf v[kRun];
#pragma unroll
        for (int r = 0; r < kRun; ++r)
            v[r] = W[i + r];
#pragma unroll
        for (int r = 0; r < kRun; ++r)
            E[i + r] = v[r];
    }
    for (; i < m; ++i)
        E[i] = W[i];
}

// determine_best_number_moments (src/moment_compaction.h:113-141): binary
// search on the rRMSE starting at m/2 with step m/4. After a step up from n
// every later n exceeds it (the later steps add up to less than s), so the
// Levinson state at n is kept (scratch columns A and B take turns) and every
// later evaluation resumes the recursion there instead of from step 1.
__device__ int determine_best_number_moments(int m, const float *moments, float error_threshold, Col E, Col X,
                                             Col A, Col B)
{
    if (error_threshold == 0.0f)
        return m;
    const float norm = 1.0f / moments[0];
    trig_to_exp_runs(m, E, X, [&](int i) { return moments[i]; });
    int n = m / 2;
    int s = m / 4;
    int n_kept = 0;
    bool inv_kept = false;
    while (s > 0)
    {
        bool inv_n = false;
        const float rrmse = evaluate_rrmse(n, m, moments, E, X, norm, A, n_kept, inv_kept, B, inv_n);
        const float d = rrmse - error_threshold;
        if (((d < 0) ? -d : d) < 1e-8f)
            return n;
        else if (rrmse > error_threshold)
        {
            n_kept = n;
            inv_kept = inv_n;
            const Col t = A;
            A = B;
            B = t;
            n += s;
        }
        else
            n -= s;
        s = s / 2;
    }
    return n;
}

// compute_num_moments (src/moment_compaction.h:147-174), one pixel per thread.
// The grid strides over the pixels in groups of blockDim.x; thread slot
// (blockIdx.x * blockDim.x + tid) owns column slot of each of the scratch's
// three arrays of gridDim.x * blockDim.x columns of m. (Lanes whose searches sit
// at different n idle in each other's loop iterations; dealing the pixels out
// by n level by level within a workgroup measured slower, the barriers
// costing more than the idling at the moment counts this search serves.)
__global__ __launch_bounds__(64) void num_moments_kernel(const float *__restrict__ dense,
                                                         const uint32_t *__restrict__ bounds, int npix, int m,
                                                         float error_threshold, uint32_t *__restrict__ counts,
                                                         cf *__restrict__ exp_scratch)
{
    extern __shared__ cf lds_cf[];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nslots = (int)gridDim.x * nt;
    // three scratch columns of m per thread: the exponential moments and two
    // Levinson states
    cf *col = exp_scratch + (size_t)blockIdx.x * nt + tid;
    const size_t span = (size_t)m * nslots;
    const Col E{col, nslots}, A{col + span, nslots}, B{col + 2 * span, nslots};
    const Col X{lds_cf + tid, nt};
    for (int base = (int)blockIdx.x * nt; base < npix; base += nslots)
    {
        const int pix = base + tid;
        if (pix >= npix)
            continue;
        const float *mo = dense + (size_t)pix * m;
        bool is_zero;
        if (bounds)
            is_zero = (bounds[pix] & 0xFFFFu) == (bounds[pix] >> 16);
        else
            is_zero = mo[0] == 0.0f;
        counts[pix] = is_zero ? 0u : (uint32_t)determine_best_number_moments(m, mo, error_threshold, E, X, A, B);
    }
}

// ---------------------------------------------------------------------------
// compaction: exclusive scan of the counts, copy into the compact layout
// ---------------------------------------------------------------------------

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v, uint32_t *red)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        red[w] = v;
    __syncthreads();
    uint32_t s = 0;
    for (int i = 0; i < kScanThreads / 64; ++i)
        s += red[i];
    return s;
}

// pass 1: the sum of each tile of counts
__global__ __launch_bounds__(kScanThreads) void scan_tile_sums(const uint32_t *__restrict__ counts, int n,
                                                               uint32_t *__restrict__ sums)
{
    __shared__ uint32_t red[kScanThreads / 64];
    const size_t base = (size_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
    for (int i = 0; i < kScanItems; ++i)
    {
        const size_t e = base + (size_t)i * kScanThreads + threadIdx.x;
        if (e < (size_t)n)
            s += counts[e];
    }
    s = block_sum_u32(s, red);
    if (threadIdx.x == 0)
        sums[blockIdx.x] = s;
}

// inclusive scan of the 256 per-thread totals in LDS (Hillis-Steele)
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *buf)
{
    const int t = threadIdx.x;
    buf[t] = v;
    __syncthreads();
    for (int o = 1; o < kScanThreads; o <<= 1)
    {
        const uint32_t add = (t >= o) ? buf[t - o] : 0u;
        __syncthreads();
        buf[t] += add;
        __syncthreads();
    }
    return buf[t] - v;
}

// pass 2: exclusive scan of the tile sums (one workgroup, chunks of 256 with
// a running carry); writes the grand total to index[n]
__global__ __launch_bounds__(kScanThreads) void scan_tile_offsets(uint32_t *__restrict__ sums, int ntiles,
                                                                  uint32_t *__restrict__ index, int n)
{
    __shared__ uint32_t buf[kScanThreads];
    uint32_t carry = 0;
    for (int c = 0; c < ntiles; c += kScanThreads)
    {
        const int i = c + threadIdx.x;
        const uint32_t v = (i < ntiles) ? sums[i] : 0u;
        const uint32_t ex = block_exclusive_scan(v, buf);
        const uint32_t chunk_total = buf[kScanThreads - 1];
        if (i < ntiles)
            sums[i] = carry + ex;
        carry += chunk_total;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        index[n] = carry;
}

// pass 3: exclusive scan inside each tile plus the tile's offset. The tile is
// staged in LDS, so counts and index may be the same array.
__global__ __launch_bounds__(kScanThreads) void scan_tile_apply(const uint32_t *counts, int n,
                                                                const uint32_t *__restrict__ offsets, uint32_t *index)
{
    __shared__ uint32_t tile[kScanTile];
    __shared__ uint32_t buf[kScanThreads];
    const int t = threadIdx.x;
    const size_t base = (size_t)blockIdx.x * kScanTile;
    for (int i = 0; i < kScanItems; ++i)
    {
        const size_t e = base + (size_t)i * kScanThreads + t;
        tile[i * kScanThreads + t] = (e < (size_t)n) ? counts[e] : 0u;
    }
    __syncthreads();
    uint32_t s = 0;
    for (int i = 0; i < kScanItems; ++i)
        s += tile[t * kScanItems + i];
    uint32_t run = offsets[blockIdx.x] + block_exclusive_scan(s, buf);
    for (int i = 0; i < kScanItems; ++i)
    {
        const uint32_t v = tile[t * kScanItems + i];
        tile[t * kScanItems + i] = run;
        run += v;
    }
    __syncthreads();
    for (int i = 0; i < kScanItems; ++i)
    {
        const size_t e = base + (size_t)i * kScanThreads + t;
        if (e < (size_t)n)
            index[e] = tile[i * kScanThreads + t];
    }
}

// compaction_copy_kernel (src/moment_image_device.cu:65-81): one thread per
// dense element, the first count(pixel) moments of each pixel are kept
__global__ __launch_bounds__(256) void compact_copy_kernel(const float *__restrict__ dense,
                                                           const uint32_t *__restrict__ index, size_t npix, int m,
                                                           float *__restrict__ out)
{
    const size_t total = npix * (size_t)m;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x)
    {
        const size_t pix = e / (size_t)m;
        const uint32_t l = (uint32_t)(e - pix * (size_t)m);
        const uint32_t start = index[pix];
        if (l < index[pix + 1] - start)
            out[start + l] = dense[e];
    }
}

// ---------------------------------------------------------------------------
// prediction coding
// ---------------------------------------------------------------------------

__device__ __forceinline__ float fmaxr(float x, float y) { return (x > y) ? x : y; }
__device__ __forceinline__ float fminr(float x, float y) { return (x < y) ? x : y; }
__device__ __forceinline__ float fclampr(float v, float lo, float hi) { return fminr(fmaxr(v, lo), hi); }

// prediction_coding_warp / prediction_coding_inv_warp (src/moment_prediction_coding.h:47-67);
// params[2(l-1)] = p1, params[2(l-1)+1] = p2 (dequantized coding parameters)
__device__ __forceinline__ float coding_warp(int l, float x, int warp, const float *params)
{
    float w;
    if (warp == MRAY_CODING_WARP_DEFAULT_TRANSFORMED)
        w = (x + params[(l - 1) * 2]) / (params[(l - 1) * 2 + 1] * 2.f);
    else
        w = x * 0.5f + 0.5f;
    return fclampr(w, 0.0f, 1.0f);
}
__device__ __forceinline__ float coding_inv_warp(int l, float x, int warp, const uint16_t *q)
{
    if (warp == MRAY_CODING_WARP_DEFAULT_TRANSFORMED)
    {
        const float p1 = q[(l - 1) * 2] / 65535.f, p2 = q[(l - 1) * 2 + 1] / 65535.f;
        return (x * p2 * 2.f) - p1;
    }
    return x * 2.f - 1.f;
}

// MomentImageHost::prediction_encode, first loop (src/moment_image.cpp:100-123):
// bias, exponential moments, encode (src/moment_prediction_coding.h:123-159)
// and transform_quantization_real (:81-93), in place. The rotation to the
// real axis depends only on code[0] = E[0], so each coefficient is written as
// soon as it is known.
__global__ __launch_bounds__(64) void prediction_encode_kernel(float *__restrict__ data,
                                                               const uint32_t *__restrict__ index, int npix, int m,
                                                               float bias)
{
    extern __shared__ cf lds_cf[];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int pix = blockIdx.x * nt + tid;
    if (pix >= npix)
        return;
    const int nm = min((int)(index[pix + 1] - index[pix]), m);
    if (nm <= 0)
        return;
    float *d = data + index[pix];
    d[0] = lerp_bias(d[0], 0.5f, bias);
    for (int i = 1; i < nm; ++i)
        d[i] = lerp_bias(d[i], 0.0f, bias);
    const Col E{lds_cf + tid, nt}, P{lds_cf + nt * m + tid, nt};
    trig_to_exp(nm, E, [&](int i) { return d[i]; });

    const cf code0 = E[0];
    const cf axis = cmk(-code0.y, code0.x);
    const float length_axis = sqrtf(cabs2(axis));
    const cf inv_axis = cdiv(cmk(length_axis, 0.0f), axis);
    P[0] = cmk(1.0f / E[0].x, 0.0f);
    bool invalid = false;
    for (int l = 1; l < nm; ++l)
    {
        cf u = cmk(0.0f, 0.0f);
        for (int k = 0; k < l; ++k)
            u = cadd(u, cmul(P[k], E[l - k]));
        d[l] = cmul(u, inv_axis).x;
        float a2 = cabs2(u);
        if (invalid || a2 >= 1.0f)
        {
            u = cmk(0.0f, 0.0f);
            a2 = 0;
            invalid = true;
        }
        levinson_update(l, P, u, 1.0f / (1.0f - a2));
    }
}

// order-preserving integer keys of floats (for atomic min / max)
__device__ __forceinline__ int float_key(float f)
{
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float key_float(int k) { return __int_as_float(k >= 0 ? k : k ^ 0x7fffffff); }

// find_coding_transform_parameters (src/coding_transform.h:37-68), min / max
// part: for every l in [1, m) the min and max (both starting at 0) of
// clamp(data[index[p] + l], -1, 1) over pixels with count >= l. A pixel whose
// count equals l contributes the first value of the next pixel, as in the
// reference; the value past the end of the data is skipped. keys holds
// (min, max) per l as float_key, initialised to float_key(0).
__global__ __launch_bounds__(256) void coding_minmax_kernel(const float *__restrict__ data,
                                                            const uint32_t *__restrict__ index, int npix, int m,
                                                            int *__restrict__ keys)
{
    extern __shared__ int lds_keys[]; // [2 (m - 1)]
    for (int i = threadIdx.x; i < 2 * (m - 1); i += blockDim.x)
        lds_keys[i] = 0;
    __syncthreads();
    const uint32_t total = index[npix];
    const int lane = threadIdx.x & 63;
    for (int base = blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < npix; base += gridDim.x * blockDim.x)
    {
        const int pix = base + lane;
        const bool valid = pix < npix;
        const uint32_t start = valid ? index[pix] : 0u;
        const int num = valid ? (int)(index[pix + 1] - start) : 0;
        for (int l = 1; l < m; ++l)
        {
            // lanes without a value contribute the neutral 0 (the start value)
            float v = 0.0f;
            if (num >= l && start + (uint32_t)l < total)
            {
                const float x = data[start + l];
                if (x == x)
                    v = fclampr(x, -1.0f, 1.0f) + 0.0f;
            }
            float lo = v, hi = v;
            for (int o = 32; o > 0; o >>= 1)
            {
                lo = fminf(lo, __shfl_xor(lo, o));
                hi = fmaxf(hi, __shfl_xor(hi, o));
            }
            if (lane == 0)
            {
                atomicMin(&lds_keys[2 * (l - 1)], float_key(lo));
                atomicMax(&lds_keys[2 * (l - 1) + 1], float_key(hi));
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * (m - 1); i += blockDim.x)
    {
        if (i & 1)
            atomicMax(&keys[i], lds_keys[i]);
        else
            atomicMin(&keys[i], lds_keys[i]);
    }
}

__global__ void fill_i32_kernel(int *__restrict__ p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// the rest of find_coding_transform_parameters (p1 = |min|, p2 = |max - min| / 2)
// and CodingParamType::quantize / dequantize (src/coding_params.h:56-78)
__global__ void coding_params_kernel(const int *__restrict__ keys, int m, uint16_t *__restrict__ q,
                                     float *__restrict__ params)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m - 1)
        return;
    const float mn = key_float(keys[2 * p]), mx = key_float(keys[2 * p + 1]);
    const float d = mx - mn;
    const float raw[2] = {(mn < 0) ? -mn : mn, ((d < 0) ? -d : d) * 0.5f};
    for (int i = 0; i < 2; ++i)
    {
        const uint16_t v = (uint16_t)ceilf(fclampr(raw[i], 0.0f, 1.0f) * 65535.f);
        q[2 * p + i] = v;
        params[2 * p + i] = v / 65535.f;
    }
}

// MomentImageHost::prediction_encode, warping loop (src/moment_image.cpp:137-151)
__global__ __launch_bounds__(256) void coding_warp_kernel(float *__restrict__ data, const uint32_t *__restrict__ index,
                                                          int npix, int m, int warp, const float *__restrict__ params)
{
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix)
        return;
    const int nm = min((int)(index[pix + 1] - index[pix]), m);
    float *d = data + index[pix];
    for (int l = 1; l < nm; ++l)
        d[l] = coding_warp(l, d[l], warp, params);
}

// prepare_moments_from_pred_coding (src/moment_preparation.h:32-53):
// transform_dequantization_real (src/moment_prediction_coding.h:98-118, the
// coefficients recomputed where decode reads them), decode (:164-206), then
// the correlations of prepareMESELagrange.
__global__ __launch_bounds__(64) void prepare_coded_kernel(const float *__restrict__ mom,
                                                           const uint32_t *__restrict__ index, int npix, int m,
                                                           int warp, const uint16_t *__restrict__ qparams,
                                                           float *__restrict__ out)
{
    extern __shared__ cf lds_cf[];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int pix = blockIdx.x * nt + tid;
    if (pix >= npix)
        return;
    const int nm = index ? min((int)(index[pix + 1] - index[pix]), m) : m;
    const size_t off = index ? (size_t)index[pix] : (size_t)pix * m;
    if (nm <= 0)
        return;
    const float *in = mom + off;
    const Col E{lds_cf + tid, nt}, P{lds_cf + nt * m + tid, nt}, R{lds_cf + 2 * nt * m + tid, nt};

    const float ph = 3.14159265f * in[0] - 1.57079633f;
    const cf code0 = crmul(2.0f * 0.0795774715f, cmk(host_cosf(ph), host_sinf(ph)));
    const cf axis = cmk(-code0.y, code0.x);
    const cf factor_axis = cdivr(axis, sqrtf(cabs2(axis)));

    E[0] = code0;
    P[0] = cmk(1.0f / E[0].x, 0.0f);
    const float eps = 1e-6f;
    bool invalid = false;
    for (int l = 1; l < nm; ++l)
    {
        cf u = crmul(coding_inv_warp(l, in[l], warp, qparams), factor_axis);
        float a2 = cabs2(u);
        if (invalid || a2 >= (1.0f - eps))
        {
            u = cmk(0.0f, 0.0f);
            a2 = 0;
            invalid = true;
        }
        cf g = cmk(0.0f, 0.0f);
        for (int k = 1; k < l; ++k)
            g = cadd(g, cmul(P[k], E[l - k]));
        E[l] = cmul(cdiv(cmk(1.0f, 0.0f), P[0]), csub(u, g));
        levinson_update(l, P, u, 1.0f / (1.0f - a2));
    }
    lagrange_from_levinson(nm, out + off, E, P, R);
}

// The coding error the quantization-curve search measures
// (compute_reference / compute_coding_error with apply_quantization of one
// moment, src/measure_coding_error.cpp:41-66, 92-182): the pixel's coded
// moments, moment q_index floored to q_bits bits (q_index < 0: none), are
// decoded as in prepare_coded_kernel and converted to trigonometric moments
// (exponentialToTrigonometricMoments, src/mese/MESE_dynamic.h:78-97).
// q_index < 0 writes those moments to ref; else err[pixel] = sqrt(sum_l
// (ref_l - t_l)^2) / ref_0 (0 for empty pixels). Grid row y measures
// q_bits = q_bits_hi - y into err + y * npix, so one launch covers the widths
// the search tries for a moment.
__global__ __launch_bounds__(64) void coding_error_kernel(const float *__restrict__ mom,
                                                          const uint32_t *__restrict__ index, int npix, int m,
                                                          int warp, const uint16_t *__restrict__ qparams,
                                                          int q_index, int q_bits_hi, float *__restrict__ ref,
                                                          float *__restrict__ err)
{
    const float q_scale = ldexpf(1.0f, q_bits_hi - (int)blockIdx.y); // 2^b, exact
    err += (size_t)blockIdx.y * npix;
    extern __shared__ cf lds_cf[];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int pix = blockIdx.x * nt + tid;
    if (pix >= npix)
        return;
    const int nm = min((int)(index[pix + 1] - index[pix]), m);
    const size_t off = index[pix];
    if (nm <= 0)
    {
        if (q_index >= 0)
            err[pix] = 0.0f;
        return;
    }
    const float *in = mom + off;
    // apply_quantization: floor(v * 2^b) / 2^b, exact with q_scale = 2^b
    auto value = [&](int l) { return (l == q_index) ? floorf(in[l] * q_scale) / q_scale : in[l]; };
    const Col E{lds_cf + tid, nt}, P{lds_cf + nt * m + tid, nt}, T{lds_cf + 2 * nt * m + tid, nt};

    const float ph = 3.14159265f * value(0) - 1.57079633f;
    const cf code0 = crmul(2.0f * 0.0795774715f, cmk(host_cosf(ph), host_sinf(ph)));
    const cf axis = cmk(-code0.y, code0.x);
    const cf factor_axis = cdivr(axis, sqrtf(cabs2(axis)));
    E[0] = code0;
    P[0] = cmk(1.0f / E[0].x, 0.0f);
    const float eps = 1e-6f;
    bool invalid = false;
    for (int l = 1; l < nm; ++l)
    {
        cf u = crmul(coding_inv_warp(l, value(l), warp, qparams), factor_axis);
        float a2 = cabs2(u);
        if (invalid || a2 >= (1.0f - eps))
        {
            u = cmk(0.0f, 0.0f);
            a2 = 0;
            invalid = true;
        }
        cf g = cmk(0.0f, 0.0f);
        for (int k = 1; k < l; ++k)
            g = cadd(g, cmul(P[k], E[l - k]));
        E[l] = cmul(cdiv(cmk(1.0f, 0.0f), P[0]), csub(u, g));
        levinson_update(l, P, u, 1.0f / (1.0f - a2));
    }

    // exponentialToTrigonometricMoments; atan2 of the float parts in double
    const cf gamma_0 = cdivr(E[0], 2.f);
    const cf t0 = cdivr(gamma_0, 0.0795774715f);
    const float zph = (float)atan2((double)t0.y, (double)t0.x);
    const float tm0 = (zph + 1.57079633f) / 3.14159265f;
    T[0] = cmk(tm0, 0.0f);
    const cf den = cmul(cmk(0.0f, (float)(2.0f * M_PI)), gamma_0);
    float sum = 0.0f;
    const float *r = ref + off;
    if (q_index >= 0)
    {
        const float d = r[0] - tm0;
        sum = d * d;
    }
    else
        ref[off] = tm0;
    for (int l = 1; l < nm; ++l)
    {
        cf t = cmk(0.0f, 0.0f);
        for (int j = 1; j < l; ++j)
        {
            const cf c = cdiv(cmk((float)(l - j), 0.0f), crmul((float)l, gamma_0));
            t = cadd(t, cmul(cmul(c, E[j]), T[l - j]));
        }
        t = csub(cdiv(E[l], den), t);
        T[l] = t;
        if (q_index >= 0)
        {
            const float d = r[l] - t.x;
            sum = sum + d * d;
        }
        else
            ref[off + l] = t.x;
    }
    if (q_index >= 0)
        err[pix] = sqrtf(sum) / r[0];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

extern size_t g_lds_limit;

// the largest of 64, 32, 16 threads per workgroup whose `cols` LDS columns of
// m complex numbers fit; 0 if none does
static int column_block(int cols, int m)
{
    for (int nt = 64; nt >= 16; nt /= 2)
        if ((size_t)cols * m * nt * sizeof(cf) <= g_lds_limit)
            return nt;
    return 0;
}

int launch_num_moments(const float *dense, const uint32_t *bounds, int W, int H, int m, float error_threshold,
                       uint32_t *counts, hipStream_t stream)
{
    const int npix = W * H;
    const int nt = column_block(1, m);
    if (!nt)
        return MRAY_ERR_UNSUPPORTED;
    if (npix <= 0)
        return 0;
    // twice the workgroups the CUs hold at once (LDS-bound), at most one per
    // pixel group; the scratch has one column of m per thread of the grid
    int device = 0, cus = 256;
    if (hipGetDevice(&device) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return MRAY_ERR_HIP;
    const size_t lds = (size_t)m * nt * sizeof(cf);
    const int per_cu = (int)std::max<size_t>(1, g_lds_limit / lds);
    const int groups = (npix + nt - 1) / nt;
    const int blocks = std::min(groups, 2 * cus * per_cu);
    cf *scratch_e = nullptr;
    if (hipMallocAsync((void **)&scratch_e, (size_t)3 * blocks * nt * m * sizeof(cf), stream) != hipSuccess)
        return MRAY_ERR_OOM;
    hipLaunchKernelGGL(num_moments_kernel, dim3(blocks), dim3(nt), lds, stream, dense, bounds, npix, m,
                       error_threshold, counts, scratch_e);
    return hipFreeAsync(scratch_e, stream) == hipSuccess ? 0 : MRAY_ERR_HIP;
}

// counts of non-empty pixels clamped to [lo, hi]; empty pixels stay empty
__global__ void clamp_counts_kernel(uint32_t *__restrict__ counts, int npix, uint32_t lo, uint32_t hi)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix)
        return;
    const uint32_t c = counts[i];
    if (c > 0u)
        counts[i] = c < lo ? lo : (c > hi ? hi : c);
}

int launch_clamp_counts(uint32_t *counts, int npix, int lo, int hi, hipStream_t stream)
{
    hipLaunchKernelGGL(clamp_counts_kernel, dim3((npix + 255) / 256), dim3(256), 0, stream, counts, npix,
                       (uint32_t)lo, (uint32_t)hi);
    return 0;
}

int launch_moment_index(const uint32_t *counts, int npix, uint32_t *index, hipStream_t stream)
{
    const int ntiles = (npix + kScanTile - 1) / kScanTile;
    uint32_t *sums = nullptr;
    if (hipMallocAsync((void **)&sums, (size_t)ntiles * sizeof(uint32_t), stream) != hipSuccess)
        return MRAY_ERR_OOM;
    hipLaunchKernelGGL(scan_tile_sums, dim3(ntiles), dim3(kScanThreads), 0, stream, counts, npix, sums);
    hipLaunchKernelGGL(scan_tile_offsets, dim3(1), dim3(kScanThreads), 0, stream, sums, ntiles, index, npix);
    hipLaunchKernelGGL(scan_tile_apply, dim3(ntiles), dim3(kScanThreads), 0, stream, counts, npix, sums, index);
    return hipFreeAsync(sums, stream) == hipSuccess ? 0 : MRAY_ERR_HIP;
}

// ---------------------------------------------------------------------------
// the index scan without allocating: its tile sums live in a file-scope device
// array, one copy per GPU, held until the process ends. Calls on one GPU share
// it, so calls that overlap on one GPU race; every call writes each sum it
// reads before reading it, so nothing has to start at zero.
// ---------------------------------------------------------------------------
constexpr int kNoAllocMaxTiles = 16384; // 64 KiB: images up to 16384 * 4096 pixels
static __device__ __align__(256) uint32_t scratch[kNoAllocMaxTiles];

namespace
{
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
uint32_t *g_scratch_ptr[kMaxDevices] = {};
} // namespace

// The device address of scratch on `device`, which must be current: resolved
// by hipGetSymbolAddress once per device (the array's name is a host-side
// stand-in, never a device address).
int noalloc_scratch_init(int device, uint32_t **out)
{
    if (device < 0 || device >= kMaxDevices)
        return MRAY_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device])
    {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
            return MRAY_ERR_HIP;
        g_scratch_ptr[device] = static_cast<uint32_t *>(ptr);
    }
    if (out)
        *out = g_scratch_ptr[device];
    return 0;
}

size_t noalloc_max_pixels() { return (size_t)kNoAllocMaxTiles * kScanTile; }

int launch_moment_index_noalloc(const uint32_t *counts, int npix, uint32_t *index, hipStream_t stream)
{
    const int ntiles = (npix + kScanTile - 1) / kScanTile;
    if (ntiles > kNoAllocMaxTiles)
        return MRAY_ERR_ARG;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess)
        return MRAY_ERR_HIP;
    uint32_t *sums = nullptr;
    const int rc = noalloc_scratch_init(device, &sums);
    if (rc)
        return rc;
    hipLaunchKernelGGL(scan_tile_sums, dim3(ntiles), dim3(kScanThreads), 0, stream, counts, npix, sums);
    hipLaunchKernelGGL(scan_tile_offsets, dim3(1), dim3(kScanThreads), 0, stream, sums, ntiles, index, npix);
    hipLaunchKernelGGL(scan_tile_apply, dim3(ntiles), dim3(kScanThreads), 0, stream, counts, npix, sums, index);
    return 0;
}

int launch_compact(const float *dense, const uint32_t *index, int W, int H, int m, float *out, hipStream_t stream)
{
    const size_t total = (size_t)W * H * m;
    const size_t blocks = std::min<size_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(compact_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, dense, index,
                       (size_t)W * H, m, out);
    return 0;
}

int launch_prediction_encode(float *data, const uint32_t *index, int W, int H, int m, int warp, float bias,
                             uint16_t *qparams, hipStream_t stream)
{
    const int npix = W * H;
    const int nt = column_block(2, m);
    if (!nt)
        return MRAY_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(prediction_encode_kernel, dim3((npix + nt - 1) / nt), dim3(nt), (size_t)2 * m * nt * sizeof(cf),
                       stream, data, index, npix, m, bias);
    float *params = nullptr;
    int *keys = nullptr;
    const size_t np = m > 1 ? (size_t)2 * (m - 1) : 1;
    if (hipMallocAsync((void **)&params, np * sizeof(float), stream) != hipSuccess)
        return MRAY_ERR_OOM;
    if (warp == MRAY_CODING_WARP_DEFAULT_TRANSFORMED && m > 1)
    {
        if (hipMallocAsync((void **)&keys, np * sizeof(int), stream) != hipSuccess)
        {
            (void)hipFreeAsync(params, stream);
            return MRAY_ERR_OOM;
        }
        // float_key(0.0f) == 0: both min and max start at 0 as in the reference
        hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, stream, keys, (int)np, 0);
        const int blocks = std::min((npix + 255) / 256, 1024);
        hipLaunchKernelGGL(coding_minmax_kernel, dim3(blocks), dim3(256), np * sizeof(int), stream, data, index, npix,
                           m, keys);
        hipLaunchKernelGGL(coding_params_kernel, dim3((m - 1 + 63) / 64), dim3(64), 0, stream, keys, m, qparams,
                           params);
        if (hipFreeAsync(keys, stream) != hipSuccess)
            return MRAY_ERR_HIP;
    }
    hipLaunchKernelGGL(coding_warp_kernel, dim3((npix + 255) / 256), dim3(256), 0, stream, data, index, npix, m, warp,
                       params);
    return hipFreeAsync(params, stream) == hipSuccess ? 0 : MRAY_ERR_HIP;
}

int launch_prepare_coded(const float *moments, const uint32_t *index, int W, int H, int m, int warp,
                         const uint16_t *qparams, float *lagrange, hipStream_t stream)
{
    const int npix = W * H;
    const int nt = column_block(3, m);
    if (!nt)
        return MRAY_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(prepare_coded_kernel, dim3((npix + nt - 1) / nt), dim3(nt), (size_t)3 * m * nt * sizeof(cf),
                       stream, moments, index, npix, m, warp, qparams, lagrange);
    return 0;
}

// create_quantization_table (src/measure_coding_error.cpp:374-437): moment 0
// at 16 bits, moment 1 at start_bits, then per moment the fewest bits (down to
// 2, from the previous moment's) whose mean coding error stays below moment
// 1's. Each measurement decodes the image on the GPU; the mean is get_mean's
// float sum over the non-empty pixels in row order, on the host.
int launch_quantization_table(const float *data, const uint32_t *index, int W, int H, int m, int warp,
                              const uint16_t *qparams, int start_bits, uint8_t *table, hipStream_t stream)
{
    const int npix = W * H;
    const int nt = column_block(3, m);
    if (!nt)
        return MRAY_ERR_UNSUPPORTED;
    std::vector<uint32_t> h_index((size_t)npix + 1);
    if (hipMemcpyAsync(h_index.data(), index, h_index.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, stream) !=
            hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return MRAY_ERR_HIP;
    size_t nonempty = 0;
    for (int p = 0; p < npix; ++p)
    {
        if (h_index[p + 1] - h_index[p] > (uint32_t)m)
            return MRAY_ERR_ARG;
        nonempty += h_index[p + 1] > h_index[p];
    }
    // widths tried per moment: the previous moment's down to 2; no width
    // exceeds start_bits (table[1]), so at most start_bits - 1 of them
    if (start_bits < 2 || start_bits > 31)
        return MRAY_ERR_ARG;
    const int max_variants = start_bits - 1;
    float *ref = nullptr, *err = nullptr;
    if (hipMallocAsync((void **)&ref, ((size_t)h_index[npix] + 1) * sizeof(float), stream) != hipSuccess)
        return MRAY_ERR_OOM;
    if (hipMallocAsync((void **)&err, (size_t)npix * max_variants * sizeof(float), stream) != hipSuccess)
    {
        (void)hipFreeAsync(ref, stream);
        return MRAY_ERR_OOM;
    }
    const dim3 grid((npix + nt - 1) / nt);
    const size_t lds = (size_t)3 * m * nt * sizeof(cf);
    hipLaunchKernelGGL(coding_error_kernel, grid, dim3(nt), lds, stream, data, index, npix, m, warp, qparams, -1, 0,
                       ref, err);
    std::vector<float> h_err((size_t)npix * max_variants);
    int rc = 0;
    // the mean errors of moment idx at widths b_hi, b_hi - 1, ..., b_hi - n + 1
    // (one launch); each mean is get_mean's float sum in pixel order
    auto mean_errors = [&](int idx, int b_hi, int n, float *mean) {
        if (n > max_variants)
            return false;
        hipLaunchKernelGGL(coding_error_kernel, dim3(grid.x, n), dim3(nt), lds, stream, data, index, npix, m, warp,
                           qparams, idx, b_hi, ref, err);
        if (hipMemcpyAsync(h_err.data(), err, (size_t)npix * n * sizeof(float), hipMemcpyDeviceToHost, stream) !=
                hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess)
            return false;
        for (int v = 0; v < n; ++v)
        {
            float s = 0.0f;
            const float *e = h_err.data() + (size_t)v * npix;
            for (int p = 0; p < npix; ++p)
                if (h_index[p + 1] > h_index[p])
                    s += e[p];
            mean[v] = s / (float)nonempty;
        }
        return true;
    };
    std::fill(table, table + m, (uint8_t)0);
    table[0] = 16;
    if (m > 1)
        table[1] = (uint8_t)start_bits;
    float threshold = 0.0f;
    if (m > 1 && !mean_errors(1, start_bits, 1, &threshold))
        rc = MRAY_ERR_HIP;
    float means[31];
    for (int idx = 2; rc == 0 && idx < m; ++idx)
    {
        // the reference measures b = table[idx-1], ..., 2 until an error
        // reaches the threshold; all of them are measured at once and the same
        // first one ends the walk
        table[idx] = table[idx - 1];
        const int b_hi = table[idx - 1], n = b_hi - 1;
        if (n <= 0)
            continue;
        if (!mean_errors(idx, b_hi, n, means))
        {
            rc = MRAY_ERR_HIP;
            break;
        }
        for (int v = 0; v < n && means[v] < threshold; ++v)
            table[idx] = (uint8_t)(b_hi - v);
    }
    if (hipFreeAsync(ref, stream) != hipSuccess || hipFreeAsync(err, stream) != hipSuccess)
        rc = rc ? rc : MRAY_ERR_HIP;
    return rc;
}

} // namespace mray
