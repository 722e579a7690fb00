// Adjoint of the bounded-MESE preparation (MIXED, FP64): given dL/dlambda,
// dL/dmoments of lambda = prepare(moments) (prepare_mixed_kernel's
// mathematics: bias lerp, Herglotz transform, Levinson recursion,
// autocorrelation, imaginary correlation; src/mese/MESE_dynamic.h:35-211).
// The reference has no gradients; this is new.
//
// Reverse mode of the algorithm itself, one thread per pixel in FP64:
//  * the forward is recomputed and its arrays kept: exponential moments E,
//    Levinson solution P, autocorrelation R and the reflection coefficient
//    u_j of every Levinson step;
//  * the Levinson recursion is swept backwards without storing its iterates:
//    each step maps the pair (P_k, P_{j-k}) to
//        a' = f (a - conj(b) u),  b' = f (b - conj(a) u),  f = 1 / (1 - |u|^2),
//    which inverts to a = a' + u conj(b'), b = b' + u conj(a'), so the
//    previous iterate is rebuilt from the current one and u_j;
//  * steps after the |u|^2 >= 1 guard fired use u = 0 (constant).
// Up to 64 moments one wavefront holds a pixel, element k of every array in
// lane k's registers (prepare_adjoint_wave_kernel). Beyond that, per-pixel
// arrays live in a global scratch laid out [array][element][thread]
// so that the lanes of a wavefront touch consecutive addresses; the grid is a
// fixed number of threads (16 wavefronts per CU) that stride over the
// pixels, which bounds the scratch to threads x 7 arrays x m x 16 bytes
// (1.2 GB at 40 moments; fewer threads are launched for smaller images).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.hpp"
#include "launch.hpp"

namespace mray
{

namespace
{

struct dc
{
    double x, y;
};
__device__ __forceinline__ dc dmk(double x, double y) { return dc{x, y}; }
__device__ __forceinline__ dc dadd(dc a, dc b) { return dc{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ dc dsub(dc a, dc b) { return dc{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ dc dmul(dc a, dc b) { return dc{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ dc dconj(dc a) { return dc{a.x, -a.y}; }
__device__ __forceinline__ dc drmul(double s, dc a) { return dc{s * a.x, s * a.y}; }
__device__ __forceinline__ double ddot(dc a, dc b) { return a.x * b.x + a.y * b.y; } // Re(conj(a) b)

// strided per-thread column of a scratch array
struct Dcol
{
    dc *p;
    size_t s;
    __device__ __forceinline__ dc &operator[](int i) const { return p[(size_t)i * s]; }
};

constexpr int kAdjThreads = 128;
constexpr int kAdjBlocks = 8192; // 16 wavefronts per CU on 256 CUs: enough to hide the scratch latency
constexpr int kArrays = 7;

__global__ __launch_bounds__(kAdjThreads) void prepare_adjoint_kernel(const float *__restrict__ mom,
                                                                      const uint32_t *__restrict__ index, int npix,
                                                                      int m, float bias,
                                                                      const float *__restrict__ grad_lag,
                                                                      float *__restrict__ grad_mom,
                                                                      dc *__restrict__ scratch)
{
    const size_t nthreads = (size_t)gridDim.x * blockDim.x;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const Dcol E{scratch + tid, nthreads}, P{scratch + (size_t)m * nthreads + tid, nthreads},
        R{scratch + (size_t)2 * m * nthreads + tid, nthreads}, U{scratch + (size_t)3 * m * nthreads + tid, nthreads},
        Eb{scratch + (size_t)4 * m * nthreads + tid, nthreads}, Pb{scratch + (size_t)5 * m * nthreads + tid, nthreads},
        Rb{scratch + (size_t)6 * m * nthreads + tid, nthreads};
    const double b64 = (double)bias;

    for (int pix = (int)tid; pix < npix; pix += (int)nthreads)
    {
        const int nm = index ? min((int)(index[pix + 1] - index[pix]), m) : m;
        const size_t off = index ? (size_t)index[pix] : (size_t)pix * m;
        if (nm <= 0)
            continue;
        const float *mo = mom + off;
        auto Tm = [&](int i) { return (1.0 - b64) * (double)mo[i] + (i == 0 ? b64 * 0.5 : 0.0); };

        // ---- forward (prepare_mixed_kernel's mathematics) ----
        const double ph0 = kPi * Tm(0) - 0.5 * kPi;
        double s0, c0;
        sincos(ph0, &s0, &c0);
        E[0] = dmk(c0 * (0.25 / kPi), s0 * (0.25 / kPi));
        for (int l = 1; l < nm; ++l)
        {
            dc acc = dmk(0.0, 0.0);
            for (int j = 0; j < l; ++j)
                acc = dadd(acc, drmul((double)(l - j) * Tm(l - j), E[j]));
            const double c = 2.0 * kPi / (double)l;
            E[l] = dmk(-acc.y * c, acc.x * c);
        }
        E[0] = drmul(2.0, E[0]);
        P[0] = dmk(1.0 / E[0].x, 0.0);
        int j_invalid = nm; // first step whose u was forced to 0
        for (int j = 1; j < nm; ++j)
        {
            dc u = dmk(0.0, 0.0);
            if (j < j_invalid)
                for (int k = 0; k < j; ++k)
                    u = dadd(u, dmul(P[k], E[j - k]));
            double a2 = u.x * u.x + u.y * u.y;
            if (j < j_invalid && a2 >= 1.0)
                j_invalid = j;
            if (j >= j_invalid)
            {
                u = dmk(0.0, 0.0);
                a2 = 0.0;
            }
            U[j] = u;
            const double f = 1.0 / (1.0 - a2);
            const double p0 = P[0].x;
            for (int k = 1; 2 * k <= j; ++k)
            {
                const int q = j - k;
                const dc a = P[k], b = P[q];
                P[k] = drmul(f, dsub(a, dmul(dconj(b), u)));
                if (q != k)
                    P[q] = drmul(f, dsub(b, dmul(dconj(a), u)));
            }
            P[j] = drmul(-f * p0, u);
            P[0] = dmk(f * p0, 0.0);
        }
        for (int i = 0; i < nm; ++i)
            P[i] = drmul(2.0 * kPi, P[i]);
        for (int k = 0; k < nm; ++k)
        {
            dc acc = dmk(0.0, 0.0);
            for (int j = 0; j < nm - k; ++j)
                acc = dadd(acc, dmul(P[j], dconj(P[j + k])));
            R[k] = acc;
        }
        const double nf = 1.0 / (kPi * P[0].x);
        auto Eh = [&](int j) { return j == 0 ? drmul(0.5, E[0]) : E[j]; }; // E with E_0 halved

        // ---- reverse ----
        for (int i = 0; i < nm; ++i)
        {
            Eb[i] = dmk(0.0, 0.0);
            Pb[i] = dmk(0.0, 0.0);
            Rb[i] = dmk(0.0, 0.0);
        }
        // lambda_k = nf C_k, C_k = sum_j R_{k+j}.x Eh_j.y + R_{k+j}.y Eh_j.x
        double nfb = 0.0;
        for (int k = 0; k < nm; ++k)
        {
            const double lb = (double)grad_lag[off + k];
            double Ck = 0.0;
            for (int j = 0; j < nm - k; ++j)
            {
                const dc r = R[k + j], e = Eh(j);
                Ck += r.x * e.y + r.y * e.x;
            }
            nfb += lb * Ck;
            const double cb = lb * nf;
            if (cb != 0.0)
                for (int j = 0; j < nm - k; ++j)
                {
                    const dc r = R[k + j], e = Eh(j);
                    Rb[k + j] = dadd(Rb[k + j], dmk(cb * e.y, cb * e.x));
                    // Eh_j.y += cb r.x, Eh_j.x += cb r.y; E_0 = 2 Eh_0 so dEh_0 = dE_0 / 2
                    const double h = j == 0 ? 0.5 : 1.0;
                    Eb[j] = dadd(Eb[j], dmk(h * cb * r.y, h * cb * r.x));
                }
        }
        // nf = 1 / (pi P0.x)
        Pb[0].x += -nfb * nf / P[0].x;
        // R_k = sum_j P_j conj(P_{j+k})
        for (int k = 0; k < nm; ++k)
        {
            const dc rb = Rb[k];
            if (rb.x == 0.0 && rb.y == 0.0)
                continue;
            for (int j = 0; j < nm - k; ++j)
            {
                const dc a = P[j], b = P[j + k];
                Pb[j] = dadd(Pb[j], dmk(rb.x * b.x - rb.y * b.y, rb.x * b.y + rb.y * b.x));
                Pb[j + k] = dadd(Pb[j + k], dmk(rb.x * a.x + rb.y * a.y, rb.x * a.y - rb.y * a.x));
            }
        }
        // P = 2 pi P_levinson
        for (int i = 0; i < nm; ++i)
        {
            P[i] = drmul(1.0 / (2.0 * kPi), P[i]);
            Pb[i] = drmul(2.0 * kPi, Pb[i]);
        }
        // Levinson, backwards
        for (int j = nm - 1; j >= 1; --j)
        {
            const dc u = U[j];
            const bool valid = j < j_invalid;
            const double a2 = u.x * u.x + u.y * u.y;
            const double f = 1.0 / (1.0 - a2);
            // fb = <Pb', dP'/df> = <Pb', P'> / f over k = 0..j
            double fb = 0.0;
            for (int k = 0; k <= j; ++k)
                fb += ddot(Pb[k], P[k]);
            fb /= f;
            const dc pjb = Pb[j];
            const double p0b = Pb[0].x;
            // previous iterate
            const double p0 = P[0].x / f;
            for (int k = 1; 2 * k <= j; ++k)
            {
                const int q = j - k;
                const dc a1 = P[k], b1 = P[q];
                P[k] = dadd(a1, dmul(u, dconj(b1)));
                if (q != k)
                    P[q] = dadd(b1, dmul(u, dconj(a1)));
            }
            P[0] = dmk(p0, 0.0);
            P[j] = dmk(0.0, 0.0);
            // ub = sum_k -f Pb'_k P_{j-k} - f p0 Pb'_j + 2 f^2 fb u
            dc ub = drmul(2.0 * f * f * fb, u);
            for (int k = 1; k < j; ++k)
                ub = dsub(ub, drmul(f, dmul(Pb[k], P[j - k])));
            ub = dsub(ub, drmul(f * p0, pjb));
            // adjoints of the previous iterate
            for (int k = 1; 2 * k <= j; ++k)
            {
                const int q = j - k;
                const dc ak = Pb[k], bq = Pb[q];
                Pb[k] = drmul(f, dsub(ak, dmul(dconj(bq), u)));
                if (q != k)
                    Pb[q] = drmul(f, dsub(bq, dmul(dconj(ak), u)));
            }
            Pb[0] = dmk(f * p0b - f * ddot(pjb, u), 0.0);
            Pb[j] = dmk(0.0, 0.0);
            if (valid)
                for (int k = 0; k < j; ++k)
                {
                    // u = sum_k P_k E_{j-k}
                    Pb[k] = dadd(Pb[k], dmul(ub, dconj(E[j - k])));
                    Eb[j - k] = dadd(Eb[j - k], dmul(ub, dconj(P[k])));
                }
        }
        // P_0 = 1 / E_0.x (E_0 doubled)
        Eb[0].x += -Pb[0].x / (E[0].x * E[0].x);
        // the recurrence used E_0 before doubling: dE0_undoubled = 2 dE0
        Eb[0] = drmul(2.0, Eb[0]);
        const dc e0 = drmul(0.5, E[0]);
        // Herglotz recurrence, backwards: E_l = i c_l sum_j (l-j) T_{l-j} E_j;
        // dL/dT_i is kept in Rb[i].x (Rb is no longer needed)
        for (int i = 0; i < nm; ++i)
            Rb[i] = dmk(0.0, 0.0);
        for (int l = nm - 1; l >= 1; --l)
        {
            const double c = 2.0 * kPi / (double)l;
            const dc accb = dmk(c * Eb[l].y, -c * Eb[l].x); // conj(i c) Eb_l
            for (int j = 0; j < l; ++j)
            {
                const double w = (double)(l - j);
                const dc ej = j == 0 ? e0 : E[j];
                Eb[j] = dadd(Eb[j], drmul(w * Tm(l - j), accb));
                Rb[l - j].x += w * ddot(accb, ej);
            }
        }
        // E_0 = (cos(ph0), sin(ph0)) / (4 pi), ph0 = pi T_0 - pi / 2
        Rb[0].x += 0.25 * (-Eb[0].x * s0 + Eb[0].y * c0);
        for (int i = 0; i < nm; ++i)
            grad_mom[off + i] = (float)((1.0 - b64) * Rb[i].x);
    }
}

// ---------------------------------------------------------------------------
// Wavefront per pixel (m <= 64): element k of every array lives in lane k, in
// registers. Partners such as P[j - k] or E[j - k] are fetched with
// cross-lane shuffles, inner products are butterfly reductions over the
// wavefront; nothing but the input and output touches memory.
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ dc wsum(dc v) { return dc{wsum(v.x), wsum(v.y)}; }
// value of lane src (any int; callers mask lanes whose src is out of range)
__device__ __forceinline__ double lane_of(double v, int src) { return __shfl(v, src & 63); }
__device__ __forceinline__ dc lane_of(dc v, int src) { return dc{lane_of(v.x, src), lane_of(v.y, src)}; }

__global__ __launch_bounds__(256) void prepare_adjoint_wave_kernel(const float *__restrict__ mom,
                                                                   const uint32_t *__restrict__ index, int npix,
                                                                   int m, float bias,
                                                                   const float *__restrict__ grad_lag,
                                                                   float *__restrict__ grad_mom)
{
    const int k = threadIdx.x & 63;
    const int pix = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (pix >= npix) // uniform per wavefront
        return;
    const int nm = index ? min((int)(index[pix + 1] - index[pix]), m) : m;
    const size_t off = index ? (size_t)index[pix] : (size_t)pix * m;
    if (nm <= 0)
        return;
    const bool act = k < nm;
    const double b64 = (double)bias;
    const double T = act ? (1.0 - b64) * (double)mom[off + k] + (k == 0 ? b64 * 0.5 : 0.0) : 0.0;
    const dc zero = dmk(0.0, 0.0);

    // ---- forward ----
    const double ph0 = kPi * lane_of(T, 0) - 0.5 * kPi;
    double s0, c0;
    sincos(ph0, &s0, &c0);
    dc E = (k == 0) ? dmk(c0 * (0.25 / kPi), s0 * (0.25 / kPi)) : zero;
    for (int l = 1; l < nm; ++l)
    {
        const double Tl = lane_of(T, l - k);
        const dc term = (k < l) ? drmul((double)(l - k) * Tl, E) : zero;
        const dc acc = wsum(term);
        if (k == l)
        {
            const double c = 2.0 * kPi / (double)l;
            E = dmk(-acc.y * c, acc.x * c);
        }
    }
    if (k == 0)
        E = drmul(2.0, E);
    const double e0x = lane_of(E.x, 0);
    dc P = (k == 0) ? dmk(1.0 / e0x, 0.0) : zero;
    dc U = zero;
    int j_invalid = nm;
    for (int j = 1; j < nm; ++j)
    {
        const dc Ejk = lane_of(E, j - k);
        dc u = (j < j_invalid) ? wsum((k < j) ? dmul(P, Ejk) : zero) : zero;
        double a2 = u.x * u.x + u.y * u.y;
        if (j < j_invalid && a2 >= 1.0)
            j_invalid = j;
        if (j >= j_invalid)
        {
            u = zero;
            a2 = 0.0;
        }
        if (k == j)
            U = u;
        const double f = 1.0 / (1.0 - a2);
        const double p0 = lane_of(P.x, 0);
        const dc Pq = lane_of(P, j - k);
        if (k >= 1 && k < j)
            P = drmul(f, dsub(P, dmul(dconj(Pq), u)));
        else if (k == j)
            P = drmul(-f * p0, u);
        else if (k == 0)
            P = dmk(f * p0, 0.0);
    }
    P = drmul(2.0 * kPi, P);
    dc R = zero;
    for (int kk = 0; kk < nm; ++kk)
    {
        const dc Pn = lane_of(P, k + kk);
        const dc r = wsum((k + kk < nm) ? dmul(P, dconj(Pn)) : zero);
        if (k == kk)
            R = r;
    }
    const double p0s = lane_of(P.x, 0);
    const double nf = 1.0 / (kPi * p0s);
    const dc Eh = (k == 0) ? drmul(0.5, E) : E; // E_0 before doubling

    // ---- reverse: imaginary correlation and normalisation ----
    const double lb = act ? (double)grad_lag[off + k] : 0.0;
    double nfb = 0.0;
    dc Rb = zero, Eb = zero;
    for (int kk = 0; kk < nm; ++kk)
    {
        const dc Rn = lane_of(R, k + kk); // R_{kk+j} on lane j
        const bool in = k + kk < nm;
        const double Ck = wsum(in ? Rn.x * Eh.y + Rn.y * Eh.x : 0.0);
        const double lbk = lane_of(lb, kk);
        nfb += lbk * Ck;
        const double cb = lbk * nf;
        // Rb_{kk+j} += cb (Eh_j.y, Eh_j.x): lane n takes Eh_{n-kk}
        const dc Ehj = lane_of(Eh, k - kk);
        if (k >= kk && k < nm)
            Rb = dadd(Rb, dmk(cb * Ehj.y, cb * Ehj.x));
        // Eh_j += cb (R_{kk+j}.y, R_{kk+j}.x); E_0 = 2 Eh_0
        if (in)
        {
            const double h = (k == 0) ? 0.5 : 1.0;
            Eb = dadd(Eb, dmk(h * cb * Rn.y, h * cb * Rn.x));
        }
    }
    dc Pb = (k == 0) ? dmk(-nfb * nf / p0s, 0.0) : zero;
    // R_kk = sum_j P_j conj(P_{j+kk})
    for (int kk = 0; kk < nm; ++kk)
    {
        const dc rb = lane_of(Rb, kk);
        const dc b = lane_of(P, k + kk);
        const dc a = lane_of(P, k - kk);
        if (k + kk < nm)
            Pb = dadd(Pb, dmk(rb.x * b.x - rb.y * b.y, rb.x * b.y + rb.y * b.x));
        if (k >= kk && k < nm)
            Pb = dadd(Pb, dmk(rb.x * a.x + rb.y * a.y, rb.x * a.y - rb.y * a.x));
    }
    P = drmul(1.0 / (2.0 * kPi), P);
    Pb = drmul(2.0 * kPi, Pb);

    // ---- reverse: Levinson, rebuilding each previous iterate ----
    for (int j = nm - 1; j >= 1; --j)
    {
        const dc u = lane_of(U, j);
        const bool valid = j < j_invalid;
        const double f = 1.0 / (1.0 - (u.x * u.x + u.y * u.y));
        const double fb = wsum((k <= j) ? ddot(Pb, P) : 0.0) / f;
        const dc pjb = lane_of(Pb, j);
        const double p0b = lane_of(Pb.x, 0);
        const double p0 = lane_of(P.x, 0) / f;
        const dc Pq = lane_of(P, j - k);
        dc Pold = P;
        if (k >= 1 && k < j)
            Pold = dadd(P, dmul(u, dconj(Pq)));
        else if (k == 0)
            Pold = dmk(p0, 0.0);
        else if (k == j)
            Pold = zero;
        const dc Poq = lane_of(Pold, j - k); // P_{j-k} of the previous iterate
        dc ub = drmul(2.0 * f * f * fb, u);
        ub = dsub(ub, drmul(f, wsum((k >= 1 && k < j) ? dmul(Pb, Poq) : zero)));
        ub = dsub(ub, drmul(f * p0, pjb));
        const dc Pbq = lane_of(Pb, j - k);
        dc Pbn = Pb;
        if (k >= 1 && k < j)
            Pbn = drmul(f, dsub(Pb, dmul(dconj(Pbq), u)));
        else if (k == 0)
            Pbn = dmk(f * p0b - f * ddot(pjb, u), 0.0);
        else if (k == j)
            Pbn = zero;
        if (valid)
        {
            // u = sum_{k<j} P_k E_{j-k}
            const dc Ejk = lane_of(E, j - k);
            if (k < j)
                Pbn = dadd(Pbn, dmul(ub, dconj(Ejk)));
            if (k >= 1 && k <= j) // lane n = j - k takes P_k = Poq
                Eb = dadd(Eb, dmul(ub, dconj(Poq)));
        }
        P = Pold;
        Pb = Pbn;
    }
    // P_0 = 1 / E_0.x (E_0 doubled), then the recurrence's undoubled E_0
    if (k == 0)
    {
        Eb.x += -Pb.x / (e0x * e0x);
        Eb = drmul(2.0, Eb);
    }
    // ---- reverse: Herglotz recurrence E_l = i c_l sum_j (l-j) T_{l-j} E_j ----
    double Tb = 0.0;
    for (int l = nm - 1; l >= 1; --l)
    {
        const dc Ebl = lane_of(Eb, l);
        const double c = 2.0 * kPi / (double)l;
        const dc accb = dmk(c * Ebl.y, -c * Ebl.x);
        const double Tl = lane_of(T, l - k);
        const dc El = lane_of(Eh, l - k); // E_{l-n} on lane n
        if (k < l)
            Eb = dadd(Eb, drmul((double)(l - k) * Tl, accb));
        if (k >= 1 && k <= l)
            Tb += (double)k * ddot(accb, El);
    }
    if (k == 0)
        Tb += 0.25 * (-Eb.x * s0 + Eb.y * c0);
    if (act)
        grad_mom[off + k] = (float)((1.0 - b64) * Tb);
}

} // namespace

int launch_prepare_adjoint(const float *moments, const uint32_t *index, int W, int H, int m, float bias,
                           const float *grad_lagrange, float *grad_moments, hipStream_t stream)
{
    const int npix = W * H;
    if (m <= 64)
    {
        hipLaunchKernelGGL(prepare_adjoint_wave_kernel, dim3((npix + 3) / 4), dim3(256), 0, stream, moments, index,
                           npix, m, bias, grad_lagrange, grad_moments);
        return 0;
    }
    const int blocks = std::min(kAdjBlocks, (npix + kAdjThreads - 1) / kAdjThreads);
    const size_t bytes = (size_t)blocks * kAdjThreads * kArrays * m * sizeof(dc);
    void *scratch = nullptr;
    if (hipMallocAsync(&scratch, bytes, stream) != hipSuccess)
        return MRAY_ERR_OOM;
    hipLaunchKernelGGL(prepare_adjoint_kernel, dim3(blocks), dim3(kAdjThreads), 0, stream, moments, index, npix, m,
                       bias, grad_lagrange, grad_moments, static_cast<dc *>(scratch));
    return hipFreeAsync(scratch, stream) == hipSuccess ? 0 : MRAY_ERR_HIP;
}

} // namespace mray
