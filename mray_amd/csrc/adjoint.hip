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

// lane value of a DPP move of a double (two 32-bit halves); lanes the move
// has no source for read 0 (bound_ctrl)
template <int CTRL, int ROW_MASK = 0xF> __device__ __forceinline__ double dpp_d(double v)
{
    int2 w = *reinterpret_cast<int2 *>(&v), o;
    o.x = __builtin_amdgcn_update_dpp(0, w.x, CTRL, ROW_MASK, 0xF, true);
    o.y = __builtin_amdgcn_update_dpp(0, w.y, CTRL, ROW_MASK, 0xF, true);
    return *reinterpret_cast<double *>(&o);
}

// sum over the 64 lanes, returned to every lane: prefix sums inside each row
// of 16 lanes by row_shr 1, 2, 4, 8, then row_bcast 15 / 31 carry the row
// totals upwards (DPP, executed by the VALU), and lane 63 holds the total
__device__ __forceinline__ double wsum(double v)
{
    v += dpp_d<0x111>(v); // row_shr:1
    v += dpp_d<0x112>(v); // row_shr:2
    v += dpp_d<0x114>(v); // row_shr:4
    v += dpp_d<0x118>(v); // row_shr:8
    v += dpp_d<0x142, 0xA>(v); // row_bcast:15 into rows 1 and 3
    v += dpp_d<0x143, 0xC>(v); // row_bcast:31 into rows 2 and 3
    int2 w = *reinterpret_cast<int2 *>(&v), o;
    o.x = __builtin_amdgcn_readlane(w.x, 63);
    o.y = __builtin_amdgcn_readlane(w.y, 63);
    return *reinterpret_cast<double *>(&o);
}
__device__ __forceinline__ dc wsum(dc v) { return dc{wsum(v.x), wsum(v.y)}; }
// value of a wave-uniform lane (v_readlane)
__device__ __forceinline__ double bcast(double v, int src)
{
    int2 w = *reinterpret_cast<int2 *>(&v), o;
    o.x = __builtin_amdgcn_readlane(w.x, src);
    o.y = __builtin_amdgcn_readlane(w.y, src);
    return *reinterpret_cast<double *>(&o);
}
__device__ __forceinline__ dc bcast(dc v, int src) { return dc{bcast(v.x, src), bcast(v.y, src)}; }

// Orders the LDS accesses of one wavefront (its lanes run in lockstep).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// Per-wavefront LDS copies of the arrays whose elements other lanes read at
// per-lane offsets (j - k, k + kk, ...): a store of lane k's element and
// masked reads of arbitrary elements.
struct WaveArrays
{
    double *T;
    dc *E, *Eh, *P, *R, *Q; // Q: scratch copy of the array being updated
};
constexpr int kWaveLdsBytes = 64 * 8 + 5 * 64 * 16;

__device__ __forceinline__ dc ld(const dc *a, int i, bool ok) { return ok ? a[i] : dc{0.0, 0.0}; }
__device__ __forceinline__ double ld(const double *a, int i, bool ok) { return ok ? a[i] : 0.0; }

__global__ __launch_bounds__(256) void prepare_adjoint_wave_kernel(const float *__restrict__ mom,
                                                                   const uint32_t *__restrict__ index, int npix,
                                                                   int m, float bias,
                                                                   const float *__restrict__ grad_lag,
                                                                   float *__restrict__ grad_mom)
{
    __shared__ __align__(16) unsigned char lds_raw[4 * kWaveLdsBytes];
    const int k = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int pix = blockIdx.x * (blockDim.x / 64) + w;
    if (pix >= npix) // uniform per wavefront
        return;
    const int nm = index ? min((int)(index[pix + 1] - index[pix]), m) : m;
    const size_t off = index ? (size_t)index[pix] : (size_t)pix * m;
    if (nm <= 0)
        return;
    unsigned char *base = lds_raw + w * kWaveLdsBytes;
    WaveArrays L;
    L.T = reinterpret_cast<double *>(base);
    L.E = reinterpret_cast<dc *>(base + 64 * 8);
    L.Eh = L.E + 64;
    L.P = L.Eh + 64;
    L.R = L.P + 64;
    L.Q = L.R + 64;

    const bool act = k < nm;
    const double b64 = (double)bias;
    const double T = act ? (1.0 - b64) * (double)mom[off + k] + (k == 0 ? b64 * 0.5 : 0.0) : 0.0;
    const dc zero = dmk(0.0, 0.0);
    L.T[k] = T;
    wave_lds_sync();

    // ---- forward ----
    const double ph0 = kPi * bcast(T, 0) - 0.5 * kPi;
    double s0, c0;
    sincos(ph0, &s0, &c0);
    dc E = (k == 0) ? dmk(c0 * (0.25 / kPi), s0 * (0.25 / kPi)) : zero;
    for (int l = 1; l < nm; ++l)
    {
        const double Tl = ld(L.T, l - k, k < l);
        const dc acc = wsum(drmul((double)(l - k) * Tl, E)); // E = 0 on lanes >= l
        if (k == l)
        {
            const double c = 2.0 * kPi / (double)l;
            E = dmk(-acc.y * c, acc.x * c);
        }
    }
    const dc Eh = E; // E_0 before doubling
    if (k == 0)
        E = drmul(2.0, E);
    L.E[k] = E;
    L.Eh[k] = Eh;
    wave_lds_sync();
    const double e0x = bcast(E.x, 0);
    dc P = (k == 0) ? dmk(1.0 / e0x, 0.0) : zero;
    dc U = zero;
    int j_invalid = nm;
    for (int j = 1; j < nm; ++j)
    {
        L.P[k] = P;
        wave_lds_sync();
        const dc Ejk = ld(L.E, j - k, k < j);
        dc u = (j < j_invalid) ? wsum(dmul(P, Ejk)) : zero;
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
        const double p0 = bcast(P.x, 0);
        const dc Pq = ld(L.P, j - k, k >= 1 && k < j);
        wave_lds_sync();
        if (k >= 1 && k < j)
            P = drmul(f, dsub(P, dmul(dconj(Pq), u)));
        else if (k == j)
            P = drmul(-f * p0, u);
        else if (k == 0)
            P = dmk(f * p0, 0.0);
    }
    P = drmul(2.0 * kPi, P);
    L.P[k] = P;
    wave_lds_sync();
    dc R = zero;
    for (int kk = 0; kk < nm; ++kk)
    {
        const dc Pn = ld(L.P, k + kk, k + kk < nm);
        const dc r = wsum(dmul(P, dconj(Pn)));
        if (k == kk)
            R = r;
    }
    L.R[k] = R;
    wave_lds_sync();
    const double p0s = bcast(P.x, 0);
    const double nf = 1.0 / (kPi * p0s);

    // ---- reverse: imaginary correlation and normalisation ----
    const double lb = act ? (double)grad_lag[off + k] : 0.0;
    double nfb = 0.0;
    dc Rb = zero, Eb = zero;
    for (int kk = 0; kk < nm; ++kk)
    {
        const bool in = k + kk < nm;
        const dc Rn = ld(L.R, k + kk, in); // R_{kk+j} on lane j
        const double Ck = wsum(Rn.x * Eh.y + Rn.y * Eh.x);
        const double lbk = bcast(lb, kk);
        nfb += lbk * Ck;
        const double cb = lbk * nf;
        // Rb_{kk+j} += cb (Eh_j.y, Eh_j.x): lane n takes Eh_{n-kk}
        const dc Ehj = ld(L.Eh, k - kk, k >= kk && k < nm);
        Rb = dadd(Rb, dmk(cb * Ehj.y, cb * Ehj.x));
        // Eh_j += cb (R_{kk+j}.y, R_{kk+j}.x); E_0 = 2 Eh_0
        const double h = (k == 0) ? 0.5 : 1.0;
        Eb = dadd(Eb, dmk(h * cb * Rn.y, h * cb * Rn.x));
    }
    dc Pb = (k == 0) ? dmk(-nfb * nf / p0s, 0.0) : zero;
    // R_kk = sum_j P_j conj(P_{j+kk}) (P is in L.P)
    for (int kk = 0; kk < nm; ++kk)
    {
        const dc rb = bcast(Rb, kk);
        const dc b = ld(L.P, k + kk, k + kk < nm);
        const dc a = ld(L.P, k - kk, k >= kk && k < nm);
        Pb = dadd(Pb, dmk(rb.x * b.x - rb.y * b.y, rb.x * b.y + rb.y * b.x));
        Pb = dadd(Pb, dmk(rb.x * a.x + rb.y * a.y, rb.x * a.y - rb.y * a.x));
    }
    P = drmul(1.0 / (2.0 * kPi), P);
    Pb = drmul(2.0 * kPi, Pb);

    // ---- reverse: Levinson, rebuilding each previous iterate ----
    dc *Pl = L.P, *Ql = L.Q, *Pbl = L.R; // L.R is free from here on
    for (int j = nm - 1; j >= 1; --j)
    {
        const dc u = bcast(U, j);
        const bool valid = j < j_invalid;
        const double f = 1.0 / (1.0 - (u.x * u.x + u.y * u.y));
        const double fb = wsum((k <= j) ? ddot(Pb, P) : 0.0) / f;
        const dc pjb = bcast(Pb, j);
        const double p0b = bcast(Pb.x, 0);
        const double p0 = bcast(P.x, 0) / f;
        Pl[k] = P;
        Pbl[k] = Pb;
        wave_lds_sync();
        const dc Pq = ld(Pl, j - k, k >= 1 && k < j);
        const dc Pbq = ld(Pbl, j - k, k >= 1 && k < j);
        dc Pold = P;
        if (k >= 1 && k < j)
            Pold = dadd(P, dmul(u, dconj(Pq)));
        else if (k == 0)
            Pold = dmk(p0, 0.0);
        else if (k == j)
            Pold = zero;
        Ql[k] = Pold;
        wave_lds_sync();
        const dc Poq = ld(Ql, j - k, k <= j); // P_{j-k} of the previous iterate
        wave_lds_sync();
        dc ub = drmul(2.0 * f * f * fb, u);
        ub = dsub(ub, drmul(f, wsum((k >= 1 && k < j) ? dmul(Pb, Poq) : zero)));
        ub = dsub(ub, drmul(f * p0, pjb));
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
            const dc Ejk = ld(L.E, j - k, k < j);
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
        const dc Ebl = bcast(Eb, l);
        const double c = 2.0 * kPi / (double)l;
        const dc accb = dmk(c * Ebl.y, -c * Ebl.x);
        const double Tl = ld(L.T, l - k, k < l);
        const dc El = ld(L.Eh, l - k, k >= 1 && k <= l); // E_{l-n} on lane n
        if (k < l)
            Eb = dadd(Eb, drmul((double)(l - k) * Tl, accb));
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
