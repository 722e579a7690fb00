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
// Up to 64 moments eight lanes hold a pixel, as in the forward kernel
// (prepare_adjoint_grp_kernel: registers per element slot, LDS for arrays read
// at other elements' offsets, DPP sums over the 8 lanes). Beyond that, one
// thread per pixel keeps its arrays in a global scratch laid out
// [array][element][thread] so that the lanes of a wavefront touch consecutive
// addresses; the grid is a fixed number of threads (16 wavefronts per CU)
// that stride over the pixels, which bounds the scratch to threads x 7
// arrays x m x 16 bytes.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.hpp"
#include "launch.hpp"
#include "mese_f64.hpp"

namespace mray
{

namespace
{

__device__ __forceinline__ dc dmk(double x, double y) { return dc{x, y}; }
__device__ __forceinline__ dc dadd(dc a, dc b) { return dc{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ dc dsub(dc a, dc b) { return dc{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ dc dmul(dc a, dc b) { return dc{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ dc dconj(dc a) { return dc{a.x, -a.y}; }
__device__ __forceinline__ dc drmul(double s, dc a) { return dc{s * a.x, s * a.y}; }
__device__ __forceinline__ double ddot(dc a, dc b) { return a.x * b.x + a.y * b.y; } // Re(conj(a) b)
// a + b c and a - b c by fused multiply-adds (the reverse passes only: their
// values are the gradient, not a replay of the reference's arithmetic)
__device__ __forceinline__ dc dfma(dc a, dc b, dc c)
{
    return dc{fma(b.x, c.x, fma(-b.y, c.y, a.x)), fma(b.x, c.y, fma(b.y, c.x, a.y))};
}
__device__ __forceinline__ dc dfms(dc a, dc b, dc c)
{
    return dc{fma(-b.x, c.x, fma(b.y, c.y, a.x)), fma(-b.x, c.y, fma(-b.y, c.x, a.y))};
}
__device__ __forceinline__ dc dsel(bool p, dc a, dc b) { return dc{p ? a.x : b.x, p ? a.y : b.y}; }

// for t in [lo, hi): acc(ld(t)) in order of t, the loads of four terms issued
// before their sums (one LDS round trip per four terms, not per term)
template <class L, class A> __device__ __forceinline__ void run4(int lo, int hi, L ld, A acc)
{
    int t = lo;
    for (; t + 4 <= hi; t += 4)
    {
        decltype(ld(t)) v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            v[i] = ld(t + i);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            acc(v[i]);
    }
    for (; t < hi; ++t)
        acc(ld(t));
}

struct dc2
{
    dc a, b;
};
struct dcd
{
    dc a;
    double d;
};
struct dcdc
{
    dc a, b;
    double d;
};

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
// Eight lanes per pixel (m <= 64), eight pixels per wavefront, as the forward
// prepare_mixed_kernel: lane g holds elements g, g + 8, ... (C slots) of the
// per-element registers (P, its adjoint Pb, Eb, Tb and the recorded
// reflection coefficients U); arrays that lanes read at other elements'
// offsets live in LDS per pixel (E, P and a work array X of 8C complex
// doubles; T overlays X at the ends). Sums over a pixel's elements reduce
// over its 8 lanes (grp8_sum), single elements come from their lane by a
// shuffle inside the group.
// ---------------------------------------------------------------------------

// the value v of lane src of this lane's group of 8
__device__ __forceinline__ double grp8_get(double v, int src) { return __shfl(v, (int)(threadIdx.x & ~7u) | src, 64); }
__device__ __forceinline__ dc grp8_get(dc v, int src) { return dc{grp8_get(v.x, src), grp8_get(v.y, src)}; }

template <int C> struct AdjRegs
{
    dc P[C], Pb[C], Eb[C], U[C];
    double Tb[C];
};

// Reverse Levinson steps j = min(8S - 1, nm - 1) down to max(1, 8(S - 1)),
// which touch the elements k <= j, i.e. slots s < S; then the octave below.
template <int S, int C>
__device__ __forceinline__ void levinson_reverse_octaves(AdjRegs<C> &R, dc *Pl, dc *Yl, const dc *E, int g,
                                                         int nm, int j_invalid)
{
    const int hi = min(8 * S - 1, nm - 1);
    for (int j = hi; j >= ((S == 1) ? 1 : 8 * (S - 1)); --j)
    {
        const dc u = grp8_get(R.U[S - 1], j & 7);
        const bool valid = j < j_invalid;
        // 1 / f = 1 - |u|^2: the forward step's factor and its reciprocal
        // (dividing by f would cost an FP64 division per use)
        const double finv = 1.0 - (u.x * u.x + u.y * u.y);
        const double f = recip_nr(finv);
        double fbs = 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s)
        {
            const double t = fma(R.Pb[s].x, R.P[s].x, fma(R.Pb[s].y, R.P[s].y, fbs));
            fbs = (g + 8 * s <= j) ? t : fbs;
        }
        const double fb = grp8_sum(fbs) * finv;
#pragma unroll
        for (int s = 0; s < S; ++s)
        {
            Pl[g + 8 * s] = R.P[s];
            Yl[g + 8 * s] = R.Pb[s];
        }
        wave_sync();
        const dc pjb = Yl[j];
        const double p0b = Yl[0].x;
        const double p0 = Pl[0].x * finv;
        // previous iterate: Pold_k = P_k + u conj(P_{j-k}) (1 <= k < j),
        // Pold_0 = p0, Pold_j = 0; its partner Poq_k = Pold_{j-k} comes from
        // the same two values, P_{j-k} + u conj(P_k). Two passes over the
        // slots, the second re-reading P_{j-k} and Pb_{j-k} from LDS, so that
        // no per-slot temporaries stay live across the step (they would take
        // the kernel past 256 VGPRs, to one wavefront per SIMD). The cases of
        // a slot (1 <= k < j, k = 0, k = j, k > j) are selected, not branched
        // on: a wavefront's lanes hold different k, so it runs every case.
        dc acc = dmk(0.0, 0.0);
#pragma unroll
        for (int s = 0; s < S; ++s)
        {
            const int k = g + 8 * s;
            const bool inner = k >= 1 && k < j;
            const dc pq = Pl[inner ? j - k : 0];
            acc = dsel(inner, dfma(acc, R.Pb[s], dfma(pq, u, dconj(R.P[s]))), acc);
        }
        acc = dc{grp8_sum(acc.x), grp8_sum(acc.y)};
        dc ub = drmul(2.0 * f * f * fb, u);
        ub = dsub(ub, drmul(f, acc));
        ub = dsub(ub, drmul(f * p0, pjb));
        const double pb0 = f * p0b - f * ddot(pjb, u);
#pragma unroll
        for (int s = 0; s < S; ++s)
        {
            const int k = g + 8 * s;
            const int r = (k >= 1 && k < j) ? j - k : 0;
            const dc pq = Pl[r], pbq = Yl[r];
            dc Pold, Poq;
            if (k >= 1 && k < j)
            {
                Pold = dfma(R.P[s], u, dconj(pq));
                Poq = dfma(pq, u, dconj(R.P[s]));
            }
            else if (k == 0)
            {
                Pold = dmk(p0, 0.0);
                Poq = dmk(0.0, 0.0);
            }
            else if (k == j)
            {
                Pold = dmk(0.0, 0.0);
                Poq = dmk(p0, 0.0);
            }
            else
            {
                Pold = R.P[s];
                Poq = dmk(0.0, 0.0);
            }
            dc Pbn = R.Pb[s];
            if (k >= 1 && k < j)
                Pbn = drmul(f, dfms(R.Pb[s], dconj(pbq), u));
            else if (k == 0)
                Pbn = dmk(pb0, 0.0);
            else if (k == j)
                Pbn = dmk(0.0, 0.0);
            if (valid)
            {
                // u = sum_{k<j} P_k E_{j-k}; P_{j-k} of the previous iterate is Poq
                if (k < j)
                    Pbn = dfma(Pbn, ub, dconj(E[j - k]));
                if (k >= 1 && k <= j)
                    R.Eb[s] = dfma(R.Eb[s], ub, dconj(Poq));
            }
            R.P[s] = Pold;
            R.Pb[s] = Pbn;
        }
        wave_sync();
    }
    if constexpr (S > 1)
        levinson_reverse_octaves<S - 1, C>(R, Pl, Yl, E, g, nm, j_invalid);
}

// Reverse Herglotz steps l = min(8S - 1, nm - 1) down to max(1, 8(S - 1)):
// E_l = i c_l sum_{j<l} (l - j) T_{l-j} E_j, c_l = 2 pi / l, touching the
// elements k <= l (slots s < S); then the octave below.
template <int S, int C>
__device__ __forceinline__ void herglotz_reverse_octaves(AdjRegs<C> &R, const double *T, const dc *E, int g, int nm)
{
    const int hi = min(8 * S - 1, nm - 1);
    for (int l = hi; l >= ((S == 1) ? 1 : 8 * (S - 1)); --l)
    {
        const dc Ebl = grp8_get(R.Eb[S - 1], l & 7);
        const double c = 2.0 * kPi * recip_nr((double)l);
        const dc accb = dmk(c * Ebl.y, -c * Ebl.x);
#pragma unroll
        for (int s = 0; s < S; ++s)
        {
            const int k = g + 8 * s;
            const int r = (k <= l) ? l - k : 0;
            const double Tl = T[r];
            dc El = E[r];
            // the recurrence ran on E_0 before its doubling
            El = dsel(r == 0, drmul(0.5, El), El);
            R.Eb[s] = dsel(k < l, dadd(R.Eb[s], drmul((double)(l - k) * Tl, accb)), R.Eb[s]);
            R.Tb[s] = (k >= 1 && k <= l) ? fma((double)k, ddot(accb, El), R.Tb[s]) : R.Tb[s];
        }
    }
    if constexpr (S > 1)
        herglotz_reverse_octaves<S - 1, C>(R, T, E, g, nm);
}

// Up to 5 slots the kernel is held to 168 VGPRs, three wavefronts per SIMD
// (at 5 slots at the cost of a 36-byte spill): 1024 x 1024 pixels at 32
// moments 6.81 -> 5.83 ms, at 40 moments 9.81 -> 9.55 ms; from 6 slots the
// spills cost more (50 moments: 15.7 -> 17.1 ms)
template <int C>
__global__ __launch_bounds__(64, (C <= 5) ? 3 : 1) void prepare_adjoint_grp_kernel(const float *__restrict__ mom,
                                                                 const uint32_t *__restrict__ index, int npix,
                                                                 int m, float bias,
                                                                 const float *__restrict__ grad_lag,
                                                                 float *__restrict__ grad_mom)
{
    constexpr int MP = 8 * C;
    extern __shared__ dc lds_adj[];
    const int lane = threadIdx.x;
    const int g = lane & 7;
    const int q = lane >> 3;
    const int pix = blockIdx.x * 8 + q;
    const bool valid = pix < npix;
    const int nm = valid ? (index ? min((int)(index[pix + 1] - index[pix]), m) : m) : 0;
    const size_t off = valid ? (index ? (size_t)index[pix] : (size_t)pix * m) : 0;
    // per-pixel stride 3 MP + 4 elements: consecutive pixels start 64 bytes
    // (16 banks) apart, so the 8-byte accesses of two pixels served together
    // do not share banks
    dc *E = lds_adj + q * (3 * MP + 4), *P = E + MP, *X = P + MP;
    double *T = reinterpret_cast<double *>(X);
    const double b64 = (double)bias;
    auto tval = [&](int e) {
        return (e < nm) ? (1.0 - b64) * (double)mom[off + e] + (e == 0 ? b64 * 0.5 : 0.0) : 0.0;
    };
    AdjRegs<C> R;

    // ---- forward, arrays kept ----
#pragma unroll
    for (int s = 0; s < C; ++s)
    {
        const int e = g + 8 * s;
        T[e] = tval(e);
        E[e] = dc{0.0, 0.0};
        P[e] = dc{0.0, 0.0};
        R.U[s] = dc{0.0, 0.0};
    }
    wave_sync();
    double s0, c0;
    sincos(kPi * T[0] - 0.5 * kPi, &s0, &c0);
    if (g == 0 && nm > 0)
        E[0] = dc{c0 * (0.25 / kPi), s0 * (0.25 / kPi)};
    wave_sync();
    herglotz_octaves<1, C>(T, E, g, nm);
    if (g == 0 && nm > 0)
        E[0] = dc{2.0 * E[0].x, 2.0 * E[0].y};
    wave_sync();
    const double e0x = (nm > 0) ? E[0].x : 1.0;
    if (g == 0 && nm > 0)
        P[0] = dc{1.0 / e0x, 0.0};
    wave_sync();
    int j_invalid = nm;
    levinson_octaves<1, C, true>(P, E, g, nm, j_invalid, R.U);
    // P <- 2 pi P (registers and LDS), autocorrelation R_kk = sum_j P_j conj(P_{j+kk})
#pragma unroll
    for (int s = 0; s < C; ++s)
    {
        const int e = g + 8 * s;
        R.P[s] = drmul(2.0 * kPi, P[e]);
    }
    wave_sync();
#pragma unroll
    for (int s = 0; s < C; ++s)
        P[g + 8 * s] = R.P[s];
    wave_sync();
    dc Rk[C];
#pragma unroll
    for (int s = 0; s < C; ++s)
    {
        const int kk = g + 8 * s;
        double rx = 0.0, ry = 0.0;
        if (kk < nm)
            run4(
                0, nm - kk, [&](int j) { return dc2{P[j], P[j + kk]}; },
                [&](const dc2 &v) {
                    rx = fma(v.a.x, v.b.x, fma(v.a.y, v.b.y, rx));
                    ry = fma(v.a.y, v.b.x, fma(-v.a.x, v.b.y, ry));
                });
        Rk[s] = dc{rx, ry};
    }
    wave_sync();
#pragma unroll
    for (int s = 0; s < C; ++s)
        X[g + 8 * s] = Rk[s]; // R
    wave_sync();
    const double p0s = (nm > 0) ? P[0].x : 1.0;
    const double nf = 1.0 / (kPi * p0s);
    wave_sync();
    // incoming gradients lb_kk, as doubles in P's space (P is in registers
    // until the reverse autocorrelation)
    double *Yd = reinterpret_cast<double *>(P);
#pragma unroll
    for (int s = 0; s < C; ++s)
    {
        const int e = g + 8 * s;
        Yd[e] = (e < nm) ? (double)grad_lag[off + e] : 0.0;
    }
    wave_sync();

    // ---- reverse: imaginary correlation and normalisation ----
    // out_kk = nf C_kk, C_kk = sum_j Im-part(R_{kk+j}, Eh_j) with Eh the
    // undoubled exponential moments (Eh_0 = E_0 / 2)
    double nfb_s = 0.0;
    dc Rb[C];
#pragma unroll
    for (int s = 0; s < C; ++s)
    {
        const int n = g + 8 * s;
        double ck = 0.0;
        dc rb = dc{0.0, 0.0}, eb = dc{0.0, 0.0};
        if (n < nm)
        {
            // over t < nm - n: C_n = sum_t (R_{n+t}, Eh_t) terms and
            // Eb_n = h_n nf sum_t lb_t (R_{n+t}.y, R_{n+t}.x), h_0 = 1/2;
            // t = 0 (Eh_0 = E_0 / 2) peeled
            {
                const dc r = X[n];
                const dc eh = drmul(0.5, E[0]);
                const double lbt = Yd[0];
                ck = fma(r.x, eh.y, r.y * eh.x);
                eb = dc{lbt * r.y, lbt * r.x};
            }
            run4(
                1, nm - n, [&](int t) { return dcdc{X[n + t], E[t], Yd[t]}; },
                [&](const dcdc &v) {
                    ck = fma(v.a.x, v.b.y, fma(v.a.y, v.b.x, ck));
                    eb.x = fma(v.d, v.a.y, eb.x);
                    eb.y = fma(v.d, v.a.x, eb.y);
                });
            const double h = (n == 0) ? 0.5 : 1.0;
            eb = drmul(h * nf, eb);
            // Rb_n = nf sum_{kk <= n} lb_kk (Eh_{n-kk}.y, Eh_{n-kk}.x), kk = n peeled
            {
                const dc eh = drmul(0.5, E[0]);
                const double lbk = Yd[n];
                rb = dc{lbk * eh.y, lbk * eh.x};
            }
            run4(
                0, n, [&](int kk) { return dcd{E[n - kk], Yd[kk]}; },
                [&](const dcd &v) {
                    rb.x = fma(v.d, v.a.y, rb.x);
                    rb.y = fma(v.d, v.a.x, rb.y);
                });
            rb = drmul(nf, rb);
            nfb_s = fma(Yd[n], ck, nfb_s);
        }
        Rb[s] = rb;
        R.Eb[s] = eb;
    }
    const double nfb = grp8_sum(nfb_s);
    wave_sync();
    // ---- reverse: autocorrelation, R_kk = sum_j P_j conj(P_{j+kk}) ----
#pragma unroll
    for (int s = 0; s < C; ++s)
    {
        X[g + 8 * s] = Rb[s];
        P[g + 8 * s] = R.P[s];
    }
    wave_sync();
#pragma unroll
    for (int s = 0; s < C; ++s)
    {
        const int j = g + 8 * s;
        dc pb = (j == 0) ? dmk(-nfb * nf / p0s, 0.0) : dmk(0.0, 0.0);
        if (j < nm)
        {
            // Rb_kk P_{j+kk} over kk < nm - j, conj(Rb_kk) P_{j-kk} over kk <= j
            run4(
                0, nm - j, [&](int kk) { return dc2{X[kk], P[j + kk]}; },
                [&](const dc2 &v) {
                    pb.x = fma(v.a.x, v.b.x, fma(-v.a.y, v.b.y, pb.x));
                    pb.y = fma(v.a.x, v.b.y, fma(v.a.y, v.b.x, pb.y));
                });
            run4(
                0, j + 1, [&](int kk) { return dc2{X[kk], P[j - kk]}; },
                [&](const dc2 &v) {
                    pb.x = fma(v.a.x, v.b.x, fma(v.a.y, v.b.y, pb.x));
                    pb.y = fma(v.a.x, v.b.y, fma(-v.a.y, v.b.x, pb.y));
                });
        }
        R.Pb[s] = drmul(2.0 * kPi, pb);
        R.P[s] = drmul(1.0 / (2.0 * kPi), R.P[s]);
    }
    wave_sync();

    // ---- reverse: Levinson, rebuilding each previous iterate ----
    levinson_reverse_octaves<C, C>(R, P, X, E, g, nm, j_invalid);
    // P_0 = 1 / E_0.x (E_0 doubled), then the recurrence's undoubled E_0
    if (g == 0)
    {
        R.Eb[0].x += -R.Pb[0].x / (e0x * e0x);
        R.Eb[0] = drmul(2.0, R.Eb[0]);
    }
    wave_sync();
    // ---- reverse: Herglotz recurrence ----
#pragma unroll
    for (int s = 0; s < C; ++s)
    {
        T[g + 8 * s] = tval(g + 8 * s);
        R.Tb[s] = 0.0;
    }
    wave_sync();
    herglotz_reverse_octaves<C, C>(R, T, E, g, nm);
    if (g == 0)
        R.Tb[0] += 0.25 * (-R.Eb[0].x * s0 + R.Eb[0].y * c0);
#pragma unroll
    for (int s = 0; s < C; ++s)
    {
        const int e = g + 8 * s;
        if (e < nm)
            grad_mom[off + e] = (float)((1.0 - b64) * R.Tb[s]);
    }
}

} // namespace

int launch_prepare_adjoint(const float *moments, const uint32_t *index, int W, int H, int m, float bias,
                           const float *grad_lagrange, float *grad_moments, hipStream_t stream)
{
    const int npix = W * H;
    if (m <= 64)
    {
        const int C = (m + 7) / 8;
        const size_t lds = (size_t)8 * (3 * 8 * C + 4) * sizeof(dc);
        const dim3 grid((npix + 7) / 8);
#define MRAY_PADJ_CASE(CC)                                                                                            \
    case CC:                                                                                                          \
        hipLaunchKernelGGL(prepare_adjoint_grp_kernel<CC>, grid, dim3(64), lds, stream, moments, index, npix, m,      \
                           bias, grad_lagrange, grad_moments);                                                        \
        return 0;
        switch (C)
        {
            MRAY_PADJ_CASE(1)
            MRAY_PADJ_CASE(2)
            MRAY_PADJ_CASE(3)
            MRAY_PADJ_CASE(4)
            MRAY_PADJ_CASE(5)
            MRAY_PADJ_CASE(6)
            MRAY_PADJ_CASE(7)
            MRAY_PADJ_CASE(8)
        default:
            break;
        }
#undef MRAY_PADJ_CASE
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
