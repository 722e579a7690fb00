// FP32 bounded-MESE building blocks with the reference's operation order
// (float_complex and its operators, src/mese/complex_algebra.h:34-158;
// src/mese/MESE_dynamic.h; src/moment_prediction_coding.h;
// src/moment_compaction.h). The kernels that must reproduce the reference CPU
// path bit for bit (PLAIN preparation, compaction counts, prediction coding)
// are built from these; compiled with -ffp-contract=off, every operation
// rounds as in the reference.
//
// Per-pixel arrays are columns of LDS (Col: element i of a thread's array at
// p[i * stride]), so consecutive lanes touch consecutive words.
//
// The reference's Levinson-style updates copy conj(P[j-k]) into a temporary
// before overwriting P[k]. Here P[k] and P[j-k] are updated together from
// their old values held in registers: the same operations on the same
// operands, without the temporary array.
#pragma once

#include "device_common.hpp"

namespace mray
{

struct cf
{
    float x, y;
};
__device__ __forceinline__ cf cmk(float x, float y) { return cf{x, y}; }
__device__ __forceinline__ cf cadd(cf a, cf b) { return cf{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cf csub(cf a, cf b) { return cf{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cf cmul(cf a, cf b) { return cf{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cf crmul(float s, cf a) { return cf{s * a.x, s * a.y}; }
__device__ __forceinline__ cf cconj(cf a) { return cf{a.x, -a.y}; }
__device__ __forceinline__ float cabs2(cf a) { return a.x * a.x + a.y * a.y; }
// operator/(float_complex, float_complex): multiply by the conjugate and 1/|d|^2
__device__ __forceinline__ cf cdiv(cf n, cf d)
{
    const float factor = 1.0f / cabs2(d);
    return cf{(n.x * d.x + n.y * d.y) * factor, (-n.x * d.y + n.y * d.x) * factor};
}
// operator/(float_complex, float)
__device__ __forceinline__ cf cdivr(cf n, float d)
{
    const float factor = 1.0f / d;
    return cf{n.x * factor, n.y * factor};
}

// cut::lerp(v, target, b)
__device__ __forceinline__ float lerp_bias(float v, float target, float b) { return (1 - b) * v + b * target; }

struct Col
{
    cf *p;
    int s;
    __device__ __forceinline__ cf &operator[](int i) const { return p[i * s]; }
};

// trigonometricToExponentialMoments, real input (src/mese/MESE_dynamic.h:35-54).
// t(i) returns trigonometric moment i.
template <class T>
__device__ __forceinline__ void trig_to_exp(int m, Col E, T t)
{
    const float ph = 3.14159265f * t(0) - 1.57079633f;
    E[0] = crmul(0.0795774715f, cmk(host_cosf(ph), host_sinf(ph)));
    for (int l = 1; l < m; ++l)
    {
        cf acc = cmk(0.0f, 0.0f);
        for (int j = 0; j < l; ++j)
            acc = cadd(acc, cmul(crmul((float)(l - j), E[j]), cmk(t(l - j), 0.0f)));
        // (2 pi / l) is evaluated in double by the reference (M_PI)
        E[l] = cmul(acc, cmk(0.0f, (float)((double)(2.0f * kPi) / (double)(float)l)));
    }
    E[0] = crmul(2.0f, E[0]);
}

// The shared update of the Levinson recursion and of prediction coding:
// P_k <- f (P_k - conj(P_{j-k}) u) for 0 < k < j, P_j <- f (-P_0 u), P_0 <- f P_0.
__device__ __forceinline__ void levinson_update(int j, Col P, cf u, float factor)
{
    const float p0 = P[0].x;
    for (int k = 1; 2 * k <= j; ++k)
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

// levinsonsAlgorithm (src/mese/MESE_dynamic.h:100-135): solves the Hermitian
// Toeplitz system with first column E for P, stopping the recursion (u = 0)
// once |u|^2 >= 1.
__device__ __forceinline__ void levinson(int m, Col P, Col E)
{
    P[0] = cmk(1.0f / E[0].x, 0.0f);
    bool invalid = false;
    for (int j = 1; j < m; ++j)
    {
        cf u = cmk(0.0f, 0.0f);
        if (!invalid)
            for (int k = 0; k < j; ++k)
                u = cadd(u, cmul(P[k], E[j - k]));
        float a2 = u.x * u.x + u.y * u.y;
        if (a2 >= 1.0f)
        {
            u = cmk(0.0f, 0.0f);
            a2 = 0;
            invalid = true;
        }
        levinson_update(j, P, u, 1.0f / (1.0f - a2));
    }
}

// computeAutocorrelation (src/mese/MESE_dynamic.h:137-147)
__device__ __forceinline__ void autocorrelation(int m, Col R, Col P)
{
    for (int k = 0; k < m; ++k)
    {
        cf acc = cmk(0.0f, 0.0f);
        for (int j = 0; j < m - k; ++j)
            acc = cadd(acc, cmul(P[j], cconj(P[j + k])));
        R[k] = acc;
    }
}

// computeImaginaryCorrelation (src/mese/MESE_dynamic.h:161-171) followed by
// the normalisation of prepareMESELagrange (:206-210): out[k] = nf * corr_k.
__device__ __forceinline__ void imaginary_correlation(int m, float *out, Col R, Col E, float nf)
{
    for (int k = 0; k < m; ++k)
    {
        float acc = 0.0f;
        for (int j = 0; j < m - k; ++j)
            acc += R[k + j].x * E[j].y + R[k + j].y * E[j].x;
        out[k] = nf * acc;
    }
}

// The tail of prepareMESELagrange / prepare_moments_from_pred_coding
// (src/mese/MESE_dynamic.h:200-210, src/moment_preparation.h:41-50): from the
// exponential moments E and the Levinson solution P to the multipliers.
__device__ __forceinline__ void lagrange_from_levinson(int m, float *out, Col E, Col P, Col R)
{
    for (int i = 0; i < m; ++i)
        P[i] = crmul(6.28318531f, P[i]);
    autocorrelation(m, R, P);
    E[0] = crmul(0.5f, E[0]);
    imaginary_correlation(m, out, R, E, 1.0f / (3.14159265f * P[0].x));
}

} // namespace mray
