// Entropy coding of compact moment images (entropy_encode / entropy_decode,
// src/moment_image_coding.cpp:98-234): moment 0 of every non-empty pixel as
// a 16-bit word, and each moment l >= 1 as its own adaptive arithmetic code
// (Amir Said's FastAC floating-point coder with an adaptive multi-symbol
// model, vendored by the reference in src/entropy_coding/arithmetic_codec.cpp)
// of the words of every pixel with more than l moments, in pixel order.
//
// The streams are independent, so one workgroup of one wavefront codes one
// stream (blockIdx.x = l; block 0 writes the moment-0 words):
//   - the 64 lanes read 64 pixels at a time, quantize their moment l and
//     ballot which of them carry a symbol;
//   - the interval arithmetic is one serial chain per stream: every lane runs
//     it on the same values (the symbol's model bounds broadcast from its
//     lane with readlane), lane 0 writes the code bytes and propagates carries;
//   - the model lives in LDS as counts and their exclusive running sums
//     (d(k) = sum_{i<k} count_i / total is formed as (1 / total) * sum, the
//     product the reference stores); its periodic update is a wave-wide scan.
//     The encoder looks up a run of symbols at once: the model is constant
//     between updates, whose positions are known in advance;
//   - the decoder finds a symbol's interval by testing 64 candidates at once:
//     the interval ends base + length d(k) increase strictly with k (d grows
//     by at least 2^-17 per symbol and length exceeds 2^-16, so the gaps are
//     far above double rounding), so the largest k with
//     base + length d(k) <= value + 2^-48 is the symbol the reference's
//     bisection finds. A table over the model's cumulative counts (rebuilt at
//     each model update) gives a symbol at or below it from a float estimate
//     of the count at the value, and within 64 of it, so one test of 64
//     candidates usually settles it; the code bytes come from a 256-byte
//     window held in the wavefront's registers.
// All double arithmetic is unfused and correctly rounded (-ffp-contract=off),
// in the reference's order, so codes and decoded values are bit-identical.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "device_common.hpp"
#include "launch.hpp"

namespace mray
{

namespace
{

constexpr double kOut = 65536.0;                          // AC__OutputFactor
constexpr double kMinLength = 1.0 / 65536.0;              // AC__MinLength
constexpr double kLsb = 1.0 / 16777216.0 / 16777216.0;    // AC__LeastSignifBit
constexpr double kLeak = 2.0 * kLsb;                      // AC__Leakage
constexpr uint32_t kMaxCount = 1u << 17;                  // DM__MaxCount
constexpr int kMaxBits = 14;                              // 2^14 symbols at most

struct EntropyTable
{
    uint8_t bits[256];
};

// quantize / dequantize (src/moment_quantization.cpp:33-63), moment l's range
// [0, 1] for moment 0 and prediction-coded images, else [-1/pi, 1/pi]
__device__ __forceinline__ void ent_range(int l, int prediction_coded, float &lo, float &hi)
{
    const bool unit = (l == 0) || prediction_coded;
    lo = unit ? 0.0f : (float)(-1.0 / M_PI);
    hi = unit ? 1.0f : (float)(1.0 / M_PI);
}

__device__ __forceinline__ uint32_t ent_quantize(float v, int bits, float lo, float hi)
{
    v = (v > lo) ? v : lo;
    v = (v < hi) ? v : hi;
    if (bits == 32)
        return __float_as_uint(v);
    const uint32_t full = (1u << bits) - 1u;
    const float norm = (v - lo) / (hi - lo);
    return (uint32_t)(norm * (float)(int32_t)full) & full;
}

__device__ __forceinline__ float ent_dequantize(uint32_t q, int bits, float lo, float hi)
{
    if (bits == 32)
    {
        float v = __uint_as_float(q);
        v = (v > lo) ? v : lo;
        return (v < hi) ? v : hi;
    }
    const float norm = (float)q / (float)(int32_t)((1u << bits) - 1u);
    return norm * (hi - lo) + lo;
}

__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ uint32_t wave_exclusive_scan_u32(uint32_t v)
{
    const int lane = threadIdx.x;
    uint32_t incl = v;
    for (int o = 1; o < 64; o <<= 1)
    {
        const uint32_t t = __shfl_up(incl, o);
        if (lane >= o)
            incl += t;
    }
    return incl - v;
}

// The adaptive model (Adaptive_Data_Model, src/entropy_coding/arithmetic_codec.cpp:838-919):
// count[A] and cum[A] (cum[k] = count_0 + ... + count_{k-1}) in LDS; total,
// cycle and until are the same in every lane.
struct Model
{
    uint32_t *count, *cum;
    uint32_t A, total, cycle, until;
    double scale; // 1 / total

    // update(): halve the counts when total would pass 2^17, recompute the
    // running sums, grow the cycle by 5/4 up to 8 (A + 6)
    __device__ void update()
    {
        const int lane = threadIdx.x;
        const uint32_t per = (A + 63) / 64;
        const uint32_t k0 = min(A, lane * per), k1 = min(A, k0 + per);
        __syncthreads();
        total += cycle;
        if (total > kMaxCount)
        {
            uint32_t part = 0;
            for (uint32_t k = k0; k < k1; ++k)
            {
                const uint32_t c = (count[k] + 1) >> 1;
                count[k] = c;
                part += c;
            }
            total = wave_sum_u32(part);
        }
        uint32_t part = 0;
        for (uint32_t k = k0; k < k1; ++k)
            part += count[k];
        uint32_t run = wave_exclusive_scan_u32(part);
        for (uint32_t k = k0; k < k1; ++k)
        {
            cum[k] = run;
            run += count[k];
        }
        scale = 1.0 / (double)total;
        cycle = (5 * cycle) >> 2;
        const uint32_t cap = (A + 6) << 3;
        if (cycle > cap)
            cycle = cap;
        until = cycle;
        __syncthreads();
    }

    // set_alphabet / reset(): counts 1, one update, then a cycle of (A + 6) / 2
    __device__ void reset(uint32_t *lds, uint32_t alphabet)
    {
        A = alphabet;
        count = lds;
        cum = lds + A;
        for (uint32_t k = threadIdx.x; k < A; k += 64)
            count[k] = 1;
        total = 0;
        cycle = A;
        update();
        cycle = until = (A + 6) >> 1;
    }

    __device__ __forceinline__ double d(uint32_t k) const { return scale * (double)cum[k]; }
};

// The decoder's lookup table over the model's cumulative counts: tab[i] is
// the symbol whose count range [cum[k], cum[k+1]) holds i * 2^kTabShift. A
// count known to within +-2 then names a symbol at or below the one decoded,
// at most 2^kTabShift + 4 symbols below it (every count is at least 1), so 64
// candidates from there hold it.
// The model's total stays at most max(2^17, 9 A + 48) < 2^18 for alphabets
// of up to 2^14 symbols: an update adds at most 8 (A + 6), and one that passes
// 2^17 halves the counts (rounding up), to at most half the sum plus A / 2.
constexpr int kTabShift = 5;
constexpr uint32_t kTabEntries = (1u << 18) >> kTabShift;

__device__ void build_table(const Model &M, uint16_t *tab)
{
    const int lane = threadIdx.x;
    const uint32_t used = (M.total >> kTabShift) + 1;
    const uint32_t per = (used + 63) / 64;
    const uint32_t i0 = min(used, lane * per), i1 = min(used, i0 + per);
    if (i0 < i1)
    {
        // the symbol holding i0 << kTabShift (cum[0] = 0), then a walk
        const uint32_t c0 = i0 << kTabShift;
        uint32_t lo = 0, hi = M.A - 1;
        while (lo < hi)
        {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (M.cum[mid] <= c0)
                lo = mid;
            else
                hi = mid - 1;
        }
        uint32_t k = lo;
        for (uint32_t i = i0; i < i1; ++i)
        {
            const uint32_t c = i << kTabShift;
            while (k + 1 < M.A && M.cum[k + 1] <= c)
                ++k;
            tab[i] = (uint16_t)k;
        }
    }
    __syncthreads();
}

// pixel p's moment count
__device__ __forceinline__ uint32_t count_of(const uint32_t *__restrict__ index, uint32_t p, uint32_t &start)
{
    start = index[p];
    return index[p + 1] - start;
}

__device__ __forceinline__ uint64_t lanes_below()
{
    const int lane = threadIdx.x;
    return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}

// the first k set bits of mask
__device__ __forceinline__ uint64_t first_bits(uint64_t mask, uint32_t k)
{
    if (k >= (uint32_t)__popcll(mask))
        return mask;
    uint64_t out = 0;
    for (uint32_t i = 0; i < k; ++i)
    {
        const uint64_t low = mask & (~mask + 1);
        out |= low;
        mask ^= low;
    }
    return out;
}

// The encoder's state (Arithmetic_Codec in mode 1): interval, and the code
// bytes written by lane 0 into out[0, cap).
struct Encoder
{
    double base, length;
    uint8_t *out;
    size_t pos, cap;
    bool overflow;

    __device__ __forceinline__ void put(uint32_t b)
    {
        if (threadIdx.x == 0)
        {
            if (pos < cap)
                out[pos] = (uint8_t)b;
            else
                overflow = true;
        }
        ++pos;
    }

    // propagate_carry (src/entropy_coding/arithmetic_codec.cpp:76-84)
    __device__ __forceinline__ void carry()
    {
        base -= 1.0;
        if (threadIdx.x == 0)
        {
            size_t p = pos - 1;
            while (p < cap && out[p] == 0xFFu)
                out[p--] = 0;
            if (p < cap)
                out[p] = (uint8_t)(out[p] + 1);
        }
    }

    // encode(data, Adaptive_Data_Model) with the symbol's d(s) and d(s + 1)
    // (:445-475), the model update done by the caller
    __device__ __forceinline__ void encode(double lo, double hi, bool last)
    {
        const double y = last ? base + length : base + length * hi;
        base += length * lo;
        length = y - base;
        if (base >= 1.0)
            carry();
        if ((length -= kLeak) <= kMinLength)
        {
            // renorm_enc_interval (:87-97)
            do
            {
                base *= kOut;
                const uint32_t a = (uint32_t)base;
                put(a >> 8);
                put(a & 0xFFu);
                base -= (double)a;
            } while ((length *= kOut) <= kMinLength);
        }
    }

    // stop_encoder (:631-675)
    __device__ void finish()
    {
        const uint32_t a = (uint32_t)(kOut * base), b = (uint32_t)(kOut * (base + length));
        int last;
        if (b - a < 2)
        {
            base += 0.5 * kMinLength;
            last = 3;
        }
        else if ((b >> 8) - (a >> 8) < 2)
        {
            base += kMinLength;
            last = 2;
        }
        else
        {
            base += 256.0 * kMinLength;
            last = 1;
        }
        if (base >= 1.0)
            carry();
        do
        {
            base *= 256.0;
            const uint32_t c = (uint32_t)base;
            put(c);
            base -= (double)c;
        } while (--last);
    }
};

// block 0: moment 0 of every non-empty pixel, as u16 words in pixel order;
// blocks l >= 1: the arithmetic code of moment l into streams + (l-1) * cap
__global__ __launch_bounds__(64) void entropy_encode_kernel(const float *__restrict__ data,
                                                            const uint32_t *__restrict__ index, uint32_t npix,
                                                            int prediction_coded, EntropyTable t,
                                                            const uint64_t *__restrict__ stream_off,
                                                            uint8_t *__restrict__ streams,
                                                            uint16_t *__restrict__ zeroth,
                                                            uint64_t *__restrict__ sizes)
{
    extern __shared__ uint32_t lds_model[];
    const int l = blockIdx.x;
    const int lane = threadIdx.x;
    float lo, hi;
    ent_range(l, prediction_coded, lo, hi);
    const int bits = t.bits[l];
    if (l == 0)
    {
        uint32_t run = 0;
        for (uint32_t p0 = 0; p0 < npix; p0 += 64)
        {
            const uint32_t p = p0 + lane;
            uint32_t start = 0, n = 0;
            if (p < npix)
                n = count_of(index, p, start);
            const bool has = n > 0;
            const uint64_t mask = __ballot(has);
            if (has)
                zeroth[run + __popcll(mask & lanes_below())] = (uint16_t)ent_quantize(data[start], bits, lo, hi);
            run += __popcll(mask);
        }
        if (lane == 0)
            sizes[0] = run;
        return;
    }

    Model M;
    M.reset(lds_model, 1u << bits);
    Encoder E{0.0, 1.0, streams + stream_off[l - 1], 0, (size_t)(stream_off[l] - stream_off[l - 1]), false};
    for (uint32_t p0 = 0; p0 < npix; p0 += 64)
    {
        const uint32_t p = p0 + lane;
        uint32_t start = 0, n = 0, sym = 0;
        if (p < npix)
            n = count_of(index, p, start);
        const bool has = n > (uint32_t)l;
        if (has)
            sym = ent_quantize(data[start + l], bits, lo, hi);
        uint64_t mask = __ballot(has);
        while (mask)
        {
            // the symbols up to the next model update share one model
            const uint64_t seg = first_bits(mask, M.until);
            const bool mine = (seg >> lane) & 1ull;
            double dlo = 0.0, dhi = 0.0;
            const bool last = sym == M.A - 1;
            if (mine)
            {
                dlo = M.d(sym);
                dhi = last ? 0.0 : M.d(sym + 1);
            }
            const uint64_t lastmask = __ballot(mine && last);
            uint64_t s = seg;
            while (s)
            {
                const int i = __ffsll((unsigned long long)s) - 1;
                s &= s - 1;
                E.encode(readlane_f64(dlo, i), readlane_f64(dhi, i), (lastmask >> i) & 1ull);
            }
            if (mine)
                atomicAdd(&M.count[sym], 1u);
            M.until -= (uint32_t)__popcll(seg);
            mask &= ~seg;
            if (M.until == 0)
                M.update();
        }
    }
    E.finish();
    if (lane == 0)
        sizes[l] = E.overflow ? ~0ull : (uint64_t)E.pos;
}

// The decoder's state (Arithmetic_Codec in mode 2); bytes past the code's
// end read as zero. Lane i holds code bytes wbase + 4i .. wbase + 4i + 3 (a
// 256-byte window over the code, refilled when the read position leaves it).
struct Decoder
{
    double base, length, value;
    const uint8_t *in;
    size_t pos, n;
    uint32_t win;
    size_t wbase;

    __device__ void load_window(size_t b)
    {
        wbase = b;
        const size_t q = b + 4 * (size_t)threadIdx.x;
        uint32_t w = 0;
#pragma unroll
        for (int r = 0; r < 4; ++r)
            w |= (q + r < n ? (uint32_t)in[q + r] : 0u) << (8 * r);
        win = w;
    }

    // code byte i (the same i in every lane)
    __device__ __forceinline__ double byte(size_t i)
    {
        if (i - wbase >= 256)
            load_window(i & ~(size_t)3);
        const uint32_t off = (uint32_t)(i - wbase);
        const uint32_t w = (uint32_t)__builtin_amdgcn_readlane((int)win, (int)(off >> 2));
        return (double)((w >> (8 * (off & 3u))) & 0xFFu);
    }

    // start_decoder (:590-604)
    __device__ void start(const uint8_t *code, size_t size)
    {
        in = code;
        n = size;
        length = 1.0;
        base = value = 0.0;
        load_window(0);
        for (size_t k = 0; k < 6; ++k)
            value = 256.0 * value + kLsb * byte(k);
        pos = 6;
    }

    // decode(Adaptive_Data_Model) (:477-519) without the model's count update
    __device__ uint32_t decode(const Model &M, const uint16_t *tab)
    {
        const double sv = value + kLsb;
        // the cumulative count at sv, (sv - base) total / length, to within a
        // tenth of a count in float (total < 2^18; the hardware reciprocal is
        // good to about an ulp); the table's symbol for two counts below it
        // is at or below the symbol decoded
        const float est = (float)(sv - base) * ((float)M.total * __builtin_amdgcn_rcpf((float)length));
        const int c_lo = (est > 2.0f) ? (int)est - 2 : 0;
        uint32_t k0 = tab[min((uint32_t)c_lo >> kTabShift, M.total >> kTabShift)];
        // the largest k >= k0 with base + length d(k) <= sv, 64 candidates a
        // round (the qualifying ones are a prefix)
        uint32_t s = 0, j = 0;
        double z = 0.0;
        for (;;)
        {
            const uint32_t k = k0 + (uint32_t)threadIdx.x;
            // candidates past the alphabet read its last symbol and do not count
            z = base + length * M.d(min(k, M.A - 1u));
            const bool le = (k < M.A) && !(z > sv);
            const uint64_t b = __ballot(le);
            if (b == 0ull)
            {
                // k0 qualifies by construction; should it not, search from
                // symbol 0, and with none qualifying take symbol 0 (as the
                // reference's bisection does)
                if (k0 == 0)
                {
                    s = 0;
                    j = 0;
                    break;
                }
                k0 = 0;
                continue;
            }
            j = 63u - (uint32_t)__clzll((unsigned long long)b);
            if (j < 63u || k0 + 63u >= M.A - 1u)
            {
                s = k0 + j;
                break;
            }
            k0 += 63u;
        }
        // x and y: lane j's and lane j + 1's ends (the same expressions)
        const double x = readlane_f64(z, (int)j);
        double y;
        if (s + 1 >= M.A)
            y = base + length;
        else if (j < 63u)
            y = readlane_f64(z, (int)j + 1);
        else
            y = base + length * M.d(s + 1);
        base = x;
        length = y - x;
        if (base >= 1.0)
        {
            base -= 1.0;
            value -= 1.0;
        }
        if ((length -= kLeak) <= kMinLength)
        {
            // renorm_dec_interval (:101-112)
            do
            {
                base *= kOut;
                const uint32_t a = (uint32_t)base;
                base -= (double)a;
                value = (kOut * value - (double)a) + kLsb * (256.0 * byte(pos) + byte(pos + 1));
                pos += 2;
            } while ((length *= kOut) <= kMinLength);
        }
        return s;
    }
};

__global__ __launch_bounds__(64) void entropy_decode_kernel(const uint8_t *__restrict__ section,
                                                            const uint64_t *__restrict__ code_off,
                                                            const uint32_t *__restrict__ index, uint32_t npix,
                                                            int prediction_coded, EntropyTable t,
                                                            float *__restrict__ data)
{
    extern __shared__ uint32_t lds_model[];
    const int l = blockIdx.x;
    const int lane = threadIdx.x;
    float lo, hi;
    ent_range(l, prediction_coded, lo, hi);
    const int bits = t.bits[l];
    if (l == 0)
    {
        // code_off[0]: the moment-0 words; code_off[1] - code_off[0] = 2 nz
        const uint8_t *words = section + code_off[0];
        const uint32_t nz = (uint32_t)((code_off[1] - code_off[0]) / 2);
        uint32_t run = 0;
        for (uint32_t p0 = 0; p0 < npix; p0 += 64)
        {
            const uint32_t p = p0 + lane;
            uint32_t start = 0, n = 0;
            if (p < npix)
                n = count_of(index, p, start);
            const bool has = n > 0;
            const uint64_t mask = __ballot(has);
            if (has)
            {
                const uint32_t k = run + __popcll(mask & lanes_below());
                const uint32_t q = (k < nz) ? (uint32_t)words[2 * k] | ((uint32_t)words[2 * k + 1] << 8) : 0u;
                data[start] = ent_dequantize(q, bits, lo, hi);
            }
            run += __popcll(mask);
        }
        return;
    }

    Model M;
    M.reset(lds_model, 1u << bits);
    uint16_t *tab = reinterpret_cast<uint16_t *>(lds_model + 2 * M.A);
    build_table(M, tab);
    // stream l: [code_off[2 l], code_off[2 l + 1])
    Decoder D;
    D.start(section + code_off[2 * l], (size_t)(code_off[2 * l + 1] - code_off[2 * l]));
    for (uint32_t p0 = 0; p0 < npix; p0 += 64)
    {
        const uint32_t p = p0 + lane;
        uint32_t start = 0, n = 0;
        if (p < npix)
            n = count_of(index, p, start);
        const bool has = n > (uint32_t)l;
        uint64_t mask = __ballot(has);
        uint32_t mine = 0;
        while (mask)
        {
            const int i = __ffsll((unsigned long long)mask) - 1;
            mask &= mask - 1;
            const uint32_t s = D.decode(M, tab);
            if (lane == i)
                mine = s;
            if (lane == 0)
                M.count[s] += 1;
            if (--M.until == 0)
            {
                M.update();
                build_table(M, tab);
            }
        }
        if (has)
            data[start + l] = ent_dequantize(mine, bits, lo, hi);
    }
}

// moment-count histogram: hist[n] = pixels with n moments (n <= 255)
__global__ void count_histogram(const uint32_t *__restrict__ index, uint32_t npix, uint32_t *__restrict__ hist)
{
    __shared__ uint32_t h[256];
    for (int i = threadIdx.x; i < 256; i += blockDim.x)
        h[i] = 0;
    __syncthreads();
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x)
        atomicAdd(&h[min(index[p + 1] - index[p], 255u)], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < 256; i += blockDim.x)
        if (h[i])
            atomicAdd(&hist[i], h[i]);
}

__global__ void fill_u32(uint32_t *p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        p[i] = 0;
}

int check_widths(int m, const uint8_t *bits, EntropyTable &t, uint32_t &max_alphabet)
{
    if (m < 1 || m > 255 || !bits)
        return MRAY_ERR_ARG;
    std::memset(&t, 0, sizeof t);
    max_alphabet = 2;
    if (bits[0] < 1 || bits[0] > 32)
        return MRAY_ERR_ARG;
    t.bits[0] = bits[0];
    for (int l = 1; l < m; ++l)
    {
        // set_alphabet accepts 2 .. 2^14 symbols (:838-855)
        if (bits[l] < 1 || bits[l] > kMaxBits)
            return MRAY_ERR_ARG;
        t.bits[l] = bits[l];
        max_alphabet = std::max(max_alphabet, 1u << bits[l]);
    }
    return 0;
}

int set_lds(const void *kernel, size_t bytes)
{
    if (bytes <= 65536)
        return 0;
    return hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess
               ? 0
               : MRAY_ERR_UNSUPPORTED;
}

void put_u32(uint8_t *p, uint32_t v) { std::memcpy(p, &v, 4); }

} // namespace

int launch_entropy_encode(const float *data, const uint32_t *index, int W, int H, int m, int prediction_coded,
                          const uint8_t *bits, uint8_t *out, size_t capacity, size_t *out_bytes, hipStream_t stream)
{
    EntropyTable t;
    uint32_t amax = 0;
    int rc = check_widths(m, bits, t, amax);
    if (rc)
        return rc;
    const uint32_t npix = (uint32_t)((size_t)W * H);
    // symbols per stream from the histogram of moment counts
    uint32_t *hist = nullptr;
    if (hipMallocAsync((void **)&hist, 256 * sizeof(uint32_t), stream) != hipSuccess)
        return MRAY_ERR_OOM;
    hipLaunchKernelGGL(fill_u32, dim3(1), dim3(256), 0, stream, hist, 256);
    hipLaunchKernelGGL(count_histogram, dim3(std::max(1u, std::min(1024u, (npix + 255) / 256))), dim3(256), 0,
                       stream, index, npix, hist);
    uint32_t h_hist[256];
    if (hipMemcpyAsync(h_hist, hist, sizeof h_hist, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        rc = MRAY_ERR_HIP;
    if (hipFreeAsync(hist, stream) != hipSuccess)
        rc = MRAY_ERR_HIP;
    if (rc)
        return rc;
    // stream l - 1 gets 4 bytes per symbol + 16, as Arithmetic_Codec::set_buffer
    // is given 4 bytes per pixel (src/moment_image_coding.cpp:101-106)
    // nsym[l] = pixels with more than l moments; a pixel with more than m
    // moments is an error of the index
    for (int n = m + 1; n < 256; ++n)
        if (h_hist[n])
            return MRAY_ERR_ARG;
    std::vector<uint64_t> off((size_t)m, 0), nsym((size_t)m, 0);
    uint64_t acc = 0;
    for (int n = m; n >= 1; --n)
    {
        acc += h_hist[n];
        nsym[(size_t)(n - 1)] = acc;
    }
    for (int l = 1; l < m; ++l)
        off[(size_t)l] = off[(size_t)l - 1] + 4 * nsym[(size_t)l] + 16;
    const size_t streams_b = (off[(size_t)m - 1] + 255) & ~(uint64_t)255;
    const size_t zeroth_b = ((size_t)npix * 2 + 255) & ~(size_t)255;
    const size_t off_b = ((size_t)m * sizeof(uint64_t) + 255) & ~(size_t)255;
    const size_t sizes_b = off_b;
    uint8_t *buf = nullptr;
    if (hipMallocAsync((void **)&buf, streams_b + zeroth_b + off_b + sizes_b, stream) != hipSuccess)
        return MRAY_ERR_OOM;
    uint8_t *streams = buf;
    uint16_t *zeroth = reinterpret_cast<uint16_t *>(buf + streams_b);
    uint64_t *d_off = reinterpret_cast<uint64_t *>(buf + streams_b + zeroth_b);
    uint64_t *d_sizes = reinterpret_cast<uint64_t *>(buf + streams_b + zeroth_b + off_b);
    const size_t lds = (size_t)amax * 2 * sizeof(uint32_t);
    rc = set_lds(reinterpret_cast<const void *>(entropy_encode_kernel), lds);
    std::vector<uint64_t> sizes((size_t)m, 0);
    if (rc == 0 &&
        hipMemcpyAsync(d_off, off.data(), (size_t)m * sizeof(uint64_t), hipMemcpyHostToDevice, stream) != hipSuccess)
        rc = MRAY_ERR_HIP;
    if (rc == 0)
    {
        hipLaunchKernelGGL(entropy_encode_kernel, dim3(m), dim3(64), lds, stream, data, index, npix,
                           prediction_coded, t, d_off, streams, zeroth, d_sizes);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(sizes.data(), d_sizes, (size_t)m * sizeof(uint64_t), hipMemcpyDeviceToHost, stream) !=
                hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess)
            rc = MRAY_ERR_HIP;
    }
    size_t total = 0;
    if (rc == 0)
    {
        for (int l = 1; l < m; ++l)
            if (sizes[(size_t)l] == ~0ull)
                rc = MRAY_ERR_ARG; // a stream longer than 4 bytes per symbol
        total = 4 + 2 * (size_t)sizes[0];
        for (int l = 1; l < m && rc == 0; ++l)
            total += 4 + (size_t)sizes[(size_t)l];
        if (out_bytes)
            *out_bytes = total;
    }
    // the section: u32 nz, nz u16 words, then per stream u32 length + code
    if (rc == 0 && out && total > capacity)
        rc = MRAY_ERR_ARG;
    if (rc == 0 && out)
    {
        put_u32(out, (uint32_t)sizes[0]);
        size_t pos = 4;
        if (sizes[0] && hipMemcpyAsync(out + pos, zeroth, 2 * (size_t)sizes[0], hipMemcpyDeviceToHost, stream) !=
                            hipSuccess)
            rc = MRAY_ERR_HIP;
        pos += 2 * (size_t)sizes[0];
        for (int l = 1; l < m && rc == 0; ++l)
        {
            put_u32(out + pos, (uint32_t)sizes[(size_t)l]);
            pos += 4;
            if (sizes[(size_t)l] && hipMemcpyAsync(out + pos, streams + off[(size_t)l - 1], (size_t)sizes[(size_t)l],
                                                   hipMemcpyDeviceToHost, stream) != hipSuccess)
                rc = MRAY_ERR_HIP;
            pos += (size_t)sizes[(size_t)l];
        }
        if (rc == 0 && hipStreamSynchronize(stream) != hipSuccess)
            rc = MRAY_ERR_HIP;
    }
    if (hipFreeAsync(buf, stream) != hipSuccess && rc == 0)
        rc = MRAY_ERR_HIP;
    return rc;
}

int launch_entropy_decode(const uint8_t *in, size_t n_in, const uint32_t *index, int W, int H, int m,
                          int prediction_coded, const uint8_t *bits, float *data, hipStream_t stream)
{
    EntropyTable t;
    uint32_t amax = 0;
    int rc = check_widths(m, bits, t, amax);
    if (rc)
        return rc;
    if (!in || n_in < 4)
        return MRAY_ERR_ARG;
    const uint32_t npix = (uint32_t)((size_t)W * H);
    // the section's layout (src/moment_image_coding.cpp:42-71): offsets of the
    // moment-0 words and of every stream
    std::vector<uint64_t> off(2 * (size_t)m, 0);
    uint32_t nz;
    std::memcpy(&nz, in, 4);
    uint64_t pos = 4;
    off[0] = pos;
    pos += 2ull * nz;
    off[1] = pos;
    if (pos > n_in)
        return MRAY_ERR_ARG;
    for (int l = 1; l < m; ++l)
    {
        if (pos + 4 > n_in)
            return MRAY_ERR_ARG;
        uint32_t size;
        std::memcpy(&size, in + pos, 4);
        pos += 4;
        if (pos + size > n_in)
            return MRAY_ERR_ARG;
        off[2 * (size_t)l] = pos;
        off[2 * (size_t)l + 1] = pos + size;
        pos += size;
    }
    const size_t sec_b = ((size_t)pos + 255) & ~(size_t)255;
    uint8_t *buf = nullptr;
    if (hipMallocAsync((void **)&buf, sec_b + off.size() * sizeof(uint64_t), stream) != hipSuccess)
        return MRAY_ERR_OOM;
    uint64_t *d_off = reinterpret_cast<uint64_t *>(buf + sec_b);
    const size_t lds = (size_t)amax * 2 * sizeof(uint32_t) + kTabEntries * sizeof(uint16_t);
    rc = set_lds(reinterpret_cast<const void *>(entropy_decode_kernel), lds);
    if (rc == 0 && (hipMemcpyAsync(buf, in, (size_t)pos, hipMemcpyHostToDevice, stream) != hipSuccess ||
                    hipMemcpyAsync(d_off, off.data(), off.size() * sizeof(uint64_t), hipMemcpyHostToDevice, stream) !=
                        hipSuccess))
        rc = MRAY_ERR_HIP;
    if (rc == 0)
    {
        hipLaunchKernelGGL(entropy_decode_kernel, dim3(m), dim3(64), lds, stream, buf, d_off, index, npix,
                           prediction_coded, t, data);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(stream) != hipSuccess)
            rc = MRAY_ERR_HIP;
    }
    if (hipFreeAsync(buf, stream) != hipSuccess && rc == 0)
        rc = MRAY_ERR_HIP;
    return rc;
}

} // namespace mray
