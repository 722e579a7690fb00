// Quantization of compact moment images into the packed bit stream of the
// moment-image files, and back (moment_quantization::quantize_prediction_coding,
// quantize, dequantize_prediction_coding, dequantize:
// src/moment_quantization.cpp:33-163, 244-360).
//
// Moment l of a pixel is a word of bits[l] bits (1..32); the words of all
// pixels follow each other in row order, least significant bit first, and a
// last partial byte is not part of the stream. A pixel's words are
// contiguous, so the stream offset of pixel p is the exclusive prefix sum of
// cum[count(p)] with cum the running sum of the widths:
//   1. scan: per-tile sums of the pixel bit counts, a 64-bit scan of the tile
//      sums, then per tile the offsets (64-bit: a 2048^2 image of 100 moments
//      at 16 bits is 6.7 Gbit); offsets[W*H] is the total
//   2. encode: the compact data is cut into chunks of kChunk consecutive
//      elements (a map gives the pixel of each chunk's first element); one
//      thread per chunk quantizes its words into a 64-bit shift register and
//      stores each full 32-bit stream word. Words inside the chunk's bit
//      range are written with plain stores, the (at most two) words shared
//      with the neighbouring chunks with atomicOr into a cleared word
//      buffer; the stream is the buffer's first bytes
//   3. decode: the stream is copied into a zero-padded word buffer (bits past
//      its end read as zero), one thread per chunk reads its words from
//      64-bit windows
// The stream copies in and out of the word buffer are DMA copies
// (hipMemcpyAsync); no hipMemsetAsync is recorded.
// Values round like the reference's float code: clamp by comparisons,
// (v - lo) / (hi - lo) and q / (2^bits - 1) correctly rounded, the product
// and sum unfused (-ffp-contract=off), float to integer by truncation.
#include <hip/hip_runtime.h>

#include <cstring>

#include "device_common.hpp"
#include "launch.hpp"

namespace mray
{

namespace
{

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;
constexpr int kChunk = 16; // elements per thread of the encode and decode

struct QuantTable
{
    uint32_t cum[257]; // cum[n] = bits[0] + ... + bits[n-1]
    uint8_t bits[256];
    int32_t n_bits;    // entries of bits given; a pixel with more moments is an error
    int32_t prediction_coded;
};

struct QuantTableLds
{
    uint32_t cum[257];
    uint8_t bits[256];
};

__device__ __forceinline__ void load_table(const QuantTable &t, QuantTableLds &s)
{
    for (int i = threadIdx.x; i < 257; i += blockDim.x)
        s.cum[i] = t.cum[i];
    for (int i = threadIdx.x; i < 256; i += blockDim.x)
        s.bits[i] = t.bits[i];
    __syncthreads();
}

// the pixel's bit count; a count beyond the table raises the error flag
__device__ __forceinline__ uint32_t pixel_bits(const uint32_t *__restrict__ index, size_t p, const QuantTable &t,
                                               const QuantTableLds &s, int *err)
{
    const uint32_t n = index[p + 1] - index[p];
    if (n > (uint32_t)t.n_bits)
    {
        atomicOr(err, 1);
        return 0;
    }
    return s.cum[n];
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t *red)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t s = 0;
    for (int i = 0; i < kThreads / 64; ++i)
        s += red[i];
    return s;
}

// exclusive scan of one value per thread (Hillis-Steele in LDS); buf[kThreads-1]
// holds the total afterwards
template <typename T> __device__ __forceinline__ T block_exclusive_scan(T v, T *buf)
{
    const int t = threadIdx.x;
    buf[t] = v;
    __syncthreads();
    for (int o = 1; o < kThreads; o <<= 1)
    {
        const T add = (t >= o) ? buf[t - o] : T(0);
        __syncthreads();
        buf[t] += add;
        __syncthreads();
    }
    return buf[t] - v;
}

// pass 1: bits of each tile of kTile pixels (at most 4096 * 8160 < 2^32)
__global__ __launch_bounds__(kThreads) void quant_tile_sums(const uint32_t *__restrict__ index, size_t npix,
                                                            QuantTable t, uint32_t *__restrict__ sums, int *err)
{
    __shared__ QuantTableLds s;
    __shared__ uint32_t red[kThreads / 64];
    load_table(t, s);
    const size_t base = (size_t)blockIdx.x * kTile;
    uint32_t sum = 0;
    for (int i = 0; i < kItems; ++i)
    {
        const size_t p = base + (size_t)i * kThreads + threadIdx.x;
        if (p < npix)
            sum += pixel_bits(index, p, t, s, err);
    }
    sum = block_sum(sum, red);
    if (threadIdx.x == 0)
        sums[blockIdx.x] = sum;
}

// what the host reads back after the scan, in one copy
struct ScanSummary
{
    uint64_t total; // bits of the stream
    uint32_t nelem; // elements of the compact image, index[W*H]
    int32_t err;    // a pixel has more moments than the table
};

// pass 2: 64-bit exclusive scan of the tile sums (one workgroup, chunks of
// kThreads with a running carry); the total goes to offsets[npix] and, with
// the element count and the error flag, to the summary
__global__ __launch_bounds__(kThreads) void quant_tile_offsets(const uint32_t *__restrict__ sums, int ntiles,
                                                               uint64_t *__restrict__ tile_off,
                                                               uint64_t *__restrict__ offsets, size_t npix,
                                                               const uint32_t *__restrict__ index,
                                                               const int *__restrict__ err,
                                                               ScanSummary *__restrict__ summary)
{
    __shared__ uint64_t buf[kThreads];
    uint64_t carry = 0;
    for (int c = 0; c < ntiles; c += kThreads)
    {
        const int i = c + threadIdx.x;
        const uint64_t v = (i < ntiles) ? sums[i] : 0u;
        const uint64_t ex = block_exclusive_scan<uint64_t>(v, buf);
        const uint64_t chunk_total = buf[kThreads - 1];
        if (i < ntiles)
            tile_off[i] = carry + ex;
        carry += chunk_total;
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        offsets[npix] = carry;
        summary->total = carry;
        summary->nelem = index[npix];
        summary->err = *err;
    }
}

// pass 3: per-pixel offsets inside each tile plus the tile's offset
__global__ __launch_bounds__(kThreads) void quant_tile_apply(const uint32_t *__restrict__ index, size_t npix,
                                                             QuantTable t, const uint64_t *__restrict__ tile_off,
                                                             uint64_t *__restrict__ offsets)
{
    __shared__ QuantTableLds s;
    __shared__ uint32_t tile[kTile];
    __shared__ uint32_t buf[kThreads];
    __shared__ int dummy_err;
    load_table(t, s);
    const int th = threadIdx.x;
    const size_t base = (size_t)blockIdx.x * kTile;
    for (int i = 0; i < kItems; ++i)
    {
        const size_t p = base + (size_t)i * kThreads + th;
        tile[i * kThreads + th] = (p < npix) ? pixel_bits(index, p, t, s, &dummy_err) : 0u;
    }
    __syncthreads();
    uint32_t sum = 0;
    for (int i = 0; i < kItems; ++i)
        sum += tile[th * kItems + i];
    uint32_t run = block_exclusive_scan<uint32_t>(sum, buf);
    const uint64_t toff = tile_off[blockIdx.x];
    __syncthreads();
    for (int i = 0; i < kItems; ++i)
    {
        const uint32_t v = tile[th * kItems + i];
        tile[th * kItems + i] = run;
        run += v;
    }
    __syncthreads();
    for (int i = 0; i < kItems; ++i)
    {
        const size_t p = base + (size_t)i * kThreads + th;
        if (p < npix)
            offsets[p] = toff + tile[i * kThreads + th];
    }
}

__device__ __forceinline__ float clampr(float v, float lo, float hi)
{
    v = (v > lo) ? v : lo;
    return (v < hi) ? v : hi;
}

// moment l's range: [0, 1] for moment 0 and for prediction-coded images,
// else [-1/pi, 1/pi] as floats of the double quotients
__device__ __forceinline__ void quant_range(uint32_t l, int prediction_coded, float &lo, float &hi)
{
    const bool unit = (l == 0) || prediction_coded;
    lo = unit ? 0.0f : (float)(-1.0 / M_PI);
    hi = unit ? 1.0f : (float)(1.0 / M_PI);
}

// quantize (src/moment_quantization.cpp:33-50)
__device__ __forceinline__ uint32_t quantize_value(float v, int bits, float lo, float hi)
{
    v = clampr(v, lo, hi);
    if (bits == 32)
        return __float_as_uint(v);
    const uint32_t full = (1u << bits) - 1u;
    const float norm = (v - lo) / (hi - lo);
    return (uint32_t)(norm * (float)(int32_t)full) & full;
}

// dequantize (src/moment_quantization.cpp:52-63)
__device__ __forceinline__ float dequantize_value(uint32_t q, int bits, float lo, float hi)
{
    if (bits == 32)
        return clampr(__uint_as_float(q), lo, hi);
    const float norm = (float)q / (float)(int32_t)((1u << bits) - 1u);
    return norm * (hi - lo) + lo;
}

__global__ __launch_bounds__(kThreads) void fill_words(uint32_t *__restrict__ w, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        w[i] = 0u;
}

// the pixel holding the first element of each chunk of kChunk elements of
// the compact data (every chunk starts in exactly one non-empty pixel)
__global__ __launch_bounds__(kThreads) void quant_chunk_map(const uint32_t *__restrict__ index, size_t npix,
                                                            uint32_t *__restrict__ chunk_pixel)
{
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix)
        return;
    const uint64_t e0 = index[p], e1 = index[p + 1];
    for (uint64_t c = (e0 + kChunk - 1) / kChunk; c * kChunk < e1; ++c)
        chunk_pixel[c] = (uint32_t)p;
}

// walks the elements of a chunk through the pixels they belong to: pixel p,
// the end of its elements, and the element's moment l
struct ElementCursor
{
    size_t p;
    uint32_t pend, l;
    __device__ __forceinline__ void at(const uint32_t *__restrict__ index, const uint32_t *__restrict__ chunk_pixel,
                                       uint64_t c, uint64_t e)
    {
        p = chunk_pixel[c];
        pend = index[p + 1];
        l = (uint32_t)(e - index[p]);
    }
    // moves past the pixels that end at element k (empty ones included);
    // k < index[W*H] bounds the walk
    __device__ __forceinline__ void settle(const uint32_t *__restrict__ index, uint64_t k)
    {
        while (k == pend)
        {
            ++p;
            pend = index[p + 1];
            l = 0;
        }
    }
};

// one thread per chunk of kChunk consecutive elements (VEC: float4 loads of
// a 16-byte aligned image): the chunk's words go into a 64-bit shift
// register, each full 32-bit stream word inside the chunk's bit range is
// stored, the first and last (shared with the neighbouring chunks) are
// or-ed into the cleared buffer
template <bool VEC>
__global__ __launch_bounds__(kThreads) void quant_encode(const float *__restrict__ data,
                                                         const uint32_t *__restrict__ index, uint64_t nelem,
                                                         QuantTable t, const uint32_t *__restrict__ chunk_pixel,
                                                         const uint64_t *__restrict__ offsets,
                                                         uint32_t *__restrict__ words)
{
    __shared__ QuantTableLds s;
    load_table(t, s);
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t e = c * kChunk;
    if (e >= nelem)
        return;
    const int cnt = (int)((nelem - e < (uint64_t)kChunk) ? (nelem - e) : (uint64_t)kChunk);
    float v[kChunk];
    if (VEC && cnt == kChunk)
    {
        const float4 *d4 = reinterpret_cast<const float4 *>(data + e);
#pragma unroll
        for (int i = 0; i < kChunk / 4; ++i)
        {
            const float4 x = d4[i];
            v[4 * i] = x.x;
            v[4 * i + 1] = x.y;
            v[4 * i + 2] = x.z;
            v[4 * i + 3] = x.w;
        }
    }
    else
    {
#pragma unroll
        for (int i = 0; i < kChunk; ++i)
            v[i] = (i < cnt) ? data[e + i] : 0.0f;
    }
    ElementCursor cur;
    cur.at(index, chunk_pixel, c, e);
    const uint64_t first = offsets[cur.p] + s.cum[cur.l];
    uint64_t w = first >> 5;
    int fill = (int)(first & 31u);
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < kChunk; ++i)
    {
        if (i < cnt)
        {
            cur.settle(index, e + i);
            float lo, hi;
            quant_range(cur.l, t.prediction_coded, lo, hi);
            const int b = s.bits[cur.l];
            acc |= (uint64_t)quantize_value(v[i], b, lo, hi) << fill;
            fill += b;
            if (fill >= 32)
            {
                if ((w << 5) >= first)
                    words[w] = (uint32_t)acc;
                else
                    atomicOr(&words[w], (uint32_t)acc);
                ++w;
                acc >>= 32;
                fill -= 32;
            }
            ++cur.l;
        }
    }
    if (fill > 0)
        atomicOr(&words[w], (uint32_t)acc);
}

// one thread per chunk of kChunk elements: each word read from a 64-bit
// window of the stream (VEC: float4 stores into a 16-byte aligned image)
template <bool VEC>
__global__ __launch_bounds__(kThreads) void quant_decode(const uint32_t *__restrict__ words,
                                                         const uint32_t *__restrict__ index, uint64_t nelem,
                                                         QuantTable t, const uint32_t *__restrict__ chunk_pixel,
                                                         const uint64_t *__restrict__ offsets,
                                                         float *__restrict__ data)
{
    __shared__ QuantTableLds s;
    load_table(t, s);
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t e = c * kChunk;
    if (e >= nelem)
        return;
    const int cnt = (int)((nelem - e < (uint64_t)kChunk) ? (nelem - e) : (uint64_t)kChunk);
    ElementCursor cur;
    cur.at(index, chunk_pixel, c, e);
    uint64_t bit = offsets[cur.p] + s.cum[cur.l];
    float v[kChunk];
#pragma unroll
    for (int i = 0; i < kChunk; ++i)
    {
        v[i] = 0.0f;
        if (i < cnt)
        {
            cur.settle(index, e + i);
            const int b = s.bits[cur.l];
            const uint64_t w = bit >> 5;
            const uint64_t win = (uint64_t)words[w] | ((uint64_t)words[w + 1] << 32);
            const uint32_t q =
                (uint32_t)((win >> (bit & 31u)) & ((b == 32) ? 0xFFFFFFFFull : ((1ull << b) - 1ull)));
            float lo, hi;
            quant_range(cur.l, t.prediction_coded, lo, hi);
            v[i] = dequantize_value(q, b, lo, hi);
            bit += b;
            ++cur.l;
        }
    }
    if (VEC && cnt == kChunk)
    {
        float4 *d4 = reinterpret_cast<float4 *>(data + e);
#pragma unroll
        for (int i = 0; i < kChunk / 4; ++i)
            d4[i] = make_float4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
    }
    else
    {
#pragma unroll
        for (int i = 0; i < kChunk; ++i)
            if (i < cnt)
                data[e + i] = v[i];
    }
}

unsigned grid_for(size_t n)
{
    return (unsigned)std::min<size_t>((n + kThreads - 1) / kThreads, 65536);
}

int make_table(const uint8_t *bits, int n_bits, int prediction_coded, QuantTable &t)
{
    if (!bits || n_bits < 0 || n_bits > 256)
        return MRAY_ERR_ARG;
    std::memset(&t, 0, sizeof t);
    t.n_bits = n_bits;
    t.prediction_coded = prediction_coded ? 1 : 0;
    for (int l = 0; l < n_bits; ++l)
    {
        if (bits[l] < 1 || bits[l] > 32)
            return MRAY_ERR_ARG;
        t.bits[l] = bits[l];
        t.cum[l + 1] = t.cum[l] + bits[l];
    }
    return 0;
}

// the pixel offsets (offsets[npix] = total bits); the total, the element
// count and the error flag are read back through a pinned host buffer
int scan_offsets(const uint32_t *index, size_t npix, const QuantTable &t, uint64_t *offsets, uint64_t &total,
                 uint64_t &nelem, hipStream_t stream)
{
    // one pinned buffer per host thread, kept until the process ends
    static thread_local ScanSummary *host = nullptr;
    if (!host && hipHostMalloc((void **)&host, sizeof(ScanSummary)) != hipSuccess)
        return MRAY_ERR_OOM;
    const int ntiles = (int)((npix + kTile - 1) / kTile); // npix >= 1
    uint8_t *scratch = nullptr;
    const size_t sums_b = ((size_t)ntiles * sizeof(uint32_t) + 255) & ~(size_t)255;
    const size_t toff_b = ((size_t)ntiles * sizeof(uint64_t) + 255) & ~(size_t)255;
    if (hipMallocAsync((void **)&scratch, sums_b + toff_b + 256, stream) != hipSuccess)
        return MRAY_ERR_OOM;
    uint32_t *sums = reinterpret_cast<uint32_t *>(scratch);
    uint64_t *toff = reinterpret_cast<uint64_t *>(scratch + sums_b);
    int *err = reinterpret_cast<int *>(scratch + sums_b + toff_b);
    ScanSummary *summary = reinterpret_cast<ScanSummary *>(scratch + sums_b + toff_b + 64);
    hipLaunchKernelGGL(fill_words, dim3(1), dim3(kThreads), 0, stream, reinterpret_cast<uint32_t *>(err), (size_t)1);
    hipLaunchKernelGGL(quant_tile_sums, dim3(ntiles), dim3(kThreads), 0, stream, index, npix, t, sums, err);
    hipLaunchKernelGGL(quant_tile_offsets, dim3(1), dim3(kThreads), 0, stream, sums, ntiles, toff, offsets, npix,
                       index, err, summary);
    hipLaunchKernelGGL(quant_tile_apply, dim3(ntiles), dim3(kThreads), 0, stream, index, npix, t, toff, offsets);
    const bool copied =
        hipMemcpyAsync(host, summary, sizeof(ScanSummary), hipMemcpyDeviceToHost, stream) == hipSuccess;
    const bool freed = hipFreeAsync(scratch, stream) == hipSuccess;
    if (!copied || !freed || hipStreamSynchronize(stream) != hipSuccess)
        return MRAY_ERR_HIP;
    if (host->err)
        return MRAY_ERR_ARG;
    total = host->total;
    nelem = host->nelem;
    return 0;
}

// dense images with one value per byte group (moment_quantization::quantize
// / dequantize of a moment vector, src/moment_quantization.cpp:361-407):
// value i = moment i % m of its pixel, (bits + 7) / 8 little-endian bytes
__global__ __launch_bounds__(kThreads) void quant_dense_encode(const float *__restrict__ data, size_t n, int m,
                                                               int bits, uint8_t *__restrict__ out)
{
    const int nb = (bits + 7) / 8;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    {
        float lo, hi;
        quant_range((uint32_t)(i % (size_t)m), 0, lo, hi);
        const uint32_t q = quantize_value(data[i], bits, lo, hi);
        for (int b = 0; b < nb; ++b)
            out[i * nb + b] = (uint8_t)(q >> (8 * b));
    }
}

__global__ __launch_bounds__(kThreads) void quant_dense_decode(const uint8_t *__restrict__ in, size_t n, int m,
                                                               int bits, float *__restrict__ data)
{
    const int nb = (bits + 7) / 8;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    {
        uint32_t q = 0;
        for (int b = 0; b < nb; ++b)
            q |= (uint32_t)in[i * nb + b] << (8 * b);
        float lo, hi;
        quant_range((uint32_t)(i % (size_t)m), 0, lo, hi);
        data[i] = dequantize_value(q, bits, lo, hi);
    }
}

} // namespace

int launch_quantize_dense(const float *data, size_t n, int m, int bits, uint8_t *out, hipStream_t stream)
{
    if (m < 1 || bits < 1 || bits > 32)
        return MRAY_ERR_ARG;
    if (n)
        hipLaunchKernelGGL(quant_dense_encode, dim3(grid_for(n)), dim3(kThreads), 0, stream, data, n, m, bits, out);
    return hipGetLastError() == hipSuccess ? 0 : MRAY_ERR_HIP;
}

int launch_dequantize_dense(const uint8_t *in, size_t n, int m, int bits, float *data, hipStream_t stream)
{
    if (m < 1 || bits < 1 || bits > 32)
        return MRAY_ERR_ARG;
    if (n)
        hipLaunchKernelGGL(quant_dense_decode, dim3(grid_for(n)), dim3(kThreads), 0, stream, in, n, m, bits, data);
    return hipGetLastError() == hipSuccess ? 0 : MRAY_ERR_HIP;
}

int launch_quantize(const float *data, const uint32_t *index, int W, int H, int prediction_coded,
                    const uint8_t *bits, int n_bits, uint8_t *stream_out, size_t capacity, size_t *stream_bytes,
                    hipStream_t stream)
{
    QuantTable t;
    int rc = make_table(bits, n_bits, prediction_coded, t);
    if (rc)
        return rc;
    const size_t npix = (size_t)W * H;
    uint64_t *offsets = nullptr;
    if (hipMallocAsync((void **)&offsets, (npix + 1) * sizeof(uint64_t), stream) != hipSuccess)
        return MRAY_ERR_OOM;
    uint64_t total = 0, nelem = 0;
    rc = scan_offsets(index, npix, t, offsets, total, nelem, stream);
    const size_t nbytes = (size_t)(total / 8);
    if (stream_bytes)
        *stream_bytes = nbytes;
    if (rc == 0 && stream_out && nbytes > capacity)
        rc = MRAY_ERR_ARG;
    if (rc == 0 && stream_out && nbytes > 0)
    {
        const size_t nwords = (size_t)((total + 31) / 32) + 1;
        const uint64_t nchunks = (nelem + kChunk - 1) / kChunk;
        const size_t words_b = (nwords * sizeof(uint32_t) + 255) & ~(size_t)255;
        uint8_t *buf = nullptr;
        if (hipMallocAsync((void **)&buf, words_b + nchunks * sizeof(uint32_t), stream) != hipSuccess)
            rc = MRAY_ERR_OOM;
        else
        {
            uint32_t *words = reinterpret_cast<uint32_t *>(buf);
            uint32_t *chunk_pixel = reinterpret_cast<uint32_t *>(buf + words_b);
            const dim3 cgrid((unsigned)((nchunks + kThreads - 1) / kThreads));
            hipLaunchKernelGGL(fill_words, dim3(grid_for(nwords)), dim3(kThreads), 0, stream, words, nwords);
            hipLaunchKernelGGL(quant_chunk_map, dim3((unsigned)((npix + kThreads - 1) / kThreads)), dim3(kThreads),
                               0, stream, index, npix, chunk_pixel);
            if (((uintptr_t)data & 15u) == 0)
                hipLaunchKernelGGL(quant_encode<true>, cgrid, dim3(kThreads), 0, stream, data, index, nelem, t,
                                   chunk_pixel, offsets, words);
            else
                hipLaunchKernelGGL(quant_encode<false>, cgrid, dim3(kThreads), 0, stream, data, index, nelem, t,
                                   chunk_pixel, offsets, words);
            // byte i of the stream is byte i of the little-endian word buffer
            if (hipMemcpyAsync(stream_out, words, nbytes, hipMemcpyDeviceToDevice, stream) != hipSuccess)
                rc = MRAY_ERR_HIP;
            if (hipFreeAsync(buf, stream) != hipSuccess)
                rc = MRAY_ERR_HIP;
        }
    }
    if (hipFreeAsync(offsets, stream) != hipSuccess && rc == 0)
        rc = MRAY_ERR_HIP;
    return rc;
}

int launch_dequantize(const uint8_t *stream_in, size_t stream_bytes, const uint32_t *index, int W, int H,
                      int prediction_coded, const uint8_t *bits, int n_bits, float *data, hipStream_t stream)
{
    QuantTable t;
    int rc = make_table(bits, n_bits, prediction_coded, t);
    if (rc)
        return rc;
    const size_t npix = (size_t)W * H;
    uint64_t *offsets = nullptr;
    if (hipMallocAsync((void **)&offsets, (npix + 1) * sizeof(uint64_t), stream) != hipSuccess)
        return MRAY_ERR_OOM;
    uint64_t total = 0, nelem = 0;
    rc = scan_offsets(index, npix, t, offsets, total, nelem, stream);
    // the stream holds every whole byte of the words
    if (rc == 0 && stream_bytes < total / 8)
        rc = MRAY_ERR_ARG;
    if (rc == 0 && total > 0)
    {
        const size_t nwords = (size_t)((total + 31) / 32) + 1;
        const uint64_t nchunks = (nelem + kChunk - 1) / kChunk;
        const size_t words_b = (nwords * sizeof(uint32_t) + 255) & ~(size_t)255;
        uint8_t *buf = nullptr;
        if (hipMallocAsync((void **)&buf, words_b + nchunks * sizeof(uint32_t), stream) != hipSuccess)
            rc = MRAY_ERR_OOM;
        else
        {
            uint32_t *words = reinterpret_cast<uint32_t *>(buf);
            uint32_t *chunk_pixel = reinterpret_cast<uint32_t *>(buf + words_b);
            const dim3 cgrid((unsigned)((nchunks + kThreads - 1) / kThreads));
            const size_t used = std::min<size_t>(stream_bytes, (size_t)((total + 7) / 8));
            // the words from the one holding byte `used` on are cleared, then
            // the stream's bytes copied over their start
            const size_t tail = used / 4;
            hipLaunchKernelGGL(fill_words, dim3(grid_for(nwords - tail)), dim3(kThreads), 0, stream, words + tail,
                               nwords - tail);
            if (used > 0 && hipMemcpyAsync(words, stream_in, used, hipMemcpyDeviceToDevice, stream) != hipSuccess)
                rc = MRAY_ERR_HIP;
            hipLaunchKernelGGL(quant_chunk_map, dim3((unsigned)((npix + kThreads - 1) / kThreads)), dim3(kThreads),
                               0, stream, index, npix, chunk_pixel);
            if (((uintptr_t)data & 15u) == 0)
                hipLaunchKernelGGL(quant_decode<true>, cgrid, dim3(kThreads), 0, stream, words, index, nelem, t,
                                   chunk_pixel, offsets, data);
            else
                hipLaunchKernelGGL(quant_decode<false>, cgrid, dim3(kThreads), 0, stream, words, index, nelem, t,
                                   chunk_pixel, offsets, data);
            if (hipFreeAsync(buf, stream) != hipSuccess)
                rc = MRAY_ERR_HIP;
        }
    }
    if (hipFreeAsync(offsets, stream) != hipSuccess && rc == 0)
        rc = MRAY_ERR_HIP;
    return rc;
}

} // namespace mray
