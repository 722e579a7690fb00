/*
 * mray_hip: C ABI of the MI355X (gfx950) moment-image library.
 *
 * Drop-in boundary for the reference's dispatch layer: where the reference's
 * CPU/CUDA wrappers (src/moment_raymarcher.cpp:168-183,329-353 ->
 * src/moment_raymarcher.cu:210-261,386-453) run generate / prepare /
 * reconstruct, a host binding calls these functions instead.
 *
 * Conventions
 *   - Every function returns an mray_status (0 = MRAY_OK); nothing throws
 *     across the ABI. mray_hip_last_error() returns a thread-local message.
 *   - Array arguments named d_* are DEVICE pointers (hipMalloc'ed, torch CUDA
 *     tensors, ...). Host pointers are only taken where stated.
 *   - stream is a hipStream_t (NULL = the default stream). Calls are
 *     asynchronous with respect to the host unless stated; parameters are
 *     passed by value to the kernels, so calls on different streams do not
 *     share state (the reference's __constant__ camera/params are gone).
 *   - Moment images: dense layout is pixel-major with moments contiguous,
 *     data[(y*W + x)*m + j]; the compact layout uses index[W*H+1] offsets
 *     (MomentImageHost, src/moment_image.h:33-98). Bounds are one uint32 per
 *     pixel, lower | upper << 16 (DensityBound, src/common.h:135-150).
 */
#ifndef MRAY_HIP_H
#define MRAY_HIP_H

#include "mray_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mray_volume_s *mray_volume;
typedef void *mray_stream;

/* Selects the device and reads its properties. Replaces init_device
 * (src/cut/cuda_common.cpp:55-59). Sets the release threshold of the
 * device's default memory pool, which the library's stream-ordered scratch
 * (hipMallocAsync) comes from, to its maximum: freed scratch stays in the
 * pool for the next call instead of being returned at each synchronisation. */
int mray_hip_init(int device);

/* Thread-local description of the last error of this thread. */
const char *mray_hip_last_error(void);

/* ---- volumes (RegularGrid, src/volume.h:31-130) ------------------------ */

/* Uploads host voxels (F32 or U8, layout [nz][ny][nx]) into a device-resident
 * volume owned by the library. Replaces RegularGridDevice's per-call upload
 * into a 3-D texture (src/volume_device.cuh:57-96); the volume stays resident
 * across generate calls. */
int mray_hip_volume_create(const void *host_voxels, int dtype, int nx, int ny, int nz, const float dom_min[3],
                           const float dom_max[3], mray_volume *out);

/* Wraps existing device voxels without copying; the caller keeps them alive
 * until mray_hip_volume_destroy. */
int mray_hip_volume_wrap(const void *d_voxels, int dtype, int nx, int ny, int nz, const float dom_min[3],
                         const float dom_max[3], mray_volume *out);

/* A particle set marched directly (generate(cam, const Particles&, ...),
 * src/moment_raymarcher.h:48-49): n particles at device positions d_x, d_y,
 * d_z with values d_v, smoothing length h, ray-marching domain [dom_min,
 * dom_max] (Particles::domain). The particles are copied into the reference's
 * uniform grid of cell size h, sorted stably by 4^3-tiled cell index
 * (Particles::create_grid, src/volume.cpp:55-81); every sample of
 * mray_hip_generate on this handle is Particles::sample (src/volume.cpp:83-94):
 * the cubic-spline weighted mean of the values of particles closer than h,
 * 0 where there are none. Synchronises the device once (the grid is sized
 * from the particles' bounding box). mray_hip_reference and
 * mray_hip_generate_backward do not take particle sets. */
int mray_hip_volume_from_particles(const float *d_x, const float *d_y, const float *d_z, const float *d_v, uint32_t n,
                                   float h, const float dom_min[3], const float dom_max[3], mray_volume *out);

int mray_hip_volume_destroy(mray_volume vol);

/* Device pointer of the volume's voxels (read-only view); for a particle set
 * the cell-sorted records (x, y, z, value) as float4. */
const void *mray_hip_volume_data(mray_volume vol);

/* ---- generate (src/moment_raymarcher.h:48-49) -------------------------- */

/* Ray marches every pixel and writes dense bounded trigonometric moments.
 * Replaces generate_moments_kernel (src/moment_raymarcher.cu:163-185) /
 * generate_ray (src/raymarch_generate.h:69-148).
 *   d_moments    W*H*m floats (dense)
 *   d_bounds     W*H uint32, written when p->transform_to_bounds (else may be NULL)
 *   d_samples    optional W*H int32: samples taken per ray (N_g), NULL to skip
 *   d_ray_range  optional W*H x 2 floats: per ray the (lo, hi) the moments
 *                were normalised with (before the 16-bit packing of
 *                d_bounds), for mray_hip_generate_backward; NULL to skip
 * form: MRAY_FORM_PLAIN or MRAY_FORM_MIXED (m up to 255 in both; MIXED splits
 * more than 64 moments into chunks of at most 56, each marched by its own
 * workgroups). Compaction and prediction coding
 * are separate calls (mray_hip_compute_num_moments, mray_hip_moment_index,
 * mray_hip_compact, mray_hip_prediction_encode). */
int mray_hip_generate(mray_volume vol, const mray_camera *cam, const mray_gen_params *p, int form, int W, int H,
                      int m, float *d_moments, uint32_t *d_bounds, int32_t *d_samples, float *d_ray_range,
                      mray_stream stream);

/* ---- reference direct volume rendering (src/moment_raymarcher.cpp:74-81) -- */

/* The reference renderer: the volume marched directly through the transfer
 * function (no moments), the ground truth of the moment-based images.
 * Replaces generate_reference(cam, tf, RegularGrid, img, ReferenceParameters)
 * (src/moment_raymarcher.cpp:74-81) -> reference_host
 * (src/moment_raymarcher.cpp:47-61) / generate_reference_device ->
 * reference_ray (src/raymarch_reference.h:93-162), without single scattering
 * (p->single_scattering with p->medium_scattering > 0 returns
 * MRAY_ERR_UNSUPPORTED).
 *   d_tf      256 float4 entries, as for mray_hip_reconstruct
 *   d_rgba    W*H float4, un-clamped float RGBA (clamp+truncate gives the
 *             reference's RGBA8)
 *   d_steps   optional W*H int32: compositing steps per ray */
int mray_hip_reference(mray_volume vol, const mray_camera *cam, const float *d_tf, const mray_ref_params *p, int W,
                       int H, float *d_rgba, int32_t *d_steps, mray_stream stream);

/* ---- prepare (src/moment_image.cpp:154-183, src/moment_image_device.cu:279-335) */

/* Bias lerp + bounded-MESE Lagrange multipliers for every pixel (dense or
 * compact via d_index, NULL = dense with m per pixel). Replaces
 * prepareMomentsKernel / prepare_moments_device. Output has the layout of the
 * input. PLAIN: FP32 (the reference arithmetic); MIXED: FP64 Herglotz
 * transform and Levinson recursion. */
int mray_hip_prepare(const float *d_moments, const uint32_t *d_index, int W, int H, int m, float bias, int form,
                     float *d_lagrange, mray_stream stream);

/* ---- reconstruct (src/moment_raymarcher.h:61-64) ----------------------- */

/* Reconstruct-and-composite march from prepared multipliers (output of
 * mray_hip_prepare, or trigonometric moments when
 * p->use_truncated_fourier). Replaces reconstruct_moments_kernel
 * (src/moment_raymarcher.cu:339-384) / reconstruct_ray
 * (src/raymarch_reconstruct.h:137-200).
 *   d_tf      256 float4 entries (the row-constant table of
 *             TransferFunction, i.e. rgb*a*s, a*s), device memory
 *   d_bounds  NULL -> bounds (0, 1)
 *   d_rgba    W*H float4, un-clamped float RGBA (the float-RGBA overload the
 *             reference lacks; clamp+truncate gives ColorImageHost's RGBA8)
 *   d_steps   optional W*H int32: compositing steps per ray (N_r) */
int mray_hip_reconstruct(const mray_camera *cam, const float dom_min[3], const float dom_max[3],
                         const float *d_lagrange, const uint32_t *d_index, const uint32_t *d_bounds, int W, int H,
                         int m, const float *d_tf, const mray_rec_params *p, int form, float *d_rgba,
                         int32_t *d_steps, mray_stream stream);

/* ---- MIXED with FP64 multipliers -------------------------------------- */

/* mray_hip_prepare (MIXED) writing the multipliers in FP64 (d_lagrange:
 * doubles, the layout of d_moments), and mray_hip_reconstruct (MIXED)
 * reading them: the pair the reference runs inside reconstruct
 * (src/moment_raymarcher.cpp:329-353 -> prepare_moments_device +
 * reconstruct_moments_kernel, src/moment_raymarcher.cu:386-453), without the
 * float rounding of the multipliers in between. At 50 moments the
 * multipliers reach |lambda| ~ 1e3 and their float rounding alone moves the
 * image by ~1e-4 RMSE; these keep it to the float storage of the moments. */
int mray_hip_prepare_f64(const float *d_moments, const uint32_t *d_index, int W, int H, int m, float bias,
                         double *d_lagrange, mray_stream stream);
int mray_hip_reconstruct_f64(const mray_camera *cam, const float dom_min[3], const float dom_max[3],
                             const double *d_lagrange, const uint32_t *d_index, const uint32_t *d_bounds, int W,
                             int H, int m, const float *d_tf, const mray_rec_params *p, float *d_rgba,
                             int32_t *d_steps, mray_stream stream);

/* ---- adjoint (no counterpart in the reference, which has no gradients) ---- */

/* Reverse mode of mray_hip_reconstruct (MIXED arithmetic) for
 * L = sum(d_grad_rgba * rgba): the gradient with respect to the multipliers
 * (d_grad_lagrange, layout of d_lagrange, overwritten; NULL to skip) and to
 * the 256 x 4 table (d_grad_tf, ADDED to, so the caller zeroes it; NULL to
 * skip). d_rgba is the forward output of mray_hip_reconstruct on the same
 * inputs. The march is recomputed per ray (no per-sample storage: memory
 * O(pixels x moments)); bounds are constants, the transfer-function lookup
 * is differentiated as the reference evaluates it (entries from the front
 * sample, weight from the back sample) and the density through the
 * fast_atan polynomial. Above 64 moments two (up to 128) or four lanes work on
 * a ray, 64 multipliers each. */
int mray_hip_reconstruct_backward(const mray_camera *cam, const float dom_min[3], const float dom_max[3],
                                  const float *d_lagrange, const uint32_t *d_index, const uint32_t *d_bounds, int W,
                                  int H, int m, const float *d_tf, const mray_rec_params *p, const float *d_rgba,
                                  const float *d_grad_rgba, float *d_grad_lagrange, float *d_grad_tf,
                                  mray_stream stream);

/* Reverse mode of mray_hip_generate for dL/dmoments (dense W*H*m):
 * dL/dvoxels ADDED to d_grad_volume (nx*ny*nz floats, the caller zeroes it).
 * Differentiates the exact integral of the reference's piecewise-linear
 * segments, map_density (where it does not clamp) and the trilinear weights;
 * the per-ray bounds are constants. d_ray_range: the forward's per-ray
 * normalisation (its d_ray_range output), or NULL to recompute it with an
 * extra pass over the samples. F32 volumes; above 64 moments the moments are
 * split into chunks of at most 56 whose passes add their shares; float
 * atomics, so the summation order (and the last bits) vary between runs. */
int mray_hip_generate_backward(mray_volume vol, const mray_camera *cam, const mray_gen_params *p, int W, int H, int m,
                               const float *d_grad_moments, const float *d_ray_range, float *d_grad_volume,
                               mray_stream stream);

/* Reverse mode of mray_hip_prepare (MIXED mathematics, FP64): given
 * dL/dlagrange, writes dL/dmoments (layout of d_moments) for the bias lerp,
 * Herglotz transform, Levinson recursion (including its |u|^2 >= 1 guard)
 * and correlations. Up to 64 moments it needs no memory beyond its
 * arguments; above 64 it uses stream-ordered scratch (hipMallocAsync) of
 * 112 x m bytes per pixel, at most 1,048,576 pixels' worth; larger images
 * are processed in turns by the same threads. */
int mray_hip_prepare_backward(const float *d_moments, const uint32_t *d_index, int W, int H, int m, float bias,
                              const float *d_grad_lagrange, float *d_grad_moments, mray_stream stream);

/* ---- compaction (src/moment_compaction.h, src/moment_image.cpp:69-94) ---- */

/* Number of moments to keep per pixel (compute_num_moments,
 * src/moment_compaction.h:147-174, and its CUDA kernel in
 * src/moment_image_device.cu): binary search on the relative RMSE of the
 * moments that the MESE predicts from the first n, against error_threshold
 * (0 keeps all m). Pixels whose bounds are equal (or, without bounds, whose
 * m_0 is 0) get 0. Bit-identical to the reference CPU path.
 *   d_moments  dense W*H*m (bound-transformed moments as generated)
 *   d_bounds   W*H or NULL;  d_counts  W*H uint32 out */
int mray_hip_compute_num_moments(const float *d_moments, const uint32_t *d_bounds, int W, int H, int m,
                                 float error_threshold, uint32_t *d_counts, mray_stream stream);

/* Exclusive scan of the counts into index[W*H+1] (index[W*H] = total moments):
 * the thrust::exclusive_scan of MomentImageDevice::compact
 * (src/moment_image_device.cu:83-110). d_index may be d_counts (W*H+1
 * entries). Uses stream-ordered scratch (hipMallocAsync). */
int mray_hip_moment_index(const uint32_t *d_counts, int W, int H, uint32_t *d_index, mray_stream stream);

/* Clamps the moment counts of non-empty pixels (count > 0) to [min_moments,
 * max_moments] in place; empty pixels keep 0. No counterpart in the
 * reference, whose compute_num_moments has no lower limit: the per-pixel
 * range of a workload (8 to 100 moments for the large noise-and-blobs
 * volume), applied between mray_hip_compute_num_moments and
 * mray_hip_moment_index. The dense moments must hold max_moments per pixel. */
int mray_hip_clamp_num_moments(uint32_t *d_counts, int W, int H, int min_moments, int max_moments,
                               mray_stream stream);

/* mray_hip_moment_index without allocating anything. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc). Same
 * result as mray_hip_moment_index. Its scratch (the scan's tile sums, 4
 * bytes per 4096 pixels) is a 64 KiB device array per GPU that the library
 * holds until the process ends, so W*H is limited to 67,108,864 pixels
 * (8192 x 8192; larger sizes return MRAY_ERR_ARG). All calls on one GPU share
 * that array: calls that overlap on one GPU (other streams, other threads)
 * race and must be serialised by the caller. */
int mray_hip_moment_index_noalloc(const uint32_t *d_counts, int W, int H, uint32_t *d_index, mray_stream stream);

/* Copies the first count(pixel) moments of each pixel of a dense image into
 * the compact layout (compaction_copy_kernel, src/moment_image_device.cu:65-81;
 * MomentImageHost::compact, src/moment_image.cpp:69-94). d_compact holds
 * index[W*H] floats. */
int mray_hip_compact(const float *d_dense, const uint32_t *d_index, int W, int H, int m, float *d_compact,
                     mray_stream stream);

/* ---- prediction coding (src/moment_image.cpp:96-152) ------------------- */

/* In-place prediction coding of a compact image (MomentImageHost::prediction_encode,
 * src/moment_image.cpp:96-152; MomentImageDevice::prediction_encode,
 * src/moment_image_device.cu:112-200): bias, exponential moments, reflection
 * coefficients rotated to the real axis (src/moment_prediction_coding.h:81-159),
 * then, for MRAY_CODING_WARP_DEFAULT_TRANSFORMED, the coding parameters
 * (find_coding_transform_parameters, src/coding_transform.h:37-68, quantized
 * as src/coding_params.h:56-66) into d_coding_params (2*(m-1) uint16: p1, p2
 * per index) and the warp into [0, 1]. Bit-identical to the reference CPU path.
 * Uses stream-ordered scratch (hipMallocAsync). */
int mray_hip_prediction_encode(float *d_data, const uint32_t *d_index, int W, int H, int m, int coding_warp,
                               float bias, uint16_t *d_coding_params, mray_stream stream);

/* Lagrange multipliers of a prediction-coded image
 * (prepare_moments_from_pred_coding, src/moment_preparation.h:32-53, as
 * MomentImageHost::prepare_moments runs it when prediction_code is set,
 * src/moment_image.cpp:154-183). FP32, bit-identical to the reference CPU
 * path. d_index NULL = dense. */
int mray_hip_prepare_coded(const float *d_moments, const uint32_t *d_index, int W, int H, int m, int coding_warp,
                           const uint16_t *d_coding_params, float *d_lagrange, mray_stream stream);

/* ---- quantization (src/moment_quantization.cpp) ----------------------- */

/* Packs a compact image into the quantized bit stream of the moment-image
 * files (moment_quantization::quantize_prediction_coding(mi, table) when
 * prediction_coded != 0, src/moment_quantization.cpp:244-283; else
 * moment_quantization::quantize(mi, bits), :300-329, as write_moment_image
 * calls them, src/moment_image_io.cpp:397-408). bits (host, n_bits entries,
 * each 1..32) is the word width of moment l: the reference's quantization
 * table for prediction-coded images (its default is 16 then
 * quantization_bits), quantization_bits for every moment otherwise. Ranges:
 * [0, 1] for prediction-coded moments and moment 0, [-1/pi, 1/pi] for the
 * others. Words follow each other least significant bit first, pixels in row
 * order, and a last partial byte is dropped as the reference's pack_bits
 * drops it; *stream_bytes receives the stream length. d_stream NULL only
 * measures. Fails with MRAY_ERR_ARG when the stream exceeds capacity or a
 * pixel has more than n_bits moments. Waits for the stream (the length is
 * read back); uses stream-ordered scratch. Bit-identical to the reference. */
int mray_hip_quantize(const float *d_data, const uint32_t *d_index, int W, int H, int prediction_coded,
                      const uint8_t *bits, int n_bits, uint8_t *d_stream, size_t capacity, size_t *stream_bytes,
                      mray_stream stream);

/* The quantization table of a prediction-coded compact image
 * (create_quantization_table, src/measure_coding_error.cpp:374-437, as
 * ::generate_moments runs it on sample_moment_image of the image when
 * quantization_bits < 32, src/main.cpp:318-327): table[0] = 16, table[1] =
 * start_bits, then for each moment the fewest bits, down to 2 and at most the
 * previous moment's, whose mean coding error stays below moment 1's at
 * start_bits. The coding error of a pixel is the relative distance of its
 * decoded trigonometric moments from the unquantized decode. Each measurement
 * decodes every pixel on the GPU; the mean is summed on the host in pixel
 * order, as the reference's CPU path sums it. table (host) receives m bytes.
 * Waits for the stream. Same table as the reference. */
int mray_hip_quantization_table(const float *d_data, const uint32_t *d_index, int W, int H, int m, int coding_warp,
                                const uint16_t *d_coding_params, int start_bits, uint8_t *table, mray_stream stream);

/* Unpacks a quantized stream into a compact image
 * (moment_quantization::dequantize_prediction_coding / dequantize,
 * src/moment_quantization.cpp:285-298, 331-360, as read_moment_image calls
 * them, src/moment_image_io.cpp:169-175); same bits and ranges as
 * mray_hip_quantize. Bits past the end of the stream (the dropped partial
 * byte) read as zero, where the reference reads past its word list. Fails
 * with MRAY_ERR_ARG when the stream is shorter than the image's words.
 * Waits for the stream; bit-identical to the reference where the stream holds
 * every word whole. */
int mray_hip_dequantize(const uint8_t *d_stream, size_t stream_bytes, const uint32_t *d_index, int W, int H,
                        int prediction_coded, const uint8_t *bits, int n_bits, float *d_data, mray_stream stream);

/* ---- error bounds and uncertainty (src/raymarch_reconstruct.h:203-367,
 *      src/error_measures.h, src/moment_raymarcher.cu:693-950) ---- */

/* Per-pixel error bounds of a moment image against its volume
 * (reconstruct_error, src/moment_raymarcher.h:73-76, as --error-bounds runs it,
 * src/main.cpp:333-367): along each pixel's ray, from t_near in steps of
 * p->step_size while t <= t_far, the reconstruction from the pixel's
 * multipliers (d_lagrange as mray_hip_prepare with MRAY_FORM_PLAIN writes
 * them; d_index for compact images, d_bounds when the image has density
 * bounds) at the phase of t against map_density of the volume sample at t;
 * the error is the maximum, the RMSE or the P^2 estimate of the
 * p->error_percentile percentile of |difference| (p->error_type). Written as
 * the uint16 e * 65535, truncated (MomentImageHost::set_error_bound). The
 * reference's FP32 arithmetic: bit-identical to its CPU path. A pixel without
 * moments reconstructs with an empty series (the reference reads an empty
 * array there); a ray of fewer than 3 samples has percentile 0 (the
 * reference reads uninitialised markers). */
int mray_hip_error_bounds(mray_volume vol, const mray_camera *cam, const float *d_lagrange, const uint32_t *d_index,
                          const uint32_t *d_bounds, int W, int H, int m, const mray_error_params *p,
                          uint16_t *d_error_bounds, mray_stream stream);

/* Uncertainty renderings (reconstruct_uncertainty_interpolation /
 * reconstruct_uncertainty_convolution, src/moment_raymarcher.h:78-86) into
 * float RGBA, from PLAIN multipliers and the W*H uint16 error bounds e:
 * mode 0 ray-marches (1 - s) lower + s upper, s = bound_interpolation,
 * lower / upper the reconstruction -/+ e clamped to the pixel's density
 * bounds, with the preintegrated 256-entry table d_tf of mray_hip_reconstruct;
 * mode 1 classifies one sample per step with the 256 x 256 RGBA table d_tf
 * (row-major, the box filter of read_from_disk_uncertainty for
 * MRAY_ERROR_MAX and MRAY_ERROR_PERCENTILE, the gaussian filter for
 * MRAY_ERROR_RMSE) at (lower, upper) for MRAY_ERROR_MAX and at (mean, e)
 * otherwise. The reference's FP32 arithmetic: bit-identical to its CPU path.
 * Single scattering is not supported. */
int mray_hip_reconstruct_uncertainty(const mray_camera *cam, const float dom_min[3], const float dom_max[3],
                                     const float *d_lagrange, const uint32_t *d_index, const uint32_t *d_bounds,
                                     const uint16_t *d_error_bounds, int W, int H, int m, const float *d_tf,
                                     const mray_rec_params *p, int mode, int error_type, float bound_interpolation,
                                     float *d_rgba, mray_stream stream);

/* The quantized values of a dense image as the LZ4 path of
 * write_static_image stores them (moment_quantization::quantize(moments,
 * num_moments, bits), src/moment_quantization.cpp:361-385, called at
 * src/moment_image_io.cpp:423-426): n values (pixel-major, m per pixel),
 * value i quantized to bits bits over [0, 1] for moment 0 (i % m == 0) and
 * [-1/pi, 1/pi] otherwise, stored as (bits + 7) / 8 little-endian bytes.
 * d_out receives n * ((bits + 7) / 8) bytes. Bit-identical to the reference. */
int mray_hip_quantize_dense(const float *d_data, size_t n, int m, int bits, uint8_t *d_out, mray_stream stream);

/* The inverse (moment_quantization::dequantize(qs, num_moments, bits),
 * src/moment_quantization.cpp:387-407): n values from n * ((bits + 7) / 8)
 * bytes. */
int mray_hip_dequantize_dense(const uint8_t *d_in, size_t n, int m, int bits, float *d_data, mray_stream stream);

/* The entropy-coded moment section of a compact image (entropy_encode,
 * src/moment_image_coding.cpp:202-217, as write_moment_image calls it with
 * entropy_coding set, src/moment_image_io.cpp:383-396): u32 count of
 * non-empty pixels, their moment 0 quantized to bits[0] bits as u16 words in
 * pixel order, then for l = 1..m-1 a u32 length and the adaptive arithmetic
 * code (Amir Said's FastAC coder with its adaptive multi-symbol model,
 * src/entropy_coding/arithmetic_codec.cpp) of moment l, quantized to bits[l]
 * bits (1..14), of every pixel with more than l moments. bits (host) holds m
 * widths; value ranges are those of mray_hip_quantize. h_out is HOST memory:
 * the section is written there when it fits capacity; *out_bytes receives its
 * length either way (h_out NULL only measures it; a too-small capacity fails
 * with MRAY_ERR_ARG). One wavefront codes each moment's stream on the GPU;
 * uses stream-ordered scratch of about 4 bytes per moment. Waits for the
 * stream. Byte-identical to the reference. */
int mray_hip_entropy_encode(const float *d_data, const uint32_t *d_index, int W, int H, int m, int prediction_coded,
                            const uint8_t *bits, uint8_t *h_out, size_t capacity, size_t *out_bytes,
                            mray_stream stream);

/* Decodes an entropy-coded moment section (entropy_decode,
 * src/moment_image_coding.cpp:219-234, as read_moment_image calls it,
 * src/moment_image_io.cpp:161-168) from HOST memory h_in (n_in bytes) into
 * the compact data of the image with index d_index; same widths and ranges
 * as mray_hip_entropy_encode. Bytes past a stream's end read as zero. Fails
 * with MRAY_ERR_ARG when a length in the section runs past n_in. Waits for
 * the stream. Bit-identical to the reference. */
int mray_hip_entropy_decode(const uint8_t *h_in, size_t n_in, const uint32_t *d_index, int W, int H, int m,
                            int prediction_coded, const uint8_t *bits, float *d_data, mray_stream stream);

/* ---- SPH particles (src/volume.cpp:55-94, src/particle_grid.{cpp,cuh}) --- */

/* Samples particles (device arrays of n positions and values) at every voxel
 * position of an nx*ny*nz grid over [dom_min, dom_max] (voxel i at
 * dom_min + i/n (dom_max - dom_min)), as Particles::sample does through its
 * uniform grid of cell size h (Particles::create_grid, src/volume.cpp:55-81;
 * Particles::sample, :83-94; InterpolateParticlesOp,
 * src/particle_interpolate.h:35-58): the cubic-spline weighted mean of the
 * values of particles closer than h. Particles are ordered by cell with a
 * stable sort, so the result is deterministic. Synchronises the stream once
 * (the cell grid is sized from the particles' bounding box); uses
 * stream-ordered scratch of about 28 bytes per particle + 8 per cell. */
int mray_hip_particles_to_grid(const float *d_x, const float *d_y, const float *d_z, const float *d_v, uint32_t n,
                               float h, const float dom_min[3], const float dom_max[3], int nx, int ny, int nz,
                               float *d_grid, mray_stream stream);

/* ---- synthetic volumes (device-side generators for the benchmark inputs) --- */

/* Mixing layer on [0,1]^3: 0.5 + 0.5 tanh((z - 0.5 - h(x,y)) / 0.01) +
 * 0.05 * value_noise_4(x,y,z); h from 16 modes (amp, theta, phi host arrays of
 * 16 doubles). n^3 floats into d_out. */
int mray_hip_synth_mixing_layer(float *d_out, int n, const double *amp, const double *theta, const double *phi,
                                uint64_t seed, mray_stream stream);

/* Marschner-Lobb on [-1,1]^3 (fM = 6, a = 0.25), n^3 floats into d_out. */
int mray_hip_synth_marschner_lobb(float *d_out, int n, mray_stream stream);

/* Noise-and-blobs uint8 volume, nx*ny*nz bytes ([z][y][x]) into d_out: at
 * p = (x, y, z) / nx, v = clamp(0.5 smoothstep(0.3, 0.7, 0.5 + 0.5 noise(p)) +
 * sum_k amps[k] exp(-|p - c_k|^2 / (2 sigmas[k]^2)), 0, 1), stored as
 * floor(255 v + 0.5); noise: 6 octaves of hashed value noise from a lattice
 * of 4 per unit. Host arrays: centres 3 * n_blobs floats (x, y, z), sigmas
 * and amps n_blobs floats, n_blobs <= 64. */
int mray_hip_synth_noise_blobs(uint8_t *d_out, int nx, int ny, int nz, const float *centres, const float *sigmas,
                               const float *amps, int n_blobs, uint64_t seed, mray_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* MRAY_HIP_H */
