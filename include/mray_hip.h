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
 * (src/cut/cuda_common.cpp:55-59). */
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

int mray_hip_volume_destroy(mray_volume vol);

/* Device pointer of the volume's voxels (read-only view). */
const void *mray_hip_volume_data(mray_volume vol);

/* ---- generate (src/moment_raymarcher.h:48-49) -------------------------- */

/* Ray marches every pixel and writes dense bounded trigonometric moments.
 * Replaces generate_moments_kernel (src/moment_raymarcher.cu:163-185) /
 * generate_ray (src/raymarch_generate.h:69-148).
 *   d_moments  W*H*m floats (dense)
 *   d_bounds   W*H uint32, written when p->transform_to_bounds (else may be NULL)
 *   d_samples  optional W*H int32: samples taken per ray (N_g), NULL to skip
 * form: MRAY_FORM_PLAIN or MRAY_FORM_MIXED. Compaction and prediction coding
 * are separate calls (mray_hip_compute_num_moments, ...). */
int mray_hip_generate(mray_volume vol, const mray_camera *cam, const mray_gen_params *p, int form, int W, int H,
                      int m, float *d_moments, uint32_t *d_bounds, int32_t *d_samples, mray_stream stream);

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

/* ---- synthetic volumes (device-side generators for the benchmark inputs) --- */

/* Mixing layer on [0,1]^3: 0.5 + 0.5 tanh((z - 0.5 - h(x,y)) / 0.01) +
 * 0.05 * value_noise_4(x,y,z); h from 16 modes (amp, theta, phi host arrays of
 * 16 doubles). n^3 floats into d_out. */
int mray_hip_synth_mixing_layer(float *d_out, int n, const double *amp, const double *theta, const double *phi,
                                uint64_t seed, mray_stream stream);

/* Marschner-Lobb on [-1,1]^3 (fM = 6, a = 0.25), n^3 floats into d_out. */
int mray_hip_synth_marschner_lobb(float *d_out, int n, mray_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* MRAY_HIP_H */
