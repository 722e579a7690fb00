// Host-side launchers implemented in the kernel translation units and called
// by the C ABI (capi.hip). Each returns 0 on success or a negative mray_status.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mray_types.h"

namespace mray
{

int launch_generate(const void *vox, int dtype, int nx, int ny, int nz, const float *dmin, const float *dmax,
                    const mray_camera &cam, const mray_gen_params &p, int form, int W, int H, int m, float *moments,
                    uint32_t *bounds, int32_t *samples, float *ray_range, hipStream_t stream);

int launch_generate_adjoint(const float *vox, int nx, int ny, int nz, const float *dmin, const float *dmax,
                            const mray_camera &cam, const mray_gen_params &p, int W, int H, int m,
                            const float *grad_moments, const float *ray_range, float *grad_volume,
                            hipStream_t stream);

int launch_prepare(const float *moments, const uint32_t *index, int W, int H, int m, float bias, int form,
                   float *lagrange, hipStream_t stream);
int launch_prepare_f64(const float *moments, const uint32_t *index, int W, int H, int m, float bias,
                       double *lagrange, hipStream_t stream);

int launch_reference(const void *vox, int dtype, int nx, int ny, int nz, const float *dmin, const float *dmax,
                     const mray_camera &cam, const mray_ref_params &p, int W, int H, const float *tf, float *rgba,
                     int32_t *steps, hipStream_t stream);

int launch_reconstruct(const mray_camera &cam, const float *dmin, const float *dmax, const float *lagrange,
                       const uint32_t *index, const uint32_t *bounds, int W, int H, int m, const float *tf,
                       const mray_rec_params &p, int form, float *rgba, int32_t *steps, hipStream_t stream);

int launch_reconstruct_f64(const mray_camera &cam, const float *dmin, const float *dmax, const double *lagrange,
                           const uint32_t *index, const uint32_t *bounds, int W, int H, int m, const float *tf,
                           const mray_rec_params &p, float *rgba, int32_t *steps, hipStream_t stream);

int launch_reconstruct_adjoint(const mray_camera &cam, const float *dmin, const float *dmax, const float *lagrange,
                               const uint32_t *index, const uint32_t *bounds, int W, int H, int m, const float *tf,
                               const mray_rec_params &p, const float *rgba, const float *grad_rgba,
                               float *grad_lagrange, float *grad_tf, hipStream_t stream);

int launch_prepare_adjoint(const float *moments, const uint32_t *index, int W, int H, int m, float bias,
                           const float *grad_lagrange, float *grad_moments, hipStream_t stream);

int launch_num_moments(const float *dense, const uint32_t *bounds, int W, int H, int m, float error_threshold,
                       uint32_t *counts, hipStream_t stream);
int launch_moment_index(const uint32_t *counts, int npix, uint32_t *index, hipStream_t stream);
int launch_clamp_counts(uint32_t *counts, int npix, int lo, int hi, hipStream_t stream);
int launch_moment_index_noalloc(const uint32_t *counts, int npix, uint32_t *index, hipStream_t stream);
int noalloc_scratch_init(int device, uint32_t **out);
size_t noalloc_max_pixels();
int launch_compact(const float *dense, const uint32_t *index, int W, int H, int m, float *out, hipStream_t stream);
int launch_prediction_encode(float *data, const uint32_t *index, int W, int H, int m, int warp, float bias,
                             uint16_t *qparams, hipStream_t stream);
int launch_prepare_coded(const float *moments, const uint32_t *index, int W, int H, int m, int warp,
                         const uint16_t *qparams, float *lagrange, hipStream_t stream);
int launch_quantization_table(const float *data, const uint32_t *index, int W, int H, int m, int warp,
                              const uint16_t *qparams, int start_bits, uint8_t *table, hipStream_t stream);
int launch_quantize(const float *data, const uint32_t *index, int W, int H, int prediction_coded,
                    const uint8_t *bits, int n_bits, uint8_t *stream_out, size_t capacity, size_t *stream_bytes,
                    hipStream_t stream);
int launch_dequantize(const uint8_t *stream_in, size_t stream_bytes, const uint32_t *index, int W, int H,
                      int prediction_coded, const uint8_t *bits, int n_bits, float *data, hipStream_t stream);
int launch_error_bounds(const void *vox, int dtype, int nx, int ny, int nz, const float *dmin, const float *dmax,
                        const mray_camera &cam, const float *lagrange, const uint32_t *index, const uint32_t *bounds,
                        int W, int H, int m, const mray_error_params &p, uint16_t *out, hipStream_t stream);
int launch_uncertainty(const mray_camera &cam, const float *dmin, const float *dmax, const float *lagrange,
                       const uint32_t *index, const uint32_t *bounds, const uint16_t *error_bounds, int W, int H, int m,
                       const float *tf, const mray_rec_params &p, int mode, int error_type, float bound_interpolation,
                       float *rgba, hipStream_t stream);
int launch_quantize_dense(const float *data, size_t n, int m, int bits, uint8_t *out, hipStream_t stream);
int launch_dequantize_dense(const uint8_t *in, size_t n, int m, int bits, float *data, hipStream_t stream);
int launch_entropy_encode(const float *data, const uint32_t *index, int W, int H, int m, int prediction_coded,
                          const uint8_t *bits, uint8_t *out, size_t capacity, size_t *out_bytes, hipStream_t stream);
int launch_entropy_decode(const uint8_t *in, size_t n_in, const uint32_t *index, int W, int H, int m,
                          int prediction_coded, const uint8_t *bits, float *data, hipStream_t stream);

int launch_particles_to_grid(const float *x, const float *y, const float *z, const float *v, uint32_t n, float h,
                             const float *dmin, const float *dmax, int nx, int ny, int nz, float *grid,
                             hipStream_t stream);

int launch_synth_mixing_layer(float *out, int n, const double *amp, const double *theta, const double *phi,
                              uint64_t seed, hipStream_t stream);
int launch_synth_marschner_lobb(float *out, int n, hipStream_t stream);
int launch_synth_noise_blobs(uint8_t *out, int nx, int ny, int nz, const float *centres, const float *sigmas,
                             const float *amps, int n_blobs, uint64_t seed, hipStream_t stream);

} // namespace mray
