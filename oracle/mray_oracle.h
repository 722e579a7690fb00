/*
 * CPU oracle for the moment-image hot path. TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's CPU algorithm
 * (TobiasRp/mray, src/...) used as the checker for the HIP kernels, for the
 * golden-vector tests and as the timed CPU baseline in bench.py. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product (mray_amd) never does.
 *
 * Two precisions are provided:
 *   - *_f32 / unsuffixed: reference-faithful FP32, reproducing the reference's
 *     operation order and its float/double promotions (including the
 *     double-precision cos/sin/exp that the reference's unqualified calls
 *     resolve to). Pinned bit-exact against the compiled reference
 *     (oracle/_ref) by tests/test_oracle_golden.py.
 *   - *_f64: the same discretisation (the reference's float sample values,
 *     16-bit bounds, loop counts, segment end points and float phase
 *     sequence) evaluated in FP64 with exact cos/sin/atan/exp. This is the
 *     "FP64 oracle" the precision targets are stated against.
 *
 * Each function cites the reference file:line it restates.
 */
#ifndef MRAY_ORACLE_H
#define MRAY_ORACLE_H

#include "../include/mray_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_volume
{
    const void *data; /* F32: float[nz][ny][nx], U8: uint8_t[nz][ny][nx], PARTICLES: orc_sph */
    int32_t dtype;    /* MRAY_DTYPE_F32 / MRAY_DTYPE_U8 */
    int32_t nx, ny, nz;
    float dmin[3], dmax[3];
} orc_volume;

/* ---- camera / rays (src/cut/raytracing.h:39-86) ---- */
void orc_eye_ray(const mray_camera *cam, float ndcx, float ndcy, float org[3], float dir[3]);
int orc_intersect_aabb(const float org[3], const float dir[3], const float bmin[3], const float bmax[3],
                       float *t_near, float *t_far);
void orc_pixel_ray(const mray_camera *cam, int W, int H, int x, int y, float org[3], float dir[3]);

/* ---- volume (src/volume.h:59-123) ---- */
float orc_volume_sample(const orc_volume *v, const float p[3]);
double orc_volume_sample_f64(const orc_volume *v, const double p[3]);

/* ---- generate (src/raymarch_generate.h:35-148, src/moment_raymarcher.cpp:121-166) ---- */
uint32_t orc_generate_ray(const orc_volume *v, const float org[3], const float dir[3], float znear, float *moments,
                          int m, const mray_gen_params *p, int32_t *n_samples);
void orc_generate_image(const orc_volume *v, const mray_camera *cam, int W, int H, int m, const mray_gen_params *p,
                        float *moments, uint32_t *bounds, int32_t *n_samples, int row0, int row_step, int nthreads);
uint32_t orc_generate_ray_f64(const orc_volume *v, const float org[3], const float dir[3], float znear,
                              double *moments, int m, const mray_gen_params *p);
void orc_generate_image_f64(const orc_volume *v, const mray_camera *cam, int W, int H, int m,
                            const mray_gen_params *p, double *moments, uint32_t *bounds, int row0, int row_step,
                            int nthreads);

/* ---- bounded MESE, FP32 (src/mese/MESE_dynamic.h, src/mese/complex_algebra.h) ----
 * complex arrays are interleaved (re, im) float pairs. */
float orc_fast_atan(float t);
void orc_trig_to_exp(int m, const float *trig, float *exp_c);
void orc_trig_to_exp_c(int m, const float *trig_c, float *exp_c);
void orc_exp_to_trig(int m, const float *exp_c, float *trig_c);
void orc_levinson(int m, const float *first_col_c, float *sol_c);
void orc_autocorrelation(int m, const float *sig_c, float *out_c);
void orc_correlation(int m, const float *lhs_c, const float *rhs_c, float *out_c);
void orc_imag_correlation(int m, const float *lhs_c, const float *rhs_c, float *out);
void orc_prepare_mese(int m, const float *trig, float *lagrange);
void orc_prepare_mese_c(int m, const float *trig_c, float *lagrange_c);
float orc_evaluate_fourier(int m, float cx, float cy, const float *coeff);
float orc_evaluate_mese(int m, float phase, const float *lagrange);
void orc_prepare_moments_std(int m, const float *moments, float bias, float *pmoments);

/* ---- bounded MESE, FP64 ----
 * complex arrays are interleaved (re, im) double pairs. */
void orc_trig_to_exp_f64(int m, const double *trig, double *exp_c);
void orc_levinson_f64(int m, const double *first_col_c, double *sol_c);
void orc_autocorrelation_f64(int m, const double *sig_c, double *out_c);
void orc_imag_correlation_f64(int m, const double *lhs_c, const double *rhs_c, double *out);
void orc_prepare_mese_f64(int m, const double *trig, double *lagrange);
void orc_prepare_moments_std_f64(int m, const double *moments, double bias, double *pmoments);
double orc_evaluate_mese_f64(int m, double phase, const double *lagrange);

/* ---- transfer function (src/transfer_function.{h,cpp}) ---- */
void orc_tf_from_rgba8(const uint8_t *rgba8, float step_size, float *tf);
void orc_tf_sample(const float *tf, float fs, float bs, float out[4]);

/* ---- reconstruct (src/raymarch_reconstruct.h:127-200, src/moment_raymarcher.cpp:283-327) ---- */
void orc_reconstruct_ray(const float org[3], const float dir[3], const float bmin[3], const float bmax[3], float znear,
                         int m, const float *pmoments, float blo, float bhi, const float *tf,
                         const mray_rec_params *p, float out[4], int32_t *n_steps);
/* ---- reference direct volume rendering (src/raymarch_reference.h:93-162,
 * src/moment_raymarcher.cpp:47-61), without single scattering ---- */
void orc_reference_ray(const orc_volume *v, const float org[3], const float dir[3], float znear, const float *tf,
                       const mray_ref_params *p, float out[4], int32_t *n_steps);
void orc_reference_image(const orc_volume *v, const mray_camera *cam, int W, int H, const float *tf,
                         const mray_ref_params *p, float *rgba, int32_t *n_steps, int row0, int row_step,
                         int nthreads);
void orc_reconstruct_image(const mray_camera *cam, const float bmin[3], const float bmax[3], int W, int H, int m,
                           const float *data, const uint32_t *index, const uint32_t *bounds, int prediction_coded,
                           int coding_warp, const float *coding_params, const float *tf, const mray_rec_params *p,
                           float *rgba, int32_t *n_steps, int row0, int row_step, int nthreads);
void orc_reconstruct_image_f64(const mray_camera *cam, const float bmin[3], const float bmax[3], int W, int H, int m,
                               const double *moments, const uint32_t *bounds, const float *tf,
                               const mray_rec_params *p, double *rgba, int32_t *n_steps, int nthreads);
/* orc_reconstruct_image_f64 on rows row0, row0 + row_step, ... only */
void orc_reconstruct_rows_f64(const mray_camera *cam, const float bmin[3], const float bmax[3], int W, int H, int m,
                              const double *moments, const uint32_t *bounds, const float *tf,
                              const mray_rec_params *rp, double *rgba, int32_t *n_steps, int row0, int row_step,
                              int nthreads);
int orc_omp_max_threads(void);
void orc_reconstruct_lagrange_f64(const mray_camera *cam, const float bmin[3], const float bmax[3], int W, int H,
                                  int m, const float *lagrange, const uint32_t *bounds, const float *tf,
                                  const mray_rec_params *p, double *rgba, int32_t *n_steps, int nthreads);

/* the FP64 march on FP64 multipliers and table, rows row0, row0 + row_step, ... */
void orc_reconstruct_lagrange64_rows_f64(const mray_camera *cam, const float bmin[3], const float bmax[3], int W,
                                         int H, int m, const double *lagrange, const uint32_t *bounds,
                                         const double *tf, const mray_rec_params *rp, double *rgba, int32_t *n_steps,
                                         int row0, int row_step, int nthreads);
void orc_reconstruct_lagrange64_f64(const mray_camera *cam, const float bmin[3], const float bmax[3], int W, int H,
                                    int m, const double *lagrange, const uint32_t *bounds, const double *tf,
                                    const mray_rec_params *rp, double *rgba, int32_t *n_steps, int nthreads);

/* ---- adjoint of the march on given multipliers (FP64, store-all reverse sweep) ---- */
void orc_reconstruct_adjoint_f64(const mray_camera *cam, const float bmin[3], const float bmax[3], int W, int H,
                                 int m, const float *lagrange, const uint32_t *bounds, const float *tf,
                                 const mray_rec_params *rp, const double *g_rgba, double *grad_lag, double *grad_tf,
                                 int nthreads);

/* ---- compaction (src/moment_compaction.h, src/moment_image.cpp:69-94) ---- */
void orc_evaluate_unknown_moments(int n, const float *exp_c, int m, float *new_exp_c);
void orc_exp_to_bounded_moments(int m, const float *exp_c, float *bounded_c, int n);
float orc_rmse(int n, int m, const float *pred_c, const float *target);
int orc_determine_best_number_moments(int m, const float *moments, float error_threshold);
void orc_compute_num_moments_image(int W, int H, int m, const float *dense, const uint32_t *bounds,
                                   float error_threshold, uint32_t *counts, int nthreads);
uint32_t orc_compact(int W, int H, int m, const float *dense, const uint32_t *counts, uint32_t *index, float *out);

/* ---- prediction coding (src/moment_prediction_coding.h, src/coding_transform.h, src/coding_params.h,
 *      src/moment_image.cpp:96-152, src/moment_preparation.h:32-53) ---- */
void orc_encode(int m, const float *exp_c, float *code_c);
void orc_decode(int m, const float *code_c, float *exp_c, float *eval_poly_c);
void orc_transform_quantization_real(int m, const float *code_c, float c0, float *out);
void orc_transform_dequantization_real(int m, const float *in, float *code_c, int warp, const float *params);
void orc_find_coding_params(int W, int H, int m, const float *data, const uint32_t *index, float *params);
void orc_quantize_coding_params(int m, const float *params, uint16_t *q);
void orc_dequantize_coding_params(int m, const uint16_t *q, float *params);
void orc_prediction_encode_image(int W, int H, int m, float *data, const uint32_t *index, int warp, float bias,
                                 uint16_t *coding_params_q, int nthreads);
void orc_prepare_moments_pred(int m, const float *moments, int warp, const float *params, float *pmoments);

/* ---- SPH gather (src/volume.cpp:55-94, src/particle_grid.{cpp,cuh}, src/particle_interpolate.h) ----
 * A particle set (orc_sph_create) is also a volume: an orc_volume with dtype
 * MRAY_DTYPE_PARTICLES and data the orc_sph samples it (Particles::sample)
 * wherever generate samples a grid. */
typedef struct orc_sph orc_sph;
orc_sph *orc_sph_create(const float *x, const float *y, const float *z, const float *v, uint32_t n, float h);
void orc_sph_destroy(orc_sph *s);
float orc_sph_sample(const orc_sph *s, const float p[3]);
float orc_cubic_spline(float q);
void orc_particles_to_grid(const float *x, const float *y, const float *z, const float *v, uint32_t n, float h,
                           const float dmin[3], const float dmax[3], int nx, int ny, int nz, float *grid,
                           int nthreads);

/* ---- quantization of compact moment images (src/moment_quantization.cpp) ----
 * bits[l] is the word width of moment l (1..32). prediction_coded selects the
 * ranges: [0, 1] for every moment (quantize_prediction_coding), else [0, 1]
 * for moment 0 and [-1/pi, 1/pi] for the others (quantize). The stream packs
 * the words of all pixels in row order, least significant bit first; a last
 * partial byte is dropped as the reference's pack_bits does. */
uint32_t orc_quantize_value(float value, int bits, float lo, float hi);
float orc_dequantize_value(uint32_t q, int bits, float lo, float hi);
size_t orc_quantize_image(int W, int H, const float *data, const uint32_t *index, int prediction_coded,
                          const uint8_t *bits, uint8_t *out);
void orc_dequantize_image(int W, int H, const uint32_t *index, int prediction_coded, const uint8_t *bits,
                          const uint8_t *in, size_t n_in, float *data);

/* ---- error bounds and uncertainty (src/raymarch_reconstruct.h:203-367,
 * src/error_measures.h, src/transfer_function.cpp:99-248) ----
 * orc_tf2d_sample samples a full 256 x 256 x RGBA table; the box (max error,
 * percentile) and gaussian (RMSE) uncertainty tables are built from a 256 x
 * RGBA8 transfer function. orc_error_bounds writes the per-pixel uint16
 * error bound (e * 65535, truncated) of reconstruct_error_host;
 * orc_uncertainty_image renders float RGBA with the interpolated bounds
 * (mode 0, tf: the preintegrated 256 x 4 table) or the bound convolution
 * (mode 1, tf: a 2D table). */
void orc_tf2d_sample(const float *tf, float fs, float bs, float out[4]);
void orc_tf_box_filter(const uint8_t *rgba8, float step_size, float *table);
void orc_tf_gaussian_filter(const uint8_t *rgba8, float step_size, float *table);
float orc_error_bound_ray(const orc_volume *v, const float org[3], const float dir[3], float znear, int m,
                          const float *pm, float blo, float bhi, const mray_error_params *ep);
void orc_error_bounds(const orc_volume *v, const mray_camera *cam, int W, int H, int m, const float *data,
                      const uint32_t *index, const uint32_t *bounds, int prediction_coded, int coding_warp,
                      const float *coding_params, const mray_error_params *ep, uint16_t *out, int nthreads);
void orc_uncertainty_image(const mray_camera *cam, const float bmin[3], const float bmax[3], int W, int H, int m,
                           const float *data, const uint32_t *index, const uint32_t *bounds,
                           const uint16_t *error_bounds, int prediction_coded, int coding_warp,
                           const float *coding_params, const float *tf, const mray_rec_params *p, int mode,
                           int error_type, float bound_interpolation, float *rgba, int nthreads);

/* ---- entropy coding (src/moment_image_coding.cpp, FastAC) ----
 * The entropy-coded moment section of a compact image: u32 count of
 * non-empty pixels, their moment 0 as u16 words (bits[0] bits), then per
 * moment l >= 1 a u32 length and the adaptive arithmetic code of the moment
 * l words (bits[l] in 1..14) of every pixel with more than l moments.
 * encode returns the length (-1 on a bad width); decode returns 0 or -1. */
long orc_entropy_encode(int W, int H, int m, const float *data, const uint32_t *index, int prediction_coded,
                        const uint8_t *bits, uint8_t *out, size_t cap);
int orc_entropy_decode(int W, int H, int m, const uint8_t *in, size_t n_in, const uint32_t *index,
                       int prediction_coded, const uint8_t *bits, float *data);

/* ---- quantization curve (src/measure_coding_error.cpp:41-192, 374-437) ----
 * Coding error of a prediction-coded compact image: every pixel decoded
 * (transform_dequantization_real, decode, exponentialToTrigonometricMoments)
 * with moment q_index floored to q_bits bits; err[pixel] = sqrt(sum_l (ref_l
 * - t_l)^2) / ref_0 against the unquantized decode (0 for empty pixels).
 * q_index < 0 writes the unquantized trigonometric moments to ref instead.
 * params: 2(m-1) dequantized coding parameters. */
void orc_coding_error(int W, int H, const float *data, const uint32_t *index, int warp, const float *params,
                      int q_index, int q_bits, float *ref, float *err);
/* create_quantization_table: table[0] = 16, table[1] = start_bits, then per
 * moment the fewest bits (down to 2) whose mean error stays below moment 1's */
void orc_quantization_table(int W, int H, int m, const float *data, const uint32_t *index, int warp,
                            const float *params, int start_bits, uint8_t *table);

#ifdef __cplusplus
}
#endif

#endif
