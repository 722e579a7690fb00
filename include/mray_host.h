/*
 * C entry points of the C++ host library (libmray_host.so) for bindings that
 * cannot call C++: the moment-image file writer and reader
 * (write_moment_image / load_moment_image, src/moment_image_io.cpp:286-313,
 * 515-530, as mray_amd/host/include/moment_image_io.h declares them), which
 * run quantization and entropy coding on the GPU through mray_hip.h.
 * Every function returns MRAY_OK or a negative mray_status; the message of
 * the last failure is mray_host_last_error().
 */
#ifndef MRAY_HOST_H
#define MRAY_HOST_H

#include "mray_types.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *mray_host_last_error(void);

/* write_moment_image (src/moment_image_io.cpp:515-530) of an image given by
 * host arrays: compact when index (W*H+1 offsets) is non-null, else dense
 * (W*H*m floats); bounds (W*H, lower | upper << 16), error_bounds (W*H) and
 * coding_params (2(m-1) quantized parameters) are written when non-null.
 * table (m entries), when non-null, first sets the prediction-coding
 * quantization table (set_prediction_coding_quantization_table), which stays
 * set for the rest of the process, as in the reference. */
int mray_host_write_moment_image(const char *path, int W, int H, int m, int prediction_coded, int coding_warp,
                                 const float dmin[3], const float dmax[3], const uint32_t *index, const float *data,
                                 size_t n_data, const uint32_t *bounds, const uint16_t *error_bounds,
                                 const uint16_t *coding_params, const uint8_t *table, int compress, int bits,
                                 int entropy_coding);

/* load_moment_image (src/moment_image_io.cpp:286-313). hdr receives W, H, m,
 * coding warp, compact, prediction coded, has bounds, has error bounds, has
 * coding params and the number of data floats; the arrays are copied where
 * the pointers are non-null (index W*H+1, data, bounds W*H packed, error
 * bounds W*H, coding params 2(m-1)). Loading a file that carries a
 * quantization table sets it, as in the reference. */
int mray_host_load_moment_image(const char *path, int64_t hdr[10], uint32_t *index, float *data, uint32_t *bounds,
                                uint16_t *error_bounds, uint16_t *coding_params);

/* The uncertainty transfer function of a 256 x RGBA8 transfer function
 * (read_from_disk_uncertainty, src/transfer_function.cpp:233-248): the box
 * filter for MRAY_ERROR_MAX and MRAY_ERROR_PERCENTILE, the gaussian filter for
 * MRAY_ERROR_RMSE, at step_size; table receives 256 x 256 x 4 floats, row y
 * (upper bound or deviation) by column x (lower bound or mean). */
int mray_host_uncertainty_tf(const uint8_t *rgba8, float step_size, int error_type, float *table);

/* The prediction-coding quantization table: returns 1 and copies its 255
 * entries into table when one is set, else 0. */
int mray_host_quantization_table(uint8_t table[255]);

#ifdef __cplusplus
}
#endif

#endif /* MRAY_HOST_H */
