/*
 * Plain-old-data types shared by the HIP C-ABI (mray_hip.h), the C++ host
 * library and the CPU oracle. No HIP, torch or C++ types appear here so that a
 * ctypes / cgo / JNI binding can mirror them field by field.
 */
#ifndef MRAY_TYPES_H
#define MRAY_TYPES_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes returned by every C-ABI entry point (never an exception). */
enum mray_status
{
    MRAY_OK = 0,
    MRAY_ERR_ARG = -1,         /* bad argument (null pointer, size out of range, ...) */
    MRAY_ERR_HIP = -2,         /* a HIP runtime call failed */
    MRAY_ERR_OOM = -3,         /* device allocation failed */
    MRAY_ERR_UNSUPPORTED = -4, /* valid request this build does not implement */
    MRAY_ERR_NOT_INIT = -5     /* mray_hip_init was not called */
};

/* Voxel types of RegularGrid (src/volume.h:33-37), and particle sets. */
enum mray_dtype
{
    MRAY_DTYPE_F32 = 0,
    MRAY_DTYPE_U8 = 1,
    MRAY_DTYPE_PARTICLES = 2 /* a particle set (Particles, src/volume.h:132-160), mray_hip_volume_from_particles */
};

/* Kernel forms. PLAIN follows the reference per-sample arithmetic (per-moment
 * sin/cos, FP32 Levinson, one thread per ray); MIXED uses the per-step phase
 * recurrence, compensated FP32 accumulation and FP64 Levinson/Herglotz. */
enum mray_form
{
    MRAY_FORM_PLAIN = 0,
    MRAY_FORM_MIXED = 1
};

/* Coding warp types (src/coding_params.h:32-39). */
enum mray_coding_warp
{
    MRAY_CODING_WARP_DEFAULT_TRANSFORMED = 0,
    MRAY_CODING_WARP_NONE = 1
};

/*
 * Camera, field for field the data read by cut::load_from_file
 * (src/cut/camera.h:39-57, src/cut/camera.cpp:30-75). transform is row-major:
 * transform[j*4+i] == Mat4f m[j][i]. znear is read only for perspective
 * cameras by the reference; set it explicitly (0) for orthographic ones.
 */
typedef struct mray_camera
{
    int32_t orthographic;
    float transform[16];
    float position[3];
    float right[3];
    float up[3];
    float dir[3];
    float scale;
    float aspect;
    float znear;
} mray_camera;

/* GenerationParameters (src/parameters.h:49-70); same names and defaults. */
typedef struct mray_gen_params
{
    float step_size;           /* 1.0 */
    float data_min;            /* 0.0 */
    float data_max;            /* 1.0 */
    float error_threshold;     /* 1.0 */
    float bias;                /* 1e-4 */
    int32_t compact_image;     /* 1 */
    int32_t prediction_coding; /* 1 */
    int32_t coding_warp;       /* 0 */
    int32_t transform_to_bounds; /* 0 */
    float transform_bound_eps; /* 0.01 */
} mray_gen_params;

/* ReconstructionParameters (src/parameters.h:231-247) without the
 * single-scattering block, which this path does not implement. */
typedef struct mray_rec_params
{
    float bias;                  /* 1e-6 */
    float medium_emission;       /* 0.1 */
    float medium_scattering;     /* 0.1 (unused without single scattering) */
    float early_ray_termination; /* 0.95 */
    float data_min;              /* 0.0 */
    float data_max;              /* 1.0 */
    float step_size;             /* 0.01 */
    int32_t use_truncated_fourier; /* 0 */
} mray_rec_params;

/* ReferenceParameters (src/parameters.h:339-350) of the reference direct
 * volume renderer, without single scattering (SingleScatteringParameters
 * .enable = false, its default). Same names and defaults. */
typedef struct mray_ref_params
{
    float step_size;             /* 1.0 */
    float medium_emission;       /* 0.1 */
    float medium_scattering;     /* 0.1 (unused without single scattering) */
    float early_ray_termination; /* 0.95 */
    float data_min;              /* 0.0 */
    float data_max;              /* 1.0 */
    int32_t single_scattering;   /* 0 (ss.enable); nonzero is not supported */
} mray_ref_params;

/* ErrorType (src/common.h:152-157) */
enum mray_error_type
{
    MRAY_ERROR_MAX = 0,
    MRAY_ERROR_RMSE = 1,
    MRAY_ERROR_PERCENTILE = 2
};

/* ErrorReconstructionParameters (src/parameters.h:124-137) without the file
 * options (compress, entropy_coding, quantization_bits), which the host
 * library takes where it writes the file. Same names and defaults. */
typedef struct mray_error_params
{
    int32_t error_type;       /* 2 (MRAY_ERROR_PERCENTILE) */
    int32_t error_percentile; /* 90 */
    float bias;               /* 1e-6 */
    float data_min;           /* 0.0 */
    float data_max;           /* 1.0 */
    float step_size;          /* 0.01 */
} mray_error_params;

/* DensityBound (src/common.h:135-150) is stored as one uint32 per pixel:
 * lower | (upper << 16). */
#define MRAY_BOUND_LO(b) ((uint16_t)((b)&0xFFFFu))
#define MRAY_BOUND_HI(b) ((uint16_t)((b) >> 16))

#ifdef __cplusplus
}
#endif

#endif /* MRAY_TYPES_H */
