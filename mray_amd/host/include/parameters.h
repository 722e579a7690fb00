// Generation / reconstruction parameters with the reference's names and
// defaults (src/parameters.h:49-115, 231-270), and the "k=v" parsers the CLI uses.
#pragma once

#include <unordered_map>

#include "common.h"

struct GenerationParameters
{
    float step_size = 1.f;
    float step_size_write_samples = 1.5f;
    float data_min = 0.0f;
    float data_max = 1.0f;
    float error_threshold = 1.f;
    float bias = 1e-4f;
    bool compact_image = true;
    bool prediction_coding = true;
    bool compress = false;          // moment-image file options: LZ4 sections,
    bool entropy_coding = false;    // arithmetic-coded moments (compact images),
    Byte quantization_bits = 32;    // and the word width of quantized moments
    int coding_warp = 0;
    int quantization_curve_sampling = 64;
    bool transform_to_bounds = false;
    float transform_bound_eps = 0.01f;
    // kernel form of this call (not in the reference, which has one CUDA
    // form): 1 = mixed (default), 0 = plain, the reference's FP32 arithmetic
    int kernel_form = 1;
};

struct ReconstructionParameters
{
    float bias = 1e-6f;
    float medium_emission = 0.1f;
    float medium_scattering = 0.1f;
    float early_ray_termination = 0.95f;
    float data_min = 0.0f;
    float data_max = 1.0f;
    float step_size = 0.01f;
    float step_size_write_samples = 1.5f;
    bool use_truncated_fourier = false;
    // kernel form of this call: 1 = mixed (default), 0 = plain
    int kernel_form = 1;
};

// ReferenceParameters (src/parameters.h:339-350) without the single
// scattering block (not part of this library)
struct ReferenceParameters
{
    float step_size = 1.f;
    float medium_emission = 0.1f;
    float medium_scattering = 0.1f;
    float early_ray_termination = 0.95f;
    float data_min = 0.0f;
    float data_max = 1.0f;
};

// ErrorReconstructionParameters (src/parameters.h:124-158); error_type:
// 0 maximum, 1 RMSE, 2 percentile (ErrorType, src/common.h:152-157)
struct ErrorReconstructionParameters
{
    int error_type = 2;
    int error_percentile = 90;
    float bias = 1e-6f;
    float data_min = 0.0f;
    float data_max = 1.0f;
    float step_size = 0.01f;
    bool compress = false;
    bool entropy_coding = false;
    Byte quantization_bits = 32;
};

// UncertaintyReconstructionParameters (src/parameters.h:319-337)
struct UncertaintyReconstructionParameters
{
    ReconstructionParameters rec;
    float bound_interpolation = 0.5f;
    int error_type = 1;
};

GenerationParameters parse_gen_params(std::unordered_map<string, string> &params);
ErrorReconstructionParameters parse_error_params(std::unordered_map<string, string> &params);
UncertaintyReconstructionParameters parse_uncertainty_rec_params(std::unordered_map<string, string> &params);
ReferenceParameters parse_ref_params(std::unordered_map<string, string> &params);
ReconstructionParameters parse_rec_params(std::unordered_map<string, string> &params);
