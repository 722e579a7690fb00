// Transfer function (src/transfer_function.{h,cpp}): 256 RGBA entries read
// from an RGBA8 .bin file, colour linearised with pow(c / 255, 2.2), then
// "preintegrated" into T[x] = (rgb * a * s, a * s) with s the step size times
// the preintegration factor. The reference stores a 256 x 256 table whose
// rows are all equal; this keeps one row.
#pragma once

#include <array>
#include <vector>

#include "common.h"

struct TransferFunction
{
    static constexpr int RESOLUTION = 256;
    std::array<Vec4f, RESOLUTION> data{};
    // the full 256 x 256 table (row y, column x at y * 256 + x) of an
    // uncertainty transfer function; empty for the preintegrated ones
    std::vector<Vec4f> table2d;
};

std::array<Vec4f, TransferFunction::RESOLUTION> read_tf(const char *filepath);
std::array<Vec4f, TransferFunction::RESOLUTION> tf_from_rgba8(const unsigned char *rgba8);
TransferFunction preintegrate(const std::array<Vec4f, TransferFunction::RESOLUTION> &tf1d, float step_size);
// read_tf + preintegrate (read_from_disk, as the CLI calls it with the
// reconstruction step size times preintegration_step_size_factor)
TransferFunction read_from_disk(const char *filepath, float step_size);
// the uncertainty transfer functions (src/transfer_function.cpp:152-248):
// box_filter averages the 1D function between x / 256 and y / 256 (bounds
// of the maximum error or a percentile), gaussian_filter weights it by the
// normal distribution of mean x / 256 and deviation y / 256 (RMSE);
// read_from_disk_uncertainty picks by error_type (mray_error_type)
TransferFunction box_filter(const std::array<Vec4f, TransferFunction::RESOLUTION> &tf1d, float step_size);
TransferFunction gaussian_filter(const std::array<Vec4f, TransferFunction::RESOLUTION> &tf1d, float step_size);
TransferFunction read_from_disk_uncertainty(const char *filepath, float step_size, int error_type);
