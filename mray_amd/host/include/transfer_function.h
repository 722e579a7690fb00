// Transfer function (src/transfer_function.{h,cpp}): 256 RGBA entries read
// from an RGBA8 .bin file, colour linearised with pow(c / 255, 2.2), then
// "preintegrated" into T[x] = (rgb * a * s, a * s) with s the step size times
// the preintegration factor. The reference stores a 256 x 256 table whose
// rows are all equal; this keeps one row.
#pragma once

#include <array>

#include "common.h"

struct TransferFunction
{
    static constexpr int RESOLUTION = 256;
    std::array<Vec4f, RESOLUTION> data{};
};

std::array<Vec4f, TransferFunction::RESOLUTION> read_tf(const char *filepath);
std::array<Vec4f, TransferFunction::RESOLUTION> tf_from_rgba8(const unsigned char *rgba8);
TransferFunction preintegrate(const std::array<Vec4f, TransferFunction::RESOLUTION> &tf1d, float step_size);
// read_tf + preintegrate (read_from_disk, as the CLI calls it with the
// reconstruction step size times preintegration_step_size_factor)
TransferFunction read_from_disk(const char *filepath, float step_size);
