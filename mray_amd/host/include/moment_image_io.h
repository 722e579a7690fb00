// Moment-image files (src/moment_image_io.cpp:255-530), byte-compatible
// with the reference in every variant: header (uint16 width, height; uint8
// moments, coding warp, flags (1 compressed, 2 compact, 4 prediction coded,
// 16 bounds, 32 error bounds, 128 coding parameters), quantization byte (bits
// 0-5 quantization_bits, 64 entropy coded, 128 table present); float3
// domain min, max), then the quantization table when set, then for compact
// images the per-pixel counts as uint8, bounds, error bounds, coding
// parameters and moments (floats, the packed quantized stream, or the
// entropy-coded section), for dense images the same without the counts and
// with the moments as floats or, compressed, LZ4 of byte-aligned quantized
// values. With compression every section is an LZ4 block behind its u32
// length (raw when LZ4 does not shrink it).
#pragma once

#include "moment_image.h"

void write_moment_image(const MomentImageHost &mi, const std::string filename, bool compress = false,
                        Byte quantizationBits = 32, bool entropy_coding = false);

MomentImageHost load_moment_image(const std::string filename);

// moment_quantization's quantization table (src/moment_quantization.cpp:165-242):
// once set, prediction-coded images are quantized with it and files carry it
// after the header
void set_prediction_coding_quantization_table(const vector<Byte> &table);
bool uses_prediction_coding_quantization_table();
vector<Byte> get_prediction_coding_quantization_table();

// create_quantization_table (src/measure_coding_error.cpp:374-437) on the
// GPU (mray_hip_quantization_table); sets the table
void create_quantization_table(const MomentImageHost &img, int start_bits);

// sample_moment_image (src/moment_image_sample.h:33-61): every
// (width / w)-th column and (height / h)-th row of a compact image. Sizes
// for which the reference's loops read or write outside the images (w != h
// with h > w, or rows past the source) throw invalid_args.
MomentImageHost sample_moment_image(const MomentImageHost &img, int w, int h);
