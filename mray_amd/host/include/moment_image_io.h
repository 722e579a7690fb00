// Moment-image files (src/moment_image_io.cpp:255-530), the subset without
// LZ4 compression, quantisation or entropy coding (compress = 0,
// quantization_bits = 32, entropy_coding = 0): byte-compatible with the
// reference for that subset. Header: uint16 width, height; uint8 moments,
// coding warp, flags (1 compressed, 2 compact, 4 prediction coded,
// 16 bounds, 32 error bounds, 128 coding parameters), quantisation byte;
// float3 domain min, max. Then, compact: per-pixel counts as uint8, bounds,
// error bounds, coding parameters, moments; dense: bounds, error bounds,
// coding parameters, moments. Other variants throw invalid_args.
#pragma once

#include "moment_image.h"

void write_moment_image(const MomentImageHost &mi, const std::string filename, bool compress = false,
                        Byte quantizationBits = 32, bool entropy_coding = false);

MomentImageHost load_moment_image(const std::string filename);
