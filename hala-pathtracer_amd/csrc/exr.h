// exr.h — OpenEXR decoding of HDRI environment maps (host code).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace hpt {

struct ExrImage {
  uint32_t width = 0;
  uint32_t height = 0;
  // height * width * 3 floats, row 0 = the data window's first (top) line.
  std::vector<float> rgb;
};

enum ExrResult { kExrOk = 0, kExrMalformed = 1, kExrUnsupported = 2 };

// Decodes a single-part scanline OpenEXR file held in memory.  With
// `pixels` false only the size is read (header and offset table).  RGB come
// from the channels named R, G and B (one missing reads as 0), else from a
// luminance channel Y copied to all three.  On failure `err` says why.
ExrResult decode_exr(const uint8_t* data, size_t size, bool pixels, ExrImage& out, std::string& err);

}  // namespace hpt
