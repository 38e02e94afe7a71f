// dh_image.hpp -- texture file decoders of the OBJ/MTL ingest (host code).
// Not part of the C ABI.
//
// Each decoder returns DH_OK, DH_E_UNSUPPORTED (a form the reference's loader
// cannot use, or one not handled) or DH_E_ARG (corrupt or oversized data),
// with a message in *err.  Pixels are 8-bit, rows top-down.  With info_only
// they set only *W, *H and *comp.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace dh {

// Largest decoded texture accepted from a file: 1 GiB (a 16384 x 16384 RGBA8
// image).
constexpr uint64_t kMaxImageBytes = 1ull << 30;

// PNG -> RGB (comp 3) or RGBA (comp 4).
int png_decode(const uint8_t* f, size_t n, std::vector<uint8_t>* img, int32_t* W, int32_t* H, int32_t* comp,
               std::string* err, bool info_only = false);

// JPEG (baseline or progressive Huffman, 8-bit samples) -> RGB (comp 3).
int jpeg_decode(const uint8_t* f, size_t n, std::vector<uint8_t>* img, int32_t* W, int32_t* H, int32_t* comp,
                std::string* err, bool info_only = false);

// Either format, chosen by the file signature.
int image_decode(const uint8_t* f, size_t n, std::vector<uint8_t>* img, int32_t* W, int32_t* H, int32_t* comp,
                 std::string* err, bool info_only = false);

}  // namespace dh
