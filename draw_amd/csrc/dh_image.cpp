// dh_image.cpp -- texture files to pixels on the host (TextureMap::
// load_from_file, scene/mod.rs:174-202), and the PNG decoder.
//
// The reference asks stb_image for a file's component count
// (stbi_info_from_reader), then decodes it to that many 8-bit channels
// (stbi_load_from_reader); 3 and 4 are accepted, anything else reaches
// unreachable!().  dh_image_info and dh_image_load are that pair over the C
// ABI; the OBJ/MTL ingest calls the same decoders.
//   PNG  8- or 16-bit RGB/RGBA and palette images, not interlaced, decoded to
//        8-bit RGB or RGBA (16-bit samples keep their high byte, as stb does);
//        grey PNGs have no 3- or 4-component form and are reported as
//        unsupported.
//   JPEG baseline and progressive Huffman files (dh_jpgdec.cpp), always RGB.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include <algorithm>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/drawhip.h"
#include "dh_image.hpp"

namespace dh {
namespace {
uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

bool read_file(const char* path, std::vector<uint8_t>* out) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  fseek(f, 0, SEEK_END);
  const long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  out->resize(n > 0 ? (size_t)n : 0);
  const size_t got = n > 0 ? fread(out->data(), 1, (size_t)n, f) : 0;
  fclose(f);
  return got == out->size();
}
}  // namespace

int png_decode(const uint8_t* f, size_t fn, std::vector<uint8_t>* img, int32_t* W, int32_t* H, int32_t* comp,
               std::string* err, bool info_only) {
  static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
  if (fn < 8 || memcmp(f, sig, 8) != 0) {
    *err = "not a PNG file";
    return DH_E_UNSUPPORTED;
  }
  uint32_t w = 0, h = 0;
  int depth = 0, ctype = 0, interlace = 0;
  std::vector<uint8_t> idat, plte, trns;
  for (size_t p = 8; p + 12 <= fn;) {
    const uint32_t n = be32(&f[p]);
    if (p + 12 + (size_t)n > fn) break;
    const char* t = reinterpret_cast<const char*>(&f[p + 4]);
    const uint8_t* d = &f[p + 8];
    if (!memcmp(t, "IHDR", 4) && n >= 13) {
      w = be32(d);
      h = be32(d + 4);
      depth = d[8];
      ctype = d[9];
      interlace = d[12];
    } else if (!memcmp(t, "PLTE", 4)) {
      plte.assign(d, d + n);
    } else if (!memcmp(t, "tRNS", 4)) {
      trns.assign(d, d + n);
    } else if (!memcmp(t, "IDAT", 4)) {
      if (!info_only) idat.insert(idat.end(), d, d + n);
    } else if (!memcmp(t, "IEND", 4)) {
      break;
    }
    p += 12 + (size_t)n;
  }
  if (w == 0 || h == 0 || w > 65535 || h > 65535) {
    *err = "bad PNG size";
    return DH_E_ARG;
  }
  int chans;
  if (ctype == 2 && (depth == 8 || depth == 16)) chans = 3;
  else if (ctype == 6 && (depth == 8 || depth == 16)) chans = 4;
  else if (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8) && !plte.empty()) chans = 1;
  else {
    *err = "PNG colour type/bit depth not handled (grey images have no 3/4-channel form)";
    return DH_E_UNSUPPORTED;
  }
  if (interlace) {
    *err = "interlaced PNG not handled";
    return DH_E_UNSUPPORTED;
  }
  const size_t bpp = (size_t)std::max(1, chans * depth / 8);           // filter unit
  const size_t stride = ((size_t)w * chans * depth + 7) / 8;           // bytes per row
  const int out_c = (ctype == 3) ? (trns.empty() ? 3 : 4) : chans;
  if ((uint64_t)w * h * (uint64_t)out_c > kMaxImageBytes) {
    *err = "PNG larger than the image limit";
    return DH_E_ARG;
  }
  if (info_only) {
    *W = (int32_t)w;
    *H = (int32_t)h;
    *comp = out_c;
    return DH_OK;
  }
  // The header's size is untrusted: refuse one that the compressed stream
  // cannot hold (deflate expands at most 1032:1) before allocating for it.
  const uint64_t raw_bytes = (uint64_t)(stride + 1) * h;
  if (raw_bytes > (uint64_t)idat.size() * 1032u + 1024u) {
    *err = "PNG size does not fit its data";
    return DH_E_ARG;
  }
  std::vector<uint8_t> raw((stride + 1) * h);
  uLongf rn = (uLongf)raw.size();
  if (uncompress(raw.data(), &rn, idat.data(), (uLong)idat.size()) != Z_OK || rn != raw.size()) {
    *err = "corrupt PNG data";
    return DH_E_ARG;
  }
  std::vector<uint8_t> cur(stride), prev(stride, 0);
  img->assign((size_t)w * h * out_c, 0);
  for (uint32_t y = 0; y < h; y++) {
    const uint8_t ft = raw[(stride + 1) * y];
    const uint8_t* src = &raw[(stride + 1) * y + 1];
    for (size_t i = 0; i < stride; i++) {
      const int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
      int v = src[i];
      switch (ft) {
        case 0: break;
        case 1: v += a; break;
        case 2: v += b; break;
        case 3: v += (a + b) >> 1; break;
        case 4: {
          const int pp = a + b - c, pa = abs(pp - a), pb = abs(pp - b), pc = abs(pp - c);
          v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
          break;
        }
        default: *err = "bad PNG filter"; return DH_E_ARG;
      }
      cur[i] = (uint8_t)v;
    }
    uint8_t* dst = &(*img)[(size_t)y * w * out_c];
    for (uint32_t x = 0; x < w; x++) {
      if (ctype == 3) {
        uint32_t idx;
        if (depth == 8) idx = cur[x];
        else {
          const uint32_t per = 8 / depth, sh = 8 - depth * (1 + x % per);
          idx = (cur[x / per] >> sh) & ((1u << depth) - 1u);
        }
        for (int k = 0; k < 3; k++) dst[x * out_c + k] = (3 * idx + k < plte.size()) ? plte[3 * idx + k] : 0;
        if (out_c == 4) dst[x * 4 + 3] = idx < trns.size() ? trns[idx] : 255;
      } else {
        for (int k = 0; k < chans; k++)
          dst[x * out_c + k] = depth == 8 ? cur[(size_t)x * chans + k] : cur[((size_t)x * chans + k) * 2];
      }
    }
    std::swap(cur, prev);
  }
  *W = (int32_t)w;
  *H = (int32_t)h;
  *comp = out_c;
  return DH_OK;
}

int image_decode(const uint8_t* f, size_t n, std::vector<uint8_t>* img, int32_t* W, int32_t* H, int32_t* comp,
                 std::string* err, bool info_only) {
  if (n >= 8 && f[0] == 0x89 && f[1] == 'P') return png_decode(f, n, img, W, H, comp, err, info_only);
  if (n >= 2 && f[0] == 0xFF && f[1] == 0xD8) return jpeg_decode(f, n, img, W, H, comp, err, info_only);
  *err = "unknown image format (PNG and JPEG are decoded)";
  return DH_E_UNSUPPORTED;
}

}  // namespace dh

namespace {
int image_file(const char* path, uint8_t* out, uint64_t cap, int32_t* W, int32_t* H, int32_t* comp, char* err,
               uint32_t err_len, bool info_only) {
  std::string e;
  int rc;
  try {
    std::vector<uint8_t> file, img;
    if (!dh::read_file(path, &file)) {
      e = std::string("cannot open ") + path;
      rc = DH_E_ARG;
    } else {
      rc = dh::image_decode(file.data(), file.size(), &img, W, H, comp, &e, info_only);
      if (rc == DH_OK && !info_only) {
        if (img.size() > cap) {
          e = "output buffer smaller than W*H*comp bytes";
          rc = DH_E_CAPACITY;
        } else {
          memcpy(out, img.data(), img.size());
        }
      }
    }
  } catch (const std::bad_alloc&) {
    e = "out of host memory";
    rc = DH_E_NOMEM;
  } catch (const std::exception& x) {
    e = x.what();
    rc = DH_E_ARG;
  }
  if (rc && err && err_len) snprintf(err, err_len, "%s", e.c_str());
  return rc;
}
}  // namespace

extern "C" int dh_image_info(const char* path, int32_t* W, int32_t* H, int32_t* comp, char* err, uint32_t err_len) {
  if (!path || !W || !H || !comp) return DH_E_ARG;
  return image_file(path, nullptr, 0, W, H, comp, err, err_len, true);
}

extern "C" int dh_image_load(const char* path, uint8_t* out, uint64_t cap, int32_t* W, int32_t* H, int32_t* comp,
                             char* err, uint32_t err_len) {
  if (!path || !out || !W || !H || !comp) return DH_E_ARG;
  return image_file(path, out, cap, W, H, comp, err, err_len, false);
}
