// dh_jpeg.cpp -- baseline JPEG writer for frame export.
//
// The reference exports frames as PNG or JPEG through stb_image_write
// (App::export_frame_as / write_img, app/mod.rs:316-379).  Its JPEG call
// passes width * 4 where stbi_write_jpg takes the quality, and stb clamps
// that to 100, so the reference's JPEG files are quality-100 files.
//
// This is a sequential baseline encoder written from ITU-T T.81 and JFIF
// 1.02: RGB -> YCbCr (JFIF equations), 8x8 blocks with the image edge
// replicated, a separable floating-point DCT-II, quantisation with the Annex K
// tables scaled by quality the way the IJG library scales them (quality 100:
// every step 1), zig-zag order, and the Annex K Huffman tables.  Chroma is not
// subsampled (every component 1x1).  The decoded pixels follow from the
// standard; the compressed bytes are not stb's.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <exception>
#include <new>
#include <vector>

#include "../../include/drawhip.h"

namespace {

// Annex K, tables K.1 and K.2, in row-major (natural) order.
const uint8_t kLumQ[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                           14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                           18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                           49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
const uint8_t kChrQ[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99,
                           24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99, 99, 99, 99,
                           99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                           99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};

// Natural index of the k-th coefficient in zig-zag order (T.81 figure A.6).
const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                             12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                             35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                             58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// Annex K.3: code counts per length 1..16 and the symbols in code order.
const uint8_t kDcLumBits[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
const uint8_t kDcChrBits[16] = {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
const uint8_t kDcVals[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
const uint8_t kAcLumBits[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
const uint8_t kAcLumVals[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07, 0x22, 0x71,
    0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72,
    0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37,
    0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59,
    0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83,
    0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3,
    0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3,
    0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2,
    0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};
const uint8_t kAcChrBits[16] = {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77};
const uint8_t kAcChrVals[162] = {
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61, 0x71, 0x13, 0x22,
    0x32, 0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33, 0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1,
    0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18, 0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36,
    0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58,
    0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a,
    0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a,
    0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba,
    0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda,
    0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};

// Canonical codes of a table (T.81 Annex C): code and length per symbol.
struct Huff {
  uint16_t code[256];
  uint8_t len[256];
};

Huff make_huff(const uint8_t bits[16], const uint8_t* vals) {
  Huff h;
  memset(&h, 0, sizeof(h));
  uint32_t code = 0, k = 0;
  for (int l = 1; l <= 16; l++) {
    for (int i = 0; i < bits[l - 1]; i++, k++) {
      h.code[vals[k]] = (uint16_t)code;
      h.len[vals[k]] = (uint8_t)l;
      code++;
    }
    code <<= 1;
  }
  return h;
}

struct BitWriter {
  std::vector<uint8_t>* out;
  uint32_t acc = 0;
  int n = 0;
  void put(uint32_t bits, int len) {
    for (int i = len - 1; i >= 0; i--) {
      acc = (acc << 1) | ((bits >> i) & 1u);
      if (++n == 8) {
        out->push_back((uint8_t)acc);
        if (acc == 0xFF) out->push_back(0x00);  // byte stuffing (T.81 F.1.2.3)
        acc = 0;
        n = 0;
      }
    }
  }
  void flush() {  // pad the last byte with 1-bits
    while (n) put(1u, 1);
  }
};

// Magnitude category of a coefficient and its appended bits (T.81 F.1.2.1).
inline int category(int v) {
  int a = v < 0 ? -v : v, s = 0;
  while (a) {
    s++;
    a >>= 1;
  }
  return s;
}
inline uint32_t extra_bits(int v, int s) { return v >= 0 ? (uint32_t)v : (uint32_t)(v + (1 << s) - 1); }

void quant_table(const uint8_t base[64], int quality, uint8_t out[64]) {
  const int scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;
  for (int i = 0; i < 64; i++) out[i] = (uint8_t)std::min(255, std::max(1, (base[i] * scale + 50) / 100));
}

void put16(std::vector<uint8_t>& o, uint32_t v) {
  o.push_back((uint8_t)(v >> 8));
  o.push_back((uint8_t)v);
}

int jpeg_encode(const uint8_t* rgba, int32_t W, int32_t H, int32_t quality, std::vector<uint8_t>* out) {
  if (!rgba || W <= 0 || H <= 0 || W > 65535 || H > 65535) return DH_E_ARG;
  quality = quality == 0 ? 90 : std::min(100, std::max(1, quality));  // stbi_write_jpg's rule
  uint8_t q[2][64];
  quant_table(kLumQ, quality, q[0]);
  quant_table(kChrQ, quality, q[1]);
  const Huff dc[2] = {make_huff(kDcLumBits, kDcVals), make_huff(kDcChrBits, kDcVals)};
  const Huff ac[2] = {make_huff(kAcLumBits, kAcLumVals), make_huff(kAcChrBits, kAcChrVals)};
  // cos((2x+1) u pi / 16) * C(u) / 2, C(0) = 1/sqrt(2), C(u>0) = 1
  double basis[8][8];
  for (int u = 0; u < 8; u++)
    for (int x = 0; x < 8; x++)
      basis[u][x] = (u == 0 ? sqrt(0.5) : 1.0) * 0.5 * cos((2 * x + 1) * u * M_PI / 16.0);

  std::vector<uint8_t>& o = *out;
  o.clear();
  o.insert(o.end(), {0xFF, 0xD8});                                                   // SOI
  o.insert(o.end(), {0xFF, 0xE0, 0, 16, 'J', 'F', 'I', 'F', 0, 1, 2, 0, 0, 1, 0, 1, 0, 0});  // APP0 JFIF 1.02
  o.insert(o.end(), {0xFF, 0xDB});                                                   // DQT, both tables
  put16(o, 2 + 2 * 65);
  for (int t = 0; t < 2; t++) {
    o.push_back((uint8_t)t);
    for (int k = 0; k < 64; k++) o.push_back(q[t][kZigzag[k]]);
  }
  o.insert(o.end(), {0xFF, 0xC0});  // SOF0: 8-bit, three components, each 1x1
  put16(o, 17);
  o.push_back(8);
  put16(o, (uint32_t)H);
  put16(o, (uint32_t)W);
  o.insert(o.end(), {3, 1, 0x11, 0, 2, 0x11, 1, 3, 0x11, 1});
  struct Tab {
    uint8_t tc_th;
    const uint8_t* bits;
    const uint8_t* vals;
    int n;
  };
  const Tab tabs[4] = {{0x00, kDcLumBits, kDcVals, 12}, {0x10, kAcLumBits, kAcLumVals, 162},
                       {0x01, kDcChrBits, kDcVals, 12}, {0x11, kAcChrBits, kAcChrVals, 162}};
  o.insert(o.end(), {0xFF, 0xC4});  // DHT
  put16(o, 2 + 4 * 17 + 2 * 12 + 2 * 162);
  for (const Tab& t : tabs) {
    o.push_back(t.tc_th);
    o.insert(o.end(), t.bits, t.bits + 16);
    o.insert(o.end(), t.vals, t.vals + t.n);
  }
  o.insert(o.end(), {0xFF, 0xDA, 0, 12, 3, 1, 0x00, 2, 0x11, 3, 0x11, 0, 63, 0});  // SOS

  BitWriter bw{&o};
  int prev_dc[3] = {0, 0, 0};
  double blk[3][64], tmp[64], coef[64];
  for (int by = 0; by < H; by += 8) {
    for (int bx = 0; bx < W; bx += 8) {
      // the MCU's three 8x8 blocks, level-shifted, edge pixels replicated
      for (int y = 0; y < 8; y++) {
        const int sy = std::min(by + y, H - 1);
        for (int x = 0; x < 8; x++) {
          const int sx = std::min(bx + x, W - 1);
          const uint8_t* p = rgba + 4ull * ((size_t)sy * (size_t)W + (size_t)sx);
          const double r = p[0], g = p[1], b = p[2];
          blk[0][y * 8 + x] = 0.299 * r + 0.587 * g + 0.114 * b - 128.0;
          blk[1][y * 8 + x] = -0.168736 * r - 0.331264 * g + 0.5 * b;
          blk[2][y * 8 + x] = 0.5 * r - 0.418688 * g - 0.081312 * b;
        }
      }
      for (int c = 0; c < 3; c++) {
        // separable DCT-II: rows, then columns
        for (int y = 0; y < 8; y++)
          for (int u = 0; u < 8; u++) {
            double s = 0.0;
            for (int x = 0; x < 8; x++) s += basis[u][x] * blk[c][y * 8 + x];
            tmp[y * 8 + u] = s;
          }
        for (int v = 0; v < 8; v++)
          for (int u = 0; u < 8; u++) {
            double s = 0.0;
            for (int y = 0; y < 8; y++) s += basis[v][y] * tmp[y * 8 + u];
            coef[v * 8 + u] = s;
          }
        const int t = c == 0 ? 0 : 1;
        int zz[64];
        for (int k = 0; k < 64; k++) {
          const double qv = coef[kZigzag[k]] / (double)q[t][kZigzag[k]];
          zz[k] = (int)(qv < 0.0 ? -floor(-qv + 0.5) : floor(qv + 0.5));
        }
        const int diff = zz[0] - prev_dc[c];
        prev_dc[c] = zz[0];
        const int s = category(diff);
        bw.put(dc[t].code[s], dc[t].len[s]);
        if (s) bw.put(extra_bits(diff, s), s);
        int run = 0;
        for (int k = 1; k < 64; k++) {
          if (zz[k] == 0) {
            run++;
            continue;
          }
          while (run > 15) {
            bw.put(ac[t].code[0xF0], ac[t].len[0xF0]);  // ZRL
            run -= 16;
          }
          const int sa = category(zz[k]);
          const int sym = (run << 4) | sa;
          bw.put(ac[t].code[sym], ac[t].len[sym]);
          bw.put(extra_bits(zz[k], sa), sa);
          run = 0;
        }
        if (run) bw.put(ac[t].code[0x00], ac[t].len[0x00]);  // EOB
      }
    }
  }
  bw.flush();
  o.insert(o.end(), {0xFF, 0xD9});  // EOI
  return DH_OK;
}

}  // namespace

extern "C" int dh_write_jpg(const char* path, const uint8_t* rgba, int32_t W, int32_t H, int32_t quality) {
  if (!path) return DH_E_ARG;
  try {
    std::vector<uint8_t> jpg;
    const int rc = jpeg_encode(rgba, W, H, quality, &jpg);
    if (rc) return rc;
    FILE* f = fopen(path, "wb");
    if (!f) return DH_E_ARG;
    const size_t w = fwrite(jpg.data(), 1, jpg.size(), f);
    const int cl = fclose(f);
    return (w == jpg.size() && cl == 0) ? DH_OK : DH_E_ARG;
  } catch (const std::bad_alloc&) {
    return DH_E_NOMEM;
  } catch (const std::exception&) {
    return DH_E_ARG;
  }
}
