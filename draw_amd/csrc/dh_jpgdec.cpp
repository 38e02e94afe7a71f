// dh_jpgdec.cpp -- JPEG texture decoding for the OBJ/MTL ingest.
//
// The reference loads texture files through stb_image
// (TextureMap::load_from_file, scene/mod.rs:174-202): stbi_info_from_reader
// gives the component count (3 for any colour JPEG, 1 for grey, which the
// reference rejects with unreachable!()), then stbi_load_from_reader decodes
// to that many channels.  stb_image comes from the `stb` crate (git branch
// fix-ubuntu-24.10-build of mororo18/stb, Cargo.toml:16), which is not part of
// the reference tree.  This decoder is written from ITU-T T.81 (Huffman
// sequential and progressive modes, Annexes B, C, F and G) and follows the
// pixel pipeline stb_image publishes for its JPEG path:
//   * coefficients dequantized as 16-bit integers (sequential: while decoding;
//     progressive: after the last scan);
//   * the integer 8x8 IDCT of the IJG "islow" method (Loeffler, Ligtenberg and
//     Moschytz), with 12-bit constants, 2 extra bits after the column pass and
//     rounding plus the +128 level shift folded into the row pass;
//   * chroma upsampled by a triangle filter for 2x horizontal and/or vertical
//     factors (others replicate);
//   * YCbCr -> RGB in 20-bit fixed point with the JFIF coefficients (1.402,
//     0.34414, 0.71414, 1.772); RGB files (component ids 'R','G','B', or an
//     Adobe marker with transform 0 and no JFIF marker) are copied; CMYK and
//     YCCK (Adobe transform 0 / 2) are multiplied by K with Blinn's exact
//     x*y/255.
// Arithmetic, lossless, hierarchical and 12-bit files are rejected, as
// stb_image rejects them.
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/drawhip.h"
#include "dh_image.hpp"

namespace dh {
namespace {

// Natural (row-major) index of the k-th coefficient in zig-zag order (T.81
// figure A.6), padded with 15 entries of 63 so that a corrupt run that
// overshoots the block lands on the last coefficient instead of outside it.
const uint8_t kDezigzag[64 + 15] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33,
                                    40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36,
                                    29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54,
                                    47, 55, 62, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63};

struct Fail {
  int rc;
  const char* msg;
};
[[noreturn]] void fail(int rc, const char* msg) { throw Fail{rc, msg}; }

// Canonical Huffman table (T.81 Annex C) with a 9-bit lookup for short codes.
struct Huff {
  static constexpr int kFast = 9;
  uint8_t fast[1 << kFast];  // symbol index + 1 for codes of <= kFast bits, 0 otherwise
  uint8_t len[256];          // code length of each symbol index
  uint8_t sym[256];
  int32_t maxcode[17];  // largest code of each length, -1 if none
  int32_t valptr[17];   // symbol index of the first code of each length minus that code
  bool defined = false;

  // counts[l - 1] codes of length l for the symbols in order (at most 256).
  void build(const uint8_t counts[16], const uint8_t* symbols) {
    memset(fast, 0, sizeof fast);
    int k = 0, code = 0;
    for (int l = 1; l <= 16; l++) {
      valptr[l] = k - code;
      for (int i = 0; i < counts[l - 1]; i++, k++, code++) {
        if (code >= (1 << l)) fail(DH_E_ARG, "corrupt JPEG: more Huffman codes than their lengths allow");
        len[k] = (uint8_t)l;
        sym[k] = symbols[k];
        if (l <= kFast)
          for (int j = code << (kFast - l); j < (code + 1) << (kFast - l); j++) fast[j] = (uint8_t)(k + 1);
      }
      maxcode[l] = counts[l - 1] ? code - 1 : -1;
      code <<= 1;
    }
    defined = true;
  }
};

// Entropy-coded bits (T.81 F.2.2.5): 0xFF00 is a data 0xFF, 0xFF fill bytes
// are skipped, any other marker ends the data.  Past the end, reads give
// zero bits, which `valid` does not count.
struct Bits {
  const uint8_t* p = nullptr;
  const uint8_t* e = nullptr;
  uint32_t buf = 0;  // next bits, most significant first
  int valid = 0;     // bits of buf that came from the data
  int marker = -1;   // the marker that ended the data, or -1
  bool nomore = false;

  void start(const uint8_t* q, const uint8_t* end) {
    p = q;
    e = end;
    buf = 0;
    valid = 0;
    marker = -1;
    nomore = false;
  }
  void fill() {
    while (valid <= 24 && !nomore) {
      if (p >= e) {
        nomore = true;
        break;
      }
      uint32_t b = *p++;
      if (b == 0xFF) {
        uint32_t c = p < e ? *p++ : 0xD9;
        while (c == 0xFF && p < e) c = *p++;
        if (c != 0) {
          marker = (int)c;
          nomore = true;
          break;
        }
      }
      buf |= b << (24 - valid);
      valid += 8;
    }
  }
  void skip(int n) {
    buf <<= n;
    valid -= n;
  }
  // n raw bits (T.81 F.2.2.4 RECEIVE); 0 once the data has run out.
  int receive(int n) {
    if (n == 0) return 0;
    if (valid < n) fill();
    if (valid < n) return 0;
    const uint32_t v = buf >> (32 - n);
    skip(n);
    return (int)v;
  }
  int bit() { return receive(1); }
  // RECEIVE then EXTEND (T.81 F.2.2.1): n bits as a signed magnitude.
  int extend(int n) {
    if (n == 0) return 0;
    if (valid < n) fill();
    if (valid < n) return 0;
    const int v = (int)(buf >> (32 - n));
    skip(n);
    return v < (1 << (n - 1)) ? v - (1 << n) + 1 : v;
  }
  // One Huffman symbol (T.81 F.2.2.3 DECODE).
  int decode(const Huff& h) {
    if (!h.defined) fail(DH_E_ARG, "scan uses an undefined Huffman table");
    if (valid < 16) fill();
    const uint32_t top = buf >> (32 - Huff::kFast);
    const int k = h.fast[top];
    if (k) {
      const int l = h.len[k - 1];
      if (l > valid) fail(DH_E_ARG, "corrupt JPEG: data ends inside a Huffman code");
      skip(l);
      return h.sym[k - 1];
    }
    for (int l = Huff::kFast + 1; l <= 16; l++) {
      const int32_t code = (int32_t)(buf >> (32 - l));
      if (code <= h.maxcode[l]) {
        if (l > valid) fail(DH_E_ARG, "corrupt JPEG: data ends inside a Huffman code");
        skip(l);
        return h.sym[h.valptr[l] + code];
      }
    }
    fail(DH_E_ARG, "corrupt JPEG: bad Huffman code");
  }
};

struct Comp {
  int id = 0, h = 1, v = 1, tq = 0;
  int x = 0, y = 0;     // samples of the component (T.81 A.1.1)
  int w2 = 0, h2 = 0;   // plane size padded to whole MCUs
  int bw = 0, bh = 0;   // blocks per row and column of the padded plane
  int td = 0, ta = 0;   // Huffman tables of the current scan
  int dc_pred = 0;
  std::vector<uint8_t> plane;
  std::vector<int16_t> coef;  // progressive: 64 per block of the padded plane
};

inline uint8_t clamp255(int x) { return (uint8_t)(x < 0 ? 0 : (x > 255 ? 255 : x)); }

// The LLM 8-point IDCT in 12-bit fixed point (constants round(c * 4096)).
// Products are formed in 64 bits and reduced to 32 bits before each shift,
// which gives the results of 32-bit wrapping arithmetic without overflow.
struct Idct1 {
  int64_t x0, x1, x2, x3, t0, t1, t2, t3;
  Idct1(int64_t s0, int64_t s1, int64_t s2, int64_t s3, int64_t s4, int64_t s5, int64_t s6, int64_t s7) {
    // even part
    int64_t p2 = s2, p3 = s6;
    int64_t p1 = (p2 + p3) * 2217;       // 0.5411961
    t2 = p1 + p3 * -7567;                // -1.847759065
    t3 = p1 + p2 * 3135;                 // 0.765366865
    p2 = s0;
    p3 = s4;
    t0 = (p2 + p3) * 4096;
    t1 = (p2 - p3) * 4096;
    x0 = t0 + t3;
    x3 = t0 - t3;
    x1 = t1 + t2;
    x2 = t1 - t2;
    // odd part
    t0 = s7;
    t1 = s5;
    t2 = s3;
    t3 = s1;
    p3 = t0 + t2;
    int64_t p4 = t1 + t3;
    p1 = t0 + t3;
    p2 = t1 + t2;
    const int64_t p5 = (p3 + p4) * 4816;  // 1.175875602
    t0 = t0 * 1223;                       // 0.298631336
    t1 = t1 * 8410;                       // 2.053119869
    t2 = t2 * 12586;                      // 3.072711026
    t3 = t3 * 6149;                       // 1.501321110
    p1 = p5 + p1 * -3685;                 // -0.899976223
    p2 = p5 + p2 * -10497;                // -2.562915447
    p3 = p3 * -8034;                      // -1.961570560
    p4 = p4 * -1597;                      // -0.390180644
    t3 += p1 + p4;
    t2 += p2 + p3;
    t1 += p2 + p4;
    t0 += p1 + p3;
  }
};
inline int32_t wrap32(int64_t v) { return (int32_t)(uint32_t)(uint64_t)v; }

void idct_block(const int16_t* d, uint8_t* out, int stride) {
  int32_t v[64];
  for (int i = 0; i < 8; i++) {
    if (!d[i + 8] && !d[i + 16] && !d[i + 24] && !d[i + 32] && !d[i + 40] && !d[i + 48] && !d[i + 56]) {
      const int32_t dc = (int32_t)d[i] * 4;
      for (int k = 0; k < 8; k++) v[i + 8 * k] = dc;
      continue;
    }
    Idct1 s(d[i], d[i + 8], d[i + 16], d[i + 24], d[i + 32], d[i + 40], d[i + 48], d[i + 56]);
    // scaled up by 2^12: back down to 2 extra bits, rounded
    s.x0 += 512;
    s.x1 += 512;
    s.x2 += 512;
    s.x3 += 512;
    v[i] = wrap32(s.x0 + s.t3) >> 10;
    v[i + 56] = wrap32(s.x0 - s.t3) >> 10;
    v[i + 8] = wrap32(s.x1 + s.t2) >> 10;
    v[i + 48] = wrap32(s.x1 - s.t2) >> 10;
    v[i + 16] = wrap32(s.x2 + s.t1) >> 10;
    v[i + 40] = wrap32(s.x2 - s.t1) >> 10;
    v[i + 24] = wrap32(s.x3 + s.t0) >> 10;
    v[i + 32] = wrap32(s.x3 - s.t0) >> 10;
  }
  for (int j = 0; j < 8; j++, out += stride) {
    const int32_t* r = v + 8 * j;
    Idct1 s(r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7]);
    // 2^12 from the constants, 2^2 from the column pass, 2^3 from the 2-D
    // normalisation: 2^17, with rounding and the +128 level shift
    const int64_t bias = 65536 + (128 << 17);
    s.x0 += bias;
    s.x1 += bias;
    s.x2 += bias;
    s.x3 += bias;
    out[0] = clamp255(wrap32(s.x0 + s.t3) >> 17);
    out[7] = clamp255(wrap32(s.x0 - s.t3) >> 17);
    out[1] = clamp255(wrap32(s.x1 + s.t2) >> 17);
    out[6] = clamp255(wrap32(s.x1 - s.t2) >> 17);
    out[2] = clamp255(wrap32(s.x2 + s.t1) >> 17);
    out[5] = clamp255(wrap32(s.x2 - s.t1) >> 17);
    out[3] = clamp255(wrap32(s.x3 + s.t0) >> 17);
    out[4] = clamp255(wrap32(s.x3 - s.t0) >> 17);
  }
}

inline int16_t mul16(int a, int b) { return (int16_t)(uint16_t)(uint32_t)(a * b); }

class Decoder {
 public:
  Decoder(const uint8_t* f, size_t n) : f_(f), n_(n) {}

  void run(bool info_only) {
    if (n_ < 2 || f_[0] != 0xFF || f_[1] != 0xD8) fail(DH_E_UNSUPPORTED, "not a JPEG file");
    pos_ = 2;
    bool frame = false;
    for (;;) {
      const int m = next_marker();
      if (m < 0 || m == 0xD9) break;  // end of data or EOI
      if (m == 0xD8) fail(DH_E_ARG, "corrupt JPEG: second SOI");
      if (m >= 0xD0 && m <= 0xD7) continue;  // a stray restart marker carries no segment
      const size_t L = seg_len();
      const uint8_t* s = f_ + pos_ + 2;
      const size_t body = L - 2;
      switch (m) {
        case 0xC0: case 0xC1: case 0xC2:
          if (frame) fail(DH_E_ARG, "corrupt JPEG: second frame header");
          frame_header(s, body, m == 0xC2);
          frame = true;
          if (info_only) return;
          break;
        case 0xC3: case 0xC5: case 0xC6: case 0xC7: case 0xC9: case 0xCA: case 0xCB: case 0xCD: case 0xCE: case 0xCF:
          fail(DH_E_UNSUPPORTED, "JPEG coding process not handled (lossless, hierarchical or arithmetic)");
        case 0xC4: huffman_tables(s, body); break;
        case 0xDB: quant_tables(s, body); break;
        case 0xDD:
          if (body < 2) fail(DH_E_ARG, "corrupt JPEG: bad DRI");
          restart_interval_ = (s[0] << 8) | s[1];
          break;
        case 0xDA: {
          if (!frame) fail(DH_E_ARG, "corrupt JPEG: scan before the frame header");
          scan_header(s, body);
          pos_ += L;
          scan();
          continue;  // pos_ is at the marker that ended the scan
        }
        case 0xDC: fail(DH_E_UNSUPPORTED, "JPEG with a DNL marker not handled");
        case 0xE0:
          if (body >= 5 && !memcmp(s, "JFIF\0", 5)) jfif_ = true;
          break;
        case 0xEE:
          // "Adobe", version (2), flags0 (2), flags1 (2), transform (1)
          if (body >= 12 && !memcmp(s, "Adobe", 5)) adobe_transform_ = s[11];
          break;
        default: break;  // other APPn, COM, ...
      }
      pos_ += L;
    }
    if (!frame) fail(DH_E_ARG, "corrupt JPEG: no frame header");
    if (!scans_) fail(DH_E_ARG, "corrupt JPEG: no scan");
    if (progressive_) finish_progressive();
  }

  void output(std::vector<uint8_t>* img) const;

  int W = 0, H = 0, ncomp = 0;

 private:
  const uint8_t* f_;
  size_t n_;
  size_t pos_ = 0;
  bool progressive_ = false, jfif_ = false;
  int adobe_transform_ = -1;
  int restart_interval_ = 0;
  int scans_ = 0;
  uint16_t q_[4][64] = {};  // natural order
  bool q_defined_[4] = {};
  Huff dc_[4], ac_[4];
  Comp c_[4];
  int hmax_ = 1, vmax_ = 1, mcux_ = 0, mcuy_ = 0;
  int rgb_ids_ = 0;
  // the current scan
  int sn_ = 0, scomp_[4] = {};
  int ss_ = 0, se_ = 63, ah_ = 0, al_ = 0;
  int eob_run_ = 0;
  Bits bits_;

  // Position of the next marker at or after pos_ (skipping anything that is
  // not one); returns its code with pos_ at its 0xFF, or -1 at the end.
  int next_marker() {
    while (pos_ + 1 < n_) {
      if (f_[pos_] == 0xFF && f_[pos_ + 1] != 0xFF && f_[pos_ + 1] != 0x00) {
        const int m = f_[pos_ + 1];
        pos_ += 2;
        return m;
      }
      pos_++;
    }
    return -1;
  }
  // Length of the segment at pos_ (after its marker), checked against the file.
  size_t seg_len() {
    if (pos_ + 2 > n_) fail(DH_E_ARG, "corrupt JPEG: truncated segment");
    const size_t L = ((size_t)f_[pos_] << 8) | f_[pos_ + 1];
    if (L < 2 || pos_ + L > n_) fail(DH_E_ARG, "corrupt JPEG: bad segment length");
    return L;
  }

  void frame_header(const uint8_t* s, size_t n, bool progressive) {
    progressive_ = progressive;
    if (n < 6) fail(DH_E_ARG, "corrupt JPEG: bad SOF");
    if (s[0] != 8) fail(DH_E_UNSUPPORTED, "JPEG sample precision other than 8 bits not handled");
    H = (s[1] << 8) | s[2];
    W = (s[3] << 8) | s[4];
    ncomp = s[5];
    if (H == 0) fail(DH_E_UNSUPPORTED, "JPEG with the height in a DNL marker not handled");
    if (W == 0) fail(DH_E_ARG, "corrupt JPEG: zero width");
    if (ncomp != 1 && ncomp != 3 && ncomp != 4) fail(DH_E_ARG, "corrupt JPEG: bad component count");
    if (n < 6 + 3 * (size_t)ncomp) fail(DH_E_ARG, "corrupt JPEG: bad SOF length");
    for (int i = 0; i < ncomp; i++) {
      Comp& c = c_[i];
      c.id = s[6 + 3 * i];
      c.h = s[7 + 3 * i] >> 4;
      c.v = s[7 + 3 * i] & 15;
      c.tq = s[8 + 3 * i];
      if (ncomp == 3 && c.id == "RGB"[i]) rgb_ids_++;
      if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4) fail(DH_E_ARG, "corrupt JPEG: bad sampling factors");
      if (c.tq > 3) fail(DH_E_ARG, "corrupt JPEG: bad quantisation table index");
      hmax_ = std::max(hmax_, c.h);
      vmax_ = std::max(vmax_, c.v);
    }
    if ((uint64_t)W * (uint64_t)H * 3u > kMaxImageBytes) fail(DH_E_ARG, "JPEG larger than the image limit");
    for (int i = 0; i < ncomp; i++)
      if (hmax_ % c_[i].h || vmax_ % c_[i].v) fail(DH_E_UNSUPPORTED, "JPEG sampling factors that do not divide");
    mcux_ = (W + 8 * hmax_ - 1) / (8 * hmax_);
    mcuy_ = (H + 8 * vmax_ - 1) / (8 * vmax_);
    for (int i = 0; i < ncomp; i++) {
      Comp& c = c_[i];
      c.x = (W * c.h + hmax_ - 1) / hmax_;
      c.y = (H * c.v + vmax_ - 1) / vmax_;
      c.w2 = mcux_ * c.h * 8;
      c.h2 = mcuy_ * c.v * 8;
      c.bw = c.w2 / 8;
      c.bh = c.h2 / 8;
    }
  }

  void allocate() {
    for (int i = 0; i < ncomp; i++) {
      Comp& c = c_[i];
      if (c.plane.empty()) c.plane.assign((size_t)c.w2 * c.h2, 0);
      if (progressive_ && c.coef.empty()) c.coef.assign((size_t)c.w2 * c.h2, 0);
    }
  }

  void quant_tables(const uint8_t* s, size_t n) {
    size_t i = 0;
    while (i < n) {
      const int pq = s[i] >> 4, tq = s[i] & 15;
      if (pq > 1 || tq > 3) fail(DH_E_ARG, "corrupt JPEG: bad DQT");
      i++;
      if (i + (pq ? 128 : 64) > n) fail(DH_E_ARG, "corrupt JPEG: bad DQT length");
      for (int k = 0; k < 64; k++) {
        q_[tq][kDezigzag[k]] = pq ? (uint16_t)((s[i] << 8) | s[i + 1]) : s[i];
        i += pq ? 2 : 1;
      }
      q_defined_[tq] = true;
    }
  }

  void huffman_tables(const uint8_t* s, size_t n) {
    size_t i = 0;
    while (i < n) {
      if (i + 17 > n) fail(DH_E_ARG, "corrupt JPEG: bad DHT length");
      const int tc = s[i] >> 4, th = s[i] & 15;
      if (tc > 1 || th > 3) fail(DH_E_ARG, "corrupt JPEG: bad DHT");
      uint8_t counts[16];
      int total = 0;
      for (int k = 0; k < 16; k++) total += counts[k] = s[i + 1 + k];
      i += 17;
      if (total > 256 || i + (size_t)total > n) fail(DH_E_ARG, "corrupt JPEG: bad DHT length");
      (tc ? ac_ : dc_)[th].build(counts, s + i);
      i += total;
    }
  }

  void scan_header(const uint8_t* s, size_t n) {
    if (n < 1) fail(DH_E_ARG, "corrupt JPEG: bad SOS");
    sn_ = s[0];
    if (sn_ < 1 || sn_ > 4 || sn_ > ncomp || n < 4 + 2 * (size_t)sn_) fail(DH_E_ARG, "corrupt JPEG: bad SOS");
    for (int i = 0; i < sn_; i++) {
      const int id = s[1 + 2 * i], t = s[2 + 2 * i];
      int k = 0;
      while (k < ncomp && c_[k].id != id) k++;
      if (k == ncomp) fail(DH_E_ARG, "corrupt JPEG: scan names an unknown component");
      c_[k].td = t >> 4;
      c_[k].ta = t & 15;
      if (c_[k].td > 3 || c_[k].ta > 3) fail(DH_E_ARG, "corrupt JPEG: bad Huffman table index");
      scomp_[i] = k;
    }
    const uint8_t* t = s + 1 + 2 * sn_;
    ss_ = t[0];
    se_ = t[1];
    ah_ = t[2] >> 4;
    al_ = t[2] & 15;
    if (progressive_) {
      if (ss_ > 63 || se_ > 63 || ss_ > se_ || ah_ > 13 || al_ > 13) fail(DH_E_ARG, "corrupt JPEG: bad SOS");
      if ((ss_ == 0) != (se_ == 0)) fail(DH_E_ARG, "corrupt JPEG: a scan mixes DC and AC");
      if (ss_ > 0 && sn_ != 1) fail(DH_E_ARG, "corrupt JPEG: an AC scan of several components");
    } else {
      if (ss_ != 0 || ah_ != 0 || al_ != 0) fail(DH_E_ARG, "corrupt JPEG: bad SOS");
      se_ = 63;
    }
  }

  // Sequential block (T.81 F.2.2): DC difference, then run/size AC codes;
  // dequantized as 16-bit values and transformed into the plane.
  void block_sequential(Comp& c, int bx, int by) {
    int16_t d[64] = {};
    const uint16_t* q = q_[c.tq];
    const int t = bits_.decode(dc_[c.td]);
    if (t > 15) fail(DH_E_ARG, "corrupt JPEG: bad DC category");
    const int diff = bits_.extend(t);
    const int64_t dc = (int64_t)c.dc_pred + diff;
    if (dc < INT16_MIN || dc > INT16_MAX) fail(DH_E_ARG, "corrupt JPEG: DC out of range");
    c.dc_pred = (int)dc;
    if (dc * q[0] < INT16_MIN || dc * q[0] > INT16_MAX) fail(DH_E_ARG, "corrupt JPEG: DC out of range");
    d[0] = (int16_t)(dc * q[0]);
    for (int k = 1; k < 64;) {
      const int rs = bits_.decode(ac_[c.ta]);
      const int r = rs >> 4, s = rs & 15;
      if (s == 0) {
        if (rs != 0xF0) break;  // end of block
        k += 16;
      } else {
        k += r;
        const int z = kDezigzag[k++];
        d[z] = mul16(bits_.extend(s), q[z]);
      }
    }
    idct_block(d, &c.plane[(size_t)by * 8 * c.w2 + (size_t)bx * 8], c.w2);
  }

  // Progressive DC scans (T.81 G.1.2.1): first scan and refinement bits.
  void block_dc(Comp& c, int16_t* d) {
    if (ah_ == 0) {
      const int t = bits_.decode(dc_[c.td]);
      if (t > 15) fail(DH_E_ARG, "corrupt JPEG: bad DC category");
      const int64_t dc = (int64_t)c.dc_pred + bits_.extend(t);
      if (dc < INT16_MIN || dc > INT16_MAX) fail(DH_E_ARG, "corrupt JPEG: DC out of range");
      c.dc_pred = (int)dc;
      d[0] = (int16_t)(uint16_t)(uint32_t)(dc * (1 << al_));
    } else if (bits_.bit()) {
      d[0] = (int16_t)(d[0] + (1 << al_));
    }
  }

  // Progressive AC scans (T.81 G.1.2.2 and G.1.2.3), over one block.
  void block_ac(Comp& c, int16_t* d) {
    const Huff& h = ac_[c.ta];
    if (ah_ == 0) {  // first scan of the band
      if (eob_run_) {
        eob_run_--;
        return;
      }
      for (int k = ss_; k <= se_;) {
        const int rs = bits_.decode(h);
        const int r = rs >> 4, s = rs & 15;
        if (s == 0) {
          if (r < 15) {  // EOBr: this block and 2^r - 1 + (r bits) more end here
            eob_run_ = (1 << r) - 1;
            if (r) eob_run_ += bits_.receive(r);
            break;
          }
          k += 16;  // ZRL
        } else {
          k += r;
          const int z = kDezigzag[k++];
          d[z] = (int16_t)(uint16_t)(uint32_t)(bits_.extend(s) * (1 << al_));
        }
      }
      return;
    }
    // refinement: a correction bit for every coefficient already nonzero,
    // new coefficients of magnitude 1 placed after the run of zero ones
    const int bit = 1 << al_;
    auto refine = [&](int16_t* p) {
      if (bits_.bit() && (*p & bit) == 0) *p = (int16_t)(*p > 0 ? *p + bit : *p - bit);
    };
    int k = ss_;
    if (eob_run_ == 0) {
      for (; k <= se_;) {
        const int rs = bits_.decode(h);
        int r = rs >> 4, s = rs & 15;
        int val = 0;
        if (s == 0) {
          if (r < 15) {
            eob_run_ = 1 << r;
            if (r) eob_run_ += bits_.receive(r);
            break;  // the rest of the band is refined as part of the run
          }
          // ZRL: skip 16 zero coefficients (refining the nonzero ones passed)
        } else {
          if (s != 1) fail(DH_E_ARG, "corrupt JPEG: bad refinement code");
          val = bits_.bit() ? bit : -bit;
        }
        while (k <= se_) {
          int16_t* p = &d[kDezigzag[k++]];
          if (*p != 0) {
            refine(p);
          } else {
            if (r == 0) {
              if (val) *p = (int16_t)val;
              break;
            }
            r--;
          }
        }
      }
    }
    if (eob_run_ > 0) {
      for (; k <= se_; k++) {
        int16_t* p = &d[kDezigzag[k]];
        if (*p != 0) refine(p);
      }
      eob_run_--;
    }
  }

  void decode_unit(Comp& c, int bx, int by) {
    if (!progressive_) {
      block_sequential(c, bx, by);
      return;
    }
    int16_t* d = &c.coef[((size_t)by * c.bw + bx) * 64];
    if (ss_ == 0) block_dc(c, d);
    else block_ac(c, d);
  }

  // One scan: MCUs of the scan's components (interleaved) or the blocks of
  // its one component (non-interleaved, T.81 A.2.2), with restart intervals.
  void scan() {
    allocate();
    if (!progressive_)
      for (int i = 0; i < sn_; i++)
        if (!q_defined_[c_[scomp_[i]].tq]) fail(DH_E_ARG, "corrupt JPEG: scan before its quantisation table");
    scans_++;
    bits_.start(f_ + pos_, f_ + n_);
    for (int i = 0; i < ncomp; i++) c_[i].dc_pred = 0;
    eob_run_ = 0;
    int todo = restart_interval_ ? restart_interval_ : 0x7FFFFFFF;
    auto restart = [&]() -> bool {  // false: the data ended (the scan is over)
      if (--todo > 0) return true;
      bits_.fill();
      const int m = bits_.marker;
      if (m < 0xD0 || m > 0xD7) return false;
      // continue after the RSTn marker
      bits_.start(bits_.p, f_ + n_);
      for (int i = 0; i < ncomp; i++) c_[i].dc_pred = 0;
      eob_run_ = 0;
      todo = restart_interval_;
      return true;
    };
    if (sn_ == 1) {
      Comp& c = c_[scomp_[0]];
      const int bw = (c.x + 7) / 8, bh = (c.y + 7) / 8;
      for (int by = 0; by < bh; by++)
        for (int bx = 0; bx < bw; bx++) {
          decode_unit(c, bx, by);
          if (!restart()) goto done;
        }
    } else {
      for (int my = 0; my < mcuy_; my++)
        for (int mx = 0; mx < mcux_; mx++) {
          for (int i = 0; i < sn_; i++) {
            Comp& c = c_[scomp_[i]];
            for (int yy = 0; yy < c.v; yy++)
              for (int xx = 0; xx < c.h; xx++) decode_unit(c, mx * c.h + xx, my * c.v + yy);
          }
          if (!restart()) goto done;
        }
    }
  done:
    // the rest of the entropy-coded data up to the next marker is skipped
    pos_ = (size_t)(bits_.p - f_);
    if (bits_.marker >= 0) pos_ -= 2;
  }

  void finish_progressive() {
    for (int i = 0; i < ncomp; i++) {
      Comp& c = c_[i];
      if (!q_defined_[c.tq]) fail(DH_E_ARG, "corrupt JPEG: no quantisation table");
      const uint16_t* q = q_[c.tq];
      const int bw = (c.x + 7) / 8, bh = (c.y + 7) / 8;
      for (int by = 0; by < bh; by++)
        for (int bx = 0; bx < bw; bx++) {
          int16_t* d = &c.coef[((size_t)by * c.bw + bx) * 64];
          for (int k = 0; k < 64; k++) d[k] = mul16(d[k], q[k]);
          idct_block(d, &c.plane[(size_t)by * 8 * c.w2 + (size_t)bx * 8], c.w2);
        }
      std::vector<int16_t>().swap(c.coef);
    }
  }
};

// Chroma upsampling of one output row: in_near is the nearer component row,
// in_far the other one of the vertical pair, w the component row's width.
uint8_t* resample_1(uint8_t*, const uint8_t* in_near, const uint8_t*, int, int) { return (uint8_t*)in_near; }
uint8_t* resample_v2(uint8_t* out, const uint8_t* in_near, const uint8_t* in_far, int w, int) {
  for (int i = 0; i < w; i++) out[i] = (uint8_t)((3 * in_near[i] + in_far[i] + 2) >> 2);
  return out;
}
uint8_t* resample_h2(uint8_t* out, const uint8_t* in, const uint8_t*, int w, int) {
  if (w == 1) {
    out[0] = out[1] = in[0];
    return out;
  }
  out[0] = in[0];
  out[1] = (uint8_t)((in[0] * 3 + in[1] + 2) >> 2);
  int i = 1;
  for (; i < w - 1; i++) {
    const int n = 3 * in[i] + 2;
    out[2 * i] = (uint8_t)((n + in[i - 1]) >> 2);
    out[2 * i + 1] = (uint8_t)((n + in[i + 1]) >> 2);
  }
  out[2 * i] = (uint8_t)((in[w - 2] * 3 + in[w - 1] + 2) >> 2);
  out[2 * i + 1] = in[w - 1];
  return out;
}
uint8_t* resample_hv2(uint8_t* out, const uint8_t* in_near, const uint8_t* in_far, int w, int) {
  if (w == 1) {
    out[0] = out[1] = (uint8_t)((3 * in_near[0] + in_far[0] + 2) >> 2);
    return out;
  }
  int t1 = 3 * in_near[0] + in_far[0];
  out[0] = (uint8_t)((t1 + 2) >> 2);
  for (int i = 1; i < w; i++) {
    const int t0 = t1;
    t1 = 3 * in_near[i] + in_far[i];
    out[2 * i - 1] = (uint8_t)((3 * t0 + t1 + 8) >> 4);
    out[2 * i] = (uint8_t)((3 * t1 + t0 + 8) >> 4);
  }
  out[2 * w - 1] = (uint8_t)((t1 + 2) >> 2);
  return out;
}
uint8_t* resample_generic(uint8_t* out, const uint8_t* in_near, const uint8_t*, int w, int hs) {
  for (int i = 0; i < w; i++)
    for (int j = 0; j < hs; j++) out[i * hs + j] = in_near[i];
  return out;
}

// x * y / 255 rounded, exactly (Blinn, "Three wrongs make a right").
inline uint8_t blinn_8x8(int x, int y) {
  const unsigned t = (unsigned)(x * y + 128);
  return (uint8_t)((t + (t >> 8)) >> 8);
}

// JFIF YCbCr -> RGB in fixed point: Y in the top 12 bits of 20 with a
// rounding half, chroma products with the coefficients rounded to 12 bits
// and shifted by 8; the Cb term of G keeps only its bits above 16.
void ycbcr_to_rgb(uint8_t* out, const uint8_t* y, const uint8_t* pcb, const uint8_t* pcr, int count, int step) {
  const int kCrR = 5743 << 8, kCrG = 2925 << 8, kCbG = 1410 << 8, kCbB = 7258 << 8;  // 1.402 0.71414 0.34414 1.772
  for (int i = 0; i < count; i++, out += step) {
    const int yf = (y[i] << 20) + (1 << 19);
    const int cr = pcr[i] - 128, cb = pcb[i] - 128;
    const int r = (yf + cr * kCrR) >> 20;
    const int g = (int)(uint32_t)((uint32_t)(yf + cr * -kCrG) + ((uint32_t)(cb * -kCbG) & 0xFFFF0000u)) >> 20;
    const int b = (yf + cb * kCbB) >> 20;
    out[0] = clamp255(r);
    out[1] = clamp255(g);
    out[2] = clamp255(b);
  }
}

void Decoder::output(std::vector<uint8_t>* img) const {
  const int n = 3;  // stb_image reports 3 components for every colour JPEG
  img->assign((size_t)W * H * n, 0);
  struct Resampler {
    uint8_t* (*fn)(uint8_t*, const uint8_t*, const uint8_t*, int, int);
    const uint8_t *line0, *line1;
    int hs, vs, w_lores, ystep, ypos;
    std::vector<uint8_t> buf;
  } rs[4];
  for (int k = 0; k < ncomp; k++) {
    Resampler& r = rs[k];
    const Comp& c = c_[k];
    r.hs = hmax_ / c.h;
    r.vs = vmax_ / c.v;
    r.ystep = r.vs >> 1;
    r.w_lores = (W + r.hs - 1) / r.hs;
    r.ypos = 0;
    r.line0 = r.line1 = c.plane.data();
    r.buf.assign((size_t)W + 3, 0);
    if (r.hs == 1 && r.vs == 1) r.fn = resample_1;
    else if (r.hs == 1 && r.vs == 2) r.fn = resample_v2;
    else if (r.hs == 2 && r.vs == 1) r.fn = resample_h2;
    else if (r.hs == 2 && r.vs == 2) r.fn = resample_hv2;
    else r.fn = resample_generic;
  }
  const bool is_rgb = ncomp == 3 && (rgb_ids_ == 3 || (adobe_transform_ == 0 && !jfif_));
  const uint8_t* row[4];
  for (int j = 0; j < H; j++) {
    uint8_t* out = img->data() + (size_t)j * W * n;
    for (int k = 0; k < ncomp; k++) {
      Resampler& r = rs[k];
      const bool y_bot = r.ystep >= (r.vs >> 1);
      row[k] = r.fn(r.buf.data(), y_bot ? r.line1 : r.line0, y_bot ? r.line0 : r.line1, r.w_lores, r.hs);
      if (++r.ystep >= r.vs) {
        r.ystep = 0;
        r.line0 = r.line1;
        if (++r.ypos < c_[k].y) r.line1 += c_[k].w2;
      }
    }
    if (ncomp == 3) {
      if (is_rgb) {
        for (int i = 0; i < W; i++) {
          out[3 * i] = row[0][i];
          out[3 * i + 1] = row[1][i];
          out[3 * i + 2] = row[2][i];
        }
      } else {
        ycbcr_to_rgb(out, row[0], row[1], row[2], W, n);
      }
    } else {  // ncomp == 4
      if (adobe_transform_ == 0) {  // CMYK
        for (int i = 0; i < W; i++) {
          const int m = row[3][i];
          out[3 * i] = blinn_8x8(row[0][i], m);
          out[3 * i + 1] = blinn_8x8(row[1][i], m);
          out[3 * i + 2] = blinn_8x8(row[2][i], m);
        }
      } else if (adobe_transform_ == 2) {  // YCCK
        ycbcr_to_rgb(out, row[0], row[1], row[2], W, n);
        for (int i = 0; i < W; i++) {
          const int m = row[3][i];
          for (int q = 0; q < 3; q++) out[3 * i + q] = blinn_8x8(255 - out[3 * i + q], m);
        }
      } else {  // YCbCr and a fourth channel, which is ignored
        ycbcr_to_rgb(out, row[0], row[1], row[2], W, n);
      }
    }
  }
}

}  // namespace

int jpeg_decode(const uint8_t* f, size_t n, std::vector<uint8_t>* img, int32_t* W, int32_t* H, int32_t* comp,
                std::string* err, bool info_only) {
  Decoder d(f, n);
  try {
    d.run(info_only);
    if (d.ncomp == 1) {
      *err = "grey JPEG: the reference's loader takes only 3- or 4-component images";
      return DH_E_UNSUPPORTED;
    }
    if (!info_only) d.output(img);
  } catch (const Fail& e) {
    *err = e.msg;
    return e.rc;
  }
  *W = d.W;
  *H = d.H;
  *comp = 3;
  return DH_OK;
}

}  // namespace dh
