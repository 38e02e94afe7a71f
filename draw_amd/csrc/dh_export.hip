// dh_export.hip -- frame export (App::export_frame_as, app/mod.rs:316-379):
// the canvas's BGRA8 frame with bytes 0 and 2 of every pixel swapped
// (app/mod.rs:347-351), rows top-down as the canvas stores them, then
// written as an 8-bit RGBA PNG (the reference uses stb_image_write).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <zlib.h>

#include <mutex>
#include <vector>

#include "dh_common.hpp"
#include "dh_internal.hpp"

namespace dh {

// One thread per 4 pixels; the tail of a row-major frame handled per pixel.
__global__ __launch_bounds__(256) void k_bgra_to_rgba(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                       uint32_t n) {
  const uint32_t i = xcd_remap(blockIdx.x, gridDim.x) * 256u + threadIdx.x;
  auto swap02 = [](uint32_t p) { return (p & 0xFF00FF00u) | ((p & 0xFFu) << 16) | ((p >> 16) & 0xFFu); };
  const uint32_t n4 = n / 4u;
  if (i < n4) {
    uint4 v = reinterpret_cast<const uint4*>(src)[i];
    v.x = swap02(v.x);
    v.y = swap02(v.y);
    v.z = swap02(v.z);
    v.w = swap02(v.w);
    reinterpret_cast<uint4*>(dst)[i] = v;
  } else if (i < n4 + (n & 3u)) {
    const uint32_t k = 4u * n4 + (i - n4);
    dst[k] = swap02(src[k]);
  }
}

hipError_t enqueue_bgra_to_rgba(const uint32_t* src, uint32_t* dst, uint32_t n, hipStream_t s) {
  const uint32_t threads = n / 4u + (n & 3u);
  if (threads == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bgra_to_rgba, dim3((threads + 255u) / 256u), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

}  // namespace dh

// ---- fixed RGBA staging for dh_canvas_read_rgba_noalloc ----------------
// The staging buffer of the export, as a file-scope device array: every GPU
// that loads this library holds its own copy (31.6 MiB, one 3840x2160 frame)
// until the process ends, so the entry point that uses it allocates nothing.
static __device__ __align__(256) uint32_t scratch[dh::kRgbaScratchPixels];

namespace {

constexpr int kMaxDevices = 64;

// Per device: the array's device address (from hipGetSymbolAddress with that
// device current; the host-side name is not a device address), the status of
// that lookup, and a lock that serialises the calls sharing the array.
struct ScratchSlot {
  std::once_flag once;
  hipError_t status = hipErrorNotInitialized;
  uint32_t* ptr = nullptr;
  std::mutex mu;
};
ScratchSlot g_scratch[kMaxDevices];

}  // namespace

namespace dh {

hipError_t rgba_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  ScratchSlot& s = g_scratch[device];
  std::call_once(s.once, [&] {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != device) e = hipSetDevice(device);
    void* p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p == nullptr) e = hipErrorInvalidSymbol;
    if (e == hipSuccess) s.ptr = static_cast<uint32_t*>(p);
    s.status = e;
    if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
  });
  return s.status;
}

hipError_t rgba_scratch(int device, uint32_t** ptr, std::mutex** mu) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  const hipError_t e = rgba_scratch_init(device);
  if (e != hipSuccess) return e;
  *ptr = g_scratch[device].ptr;
  *mu = &g_scratch[device].mu;
  return hipSuccess;
}

}  // namespace dh

namespace {

void put_be32(std::vector<uint8_t>& out, uint32_t v) {
  out.push_back((uint8_t)(v >> 24));
  out.push_back((uint8_t)(v >> 16));
  out.push_back((uint8_t)(v >> 8));
  out.push_back((uint8_t)v);
}

void put_chunk(std::vector<uint8_t>& out, const char type[4], const uint8_t* data, size_t n) {
  put_be32(out, (uint32_t)n);
  const size_t start = out.size();
  out.insert(out.end(), type, type + 4);
  if (n) out.insert(out.end(), data, data + n);
  const uLong crc = crc32(0L, out.data() + start, (uInt)(n + 4));
  put_be32(out, (uint32_t)crc);
}

// PNG (8-bit RGBA, no interlace; every row filter 0) of W*H pixels, rows
// top-down, into `out`.
int png_encode(const uint8_t* rgba, int32_t W, int32_t H, std::vector<uint8_t>* out) {
  if (!rgba || W <= 0 || H <= 0) return DH_E_ARG;
  const size_t row = 4ull * (size_t)W;
  std::vector<uint8_t> raw((row + 1) * (size_t)H);
  for (int32_t y = 0; y < H; y++) {
    raw[(row + 1) * y] = 0;
    memcpy(&raw[(row + 1) * y + 1], rgba + row * y, row);
  }
  uLongf zn = compressBound((uLong)raw.size());
  std::vector<uint8_t> z(zn);
  if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), 6) != Z_OK) return DH_E_ARG;
  out->clear();
  const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
  out->insert(out->end(), sig, sig + 8);
  std::vector<uint8_t> ihdr;
  put_be32(ihdr, (uint32_t)W);
  put_be32(ihdr, (uint32_t)H);
  const uint8_t rest[5] = {8, 6, 0, 0, 0};  // bit depth 8, colour type RGBA, deflate, filter 0, no interlace
  ihdr.insert(ihdr.end(), rest, rest + 5);
  put_chunk(*out, "IHDR", ihdr.data(), ihdr.size());
  put_chunk(*out, "IDAT", z.data(), zn);
  put_chunk(*out, "IEND", nullptr, 0);
  return DH_OK;
}

}  // namespace

extern "C" int dh_write_png(const char* path, const uint8_t* rgba, int32_t W, int32_t H) {
  if (!path) return DH_E_ARG;
  std::vector<uint8_t> png;
  const int rc = png_encode(rgba, W, H, &png);
  if (rc) return rc;
  FILE* f = fopen(path, "wb");
  if (!f) return DH_E_ARG;
  const size_t w = fwrite(png.data(), 1, png.size(), f);
  const int cl = fclose(f);
  return (w == png.size() && cl == 0) ? DH_OK : DH_E_ARG;
}
