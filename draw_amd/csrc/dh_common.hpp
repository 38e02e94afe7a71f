// dh_common.hpp -- shared types and exact-arithmetic helpers for libdrawhip.
//
// Every floating-point helper here reproduces one operation sequence of the
// reference renderer (mororo18/draw) in IEEE binary32, in the reference's
// order.  The library is compiled with -ffp-contract=off, correctly rounded
// f32 division/sqrt and f32 denormals preserved, so each `a*b + c` below is two
// roundings, as in rustc's output.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/drawhip.h"

typedef dh_frame_params FrameParamsPod;

namespace dh {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kSkip = 0xFFFFFFFEu;  // a pixel another shard shades (dh_shade_shard)
constexpr int kShardShift = 29;          // packed visibility id: shard << 29 | clipped-triangle id
constexpr uint32_t kShardMax = 8;
constexpr float kEps = 0.0000001f;  // linalg.rs:6

// Setup record flags.
enum : uint32_t {
  kTieA = 1u, kTieB = 2u, kTieG = 4u,  // f * f_outside > 0 (canvas.rs:678-680)
  kFast = 8u,                          // all normalizers finite and nonzero
  kDrop = 16u,                         // provably writes no pixel: not binned
  kUvSafe = 32u,                       // every covered fragment's texel offset is in range
};

// Status words written by the kernels and read back once per frame.
constexpr uint32_t kFragWords = 32;

struct Status {
  uint32_t n_out;     // clipped triangles emitted (may exceed capacity)
  uint32_t n_pairs;   // (tile, triangle) bin entries (may exceed capacity)
  uint32_t overflow;  // bit0: triangle records, bit1: bin entries
  uint32_t error;     // bit0: texture slice out of range (scene/mod.rs:167 panic)
  uint32_t n_simple;  // input triangles emitted unclipped (inside the near and far planes)
  uint32_t n_clip;    // input triangles that emit after near/far clipping
  uint32_t n_wide;    // triangles on the wide-triangle bin list
  // covered fragments submitted to the visibility test: the sum of these
  // words, tile t adding to word t % kFragWords (one address for every tile
  // of a large batch serializes their atomics in memory)
  uint32_t frags[kFragWords];
};

// Material = the reference's `Texture` (scene/mod.rs:206-216) with device maps.
struct DevMaterial {
  float ka[3], kd[3], ks[3];
  float alpha;
  int32_t power;
  int32_t pad;
  const uint8_t* kd_img;
  const uint8_t* ka_img;
  int32_t kd_w, kd_h, kd_comp;
  int32_t ka_w, ka_h, ka_comp;
};

// One draw = one IndexedMesh of one Object, in submission order.
constexpr uint32_t kDrawNoCull = 1u;
// Per-triangle vertex data of an opaque mesh, built once when the mesh is
// created, in one allocation at soa:
//  - streams: stream s of triangle t at soa[s * n_tri + t]; streams 0-8 hold
//    the positions of vertices a, b, c (x, y, z each), 9-17 their normals,
//    18-23 their texture coordinates (u, v each).  A wave whose triangles are
//    consecutive reads each value with one coalesced load instead of
//    gathering through the index triplets;
//  - records: the same 24 values of triangle t as one 128-byte record (then 32 bytes
//    of padding), 128-byte aligned, at float offset
//    mesh_records_offset(n_tri) + 32 t.  A wave whose triangles lie far apart
//    in the mesh (survivors of culling scattered among culled ones) reads one
//    cache line per triangle there, where each stream load would fetch a
//    line per lane for one value.
constexpr uint32_t kMeshStreams = 24;
constexpr uint32_t kMeshRecFloats = 32;
__host__ __device__ inline uint64_t mesh_records_offset(uint32_t n_tri) {
  return ((uint64_t)kMeshStreams * n_tri + (kMeshRecFloats - 1u)) & ~(uint64_t)(kMeshRecFloats - 1u);
}
__host__ __device__ inline uint64_t mesh_stream_floats(uint32_t n_tri) {
  return mesh_records_offset(n_tri) + (uint64_t)kMeshRecFloats * n_tri;
}

struct DevDraw {
  const float* pos;       // xyz
  const float* nrm;       // xyz
  const float* uv;        // uv
  const uint32_t* tri;    // 9 per triangle (pos a,b,c; uv a,b,c; nrm a,b,c)
  const float* soa;       // opaque meshes: vertex streams and records (mesh_stream_floats), else null
  uint32_t n_tri;
  uint32_t tri_begin;     // first input-triangle index of this draw
  uint32_t material;      // slot in the material table
  uint32_t view;          // view slot (batched rendering); 0 for one frame
  uint32_t flags;         // kDrawNoCull: transparent meshes skip back-face culling
  uint32_t group;         // object group (submission order of Scene::render's object loop)
};

// Per-view raster geometry (one entry for a single frame).
struct DevView {
  FrameParamsPod fr;      // uniforms of this view
  uint32_t tile_base;     // first global tile index of this view
  uint32_t ntx, nty;      // tiles per row / column
  uint32_t px_base;       // first pixel of this view in the visibility buffer
  uint32_t* out_bgra;     // canvas frame, W*H BGRA8 words, rows top-down
  float* out_depth;       // canvas depth, W*H, rows bottom-up
  uint32_t* out_winner;   // canvas winner ids (may be null), rows bottom-up
  uint32_t tri_begin;     // first input triangle of this view
  uint32_t id_base;       // first clipped-triangle id of this view (set on the device)
  uint32_t depth_base;    // float bits of half the camera's near-plane distance (near-first bin order)
};

// Raster record of one clipped triangle: the edge-function constants of the
// snapped vertices (canvas.rs:597-616), normalizers, depths, scan rectangle.
struct alignas(16) SetupRec {
  float ea[4];            // edge bc (alpha): A, B, C, D
  float eb[4];            // edge ca (beta)
  float eg[4];            // edge ab (gamma)
  float f[4];             // f_alpha, f_beta, f_gama, (unused)
  float d[4];             // a_depth, b_depth, c_depth, (unused)
  uint32_t rx, ry;        // x0 | x1 << 16, y0 | y1 << 16 (inclusive scan rectangle)
  uint32_t flags, view;
};
static_assert(sizeof(SetupRec) == 96, "SetupRec layout");

// Shading attributes of one clipped triangle.
struct alignas(16) ShadeRec {
  float uv[6];            // uA vA uB vB uC vC
  uint32_t material, group;  // group: the draw's object group
  float n[9], l[9], h[9]; // per-vertex normal, light, halfway
  float pad1;
};
static_assert(sizeof(ShadeRec) == 144, "ShadeRec layout");

// ------------------------------------------------------------ f32 helpers
struct V3 { float x, y, z; };
__host__ __device__ inline V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
__host__ __device__ inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline V3 muls(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline V3 divs(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
// Vec3::dot linalg.rs:182-184
__host__ __device__ inline float dot(V3 a, V3 b) { float s = a.x * b.x + a.y * b.y; return s + a.z * b.z; }
// Vec3::norm linalg.rs:167-171
__host__ __device__ inline float norm(V3 a) { float s = a.x * a.x + a.y * a.y; s = s + a.z * a.z; return sqrtf(s); }
__host__ __device__ inline V3 normalized(V3 a) { return divs(a, norm(a)); }
// Vec3::cross linalg.rs:186-200
__host__ __device__ inline V3 cross(V3 a, V3 b) {
  return v3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}

// Edge function with precomputed constants: ((A*x + B*y) + C) - D, the order
// of the closures at canvas.rs:597-616 (A = p.y-q.y, B = q.x-p.x, C = p.x*q.y,
// D = q.x*p.y).
__host__ __device__ inline float edge_eval(const float e[4], float x, float y) {
  return ((e[0] * x + e[1] * y) + e[2]) - e[3];
}

// Coverage of one pixel for a FAST triangle, decided without division.  Edge
// values are integers here (snapped vertices are floor() results and pixel
// centres are integers), so for a finite nonzero normalizer f the quotient
// e/f is NaN iff e is NaN, zero iff e is zero, and otherwise nonzero with the
// sign of e*f (|e/f| >= 1/FLT_MAX is representable as a denormal).  Hence
// alpha >= 0 && (alpha > 0 || tie) <=> e*f > 0 || (e == 0 && tie).
__device__ inline bool covered_fast(float e, float f, bool tie) { return (e * f > 0.0f) || (e == 0.0f && tie); }

// Rust `f32 as usize` on a 64-bit target: NaN and negatives -> 0, saturating.
__host__ __device__ inline uint64_t f2usize(float v) {
  if (!(v > 0.0f)) return 0;
  if (v >= 18446744073709551616.0f) return ~0ull;
  return (uint64_t)v;
}
// Rust `f32 as u8`.
__host__ __device__ inline uint32_t f2u8(float v) {
  if (!(v > 0.0f)) return 0;
  if (v >= 255.0f) return 255;
  return (uint32_t)v;
}

// Byte offset of TextureMap::get_rgb_slice (scene/mod.rs:154-168) with Rust
// release-mode wrapping usize arithmetic; returns false where the reference's
// bounds-checked slice would panic.
__host__ __device__ inline bool rgb_offset(int32_t w, int32_t h, int32_t comp, float u, float v, uint64_t* off) {
  float fw = (float)w, fh = (float)h;
  uint64_t ui = f2usize(floorf(u * (fw - 1.0f)));
  uint64_t vi = ((uint64_t)h - 1u) - f2usize(floorf(v * (fh - 1.0f)));
  uint64_t o = (vi * (uint64_t)w + ui) * (uint64_t)comp;
  uint64_t end = o + 3u;
  uint64_t len = (uint64_t)w * (uint64_t)h * (uint64_t)comp;
  *off = o;
  return !(o > end || end > len);
}

// Order-preserving 32-bit key of a depth that is not NaN; +0 and -0 share a
// key (the strict `<` of canvas.rs:923 treats them as equal).
__device__ inline uint32_t depth_key(float d) {
  float c = d + 0.0f;  // -0 -> +0, every other value unchanged
  uint32_t b = __float_as_uint(c);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Views are stored in increasing tile_base / px_base order: the view owning a
// global tile or pixel index is the last one whose base is <= it.  A whole
// wave searches at once, for a t (or p0) that every lane holds: each round
// the 64 lanes load evenly spaced candidates and a ballot keeps the span from
// the last one at or below t to the next, so two rounds of loads cover up to
// 4096 views where a binary search takes log2(n) dependent loads.  Every lane
// of the wave must be active.
template <typename Key>
__device__ inline uint32_t wave_find_view(const DevView* views, uint32_t n, uint32_t t, Key key) {
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t lo = 0, len = n;  // views[lo] is at or below t (views[0] starts at 0)
  while (len > 1) {
    const uint32_t step = (len + 63u) >> 6;
    const uint32_t j = lane * step;
    const bool le = j < len && key(views[lo + j]) <= t;
    const uint32_t k = (uint32_t)__popcll(__ballot(le)) - 1u;  // lane 0 is always set
    lo += k * step;
    len = min(step, len - k * step);
  }
  return lo;
}
__device__ inline uint32_t wave_find_view_tile(const DevView* views, uint32_t n, uint32_t t) {
  return wave_find_view(views, n, t, [](const DevView& v) { return v.tile_base; });
}
// View of pixel p >= p0, where p0 (the same in every lane) is the wave's
// first pixel: the view of p0, then the (rare) view starts between p0 and p.
__device__ inline uint32_t wave_find_view_px(const DevView* views, uint32_t n, uint32_t p0, uint32_t p) {
  uint32_t v = wave_find_view(views, n, p0, [](const DevView& w) { return w.px_base; });
  while (v + 1u < n && views[v + 1u].px_base <= p) v++;
  return v;
}

// XCD-aware block order: workgroups are dispatched round-robin over the 8
// XCDs, so give each XCD one contiguous run of logical indices (neighbouring
// tiles share triangles and their records then share one L2).
__device__ inline uint32_t xcd_remap(uint32_t b, uint32_t nb) {
  const uint32_t xcd = b & 7u, k = b >> 3;
  const uint32_t per = nb >> 3, rem = nb & 7u;
  return (xcd < rem) ? xcd * (per + 1u) + k : rem * (per + 1u) + (xcd - rem) * per + k;
}

// f32::powi with a non-negative exponent (binary exponentiation, low bit
// first; power 2 is x*x as at canvas.rs:739).
__host__ __device__ inline float powi_f(float x, int p) {
  float r = 1.0f;
  bool first = true;
  while (p) {
    if (p & 1) { r = first ? x : r * x; first = false; }
    p >>= 1;
    if (p) x = x * x;
  }
  return r;
}

}  // namespace dh
