// dh_kernels.hip -- gfx950 kernels of the per-frame raster path.
//
// Pipeline (one stream, no host synchronisation inside a frame):
//   1 k_geom_count   per input triangle: back-face cull, frustum clip count
//                    (scene/mod.rs:1016-1033, 43-90, 641-746), positions only,
//                    and whether it crosses the near or far plane
//   2 scan, k_geom_compact  block counts -> ordered output offsets (submission
//                    order) and two dense lists of emitting input triangles:
//                    unclipped, clipped
//   3 k_geom_emit    per listed triangle: per-vertex lighting (917-926), clip
//                    with attributes (clipped list only), projection
//                    (1047-1063), raster setup (canvas.rs:585-666) -> SetupRec +
//                    ShadeRec at its output id; with ordered bins also the
//                    conservative tile count, and the near-first depth key
//   4 k_wide_list, k_bin_count_wide  triangles spanning many tiles: listed,
//                    then counted a wave each (k_bin_count counts per tile for
//                    unordered bins)
//   5 near first     stable 8-bit radix pass of the triangles by depth key,
//                    their counts gathered in that order
//   6 scan           -> bin entry offsets
//   7 k_bin_emit(_wide)  (tile, triangle) pairs in that order
//   8 radix sort     stable LSD sort of the pairs by tile -> per-tile lists,
//                    near first (or in submission order)
//   9 tile ranges    [start, end) of every tile's list: written by the sort's
//                    scatter from its scanned histogram when the tile key
//                    takes one pass, else by k_tile_ranges from the sorted keys
//  10 k_raster       one workgroup per screen tile: coverage and the strict
//                    depth test (canvas.rs:668-682, 923) into an LDS visibility
//                    buffer of (depth key << 32 | triangle id); the result does
//                    not depend on list order
//  11 k_shade        deferred Phong + texture shading of each pixel's winner
//                    (canvas.rs:684-745), BGRA8 + f32 depth
#include <math.h>
#include <stdio.h>

#include <algorithm>

#include "dh_common.hpp"
#include "dh_pipeline.hpp"

namespace dh {

// ===================================================================== common
__device__ inline uint32_t find_draw(const DevDraw* draws, uint32_t n, uint32_t t) {
  uint32_t lo = 0, hi = n;  // last draw with tri_begin <= t
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (draws[mid].tri_begin <= t) lo = mid; else hi = mid;
  }
  return lo;
}

// The draw of input triangle t for every active lane of a wave (every lane
// calls it).  The lanes' triangles increase with the lane (ordered lists), so
// the draw of the first active lane's triangle, found with scalar loads, is
// every lane's whenever the last active lane's triangle lies before the next
// draw; otherwise each lane searches.
__device__ inline uint32_t wave_find_draw(const DevDraw* draws, uint32_t n, uint32_t t, bool active) {
  const uint64_t m = __ballot(active);
  if (!m) return 0u;
  const uint32_t t0 = (uint32_t)__builtin_amdgcn_readlane((int)t, __builtin_ctzll(m));
  const uint32_t t1 = (uint32_t)__builtin_amdgcn_readlane((int)t, 63 - __builtin_clzll(m));
  const uint32_t d0 = find_draw(draws, n, t0);
  if (d0 + 1u >= n || t1 < draws[d0 + 1u].tri_begin) return d0;
  return find_draw(draws, n, t);
}

__device__ inline V3 ld3(const float* a, uint32_t i) {
  const float* p = a + 3ull * i;
  return v3(p[0], p[1], p[2]);
}

// Vertex positions of a draw's triangle lt (a, b, c): from the mesh's
// per-triangle streams when it has them, else through its index triplet.
__device__ inline void tri_positions(const DevDraw& dr, uint32_t lt, V3& a, V3& b, V3& c) {
  if (dr.soa) {
    const uint64_t n = dr.n_tri;
    const float* s = dr.soa + lt;
    a = v3(s[0], s[n], s[2 * n]);
    b = v3(s[3 * n], s[4 * n], s[5 * n]);
    c = v3(s[6 * n], s[7 * n], s[8 * n]);
  } else {
    const uint32_t* ti = dr.tri + 9ull * lt;
    a = ld3(dr.pos, ti[0]);
    b = ld3(dr.pos, ti[1]);
    c = ld3(dr.pos, ti[2]);
  }
}

// Whether a wave reads its triangles' vertex records rather than the streams:
// when the input triangles of its active lanes span more than `span`, a
// stream load fetches several cache lines for few values each, and one
// 128-byte record per triangle moves fewer bytes.  The 24 streams of a wave
// spanning S triangles move about 1.5 S bytes per triangle, a record 128, but
// a record costs each lane its own lines: kSparseSpan (eight triangles per
// lane) by default, and kSparseSpanDense on frames of at least one input
// triangle per pixel, whose micro-triangles' culling survivors interleave
// finely (about 45% of them, a span near 145).  The choice only selects which
// copy of the same values is read.  Every lane of the wave calls it.
constexpr uint32_t kSparseSpan = 512u, kSparseSpanDense = 96u;
__device__ inline bool wave_sparse(bool active, uint32_t t, uint32_t span) {
  const uint64_t m = __ballot(active);
  if (!m) return false;
  const int first = __builtin_ctzll(m), last = 63 - __builtin_clzll(m);
  const uint32_t t0 = (uint32_t)__builtin_amdgcn_readlane((int)t, first);
  const uint32_t t1 = (uint32_t)__builtin_amdgcn_readlane((int)t, last);
  return t1 - t0 > span;
}

// Positions, normals and texture coordinates of the vertices of triangle lt:
// from the mesh's records (sparse) or streams when it has them, else through
// its index triplet.
__device__ inline void tri_vertices(const DevDraw& dr, uint32_t lt, bool sparse, V3 (&P)[3], V3 (&nrm)[3],
                                    float (&u)[3], float (&v)[3]) {
  if (dr.soa && sparse) {
    const uint4* r = reinterpret_cast<const uint4*>(dr.soa + mesh_records_offset(dr.n_tri)) +
                     (uint64_t)(kMeshRecFloats / 4u) * lt;
    float f[24];
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const uint4 x = r[q];
      f[4 * q] = __uint_as_float(x.x);
      f[4 * q + 1] = __uint_as_float(x.y);
      f[4 * q + 2] = __uint_as_float(x.z);
      f[4 * q + 3] = __uint_as_float(x.w);
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
      P[k] = v3(f[3 * k], f[3 * k + 1], f[3 * k + 2]);
      nrm[k] = v3(f[9 + 3 * k], f[10 + 3 * k], f[11 + 3 * k]);
      u[k] = f[18 + 2 * k];
      v[k] = f[19 + 2 * k];
    }
  } else if (dr.soa) {
    const uint64_t n = dr.n_tri;
    const float* s = dr.soa + lt;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      P[k] = v3(s[(3 * k) * n], s[(3 * k + 1) * n], s[(3 * k + 2) * n]);
      nrm[k] = v3(s[(9 + 3 * k) * n], s[(10 + 3 * k) * n], s[(11 + 3 * k) * n]);
      u[k] = s[(18 + 2 * k) * n];
      v[k] = s[(19 + 2 * k) * n];
    }
  } else {
    const uint32_t* ti = dr.tri + 9ull * lt;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      P[k] = ld3(dr.pos, ti[k]);
      nrm[k] = ld3(dr.nrm, ti[6 + k]);
      const float* uv = dr.uv + 2ull * ti[3 + k];
      u[k] = uv[0];
      v[k] = uv[1];
    }
  }
}

// Builds a mesh's vertex streams and records (once, when the mesh is created).
__global__ __launch_bounds__(256) void k_mesh_streams(const uint32_t* __restrict__ tri, uint32_t n_tri,
                                                      const float* __restrict__ pos, const float* __restrict__ nrm,
                                                      const float* __restrict__ uv, float* __restrict__ soa) {
  const uint64_t n = n_tri;
  uint4* rec = reinterpret_cast<uint4*>(soa + mesh_records_offset(n_tri));
  for (uint32_t t = blockIdx.x * 256u + threadIdx.x; t < n_tri; t += gridDim.x * 256u) {
    const uint32_t* q = tri + 9ull * t;
    float f[kMeshRecFloats];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float* p = pos + 3ull * q[k];
      const float* m = nrm + 3ull * q[6 + k];
      const float* w = uv + 2ull * q[3 + k];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        f[3 * k + j] = p[j];
        f[9 + 3 * k + j] = m[j];
      }
      f[18 + 2 * k] = w[0];
      f[19 + 2 * k] = w[1];
    }
#pragma unroll
    for (int k = 24; k < (int)kMeshRecFloats; k++) f[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < (int)kMeshStreams; k++) soa[(uint64_t)k * n + t] = f[k];
#pragma unroll
    for (int k = 0; k < (int)kMeshRecFloats / 4; k++)
      rec[(uint64_t)(kMeshRecFloats / 4u) * t + k] = make_uint4(
          __float_as_uint(f[4 * k]), __float_as_uint(f[4 * k + 1]), __float_as_uint(f[4 * k + 2]),
          __float_as_uint(f[4 * k + 3]));
  }
}

// Block of 256 threads (4 waves): exclusive scan of one value per thread.
__device__ inline uint32_t block_exclusive_scan256(uint32_t v, uint32_t* s_tmp /*[4]*/, uint32_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_tmp[wave] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < 4; w++) {
    uint32_t t = s_tmp[w];
    if (w < wave) before += t;
    all += t;
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

// =================================================================== clipping
// Clip vertex with positions only (for counting) and with all attributes.
struct PV {
  V3 p;
};
struct FV {
  V3 p;
  float d;
  V3 n, l, h;
  float u, v;
};

// VertexAttributes a + (c - a) * t, canvas.rs:242-291 (screen_coord is still
// (0,0) at this stage and is overwritten by the projection).
__device__ inline PV lerpv(const PV& a, const PV& c, float t) {
  PV r;
  r.p = add(a.p, muls(sub(c.p, a.p), t));
  return r;
}
__device__ inline FV lerpv(const FV& a, const FV& c, float t) {
  FV r;
  r.p = add(a.p, muls(sub(c.p, a.p), t));
  r.d = a.d + ((c.d - a.d) * t);
  r.n = add(a.n, muls(sub(c.n, a.n), t));
  r.l = add(a.l, muls(sub(c.l, a.l), t));
  r.h = add(a.h, muls(sub(c.h, a.h), t));
  r.u = a.u + ((c.u - a.u) * t);
  r.v = a.v + ((c.v - a.v) * t);
  return r;
}

// Member-wise select (keeps the clip vertices in registers).
__device__ inline V3 sel3(bool q, V3 x, V3 y) { return v3(q ? x.x : y.x, q ? x.y : y.y, q ? x.z : y.z); }
__device__ inline PV selv(bool q, const PV& x, const PV& y) {
  PV r;
  r.p = sel3(q, x.p, y.p);
  return r;
}
__device__ inline FV selv(bool q, const FV& x, const FV& y) {
  FV r;
  r.p = sel3(q, x.p, y.p);
  r.d = q ? x.d : y.d;
  r.n = sel3(q, x.n, y.n);
  r.l = sel3(q, x.l, y.l);
  r.h = sel3(q, x.h, y.h);
  r.u = q ? x.u : y.u;
  r.v = q ? x.v : y.v;
  return r;
}

__device__ inline float plane_f(const float* pl, V3 p) { return dot(v3(pl[0], pl[1], pl[2]), p) + pl[3]; }

// ViewPlane::at_least_partially_visible, scene/mod.rs:641-660.
__device__ inline bool part_visible(const float* pl, V3 a, V3 b, V3 c) {
  float fa = plane_f(pl, a), fb = plane_f(pl, b), fc = plane_f(pl, c);
  return !(fa <= 0.0f && fb <= 0.0f && fc <= 0.0f);
}

// ViewPlane::clip, scene/mod.rs:662-746.  First output (o0,o1,o2), second
// output (p0,p1,p2); returns the count.
template <class V>
__device__ inline int clip_plane(const float* pl, V a, V b, V c, V& o0, V& o1, V& o2, V& p0, V& p1, V& p2) {
  float fa = plane_f(pl, a.p), fb = plane_f(pl, b.p), fc = plane_f(pl, c.p);
  if (fa > 0.0f && fb > 0.0f && fc > 0.0f) {
    o0 = a; o1 = b; o2 = c;
    return 1;
  }
  if (fa <= 0.0f && fb <= 0.0f && fc <= 0.0f) return 0;
  // rotate so that c is the vertex alone on its side: r1 (a,b,c) <- (c,a,b),
  // r2 (a,b,c) <- (b,c,a)
  const bool r1 = fa * fc >= 0.0f;
  const bool r2 = !r1 && (fb * fc >= 0.0f);
  const V A = selv(r1, c, selv(r2, b, a));
  const V B = selv(r1, a, selv(r2, c, b));
  const V C = selv(r1, b, selv(r2, a, c));
  fc = r1 ? fb : (r2 ? fa : fc);
  a = A;
  b = B;
  c = C;
  const V3 n = v3(pl[0], pl[1], pl[2]);
  const float ta = plane_f(pl, a.p) / dot(n, sub(a.p, c.p)) - kEps;
  const V na = lerpv(a, c, ta);
  const float tb = plane_f(pl, b.p) / dot(n, sub(b.p, c.p)) - kEps;
  const V nb = lerpv(b, c, tb);
  // one vertex outside (f_c <= 0): [a, A', B'] and [a, b, B']; else [c, A', B']
  const bool two = fc <= 0.0f;
  o0 = selv(two, a, c);
  o1 = na;
  o2 = nb;
  p0 = a;
  p1 = b;
  p2 = nb;
  return two ? 2 : 1;
}

// Back-face test (scene/mod.rs:1016-1020) and Triangle::clip_against_planes
// (43-90): lateral trivial reject, then near and far full clipping.  Calls
// emit(k, v0, v1, v2) for each output in order; returns the count.
template <class V, class Emit>
__device__ inline int cull_and_clip(const FrameParamsPod& fr, const V& a, const V& b, const V& c, bool cull,
                                    Emit emit) {
  if (cull) {  // the transparent pass does not cull (scene/mod.rs:1194-1203)
    const V3 cam = v3(fr.cam[0], fr.cam[1], fr.cam[2]);
    const V3 nrm = cross(sub(b.p, a.p), sub(c.p, b.p));                             // Triangle::calc_normal
    const V3 ctr = divs(add(add(add(v3(0.0f, 0.0f, 0.0f), a.p), b.p), c.p), 3.0f);  // get_center
    if (dot(sub(cam, ctr), nrm) <= 0.0f) return 0;
  }
  if (!(part_visible(fr.planes[2], a.p, b.p, c.p) && part_visible(fr.planes[3], a.p, b.p, c.p) &&
        part_visible(fr.planes[4], a.p, b.p, c.p) && part_visible(fr.planes[5], a.p, b.p, c.p)))
    return 0;
  V n0, n1, n2, m0, m1, m2;
  const int nn = clip_plane(fr.planes[0], a, b, c, n0, n1, n2, m0, m1, m2);
  int cnt = 0;
  if (nn >= 1) {
    V f0, f1, f2, g0, g1, g2;
    const int nf = clip_plane(fr.planes[1], n0, n1, n2, f0, f1, f2, g0, g1, g2);
    if (nf >= 1) emit(cnt++, f0, f1, f2);
    if (nf >= 2) emit(cnt++, g0, g1, g2);
  }
  if (nn >= 2) {
    V f0, f1, f2, g0, g1, g2;
    const int nf = clip_plane(fr.planes[1], m0, m1, m2, f0, f1, f2, g0, g1, g2);
    if (nf >= 1) emit(cnt++, f0, f1, f2);
    if (nf >= 2) emit(cnt++, g0, g1, g2);
  }
  return cnt;
}

// ====================================================================== setup
// Rectangle::from_coords + Rectangle::clip along one axis (canvas.rs:315-350,
// 646-658).
__device__ inline void rect_axis(uint64_t a0, uint64_t a1, uint64_t smax, uint32_t* lo, uint32_t* hi) {
  uint64_t mn = a0 < a1 ? a0 : a1, mx = a0 < a1 ? a1 : a0;
  mx = mx < smax ? mx : smax;
  if (mn > mx) { mn = 0; mx = 0; }
  *lo = (uint32_t)mn;
  *hi = (uint32_t)mx;
}

// Projection (scene/mod.rs:1047-1058: Matrix4 * Vec4 then as_vec2()/w),
// snapping (canvas.rs:585-587, 896-904), edge constants, normalizers,
// tie-break products (660-666) and the scan rectangle (640-658).  Writes the
// SetupRec and ShadeRec of one clipped triangle with 16-byte stores.
// TINY_DROP: drop a FAST triangle of at most 3x3 pixels that covers none of
// them (the unclipped emits; the clipper's registers leave no room for it).
template <bool TINY_DROP>
__device__ inline void make_setup(const FrameParamsPod& fr, const FV& A, const FV& B, const FV& C, uint32_t material,
                                  const DevMaterial& mat, uint32_t view, uint32_t group, uint4 (&ro)[6],
                                  uint4 (&so)[9], SetupRec& rec) {
  const float* M = fr.M;
  float X[3], Y[3];
  const V3 P[3] = {A.p, B.p, C.p};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float o0 = 0.0f + M[0] * P[k].x;
    o0 = o0 + M[1] * P[k].y;
    o0 = o0 + M[2] * P[k].z;
    o0 = o0 + M[3] * 1.0f;
    float o1 = 0.0f + M[4] * P[k].x;
    o1 = o1 + M[5] * P[k].y;
    o1 = o1 + M[6] * P[k].z;
    o1 = o1 + M[7] * 1.0f;
    float o3 = 0.0f + M[12] * P[k].x;
    o3 = o3 + M[13] * P[k].y;
    o3 = o3 + M[14] * P[k].z;
    o3 = o3 + M[15] * 1.0f;
    const float sx = o0 / o3, sy = o1 / o3;
    X[k] = floorf((sx + (-fr.off_x)) + 0.5f);
    Y[k] = floorf((sy + (-fr.off_y)) + 0.5f);
  }
  // f_bc (alpha), f_ca (beta), f_ab (gamma)
  const float ea[4] = {Y[1] - Y[2], X[2] - X[1], X[1] * Y[2], X[2] * Y[1]};
  const float eb[4] = {Y[2] - Y[0], X[0] - X[2], X[2] * Y[0], X[0] * Y[2]};
  const float eg[4] = {Y[0] - Y[1], X[1] - X[0], X[0] * Y[1], X[1] * Y[0]};
  const float fa = edge_eval(ea, X[0], Y[0]);
  const float fb = edge_eval(eb, X[1], Y[1]);
  const float fg = edge_eval(eg, X[2], Y[2]);
  const float oa = edge_eval(ea, -1.0f, -1.0f);
  const float ob = edge_eval(eb, -1.0f, -1.0f);
  const float og = edge_eval(eg, -1.0f, -1.0f);
  uint32_t flags = 0;
  if (fa * oa > 0.0f) flags |= kTieA;
  if (fb * ob > 0.0f) flags |= kTieB;
  if (fg * og > 0.0f) flags |= kTieG;
  // bbox: min/max closures (NaN skipped), `as usize`, clipped to the screen
  float mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (X[k] < mnx) mnx = X[k];
    if (Y[k] < mny) mny = Y[k];
    if (X[k] > mxx) mxx = X[k];
    if (Y[k] > mxy) mxy = Y[k];
  }
  uint32_t x0, x1, y0, y1;
  rect_axis(f2usize(mnx), f2usize(mxx), (uint64_t)fr.W - 1u, &x0, &x1);
  rect_axis(f2usize(mny), f2usize(mxy), (uint64_t)fr.H - 1u, &y0, &y1);
  const bool fin = isfinite(fa) && isfinite(fb) && isfinite(fg);
  if (fin && fa != 0.0f && fb != 0.0f && fg != 0.0f) flags |= kFast;
  // (0) A FAST triangle whose scan rectangle is at most 3x3 pixels and holds
  // no covered pixel: the raster's own test (covered_fast at each pixel of
  // the rectangle) settled here, where the triangle is still in registers.
  if (TINY_DROP && (flags & kFast) && x1 - x0 <= 2u && y1 - y0 <= 2u) {
    bool any = false;
#pragma unroll
    for (uint32_t k = 0; k < 9u; k++) {
      const uint32_t px = x0 + k % 3u, py = y0 + k / 3u;
      const float xf = (float)px, yf = (float)py;
      const bool cov = covered_fast(edge_eval(ea, xf, yf), fa, flags & kTieA) &&
                       covered_fast(edge_eval(eb, xf, yf), fb, flags & kTieB) &&
                       covered_fast(edge_eval(eg, xf, yf), fg, flags & kTieG);
      any |= cov && px <= x1 && py <= y1;
    }
    if (!any) flags |= kDrop;
  }
  // Exact drops.  (1) A NaN normalizer makes that barycentric NaN at every
  // pixel, which fails `>= 0`.  (2) If all three normalizers are +0 and every
  // edge evaluation over the scan rectangle and the vertices is exact (all
  // partial sums below 2^24), the three edge values sum to 0 at every pixel;
  // coverage would need all three > 0 (x/+0 = +inf), so no pixel is covered.
  if (isnan(fa) || isnan(fb) || isnan(fg)) flags |= kDrop;
  // `exact`: every edge evaluation at |x|,|y| <= R (the vertices and the scan
  // rectangle) has all partial sums below 2^24 in magnitude, so with integer
  // operands each one is exact and e_bc + e_ca + e_ab = f_alpha = f_beta =
  // f_gama (twice the signed area) at every pixel.  S = |A|R + |B|R + |C| +
  // |D| bounds every partial sum; its float evaluation (4 roundings to
  // nearest) is at least S (1 - 2^-24)^4, so S_f < 16777000 gives S < 2^24.
  // (A triangle this leaves out is only handled by the general paths.)
  float R = (float)(x1 > y1 ? x1 : y1);
#pragma unroll
  for (int k = 0; k < 3; k++) R = fmaxf(fmaxf(R, fabsf(X[k])), fabsf(Y[k]));
  auto exact_edge = [R](const float e[4]) {
    const float S = ((fabsf(e[0]) * R + fabsf(e[1]) * R) + fabsf(e[2])) + fabsf(e[3]);
    return S < 16777000.0f;
  };
  const bool exact = isfinite(R) && exact_edge(ea) && exact_edge(eb) && exact_edge(eg);
  if (exact && __float_as_uint(fa) == 0u && __float_as_uint(fb) == 0u && __float_as_uint(fg) == 0u) flags |= kDrop;
  // Texture-index safety.  With `exact`, a covered pixel has e_i / f >= 0 and
  // sum e_i / f = 1, so alpha + beta + gama <= 1 + 2^-24 after rounding; with
  // all six uv in [0,1] the interpolated u, v lie in [0, 1 + 5*2^-24], and for
  // maps of at most 16384 texels per side floor(u*(w-1)) <= w-1 and
  // floor(v*(h-1)) <= h-1: get_rgb_slice (scene/mod.rs:154-168) stays in range.
  auto unit = [](float t) { return t >= 0.0f && t <= 1.0f; };
  if ((flags & kFast) && exact && unit(A.u) && unit(A.v) && unit(B.u) && unit(B.v) && unit(C.u) && unit(C.v) &&
      mat.kd_w <= 16384 && mat.kd_h <= 16384 && mat.ka_w <= 16384 && mat.ka_h <= 16384)
    flags |= kUvSafe;
  // Depth lower bound for early rejection.  For a UV-safe triangle (exact
  // edge arithmetic) a covered pixel has alpha + beta + gama >= 1 - 2^-24 with
  // non-negative weights, so with non-negative vertex depths its depth is at
  // least d_min * (1 - 2^-24)^4 > zlb = RN(d_min * (1 - 2^-20)).  A fragment
  // may be skipped only when zlb is strictly above a depth already recorded at
  // its pixel, and only on UV-safe triangles (whose texture reads cannot fail).
  float zlb = -INFINITY;
  {
    const float dmin = fminf(fminf(A.d, B.d), C.d);
    if ((flags & kUvSafe) && A.d >= 0.0f && B.d >= 0.0f && C.d >= 0.0f && dmin >= 1e-30f)
      zlb = dmin * (1.0f - 9.5367431640625e-07f);
    if ((flags & kUvSafe) && zlb >= fr.clear_depth) flags |= kDrop;  // no pixel can pass `<` the clear depth
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    rec.ea[k] = ea[k];
    rec.eb[k] = eb[k];
    rec.eg[k] = eg[k];
  }
  rec.f[0] = fa; rec.f[1] = fb; rec.f[2] = fg; rec.f[3] = 0.0f;
  rec.d[0] = A.d; rec.d[1] = B.d; rec.d[2] = C.d; rec.d[3] = zlb;
  rec.rx = x0 | (x1 << 16);
  rec.ry = y0 | (y1 << 16);
  rec.flags = flags;
  rec.view = view;
  const uint4* ri = reinterpret_cast<const uint4*>(&rec);
#pragma unroll
  for (int k = 0; k < 6; k++) ro[k] = ri[k];
  // shading attributes: uv[6], material, pad, n[9], l[9], h[9], pad
  auto u4 = [](float a, float b, float c, float d) {
    return make_uint4(__float_as_uint(a), __float_as_uint(b), __float_as_uint(c), __float_as_uint(d));
  };
  so[0] = u4(A.u, A.v, B.u, B.v);
  so[1] = u4(C.u, C.v, __uint_as_float(material), __uint_as_float(group));
  so[2] = u4(A.n.x, A.n.y, A.n.z, B.n.x);
  so[3] = u4(B.n.y, B.n.z, C.n.x, C.n.y);
  so[4] = u4(C.n.z, A.l.x, A.l.y, A.l.z);
  so[5] = u4(B.l.x, B.l.y, B.l.z, C.l.x);
  so[6] = u4(C.l.y, C.l.z, A.h.x, A.h.y);
  so[7] = u4(A.h.z, B.h.x, B.h.y, B.h.z);
  so[8] = u4(C.h.x, C.h.y, C.h.z, 0.0f);
}

// Stores one N-uint4 record per lane at dst[idx] (record units) for the lanes
// with ok and not skip (skip: a record nothing will read, whose id is still
// taken).  When the ok lanes are a prefix of the wave with consecutive idx (a
// dense run of output ids), the records go through the wave's LDS area
// stage[64 * N] and out in coalesced 16-byte lanes, leaving the skipped
// records' memory untouched; else each lane stores its own.  Every lane of the
// wave must call it.
template <int N>
__device__ inline void wave_store(uint4* stage, const uint4 (&v)[N], bool ok, uint32_t idx, uint4* dst,
                                  bool skip = false) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t m = __ballot(ok);
  const uint64_t sm = __ballot(ok && skip);
  if (m == sm) return;
  const uint32_t cnt = (uint32_t)__popcll(m);
  const uint32_t first = (uint32_t)__builtin_amdgcn_readfirstlane((int)idx);
  const bool prefix = m == (cnt == 64u ? ~0ull : ((1ull << cnt) - 1ull));
  if (prefix && __ballot(ok && idx != first + lane) == 0ull) {
    if (ok && !skip) {
#pragma unroll
      for (int q = 0; q < N; q++) stage[lane * N + q] = v[q];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    uint4* out = dst + (uint64_t)first * N;
    for (uint32_t k = lane; k < cnt * N; k += 64u)
      if (!((sm >> (k / (uint32_t)N)) & 1ull)) out[k] = stage[k];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  } else if (ok && !skip) {
#pragma unroll
    for (int q = 0; q < N; q++) dst[(uint64_t)idx * N + q] = v[q];
  }
}

// =================================================================== binning
// Conservative tile sets of a clipped triangle.  For a FAST triangle every
// edge value at an integer pixel is e = fl(fl(fl(fl(Ax)+fl(By))+C)-D) with
// |e - (Ax+By+C-D)| <= 4.01 u S, S = |A|x + |B|y + |C| + |D|; a pixel can only
// be covered when sign(f) * e >= 0 on all three edges.  Each edge's
// exact-arithmetic function, widened by that bound (plus 1), therefore gives
// a superset of the tiles holding covered pixels: per tile for triangles of a
// few tiles (tile_may_cover), per tile row for wide ones (spans_row).  Other
// triangles, and those within one tile, are binned to every tile of their
// scan rectangle.
struct Spans {
  int cx0, cx1, cy0, cy1;
  bool exact;
  float fa[3], fb[3], fk[3], fm[3];  // per-tile corner test (tile_may_cover)
  double a[3], b[3], k[3], m[3];     // per-row spans of wide triangles (spans_row)
};

// Bounds and the per-tile test's constants.  The test evaluates
// G_f = fl(fl(fl(a xc) + fl(b yc)) + fl(s (C - D))) at the tile corner
// (xc, yc) maximising the edge's exact function G = a x + b y + s (C - D).
// With M = |A| (x1 + TS) + |B| (y1 + TS) + |C| + |D|, which bounds the
// magnitudes at every corner of a tile of the rectangle, |G_f - G| <= 3.02 u M
// (u = 2^-24).  A covered pixel p of the tile has G >= g(p) >= -4.01 u M (the
// pixel evaluation bound above), so G_f >= -7.03 u M; rejecting only when
// G_f < -(1 + 9 u M) keeps every tile with a covered pixel.
__device__ inline void spans_init(const SetupRec& r, int TS, Spans& g) {
  const uint32_t x0 = r.rx & 0xFFFFu, x1 = r.rx >> 16, y0 = r.ry & 0xFFFFu, y1 = r.ry >> 16;
  g.cx0 = (int)(x0 / TS); g.cx1 = (int)(x1 / TS);
  g.cy0 = (int)(y0 / TS); g.cy1 = (int)(y1 / TS);
  g.exact = (r.flags & kFast) != 0;
  if (g.cx0 == g.cx1 && g.cy0 == g.cy1) {  // one tile: binned there without edge tests
    g.exact = false;
    return;
  }
  const float* es[3] = {r.ea, r.eb, r.eg};
  const float xm = (float)x1 + (float)TS, ym = (float)y1 + (float)TS;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const float A = es[i][0], B = es[i][1], C = es[i][2], D = es[i][3];
    const float s = r.f[i] > 0.0f ? 1.0f : -1.0f;
    const float M = ((fabsf(A) * xm + fabsf(B) * ym) + fabsf(C)) + fabsf(D);
    g.fa[i] = s * A;
    g.fb[i] = s * B;
    g.fk[i] = s * (C - D);
    g.fm[i] = 1.0f + 5.4e-7f * M;  // 1 + 9 u M, rounded up
    if (!isfinite(M)) g.exact = false;
  }
}

// The per-row constants of a wide triangle (after spans_init), in double.
__device__ inline void spans_init_rows(const SetupRec& r, Spans& g) {
  if (!g.exact) return;
  const uint32_t x1 = r.rx >> 16, y1 = r.ry >> 16;
  const float* es[3] = {r.ea, r.eb, r.eg};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double A = es[i][0], B = es[i][1], C = es[i][2], D = es[i][3];
    const double s = r.f[i] > 0.0f ? 1.0 : -1.0;
    const double S = fabs(A) * (double)x1 + fabs(B) * (double)y1 + fabs(C) + fabs(D);
    g.a[i] = s * A;
    g.b[i] = s * B;
    g.k[i] = s * (C - D);
    g.m[i] = 4.02 * 5.9604644775390625e-08 * S + 1.0 + 1e-9 * S;
    if (!isfinite(S)) g.exact = false;
  }
}

// Columns [lo, hi] of tile row ty that may hold covered pixels (lo > hi: none).
__device__ inline void spans_row(const Spans& g, int TS, int ty, int* lo, int* hi) {
  int l = g.cx0, h = g.cx1;
  if (g.exact) {
    const double ya = (double)(ty * TS), yb = (double)(ty * TS + TS - 1);
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const double K = g.k[i] + fmax(g.b[i] * ya, g.b[i] * yb);
      const double a = g.a[i];
      if (a > 0.0) {
        // need a*(c*TS + TS-1) + K >= -m
        double xl = (-g.m[i] - K) / a - 1.0;
        double c = (xl - (double)(TS - 1)) / (double)TS;
        c = fmin(fmax(c, -1.0), 1.0e6);
        const int ci = (int)ceil(c);
        l = max(l, ci);
      } else if (a < 0.0) {
        // need a*(c*TS) + K >= -m
        double xr = (K + g.m[i]) / (-a) + 1.0;
        double c = xr / (double)TS;
        c = fmin(fmax(c, -2.0), 1.0e6);
        const int ci = (int)floor(c);
        h = min(h, ci);
      } else if (K < -g.m[i]) {
        h = l - 1;
      }
    }
  }
  *lo = l;
  *hi = h;
}

// Work split: a triangle whose tile range is more than kWideSpan tiles tall
// or wide is binned by a whole wave (lanes over rows, then over entries); the
// rest by one thread each (at most kWideSpan^2 entries).  The split only
// balances work.
constexpr int kWideSpan = 4;

__device__ inline bool bin_is_wide(const Spans& g) {
  return g.cy1 - g.cy0 + 1 > kWideSpan || g.cx1 - g.cx0 + 1 > kWideSpan;
}

__device__ inline uint64_t lanemask_lt_u64(int lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }

// Tile (tx, ty) may hold a covered pixel: every edge's function, widened by
// fm (see spans_init), reaches 0 at the tile corner maximising it.
__device__ inline bool tile_may_cover(const Spans& g, int TS, int tx, int ty) {
  if (!g.exact) return true;
  const float xa = (float)(tx * TS), xb = (float)(tx * TS + TS - 1);
  const float ya = (float)(ty * TS), yb = (float)(ty * TS + TS - 1);
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const float v = (g.fa[i] * (g.fa[i] > 0.0f ? xb : xa) + g.fb[i] * (g.fb[i] > 0.0f ? yb : ya)) + g.fk[i];
    if (v < -g.fm[i]) return false;
  }
  return true;
}

// Calls f(e, ty, tx) for every tile of a triangle that is not wide (at most
// kWideSpan x kWideSpan tiles in its rectangle), e numbering the tiles in
// row-major order; returns the tile count.
template <class F>
__device__ inline uint32_t thread_tiles(const Spans& g, int TS, F f) {
  uint32_t e = 0;
  for (int ty = g.cy0; ty <= g.cy1; ty++)
    for (int tx = g.cx0; tx <= g.cx1; tx++)
      if (tile_may_cover(g, TS, tx, ty)) f(e++, ty, tx);
  return e;
}

// Wave version: lanes take tile rows, a wave scan places each row's span and
// the entries are spread over the lanes.  s_lo/s_off: 64-entry LDS scratch of
// this wave.  Returns the tile count (wave-uniform).
template <class F>
__device__ inline uint32_t wave_tiles(const Spans& g, int TS, int lane, int* s_lo, uint32_t* s_off, F f) {
  uint32_t o = 0;
  for (int row0 = g.cy0; row0 <= g.cy1; row0 += 64) {
    const int ty = row0 + lane;
    int lo = 0, hi = -1;
    if (ty <= g.cy1) spans_row(g, TS, ty, &lo, &hi);
    const uint32_t c = hi >= lo ? (uint32_t)(hi - lo + 1) : 0u;
    uint32_t x = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)x, d, 64);
      if (lane >= d) x += y;
    }
    const uint32_t total = (uint32_t)__shfl((int)x, 63, 64);
    s_lo[lane] = lo;
    s_off[lane] = x - c;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (uint32_t e = (uint32_t)lane; e < total; e += 64) {
      // the row of entry e: the last row whose offset is <= e (rows without
      // entries share their successor's offset, so the last one has entries)
      int k = 0;
#pragma unroll
      for (int step = 32; step >= 1; step >>= 1)
        if (s_off[k + step] <= e) k += step;
      f(o + e, row0 + k, s_lo[k] + (int)(e - s_off[k]));
    }
    o += total;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  return o;
}

// Ordered bins keep, per clipped triangle, a bin box: x the tile key of the
// corner (cx0, cy0) of its tile box, y its tiles as bits 4 * (ty - cy0) +
// (tx - cx0) (low 16 bits, within the at most 4 x 4 box) and the view's tiles
// per row (high 16 bits), so that the pair emit finds every tile key from
// these 8 bytes alone; x = kWideBox for a wide triangle, whose tiles a wave
// counts and emits later.  They live in the flag buffer after the flag bytes
// (12 bytes per clipped triangle in all).
constexpr uint32_t kWideBox = 0xFFFFFFFFu;
__device__ __host__ inline uint2* bin_boxes(void* big, uint32_t cap_tri) {
  return reinterpret_cast<uint2*>(reinterpret_cast<uint8_t*>(big) + ((cap_tri + 15u) & ~15u));
}

// Ordered bins, one clipped triangle of view vw: its flag byte (0 binned here,
// 1 wide, 2 dropped: no tiles, and its records were not stored), its bin box
// and, when it is not wide, its tile count.
__device__ inline void bin_count_one(const SetupRec& r, int TS, uint32_t i, const DevView& vw,
                                     uint32_t* __restrict__ bincnt, uint8_t* __restrict__ wflag,
                                     uint2* __restrict__ boxes) {
  uint32_t c = 0, mask = 0, key0 = 0;
  bool is_wide = false;
  if (!(r.flags & kDrop)) {
    Spans g;
    spans_init(r, TS, g);
    is_wide = bin_is_wide(g);
    key0 = vw.tile_base + (uint32_t)g.cy0 * vw.ntx + (uint32_t)g.cx0;
    if (!is_wide)
      c = thread_tiles(g, TS, [&](uint32_t, int ty, int tx) { mask |= 1u << (4 * (ty - g.cy0) + (tx - g.cx0)); });
  }
  wflag[i] = is_wide ? 1u : ((r.flags & kDrop) ? 2u : 0u);
  boxes[i] = make_uint2(is_wide ? kWideBox : key0, mask | (vw.ntx << 16));
  if (!is_wide) bincnt[i] = c;
}

// The pairs of a triangle that is not wide from its bin box, its e-th tile at
// position o + e (tiles in the order of the count pass).
__device__ inline void emit_box_pairs(uint2 box, uint32_t i, uint32_t o, uint32_t cap_pairs,
                                      uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t ntx = box.y >> 16;
  uint32_t mask = box.y & 0xFFFFu;
  while (mask) {
    const uint32_t b = (uint32_t)__builtin_ctz(mask);
    mask &= mask - 1u;
    if (o < cap_pairs) {
      keys[o] = box.x + (b >> 2) * ntx + (b & 3u);
      vals[o] = i;
    }
    o++;
  }
}

// ================================================================== geometry
// Near-first binning order: an 8-bit key of the nearest vertex depth, 16
// steps per octave above depth_base (the float bits of half the near-plane
// distance), so 16 octaves of depth; monotonic for positive depths, others
// sort first.  Only the speed of the raster depends on it.
constexpr int kDepthKeyBits = 8;
__device__ inline uint32_t depth_order_key(const SetupRec& r, uint32_t depth_base) {
  const float d = fminf(fminf(r.d[0], r.d[1]), r.d[2]);
  const uint32_t b = __float_as_uint(d);
  if (!(d > 0.0f) || b <= depth_base) return 0u;
  return min((b - depth_base) >> 19, (1u << kDepthKeyBits) - 1u);
}

// Small single-view frames: at most kSmallBinTris input triangles and
// kSmallBinTiles tiles take single-workgroup kernels for the geometry lists
// (k_geom_small) and the binning (k_bin_small), and one launch for both emits.
constexpr uint32_t kSmallBinTiles = 4096;  // LDS counters and cursors: 32 KB
constexpr uint32_t kSmallBinTris = 16384;  // LDS count bytes: 16 KB

// cnt[t]: clipped-triangle count (bits 0-2, at most 4) | kCntClip when the
// triangle crosses the near or far plane.
constexpr uint32_t kCntMask = 7u, kCntClip = 8u;

__device__ inline bool inside_plane(const float* pl, V3 a, V3 b, V3 c) {
  return plane_f(pl, a) > 0.0f && plane_f(pl, b) > 0.0f && plane_f(pl, c) > 0.0f;
}

// Block-scan word: clipped triangles << 18 | unclipped emitters << 9 |
// clipped emitters (a block has at most 256 * 4 of the first, 256 of each other).
__device__ inline uint32_t pack_geom(uint32_t c, bool clip) {
  return (c << 18) | (c && !clip ? (1u << 9) : 0u) | (c && clip ? 1u : 0u);
}

__global__ __launch_bounds__(256) void k_geom_count(const DevDraw* __restrict__ draws, uint32_t n_draws,
                                                    const DevView* __restrict__ views, uint32_t n_in,
                                                    uint8_t* __restrict__ cnt, uint32_t* __restrict__ blk,
                                                    uint32_t* __restrict__ blk_simple, uint32_t* __restrict__ blk_clip) {
  __shared__ uint32_t s_tmp[4];
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  uint32_t c = 0;
  bool clip = false;
  const uint32_t d = wave_find_draw(draws, n_draws, t, t < n_in);
  if (t < n_in) {
    const DevDraw& dr = draws[d];
    const FrameParamsPod& fr = views[dr.view].fr;
    PV a, b, cc;
    tri_positions(dr, t - dr.tri_begin, a.p, b.p, cc.p);
    c = (uint32_t)cull_and_clip(fr, a, b, cc, !(dr.flags & kDrawNoCull), [](int, const PV&, const PV&, const PV&) {});
    // ViewPlane::clip keeps a triangle unchanged exactly when its three
    // vertices are strictly inside the plane (clip_plane's first branch)
    if (c) clip = !(inside_plane(fr.planes[0], a.p, b.p, cc.p) && inside_plane(fr.planes[1], a.p, b.p, cc.p));
    cnt[t] = (uint8_t)(c | (clip ? kCntClip : 0u));
  }
  uint32_t total;
  block_exclusive_scan256(pack_geom(c, clip), s_tmp, &total);
  if (threadIdx.x == 0) {
    blk[blockIdx.x] = total >> 18;
    blk_simple[blockIdx.x] = (total >> 9) & 0x1FFu;
    blk_clip[blockIdx.x] = total & 0x1FFu;
  }
}

// Exclusive scan of three values per thread over a block of 1024 threads,
// packed 22 | 21 | 21 bits into 64 (a block has at most 4096 of the first,
// 1024 of the others); *total gets the three totals, packed alike.
__device__ inline uint64_t block_exclusive_scan1024x3(uint64_t v, uint64_t* s_w /*[16]*/, uint64_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_w[wave] = x;
  __syncthreads();
  uint64_t before = 0, all = 0;
  for (int w = 0; w < 16; w++) {
    if (w < wave) before += s_w[w];
    all += s_w[w];
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

// k_geom_count, the scan of its block counts and k_geom_compact for a small
// single-view frame (at most kSmallBinTris input triangles), in one
// workgroup of 1024 threads: the same cnt, the same per-256-triangle output
// offsets in blk, the same status counts and the same two ordered lists,
// without two launches and their gaps.
__global__ __launch_bounds__(1024) void k_geom_small(const DevDraw* __restrict__ draws, uint32_t n_draws,
                                                     const DevView* __restrict__ views, uint32_t n_in,
                                                     uint8_t* __restrict__ cnt, uint32_t* __restrict__ blk,
                                                     Status* __restrict__ status, uint32_t cap_tri,
                                                     uint2* __restrict__ surv) {
  __shared__ uint64_t s_w[16];
  __shared__ uint8_t s_cnt[kSmallBinTris];
  const uint32_t tid = threadIdx.x;
  auto field = [](uint64_t x, int k) { return (uint32_t)((x >> (k == 0 ? 42 : (k == 1 ? 21 : 0))) & 0x1FFFFFu); };
  auto pack = [](uint32_t c, bool clip) {
    return ((uint64_t)c << 42) | ((uint64_t)(c && !clip) << 21) | (uint64_t)(c && clip);
  };
  // counts, output offsets, the unclipped emitters' entries
  uint32_t run_o = 0, run_s = 0, run_c = 0;
  for (uint32_t c0 = 0; c0 < n_in; c0 += 1024u) {
    const uint32_t t = c0 + tid;
    uint32_t c = 0;
    bool clip = false;
    const uint32_t d = wave_find_draw(draws, n_draws, t, t < n_in);
    if (t < n_in) {
      const DevDraw& dr = draws[d];
      const FrameParamsPod& fr = views[dr.view].fr;
      PV a, b, cc;
      tri_positions(dr, t - dr.tri_begin, a.p, b.p, cc.p);
      c = (uint32_t)cull_and_clip(fr, a, b, cc, !(dr.flags & kDrawNoCull),
                                  [](int, const PV&, const PV&, const PV&) {});
      if (c) clip = !(inside_plane(fr.planes[0], a.p, b.p, cc.p) && inside_plane(fr.planes[1], a.p, b.p, cc.p));
      const uint8_t q = (uint8_t)(c | (clip ? kCntClip : 0u));
      cnt[t] = q;
      s_cnt[t] = q;
    }
    uint64_t tot;
    const uint64_t ex = block_exclusive_scan1024x3(pack(c, clip), s_w, &tot);
    const uint32_t first = run_o + field(ex, 0);
    if (t < n_in && (t & 255u) == 0u) blk[t >> 8] = first;
    if (c && !clip) surv[run_s + field(ex, 1)] = make_uint2(t, first);
    run_o += field(tot, 0);
    run_s += field(tot, 1);
    run_c += field(tot, 2);
  }
  // the clipped emitters' entries, after the unclipped ones
  if (run_c) {
    uint32_t r_o = 0, r_c = 0;
    for (uint32_t c0 = 0; c0 < n_in; c0 += 1024u) {
      const uint32_t t = c0 + tid;
      const uint32_t q = t < n_in ? (uint32_t)s_cnt[t] : 0u;
      const uint32_t c = q & kCntMask;
      const bool clip = (q & kCntClip) != 0u;
      uint64_t tot;
      const uint64_t ex = block_exclusive_scan1024x3(pack(c, clip), s_w, &tot);
      if (c && clip) surv[run_s + r_c + field(ex, 2)] = make_uint2(t, r_o + field(ex, 0));
      r_o += field(tot, 0);
      r_c += field(tot, 2);
    }
  }
  if (tid == 0) {
    status->n_out = run_o;
    status->n_simple = run_s;
    status->n_clip = run_c;
    if (run_o > cap_tri) atomicOr(&status->overflow, 1u);
  }
}

// Ordered lists of the input triangles that emit, (triangle, first output
// id): the unclipped ones in surv[0, n_simple), the clipped ones after them.
__global__ __launch_bounds__(256) void k_geom_compact(uint32_t n_in, const uint8_t* __restrict__ cnt,
                                                      const uint32_t* __restrict__ blk_off,
                                                      const uint32_t* __restrict__ simple_off,
                                                      const uint32_t* __restrict__ clip_off,
                                                      const Status* __restrict__ status, uint2* __restrict__ surv) {
  __shared__ uint32_t s_tmp[4];
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  const uint32_t q = t < n_in ? (uint32_t)cnt[t] : 0u;
  const uint32_t c = q & kCntMask;
  const bool clip = (q & kCntClip) != 0;
  uint32_t total;
  const uint32_t local = block_exclusive_scan256(pack_geom(c, clip), s_tmp, &total);
  if (!c) return;
  const uint2 rec = make_uint2(t, blk_off[blockIdx.x] + (local >> 18));
  if (clip)
    surv[status->n_simple + clip_off[blockIdx.x] + (local & 0x1FFu)] = rec;
  else
    surv[simple_off[blockIdx.x] + ((local >> 9) & 0x1FFu)] = rec;
}



// One thread per emitting input triangle of one list (dense: culled
// triangles take no lane).  The unclipped list (CLIP false) skips the
// clipper, whose attribute registers would otherwise size the whole kernel.
template <bool BIN>
__device__ __forceinline__ void geom_emit_clipped(uint32_t j, const DevDraw* __restrict__ draws, uint32_t n_draws,
                                                  const DevView* __restrict__ views, const uint2* __restrict__ surv,
                                                  const DevMaterial* __restrict__ mats, SetupRec* __restrict__ setup,
                                                  ShadeRec* __restrict__ shade, uint32_t cap_tri, Status* __restrict__ status,
                                                  int TS, uint32_t* __restrict__ bincnt, uint8_t* __restrict__ wflag,
                                                  uint32_t* __restrict__ dkey, uint32_t span) {
  const uint2 sv = surv[status->n_simple + j];
  const uint32_t t = sv.x, base = sv.y;
  // the lanes of the wave still in the loop
  const bool sparse = wave_sparse(true, t, span);
  const DevDraw& dr = draws[wave_find_draw(draws, n_draws, t, true)];
  const FrameParamsPod& fr = views[dr.view].fr;
  const uint32_t lt = t - dr.tri_begin;
  const V3 cam = v3(fr.cam[0], fr.cam[1], fr.cam[2]);
  const V3 L = v3(fr.light[0], fr.light[1], fr.light[2]);
  FV v[3];
  V3 P[3], N[3];
  float U[3], W[3];
  tri_vertices(dr, lt, sparse, P, N, U, W);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const V3 p = P[k];
    v[k].p = p;
    // per-vertex visual info, scene/mod.rs:920-925
    const V3 eye_dir = sub(p, cam);
    const V3 light = normalized(sub(p, L));
    const V3 eye = normalized(eye_dir);
    v[k].d = norm(eye_dir);
    v[k].l = light;
    v[k].h = normalized(add(light, eye));
    v[k].n = N[k];
    v[k].u = U[k];
    v[k].v = W[k];
  }
  const uint32_t material = dr.material, view = dr.view, group = dr.group;
  auto out = [&](int k, const FV& A, const FV& B, const FV& C) {
    const uint32_t idx = base + (uint32_t)k;
    if (idx >= cap_tri) {
      atomicOr(&status->overflow, 1u);
      return;
    }
    SetupRec rec;
    uint4 ro[6], so[9];
    make_setup<false>(fr, A, B, C, material, mats[material], view, group, ro, so, rec);
    // with ordered bins nothing reads a dropped triangle's records
    if (!BIN || !(rec.flags & kDrop)) {
#pragma unroll
      for (int q = 0; q < 6; q++) reinterpret_cast<uint4*>(setup + idx)[q] = ro[q];
#pragma unroll
      for (int q = 0; q < 9; q++) reinterpret_cast<uint4*>(shade + idx)[q] = so[q];
    }
    if (BIN) bin_count_one(rec, TS, idx, views[view], bincnt, wflag, bin_boxes(wflag, cap_tri));
    if (dkey) dkey[idx] = depth_order_key(rec, views[view].depth_base);
  };
  cull_and_clip(fr, v[0], v[1], v[2], !(dr.flags & kDrawNoCull), out);
}

// One unclipped list entry per lane (every lane of the wave calls it; lanes
// past the list end with active false): one record each at the entry's output
// id, stored through the wave's LDS area stage[64 * 9].
template <bool BIN>
__device__ __forceinline__ void geom_emit_simple(uint32_t j, bool active, uint4* stage, const DevDraw* __restrict__ draws,
                                                 uint32_t n_draws, const DevView* __restrict__ views,
                                                 const uint2* __restrict__ surv, const DevMaterial* __restrict__ mats,
                                                 SetupRec* __restrict__ setup, ShadeRec* __restrict__ shade, uint32_t cap_tri,
                                                 Status* __restrict__ status, int TS, uint32_t* __restrict__ bincnt,
                                                 uint8_t* __restrict__ wflag, uint32_t* __restrict__ dkey, uint32_t span) {
  uint4 ro[6], so[9];
  uint32_t idx = 0;
  bool ok = false, skip = false;
  const uint2 sv = active ? surv[j] : make_uint2(0u, 0u);
  const bool sparse = wave_sparse(active, sv.x, span);
  const uint32_t d = wave_find_draw(draws, n_draws, sv.x, active);
  if (active) {
    const uint32_t t = sv.x;
    idx = sv.y;
    const DevDraw& dr = draws[d];
    const FrameParamsPod& fr = views[dr.view].fr;
    const uint32_t lt = t - dr.tri_begin;
    const V3 cam = v3(fr.cam[0], fr.cam[1], fr.cam[2]);
    const V3 L = v3(fr.light[0], fr.light[1], fr.light[2]);
    FV v[3];
    V3 P[3], N[3];
    float U[3], W[3];
    tri_vertices(dr, lt, sparse, P, N, U, W);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const V3 p = P[k];
      v[k].p = p;
      // per-vertex visual info, scene/mod.rs:920-925
      const V3 eye_dir = sub(p, cam);
      const V3 light = normalized(sub(p, L));
      const V3 eye = normalized(eye_dir);
      v[k].d = norm(eye_dir);
      v[k].l = light;
      v[k].h = normalized(add(light, eye));
      v[k].n = N[k];
      v[k].u = U[k];
      v[k].v = W[k];
    }
    if (idx >= cap_tri) {
      atomicOr(&status->overflow, 1u);
    } else {
      // passed the back-face and lateral tests in k_geom_count
      SetupRec rec;
      make_setup<true>(fr, v[0], v[1], v[2], dr.material, mats[dr.material], dr.view, dr.group, ro, so, rec);
      ok = true;
      // with ordered bins nothing reads a dropped triangle's records
      skip = BIN && (rec.flags & kDrop);
      if (BIN) bin_count_one(rec, TS, idx, views[dr.view], bincnt, wflag, bin_boxes(wflag, cap_tri));
      if (dkey) dkey[idx] = depth_order_key(rec, views[dr.view].depth_base);
    }
  }
  wave_store<6>(stage, ro, ok, idx, reinterpret_cast<uint4*>(setup), skip);
  wave_store<9>(stage, so, ok, idx, reinterpret_cast<uint4*>(shade), skip);
}

// A grid-stride loop over one list (its length is on the device), so that the
// grid can be capped well below one thread per input triangle.  The unclipped
// list goes a wave at a time (its records leave through LDS).
template <bool BIN, bool CLIP>
__global__ __launch_bounds__(256) void k_geom_emit(const DevDraw* __restrict__ draws, uint32_t n_draws,
                                                   const DevView* __restrict__ views,
                                                   const uint2* __restrict__ surv,
                                                   const DevMaterial* __restrict__ mats, SetupRec* __restrict__ setup,
                                                   ShadeRec* __restrict__ shade, uint32_t cap_tri,
                                                   Status* __restrict__ status, int TS, uint32_t* __restrict__ bincnt,
                                                   uint8_t* __restrict__ wflag, uint32_t* __restrict__ dkey, uint32_t span) {
  if constexpr (CLIP) {
    const uint32_t n = status->n_clip;
    for (uint32_t j = blockIdx.x * 256u + threadIdx.x; j < n; j += gridDim.x * 256u)
      geom_emit_clipped<BIN>(j, draws, n_draws, views, surv, mats, setup, shade, cap_tri, status, TS, bincnt, wflag,
                             dkey, span);
  } else {
    __shared__ uint4 s_stage[4][64 * 9];
    const uint32_t n = status->n_simple;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t j0 = blockIdx.x * 256u + wave * 64u; j0 < n; j0 += gridDim.x * 256u)
      geom_emit_simple<BIN>(j0 + lane, j0 + lane < n, s_stage[wave], draws, n_draws, views, surv, mats, setup, shade,
                            cap_tri, status, TS, bincnt, wflag, dkey, span);
  }
}

// Both emitter lists of a small frame in one launch on the stream (for a few
// thousand triangles the fork to a second stream and the join back cost more
// than the emits): of its g_clip + g_simple blocks, g_clip run the clipped
// list and the others the unclipped one, interleaved in proportion (block b
// is a clipped block when floor((b + 1) g_clip / G) > floor(b g_clip / G)).
// No occupancy floor: the grid (at most 128 blocks) fills a fraction of the
// chip, so registers go to the clipped path (167-172 VGPRs, no spills)
// rather than to waves that would have nothing to run.
template <bool BIN>
__global__ __launch_bounds__(256) void k_geom_emit_both(const DevDraw* __restrict__ draws, uint32_t n_draws,
                                                        const DevView* __restrict__ views,
                                                        const uint2* __restrict__ surv,
                                                        const DevMaterial* __restrict__ mats,
                                                        SetupRec* __restrict__ setup, ShadeRec* __restrict__ shade,
                                                        uint32_t cap_tri, Status* __restrict__ status, int TS,
                                                        uint32_t* __restrict__ bincnt, uint8_t* __restrict__ wflag,
                                                        uint32_t* __restrict__ dkey, uint32_t g_clip) {
  __shared__ uint4 s_stage[4][64 * 9];
  const uint64_t G = gridDim.x, b = blockIdx.x;
  const uint32_t c0 = (uint32_t)(b * g_clip / G), c1 = (uint32_t)((b + 1) * g_clip / G);
  if (c1 > c0) {
    const uint32_t n = status->n_clip;
    for (uint32_t j = c0 * 256u + threadIdx.x; j < n; j += g_clip * 256u)
      geom_emit_clipped<BIN>(j, draws, n_draws, views, surv, mats, setup, shade, cap_tri, status, TS, bincnt, wflag,
                             dkey, kSparseSpan);
  } else {
    const uint32_t ub = (uint32_t)b - c0, g_simple = (uint32_t)G - g_clip;
    const uint32_t n = status->n_simple;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t j0 = ub * 256u + wave * 64u; j0 < n; j0 += g_simple * 256u)
      geom_emit_simple<BIN>(j0 + lane, j0 + lane < n, s_stage[wave], draws, n_draws, views, surv, mats, setup, shade,
                            cap_tri, status, TS, bincnt, wflag, dkey, kSparseSpan);
  }
}

// First clipped-triangle id of each view: the exclusive prefix of the clip
// counts at the view's first input triangle (blk_off holds the scanned block
// sums).  Shading subtracts it, so a view's winner ids do not depend on the
// other views of a batch.
__global__ void k_view_id_base(DevView* __restrict__ views, uint32_t n_views, uint32_t n_in,
                               const uint8_t* __restrict__ cnt, const uint32_t* __restrict__ blk_off,
                               const Status* __restrict__ status) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n_views) return;
  const uint32_t t0 = views[v].tri_begin;
  uint32_t base;
  if (t0 >= n_in) {
    base = status->n_out;
  } else {
    const uint32_t b = t0 >> 8;
    base = blk_off[b];
    for (uint32_t t = b << 8; t < t0; t++) base += cnt[t] & kCntMask;
  }
  views[v].id_base = base;
}

// ====================================================================== scan
// Generic exclusive scan of u32 with the element count possibly on the device.
constexpr int kScanG = 1024;  // partial sums per scan

__device__ inline uint32_t scan_count(const uint32_t* n_ptr, uint32_t n_const, uint32_t cap) {
  if (!n_ptr) return n_const;
  uint32_t n = *n_ptr;
  return n < cap ? n : cap;
}
__device__ inline uint32_t scan_chunk(uint32_t n) { return (((n + kScanG - 1u) / kScanG) + 1023u) & ~1023u; }

// The scans' input: in[i], or in[idx[i]] (a gather fused into the scan).
__device__ inline uint32_t scan_in(const uint32_t* in, const uint32_t* idx, uint32_t i) {
  return idx ? in[idx[i]] : in[i];
}

__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* __restrict__ in, const uint32_t* n_ptr,
                                                     uint32_t n_const, uint32_t cap, uint32_t* __restrict__ partials,
                                                     const uint32_t* __restrict__ idx) {
  __shared__ uint32_t s_tmp[4];
  const uint32_t n = scan_count(n_ptr, n_const, cap);
  const uint32_t C = scan_chunk(n);
  const uint32_t b0 = blockIdx.x * C;
  const uint32_t b1 = min(b0 + C, n);
  uint32_t s = 0;
  for (uint32_t i = b0 + threadIdx.x; i < b1; i += 256) s += scan_in(in, idx, i);
  uint32_t total;
  block_exclusive_scan256(s, s_tmp, &total);
  if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// Second pass: each block takes the exclusive prefix of the partial sums
// before it (at most kScanG values, read straight from the partials) and
// scans its chunk; block 0 also publishes the total and the overflow flag.
// in and out may be the same array (in-place scans): no __restrict__ on them.
__global__ __launch_bounds__(256) void k_scan_apply(const uint32_t* in, uint32_t* out,
                                                    const uint32_t* n_ptr, uint32_t n_const, uint32_t cap,
                                                    const uint32_t* __restrict__ partials, uint32_t* total_out,
                                                    uint32_t cap_check, uint32_t* overflow_word,
                                                    uint32_t overflow_bit, const uint32_t* __restrict__ idx) {
  __shared__ uint32_t s_tmp[4];
  const uint32_t n = scan_count(n_ptr, n_const, cap);
  const uint32_t C = scan_chunk(n);
  const uint32_t b0 = blockIdx.x * C;
  const uint32_t b1 = min(b0 + C, n);
  const uint32_t upto = blockIdx.x == 0 ? (uint32_t)kScanG : blockIdx.x;
  uint32_t p = 0;
  for (uint32_t i = threadIdx.x; i < upto; i += 256) p += partials[i];
  uint32_t sum;
  block_exclusive_scan256(p, s_tmp, &sum);
  uint32_t running = 0;
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) {
      if (total_out) *total_out = sum;
      if (overflow_word && sum > cap_check) atomicOr(overflow_word, overflow_bit);
    }
  } else {
    running = sum;
  }
  for (uint32_t tb = b0; tb < b1; tb += 1024) {
    const uint32_t i0 = tb + threadIdx.x * 4u;
    uint32_t v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = (i0 + k < b1) ? scan_in(in, idx, i0 + k) : 0u;
    const uint32_t s4 = v[0] + v[1] + v[2] + v[3];
    uint32_t total;
    uint32_t ex = block_exclusive_scan256(s4, s_tmp, &total) + running;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (i0 + k < b1) out[i0 + k] = ex;
      ex += v[k];
    }
    running += total;
  }
}

// Scans of at most 4 * 1024 elements whose count the host knows: one block.
// Up to three independent scans of n <= 4096 values, one block each.
struct ScanSeg {
  const uint32_t* in;
  uint32_t* out;
  uint32_t* total_out;      // may be null
  uint32_t* overflow_word;  // may be null
  uint32_t cap_check, overflow_bit;
};
struct ScanSegs {
  ScanSeg s[3];
};

__global__ __launch_bounds__(1024) void k_scan_small(ScanSegs segs, uint32_t n) {
  const ScanSeg& g = segs.s[blockIdx.x];
  const uint32_t* in = g.in;  // may be out (in-place scans)
  uint32_t* out = g.out;
  uint32_t* total_out = g.total_out;
  uint32_t* overflow_word = g.overflow_word;
  const uint32_t cap_check = g.cap_check, overflow_bit = g.overflow_bit;
  __shared__ uint32_t s_w[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t i0 = threadIdx.x * 4u;
  uint32_t v[4];
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = (i0 + k < n) ? in[i0 + k] : 0u;
  const uint32_t s4 = v[0] + v[1] + v[2] + v[3];
  uint32_t x = s4;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_w[wave] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (int w = 0; w < 16; w++) {
    if (w < wave) before += s_w[w];
    all += s_w[w];
  }
  uint32_t ex = before + x - s4;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (i0 + k < n) out[i0 + k] = ex;
    ex += v[k];
  }
  if (threadIdx.x == 0) {
    if (total_out) *total_out = all;
    if (overflow_word && all > cap_check) atomicOr(overflow_word, overflow_bit);
  }
}

// out = exclusive scan of in (or of in[idx[i]] when idx is given).
static void scan_u32(const uint32_t* in, uint32_t* out, const uint32_t* n_ptr, uint32_t n_const, uint32_t cap,
                     uint32_t* partials, uint32_t* total_out, uint32_t cap_check, uint32_t* overflow_word,
                     uint32_t overflow_bit, hipStream_t s, const uint32_t* idx = nullptr) {
  if (!n_ptr && n_const <= 4096u && !idx) {
    ScanSegs g{};
    g.s[0] = ScanSeg{in, out, total_out, overflow_word, cap_check, overflow_bit};
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, s, g, n_const);
    return;
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3(kScanG), dim3(256), 0, s, in, n_ptr, n_const, cap, partials, idx);
  hipLaunchKernelGGL(k_scan_apply, dim3(kScanG), dim3(256), 0, s, in, out, n_ptr, n_const, cap, partials, total_out,
                     cap_check, overflow_word, overflow_bit, idx);
}

// Pass 1 of unordered bins, one thread per clipped triangle: flag wide
// triangles (wflag, one byte each) and count the tiles of the others per tile
// with atomics.  (Ordered bins take the per-triangle counts from k_geom_emit.)
__global__ __launch_bounds__(256) void k_bin_count(const SetupRec* __restrict__ setup, const DevView* __restrict__ views,
                                                   const Status* status, uint32_t cap_tri, int TS,
                                                   uint32_t* __restrict__ tile_cnt, uint8_t* __restrict__ wflag) {
  const uint32_t n = min(status->n_out, cap_tri);
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    const SetupRec r = setup[i];
    bool is_wide = false;
    if (!(r.flags & kDrop)) {
      Spans g;
      spans_init(r, TS, g);
      is_wide = bin_is_wide(g);
      if (!is_wide) {
        const DevView& vw = views[r.view];
        thread_tiles(g, TS, [&](uint32_t, int ty, int tx) {
          atomicAdd(&tile_cnt[vw.tile_base + (uint32_t)ty * vw.ntx + (uint32_t)tx], 1u);
        });
      }
    }
    wflag[i] = is_wide ? 1u : 0u;
  }
}

// The wide-triangle list: the wide triangles (flag 1) of each 4096-triangle chunk,
// 16 flags per thread, placed by a block scan; one atomic per chunk reserves
// the chunk's run of the list.  The list order does not matter: a triangle's
// bin entries go to its own scanned offset (or through per-tile atomics for
// unordered bins).
__global__ __launch_bounds__(256) void k_wide_list(const uint8_t* __restrict__ wflag, Status* status,
                                                   uint32_t cap_tri, uint32_t* __restrict__ list,
                                                   uint32_t* __restrict__ zero_dst, uint32_t zero_words) {
  __shared__ uint32_t s_tmp[4];
  __shared__ uint32_t s_base;
  // also clears the tile ranges that k_tile_ranges fills later
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < zero_words; i += gridDim.x * 256u) zero_dst[i] = 0u;
  const uint32_t n = min(status->n_out, cap_tri);
  for (uint32_t c0 = blockIdx.x * 4096u; c0 < n; c0 += gridDim.x * 4096u) {
    const uint32_t t0 = c0 + threadIdx.x * 16u;
    uint32_t mask = 0;
    if (t0 + 16u <= n) {
      const uint4 f = *reinterpret_cast<const uint4*>(wflag + t0);  // t0 is a multiple of 16
      const uint32_t w[4] = {f.x, f.y, f.z, f.w};
#pragma unroll
      for (int k = 0; k < 16; k++)
        if (((w[k >> 2] >> (8 * (k & 3))) & 0xFFu) == 1u) mask |= 1u << k;
    } else {
      for (uint32_t k = 0; t0 + k < n; k++)
        if (wflag[t0 + k] == 1u) mask |= 1u << k;
    }
    uint32_t total;
    const uint32_t local = block_exclusive_scan256((uint32_t)__popc(mask), s_tmp, &total);
    if (threadIdx.x == 0 && total) s_base = atomicAdd(&status->n_wide, total);
    __syncthreads();
    uint32_t o = s_base + local;
    while (mask) {
      const uint32_t k = (uint32_t)__builtin_ctz(mask);
      mask &= mask - 1u;
      list[o++] = t0 + k;
    }
    __syncthreads();
  }
}

// Visits the listed (wide) triangles, one wave per triangle.
template <class F>
__device__ inline void for_each_wide(const uint32_t* __restrict__ list, const Status* status, F f) {
  const uint32_t gw = (blockIdx.x * 256u + threadIdx.x) >> 6, nwaves = gridDim.x * 4u;
  const uint32_t n = status->n_wide;
  for (uint32_t w = gw; w < n; w += nwaves) f((uint32_t)__builtin_amdgcn_readfirstlane(list[w]));
}

template <bool ORDERED>
__global__ __launch_bounds__(256) void k_bin_count_wide(const SetupRec* __restrict__ setup,
                                                        const DevView* __restrict__ views, const Status* status,
                                                        int TS, const uint32_t* __restrict__ wide_list,
                                                        uint32_t* __restrict__ bincnt, uint32_t* __restrict__ tile_cnt) {
  __shared__ int s_lo[4][64];
  __shared__ uint32_t s_off[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for_each_wide(wide_list, status, [&](uint32_t i) {
    const SetupRec r = setup[i];
    const DevView& vw = views[r.view];
    Spans g;
    spans_init(r, TS, g);
    spans_init_rows(r, g);
    const uint32_t c = wave_tiles(g, TS, lane, s_lo[wave], s_off[wave], [&](uint32_t, int ty, int tx) {
      if (!ORDERED) atomicAdd(&tile_cnt[vw.tile_base + (uint32_t)ty * vw.ntx + (uint32_t)tx], 1u);
    });
    if (ORDERED && lane == 0) bincnt[i] = c;
  });
}

// Pass 2.  Ordered: (tile key, triangle) pairs at the triangle's scanned
// offset, from its bin box, for the stable radix sort.  Unordered: the
// triangle id goes straight into its tile's list at tile_start +
// atomicAdd(cursor).
//
// Unordered pairs of one triangle that is not wide.
__device__ inline void bin_emit_tri_unordered(const SetupRec* __restrict__ setup, const DevView* __restrict__ views,
                                              int TS, uint32_t i, const uint32_t* __restrict__ tile_start,
                                              uint32_t* __restrict__ tile_cur, uint32_t cap_pairs,
                                              uint32_t* __restrict__ vals) {
  auto put = [&](const DevView& vw, int ty, int tx) {
    const uint32_t key = vw.tile_base + (uint32_t)ty * vw.ntx + (uint32_t)tx;
    const uint32_t pos = tile_start[key] + atomicAdd(&tile_cur[key], 1u);
    if (pos < cap_pairs) vals[pos] = i;
  };
  // the record's last 16 bytes first: a triangle within one tile needs no edges
  const uint4 h = reinterpret_cast<const uint4*>(setup + i)[5];
  if (h.z & kDrop) return;
  const uint32_t cx0 = (h.x & 0xFFFFu) / (uint32_t)TS, cx1 = (h.x >> 16) / (uint32_t)TS;
  const uint32_t cy0 = (h.y & 0xFFFFu) / (uint32_t)TS, cy1 = (h.y >> 16) / (uint32_t)TS;
  if (cx0 == cx1 && cy0 == cy1) {
    put(views[h.w], (int)cy0, (int)cx0);
    return;
  }
  const SetupRec r = setup[i];
  Spans g;
  spans_init(r, TS, g);
  const DevView& vw = views[r.view];
  thread_tiles(g, TS, [&](uint32_t, int ty, int tx) { put(vw, ty, tx); });
}

template <bool ORDERED>
__global__ __launch_bounds__(256) void k_bin_emit(const SetupRec* __restrict__ setup, const DevView* __restrict__ views,
                                                  const Status* status, uint32_t cap_tri, int TS,
                                                  const uint8_t* __restrict__ wflag,
                                                  const uint32_t* __restrict__ binoff,
                                                  const uint32_t* __restrict__ tile_start,
                                                  uint32_t* __restrict__ tile_cur, uint32_t cap_pairs,
                                                  uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t n = min(status->n_out, cap_tri);
  const uint2* boxes = bin_boxes(const_cast<uint8_t*>(wflag), cap_tri);
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    if constexpr (ORDERED) {
      const uint2 box = boxes[i];
      if (box.x != kWideBox) emit_box_pairs(box, i, binoff[i], cap_pairs, keys, vals);
    } else if (wflag[i] == 0u) {
      bin_emit_tri_unordered(setup, views, TS, i, tile_start, tile_cur, cap_pairs, vals);
    }
  }
}

// Near-first pairs with the scan of their counts fused in: block b takes
// chunk b of the depth order (the chunks of k_scan_reduce, whose sums are in
// partials), scans the counts bincnt[order[j]] from the sums before it, and
// emits each triangle's pairs at its offset from its bin box (a wide
// triangle's offset goes to binoff for k_bin_emit_wide).  The ids, counts and
// boxes of up to four rounds of 256 are loaded before the first round's scan,
// so that a block waits on two dependent loads per four rounds.  Block 0
// publishes the pair total and the overflow flag.
__global__ __launch_bounds__(256) void k_bin_emit_near(Status* status, uint32_t cap_tri,
                                                       const uint8_t* __restrict__ wflag,
                                                       uint32_t* __restrict__ binoff, const uint32_t* __restrict__ order,
                                                       const uint32_t* __restrict__ bincnt,
                                                       const uint32_t* __restrict__ partials, uint32_t cap_pairs,
                                                       uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  __shared__ uint32_t s_tmp[4];
  constexpr int kRounds = 4;
  const uint32_t n = min(status->n_out, cap_tri);
  const uint32_t C = scan_chunk(n);
  const uint32_t b0 = blockIdx.x * C, b1 = min(b0 + C, n);
  const uint2* boxes = bin_boxes(const_cast<uint8_t*>(wflag), cap_tri);
  const uint32_t upto = blockIdx.x == 0 ? (uint32_t)kScanG : blockIdx.x;
  uint32_t p = 0;
  for (uint32_t k = threadIdx.x; k < upto; k += 256) p += partials[k];
  uint32_t sum;
  block_exclusive_scan256(p, s_tmp, &sum);
  uint32_t running = 0;
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) {
      status->n_pairs = sum;
      if (sum > cap_pairs) atomicOr(&status->overflow, 2u);
    }
  } else {
    running = sum;
  }
  for (uint32_t tg = b0; tg < b1; tg += 256u * kRounds) {
    uint32_t id[kRounds], c[kRounds];
    uint2 box[kRounds];
#pragma unroll
    for (int k = 0; k < kRounds; k++) {
      const uint32_t j = tg + 256u * k + threadIdx.x;
      id[k] = j < b1 ? order[j] : 0u;
    }
#pragma unroll
    for (int k = 0; k < kRounds; k++) {
      const bool act = tg + 256u * k + threadIdx.x < b1;
      c[k] = act ? bincnt[id[k]] : 0u;
      box[k] = act ? boxes[id[k]] : make_uint2(0u, 0u);
    }
#pragma unroll
    for (int k = 0; k < kRounds; k++) {
      uint32_t tot;
      const uint32_t o = block_exclusive_scan256(c[k], s_tmp, &tot) + running;
      running += tot;
      if (tg + 256u * k + threadIdx.x >= b1) continue;
      if (box[k].x == kWideBox)
        binoff[id[k]] = o;
      else
        emit_box_pairs(box[k], id[k], o, cap_pairs, keys, vals);
    }
  }
}

template <bool ORDERED>
__global__ __launch_bounds__(256) void k_bin_emit_wide(const SetupRec* __restrict__ setup,
                                                       const DevView* __restrict__ views, const Status* status,
                                                       const uint32_t* __restrict__ wide_list, int TS,
                                                       const uint32_t* __restrict__ binoff,
                                                       const uint32_t* __restrict__ tile_start,
                                                       uint32_t* __restrict__ tile_cur, uint32_t cap_pairs,
                                                       uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  __shared__ int s_lo[4][64];
  __shared__ uint32_t s_off[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for_each_wide(wide_list, status, [&](uint32_t i) {
    const SetupRec r = setup[i];
    const DevView& vw = views[r.view];
    Spans g;
    spans_init(r, TS, g);
    spans_init_rows(r, g);
    const uint32_t o = ORDERED ? binoff[i] : 0u;
    wave_tiles(g, TS, lane, s_lo[wave], s_off[wave], [&](uint32_t e, int ty, int tx) {
      const uint32_t key = vw.tile_base + (uint32_t)ty * vw.ntx + (uint32_t)tx;
      const uint32_t pos = ORDERED ? o + e : tile_start[key] + atomicAdd(&tile_cur[key], 1u);
      if (pos < cap_pairs) {
        if (ORDERED) keys[pos] = key;
        vals[pos] = i;
      }
    });
  });
}

// Binning of a small single-view frame in one workgroup of 1024 threads,
// in place of the general path's dozen launches (which a frame of a few
// thousand triangles spends mostly starting and draining): per-tile counts
// with LDS atomics from each triangle's bin box (a wave per wide triangle,
// as k_bin_count_wide), their scan into the tile ranges, then each
// triangle's id written into its tiles' lists at LDS-atomic cursors.  The
// lists are unordered, which changes no result: a tile's visibility is the
// minimum of its fragments' keys whatever the order (the transparent pass,
// which blends in list order, never comes here).  Pairs past cap_pairs are
// counted, flagged (overflow bit 2) and not written; ranges end at cap_pairs.

__global__ __launch_bounds__(1024) void k_bin_small(const SetupRec* __restrict__ setup,
                                                    const DevView* __restrict__ views, Status* __restrict__ status,
                                                    uint32_t cap_tri, int TS, const uint8_t* __restrict__ wflag,
                                                    uint32_t* __restrict__ wide_list, uint32_t n_tiles,
                                                    uint32_t cap_pairs, uint32_t* __restrict__ vals,
                                                    uint2* __restrict__ ranges) {
  __shared__ uint32_t s_cnt[kSmallBinTiles];
  __shared__ uint32_t s_cur[kSmallBinTiles];
  __shared__ int s_lo[16][64];
  __shared__ uint32_t s_off[16][64];
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_nwide;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (uint32_t k = (uint32_t)tid; k < n_tiles; k += 1024u) s_cnt[k] = 0u;
  if (tid == 0) s_nwide = 0u;
  __syncthreads();
  const uint32_t n = min(status->n_out, cap_tri);
  const uint2* boxes = bin_boxes(const_cast<uint8_t*>(wflag), cap_tri);
  const DevView& vw = views[0];
  // one pass over the triangles' tiles: f(tile key) for each
  auto each_tile = [&](auto f) {
    for (uint32_t i = (uint32_t)tid; i < n; i += 1024u) {
      if (wflag[i] != 0u) continue;  // wide (below) or dropped
      const uint2 box = boxes[i];
      const uint32_t ntx = box.y >> 16;
      uint32_t mask = box.y & 0xFFFFu;
      while (mask) {
        const uint32_t b = (uint32_t)__builtin_ctz(mask);
        mask &= mask - 1u;
        f(i, box.x + (b >> 2) * ntx + (b & 3u));
      }
    }
  };
  auto each_wide_tile = [&](uint32_t nw, auto f) {
    for (uint32_t w = (uint32_t)wave; w < nw; w += 16u) {
      const uint32_t i = (uint32_t)__builtin_amdgcn_readfirstlane((int)wide_list[w]);
      const SetupRec r = setup[i];
      Spans g;
      spans_init(r, TS, g);
      spans_init_rows(r, g);
      wave_tiles(g, TS, lane, s_lo[wave], s_off[wave], [&](uint32_t, int ty, int tx) {
        f(i, vw.tile_base + (uint32_t)ty * vw.ntx + (uint32_t)tx);
      });
    }
  };
  // counts, and the list of wide triangles
  for (uint32_t i = (uint32_t)tid; i < n; i += 1024u)
    if (wflag[i] == 1u) wide_list[atomicAdd(&s_nwide, 1u)] = i;
  each_tile([&](uint32_t, uint32_t key) { atomicAdd(&s_cnt[key], 1u); });
  __syncthreads();
  const uint32_t nw = s_nwide;
  each_wide_tile(nw, [&](uint32_t, uint32_t key) { atomicAdd(&s_cnt[key], 1u); });
  __syncthreads();
  // exclusive scan of the counts, 1024 tiles at a time with a carry
  uint32_t carry = 0;
  for (uint32_t k0 = 0; k0 < n_tiles; k0 += 1024u) {
    const uint32_t k = k0 + (uint32_t)tid;
    const uint32_t c = k < n_tiles ? s_cnt[k] : 0u;
    uint32_t x = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < 16; w++) {
      if (w < wave) before += s_w[w];
      all += s_w[w];
    }
    __syncthreads();
    const uint32_t start = carry + before + x - c;
    if (k < n_tiles) {
      s_cur[k] = start;
      ranges[k] = make_uint2(start, min(start + c, cap_pairs));
    }
    carry += all;
  }
  if (tid == 0) {
    status->n_pairs = carry;
    if (carry > cap_pairs) atomicOr(&status->overflow, 2u);
  }
  __syncthreads();
  // the lists
  auto put = [&](uint32_t i, uint32_t key) {
    const uint32_t pos = atomicAdd(&s_cur[key], 1u);
    if (pos < cap_pairs) vals[pos] = i;
  };
  each_tile(put);
  each_wide_tile(nw, put);
}

__global__ __launch_bounds__(256) void k_ranges_from_counts(const uint32_t* __restrict__ tile_start,
                                                            const uint32_t* __restrict__ tile_cnt, uint32_t n_tiles,
                                                            uint32_t cap_pairs, uint2* __restrict__ ranges) {
  for (uint32_t t = blockIdx.x * 256u + threadIdx.x; t < n_tiles; t += gridDim.x * 256u) {
    const uint32_t s0 = min(tile_start[t], cap_pairs);
    ranges[t] = make_uint2(s0, min(s0 + tile_cnt[t], cap_pairs));
  }
}

// ================================================================ radix sort
// Stable LSD radix sort of (tile key, triangle id) pairs with BITS-bit digits
// (one pass when the tile key fits in 11 bits, else 8-bit passes).  Each wave
// owns one contiguous sub-chunk of the input; k_radix_hist counts its digits,
// the scan over (digit, wave) turns the counts into output bases, and
// k_radix_scatter ranks the wave's items in input order (ballot match) and
// advances its own LDS bases -- no block barrier inside the scatter.  Equal
// keys keep the triangle order.  The chunks of one XCD are contiguous, so a
// tile's output run is written by few XCDs: with ~1 item per (wave, tile),
// interleaved chunks would make every XCD's L2 write back partial lines of
// every run.
constexpr int kRadixG = 256;               // blocks of 4 waves
constexpr int kRadixW = kRadixG * 4;       // waves
constexpr int kRadixU = 8;                 // 64-item groups whose loads a wave issues together
static_assert(2048u * kRadixW == kRadixHistEntries, "histogram size");

__device__ inline uint32_t radix_chunk(uint32_t n) { return (((n + kRadixW - 1u) / kRadixW) + 63u) & ~63u; }

template <int BITS>
__global__ __launch_bounds__(256) void k_radix_hist(const uint32_t* __restrict__ keys, const uint32_t* n_ptr,
                                                    uint32_t cap, int shift, uint32_t* __restrict__ hist) {
  constexpr uint32_t ND = 1u << BITS;
  __shared__ uint32_t h[4][ND];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t n = min(*n_ptr, cap);
  const uint32_t C = radix_chunk(n);
  const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t gw = lb * 4u + (uint32_t)wave;
  const uint32_t b0 = min(gw * C, n), b1 = min(b0 + C, n);
  for (uint32_t d = (uint32_t)tid; d < ND; d += 256)
#pragma unroll
    for (int w = 0; w < 4; w++) h[w][d] = 0;
  __syncthreads();
  // kRadixU loads in flight per lane before their LDS increments
  for (uint32_t tb = b0; tb < b1; tb += 64u * kRadixU) {
    uint32_t k[kRadixU];
#pragma unroll
    for (int u = 0; u < kRadixU; u++) {
      const uint32_t i = tb + 64u * (uint32_t)u + (uint32_t)lane;
      k[u] = i < b1 ? keys[i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kRadixU; u++)
      if (tb + 64u * (uint32_t)u + (uint32_t)lane < b1) atomicAdd(&h[wave][(k[u] >> shift) & (ND - 1u)], 1u);
  }
  __syncthreads();
  for (uint32_t d = (uint32_t)tid; d < ND; d += 256)
#pragma unroll
    for (int w = 0; w < 4; w++) hist[d * kRadixW + lb * 4u + (uint32_t)w] = h[w][d];
}

// KEYS_OUT false: only the values are written (a last pass whose sorted keys
// nobody reads).  vin null: the values are the input positions.  ranges (may be null): [start, end) of each of the first
// n_ranges digits in the output, straight from the scanned histogram -- the
// per-tile lists when this pass sorts by the whole key.
template <int BITS, bool KEYS_OUT>
__global__ __launch_bounds__(256) void k_radix_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                       uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                       const uint32_t* n_ptr, uint32_t cap, int shift,
                                                       const uint32_t* __restrict__ hist_off,
                                                       uint2* __restrict__ ranges, uint32_t n_ranges) {
  constexpr uint32_t ND = 1u << BITS;
  __shared__ uint32_t run[4][ND];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t n = min(*n_ptr, cap);
  const uint32_t C = radix_chunk(n);
  const uint32_t gw = xcd_remap(blockIdx.x, gridDim.x) * 4u + (uint32_t)wave;
  const uint32_t b0 = min(gw * C, n), b1 = min(b0 + C, n);
  if (ranges)
    for (uint32_t d = blockIdx.x * 256u + (uint32_t)tid; d < n_ranges; d += gridDim.x * 256u)
      ranges[d] = make_uint2(hist_off[d * kRadixW], d + 1u < ND ? hist_off[(d + 1u) * kRadixW] : n);
  for (uint32_t d = (uint32_t)lane; d < ND; d += 64) run[wave][d] = hist_off[d * kRadixW + gw];
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (uint32_t tb = b0; tb < b1; tb += 64u * kRadixU) {
    // kRadixU groups of 64 items loaded up front, then ranked in order
    uint32_t kk[kRadixU], vv[kRadixU];
#pragma unroll
    for (int u = 0; u < kRadixU; u++) {
      const uint32_t i = tb + 64u * (uint32_t)u + (uint32_t)lane;
      kk[u] = i < b1 ? kin[i] : 0u;
      vv[u] = i < b1 ? (vin ? vin[i] : i) : 0u;
    }
#pragma unroll
    for (int u = 0; u < kRadixU; u++) {
      const bool valid = tb + 64u * (uint32_t)u + (uint32_t)lane < b1;
      const uint32_t key = kk[u];
      const uint32_t dgt = (key >> shift) & (ND - 1u);
      uint64_t peers = __ballot(valid);
#pragma unroll
      for (int bit = 0; bit < BITS; bit++) {
        const bool on = (dgt >> bit) & 1u;
        const uint64_t m = __ballot(on);
        peers &= on ? m : ~m;
      }
      const uint32_t rank = (uint32_t)__popcll(peers & lt);
      uint32_t base = 0;
      if (valid) base = run[wave][dgt];
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if (valid) {
        if (KEYS_OUT) kout[base + rank] = key;
        vout[base + rank] = vv[u];
        if (rank == 0) run[wave][dgt] = base + (uint32_t)__popcll(peers);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
  }
}

// One-pass sort by keys below DIG (256 or 2048) whose order within a key is
// free: the near-first depth order and tile lists, whose entries only need to
// stay roughly in depth order (a tile's visibility does not depend on the
// order of its list).  Each block owns one contiguous chunk of the input
// (chunks contiguous per XCD, as above), so the histogram has DIG x 256
// entries instead of DIG per wave, and its items take their output slots
// with LDS atomic increments of their key's cursor: a key's entries come
// chunk by chunk in input order, in any order within a chunk.
constexpr int kRadixBG = 256;   // blocks (chunks) of the unordered one-pass sort
constexpr int kRadixBT = 1024;  // threads per chunk: 16 waves, one block per CU
static_assert(2048u * kRadixBG <= kRadixHistEntries, "histogram size");

__device__ inline uint32_t radix_chunk_blk(uint32_t n) { return (((n + kRadixBG - 1u) / kRadixBG) + 255u) & ~255u; }

template <int DIG>
__global__ __launch_bounds__(kRadixBT) void k_radix_hist_blk(const uint32_t* __restrict__ keys, const uint32_t* n_ptr,
                                                             uint32_t cap, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[DIG];
  const int tid = threadIdx.x;
  const uint32_t n = min(*n_ptr, cap);
  const uint32_t C = radix_chunk_blk(n);
  const uint32_t cb = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t b0 = min(cb * C, n), b1 = min(b0 + C, n);
  for (int d = tid; d < DIG; d += kRadixBT) h[d] = 0u;
  __syncthreads();
  for (uint32_t tb = b0; tb < b1; tb += (uint32_t)kRadixBT * kRadixU) {
    uint32_t k[kRadixU];
#pragma unroll
    for (int u = 0; u < kRadixU; u++) {
      const uint32_t i = tb + (uint32_t)kRadixBT * (uint32_t)u + (uint32_t)tid;
      k[u] = i < b1 ? keys[i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kRadixU; u++)
      if (tb + (uint32_t)kRadixBT * (uint32_t)u + (uint32_t)tid < b1) atomicAdd(&h[k[u] & (DIG - 1u)], 1u);
  }
  __syncthreads();
  for (int d = tid; d < DIG; d += kRadixBT) hist[(uint32_t)d * kRadixBG + cb] = h[d];
}

// vout[slot] = vin[i] (vin null: i itself) for the chunk's items, slots from
// the scanned histogram (hist_off); ranges (may be null): [start, end) of each
// of the first n_ranges keys in the output.
template <int DIG>
__global__ __launch_bounds__(kRadixBT) void k_radix_scatter_blk(const uint32_t* __restrict__ kin,
                                                                const uint32_t* __restrict__ vin,
                                                                uint32_t* __restrict__ vout, const uint32_t* n_ptr,
                                                                uint32_t cap, const uint32_t* __restrict__ hist_off,
                                                                uint2* __restrict__ ranges, uint32_t n_ranges) {
  __shared__ uint32_t run[DIG];
  const int tid = threadIdx.x;
  const uint32_t n = min(*n_ptr, cap);
  const uint32_t C = radix_chunk_blk(n);
  const uint32_t cb = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t b0 = min(cb * C, n), b1 = min(b0 + C, n);
  if (ranges)
    for (uint32_t d = blockIdx.x * (uint32_t)kRadixBT + (uint32_t)tid; d < n_ranges; d += gridDim.x * (uint32_t)kRadixBT)
      ranges[d] = make_uint2(hist_off[d * kRadixBG], d + 1u < (uint32_t)DIG ? hist_off[(d + 1u) * kRadixBG] : n);
  for (int d = tid; d < DIG; d += kRadixBT) run[d] = hist_off[(uint32_t)d * kRadixBG + cb];
  __syncthreads();
  for (uint32_t tb = b0; tb < b1; tb += (uint32_t)kRadixBT * kRadixU) {
    uint32_t k[kRadixU], v[kRadixU];
#pragma unroll
    for (int u = 0; u < kRadixU; u++) {
      const uint32_t i = tb + (uint32_t)kRadixBT * (uint32_t)u + (uint32_t)tid;
      k[u] = i < b1 ? kin[i] : 0u;
      v[u] = i < b1 ? (vin ? vin[i] : i) : 0u;
    }
#pragma unroll
    for (int u = 0; u < kRadixU; u++)
      if (tb + (uint32_t)kRadixBT * (uint32_t)u + (uint32_t)tid < b1) vout[atomicAdd(&run[k[u] & (DIG - 1u)], 1u)] = v[u];
  }
}

// A pass's first kernel: its Status words cleared and (opaque pass) the
// draw/view tables copied in from pinned host memory.
__global__ __launch_bounds__(256) void k_frame_begin(uint32_t* __restrict__ status, uint32_t status_words,
                                                     const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                     uint32_t n_words) {
  if (blockIdx.x == 0 && threadIdx.x < status_words) status[threadIdx.x] = 0u;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n_words; i += gridDim.x * 256u) dst[i] = src[i];
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

__global__ __launch_bounds__(256) void k_tile_ranges(const uint32_t* __restrict__ keys, const uint32_t* n_ptr,
                                                     uint32_t cap, uint2* __restrict__ ranges) {
  const uint32_t n = min(*n_ptr, cap);
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    const uint32_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) ranges[k].x = i;
    if (i + 1 == n || keys[i + 1] != k) ranges[k].y = i + 1;
  }
}

// ======================================================== transparent pass
// First opaque clipped-triangle id of each object group (the same prefix as
// k_view_id_base, at the group's first opaque input triangle).
__global__ void k_group_base(const uint32_t* __restrict__ grp_first, uint32_t n_groups, uint32_t n_in,
                             const uint8_t* __restrict__ cnt, const uint32_t* __restrict__ blk_off,
                             const Status* __restrict__ status, uint32_t* __restrict__ grp_base) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_groups) return;
  const uint32_t t0 = grp_first[g];
  uint32_t base;
  if (t0 >= n_in) {
    base = status->n_out;
  } else {
    const uint32_t b = t0 >> 8;
    base = blk_off[b];
    for (uint32_t t = b << 8; t < t0; t++) base += cnt[t] & kCntMask;
  }
  grp_base[g] = base;
}

// Sort keys of a transparent mesh's triangles (scene/mod.rs:1100-1115): the
// distance from the camera to the centroid (a + b + c) / 3, ordered by
// f32::total_cmp reversed.  key = ~(order-preserving bits of the distance),
// so an ascending stable sort gives the descending total order with ties in
// the current triangle order.
__global__ __launch_bounds__(256) void k_tsort_keys(const uint32_t* __restrict__ tri, uint32_t n,
                                                    const float* __restrict__ pos, float cx, float cy, float cz,
                                                    uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const uint32_t* t9 = tri + 9ull * i;
  const V3 a = ld3(pos, t9[0]), b = ld3(pos, t9[1]), c = ld3(pos, t9[2]);
  const V3 ctr = divs(add(add(a, b), c), 3.0f);
  const float d = norm(sub(ctr, v3(cx, cy, cz)));  // Vec3::dist, linalg.rs:177-180
  const uint32_t u = __float_as_uint(d);
  const uint32_t k = (u & 0x80000000u) ? ~u : (u | 0x80000000u);  // total order as unsigned
  keys[i] = ~k;
  vals[i] = i;
}

// out[i] = tri[order[i]] (9 indices per triangle).
__global__ __launch_bounds__(256) void k_permute_tris(const uint32_t* __restrict__ tri,
                                                      const uint32_t* __restrict__ order, uint32_t n,
                                                      uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const uint32_t* src = tri + 9ull * order[i];
  uint32_t* dst = out + 9ull * i;
#pragma unroll
  for (int k = 0; k < 9; k++) dst[k] = src[k];
}

// ============================================================= orchestration
static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

void enqueue_mesh_streams(const uint32_t* tri, uint32_t n_tri, const float* pos, const float* nrm, const float* uv,
                          float* soa, hipStream_t s) {
  if (n_tri == 0) return;
  hipLaunchKernelGGL(k_mesh_streams, dim3(std::min<uint32_t>(cdiv(n_tri, 256), 65535u)), dim3(256), 0, s, tri, n_tri,
                     pos, nrm, uv, soa);
}

// Stable LSD radix sort of (keys[0], vals[0]) by the low key_bits bits of the
// keys (vals[0] may be null: the values are then the input positions), *n_ptr
// items (at most cap): one 8-bit pass when key_bits <= 8, one
// 11-bit pass when key_bits <= 11, else
// 8-bit passes.  Returns the index of the buffer pair holding the result.
// keys_out false: the sorted keys are not written (the values are).  ranges:
// when the sort is one pass (key_bits <= 11), [start, end) of each key below
// n_ranges in the sorted order; not written otherwise.
int enqueue_radix_sort(uint32_t* const keys[2], uint32_t* const vals[2], const uint32_t* n_ptr, uint32_t cap,
                       int key_bits, uint32_t* hist, uint32_t* partials, hipStream_t s, bool keys_out,
                       uint2* ranges, uint32_t n_ranges, bool stable) {
  if (!stable && !keys_out && key_bits <= 11) {
    if (key_bits <= 8) {
      const uint32_t nh = 256u * kRadixBG;
      hipLaunchKernelGGL(k_radix_hist_blk<256>, dim3(kRadixBG), dim3(kRadixBT), 0, s, keys[0], n_ptr, cap, hist);
      scan_u32(hist, hist, nullptr, nh, nh, partials, nullptr, 0, nullptr, 0, s);
      hipLaunchKernelGGL(k_radix_scatter_blk<256>, dim3(kRadixBG), dim3(kRadixBT), 0, s, keys[0], vals[0], vals[1],
                         n_ptr, cap, hist, ranges, ranges ? n_ranges : 0u);
    } else {
      const uint32_t nh = 2048u * kRadixBG;
      hipLaunchKernelGGL(k_radix_hist_blk<2048>, dim3(kRadixBG), dim3(kRadixBT), 0, s, keys[0], n_ptr, cap, hist);
      scan_u32(hist, hist, nullptr, nh, nh, partials, nullptr, 0, nullptr, 0, s);
      hipLaunchKernelGGL(k_radix_scatter_blk<2048>, dim3(kRadixBG), dim3(kRadixBT), 0, s, keys[0], vals[0], vals[1],
                         n_ptr, cap, hist, ranges, ranges ? n_ranges : 0u);
    }
    return 1;
  }
  int cur = 0;
  // one 8-bit pass for keys of up to 8 bits, one 11-bit pass up to 11 bits
  const bool one_pass = key_bits > 8 && key_bits <= 11;
  const int step = one_pass ? 11 : 8;
  for (int shift = 0; shift < key_bits; shift += step) {
    const uint32_t nh = (one_pass ? 2048u : 256u) * kRadixW;
    const bool last = shift + step >= key_bits;
    const bool kout = !last || keys_out;
    uint2* const rg = (last && shift == 0) ? ranges : nullptr;
    const uint32_t nr = rg ? n_ranges : 0u;
    if (one_pass)
      hipLaunchKernelGGL(k_radix_hist<11>, dim3(kRadixG), dim3(256), 0, s, keys[cur], n_ptr, cap, shift, hist);
    else
      hipLaunchKernelGGL(k_radix_hist<8>, dim3(kRadixG), dim3(256), 0, s, keys[cur], n_ptr, cap, shift, hist);
    scan_u32(hist, hist, nullptr, nh, nh, partials, nullptr, 0, nullptr, 0, s);
    if (one_pass) {
      if (kout)
        hipLaunchKernelGGL((k_radix_scatter<11, true>), dim3(kRadixG), dim3(256), 0, s, keys[cur], vals[cur],
                           keys[cur ^ 1], vals[cur ^ 1], n_ptr, cap, shift, hist, rg, nr);
      else
        hipLaunchKernelGGL((k_radix_scatter<11, false>), dim3(kRadixG), dim3(256), 0, s, keys[cur], vals[cur],
                           keys[cur ^ 1], vals[cur ^ 1], n_ptr, cap, shift, hist, rg, nr);
    } else {
      if (kout)
        hipLaunchKernelGGL((k_radix_scatter<8, true>), dim3(kRadixG), dim3(256), 0, s, keys[cur], vals[cur],
                           keys[cur ^ 1], vals[cur ^ 1], n_ptr, cap, shift, hist, rg, nr);
      else
        hipLaunchKernelGGL((k_radix_scatter<8, false>), dim3(kRadixG), dim3(256), 0, s, keys[cur], vals[cur],
                           keys[cur ^ 1], vals[cur ^ 1], n_ptr, cap, shift, hist, rg, nr);
    }
    cur ^= 1;
  }
  return cur;
}

// Fork: work launched on the returned stream runs beside the main stream's
// (the side stream waits for everything enqueued on `s` so far).  Without a
// side stream it is `s` itself.
static hipStream_t fork(hipStream_t s, const PipeArgs& a, int i) {
  if (!a.side) return s;
  (void)hipEventRecord(a.xev[i], s);
  (void)hipStreamWaitEvent(a.side, a.xev[i], 0);
  return a.side;
}
// Join: the main stream waits for everything enqueued on the side stream.
static void join(hipStream_t s, const PipeArgs& a, int i) {
  if (!a.side) return;
  (void)hipEventRecord(a.xev[i], a.side);
  (void)hipStreamWaitEvent(s, a.xev[i], 0);
}

// Status reset, geometry, view bases and binning of one pipeline pass;
// returns the arguments the per-tile stages read (tile lists in b.vals[0]).
PipeArgs enqueue_front(const PipeArgs& a, hipStream_t s) {
  // near-first lists keep the depth sort in keys[1] / vals[1] (behind the
  // wide-triangle list), which needs cap_pairs >= 3 cap_tri
  const bool depth_sorted = a.ordered_bins && a.depth_order && (uint64_t)a.cap_pairs >= 3ull * a.cap_tri;
  // a small single-view frame is binned by one workgroup (k_bin_small)
  const bool small_bins = depth_sorted && a.n_views == 1 && a.n_tiles <= kSmallBinTiles &&
                          a.n_tri_in <= kSmallBinTris;
  if (a.ev[0]) (void)hipEventRecord(a.ev[0], s);
  hipLaunchKernelGGL(k_frame_begin, dim3(a.tables_src ? std::min<uint32_t>(cdiv(a.tables_words, 1024), 64u) : 1u),
                     dim3(256), 0, s, reinterpret_cast<uint32_t*>(a.status), (uint32_t)(sizeof(Status) / 4),
                     a.tables_src, a.tables_dst, a.tables_src ? a.tables_words : 0u);
  const uint32_t n_blk = cdiv(a.n_tri_in, 256);
  // a small single-view frame: one workgroup builds the emitter lists, and
  // both emits share one launch on this stream
  const bool small_geom = a.n_views == 1 && a.n_tri_in <= kSmallBinTris;
  // 1-3 geometry with ordered compaction
  if (n_blk && small_geom) {
    hipLaunchKernelGGL(k_geom_small, dim3(1), dim3(1024), 0, s, a.draws, a.n_draws, a.views, a.n_tri_in, a.cnt, a.blk,
                       a.status, a.cap_tri, a.surv);
    uint8_t* wf = reinterpret_cast<uint8_t*>(a.big);
    uint32_t* dk = depth_sorted && !small_bins ? a.keys[1] + a.cap_tri : nullptr;
    const uint32_t g_simple = n_blk, g_clip = n_blk;
    if (a.ordered_bins)
      hipLaunchKernelGGL(k_geom_emit_both<true>, dim3(g_clip + g_simple), dim3(256), 0, s, a.draws, a.n_draws,
                         a.views, a.surv, a.mats, a.setup, a.shade, a.cap_tri, a.status, a.TS, a.bincnt, wf, dk,
                         g_clip);
    else
      hipLaunchKernelGGL(k_geom_emit_both<false>, dim3(g_clip + g_simple), dim3(256), 0, s, a.draws, a.n_draws,
                         a.views, a.surv, a.mats, a.setup, a.shade, a.cap_tri, a.status, a.TS, nullptr, nullptr,
                         nullptr, g_clip);
  } else if (n_blk) {
    uint32_t* blk_simple = a.blk + n_blk + 1;
    uint32_t* blk_clip = a.blk + 2 * (n_blk + 1);
    hipLaunchKernelGGL(k_geom_count, dim3(n_blk), dim3(256), 0, s, a.draws, a.n_draws, a.views, a.n_tri_in, a.cnt,
                       a.blk, blk_simple, blk_clip);
    if (n_blk <= 4096u) {
      ScanSegs g{};
      g.s[0] = ScanSeg{a.blk, a.blk, &a.status->n_out, &a.status->overflow, a.cap_tri, 1u};
      g.s[1] = ScanSeg{blk_simple, blk_simple, &a.status->n_simple, nullptr, 0u, 0u};
      g.s[2] = ScanSeg{blk_clip, blk_clip, &a.status->n_clip, nullptr, 0u, 0u};
      hipLaunchKernelGGL(k_scan_small, dim3(3), dim3(1024), 0, s, g, n_blk);
    } else {
      scan_u32(a.blk, a.blk, nullptr, n_blk, n_blk, a.partials, &a.status->n_out, a.cap_tri, &a.status->overflow,
               1u, s);
      scan_u32(blk_simple, blk_simple, nullptr, n_blk, n_blk, a.partials, &a.status->n_simple, 0u, nullptr, 0u, s);
      scan_u32(blk_clip, blk_clip, nullptr, n_blk, n_blk, a.partials, &a.status->n_clip, 0u, nullptr, 0u, s);
    }
    hipLaunchKernelGGL(k_geom_compact, dim3(n_blk), dim3(256), 0, s, a.n_tri_in, a.cnt, a.blk, blk_simple, blk_clip,
                       a.status, a.surv);
    // ordered bins: the per-triangle tile counts and wide flags come with the
    // records, and for near-first lists each record's depth key
    uint8_t* wf = reinterpret_cast<uint8_t*>(a.big);
    uint32_t* dk = depth_sorted && !small_bins ? a.keys[1] + a.cap_tri : nullptr;
    // the unclipped and clipped emits write disjoint outputs: the clipped
    // one runs beside the other on the side stream.  Capped grids (the lists'
    // lengths are on the device): 2048 and 1024 blocks, a grid-stride loop
    // beyond that.
    const uint32_t g_simple = std::min<uint32_t>(n_blk, 2048u), g_clip = std::min<uint32_t>(n_blk, 1024u);
    const uint32_t span = (uint64_t)a.n_tri_in >= (uint64_t)a.n_pixels ? kSparseSpanDense : kSparseSpan;
    const hipStream_t s2 = fork(s, a, 0);
    // (the clipped emit, the longer one per thread, starts first on the main
    // stream; the unclipped one waits for the fork on the side stream)
    if (a.ordered_bins) {
      hipLaunchKernelGGL((k_geom_emit<true, true>), dim3(g_clip), dim3(256), 0, s, a.draws, a.n_draws, a.views,
                         a.surv, a.mats, a.setup, a.shade, a.cap_tri, a.status, a.TS, a.bincnt, wf, dk, span);
      hipLaunchKernelGGL((k_geom_emit<true, false>), dim3(g_simple), dim3(256), 0, s2, a.draws, a.n_draws, a.views,
                         a.surv, a.mats, a.setup, a.shade, a.cap_tri, a.status, a.TS, a.bincnt, wf, dk, span);
    } else {
      hipLaunchKernelGGL((k_geom_emit<false, true>), dim3(g_clip), dim3(256), 0, s, a.draws, a.n_draws, a.views,
                         a.surv, a.mats, a.setup, a.shade, a.cap_tri, a.status, a.TS, nullptr, nullptr, nullptr, span);
      hipLaunchKernelGGL((k_geom_emit<false, false>), dim3(g_simple), dim3(256), 0, s2, a.draws, a.n_draws, a.views,
                         a.surv, a.mats, a.setup, a.shade, a.cap_tri, a.status, a.TS, nullptr, nullptr, nullptr, span);
    }
    join(s, a, 1);
  }
  if (a.n_views > 1)
    hipLaunchKernelGGL(k_view_id_base, dim3(cdiv(a.n_views, 256)), dim3(256), 0, s, const_cast<DevView*>(a.views),
                       a.n_views, a.n_tri_in, a.cnt, a.blk, a.status);
  if (a.ev[1]) (void)hipEventRecord(a.ev[1], s);
  // 4-8 binning
  const uint32_t gb = std::min<uint32_t>(cdiv(a.cap_tri, 256), 4096u);
  uint8_t* wflag = reinterpret_cast<uint8_t*>(a.big);
  // the wide-triangle list lives in keys[1] until the radix sort needs it
  // (cap_pairs >= cap_tri); a wave per listed triangle, up to 16384 waves
  uint32_t* wide_list = a.keys[1];
  const uint32_t gl = std::min<uint32_t>(cdiv(a.cap_tri, 4096), 1024u);
  const uint32_t gw = std::min<uint32_t>(cdiv(a.cap_tri, 64), 4096u);
  PipeArgs b = a;
  if (small_bins) {
    hipLaunchKernelGGL(k_bin_small, dim3(1), dim3(1024), 0, s, a.setup, a.views, a.status, a.cap_tri, a.TS, wflag,
                       wide_list, a.n_tiles, a.cap_pairs, a.vals[0], a.ranges);
  } else if (a.ordered_bins) {
    // the wide-triangle list and counts (keys[1][0, cap_tri), bincnt) beside
    // the depth sort (keys[1] + cap_tri .., vals[1], hist, partials)
    const hipStream_t s2 = depth_sorted ? fork(s, a, 2) : s;
    hipLaunchKernelGGL(k_wide_list, dim3(gl), dim3(256), 0, s2, wflag, a.status, a.cap_tri, wide_list,
                       reinterpret_cast<uint32_t*>(a.ranges), 2u * a.n_tiles);
    hipLaunchKernelGGL(k_bin_count_wide<true>, dim3(gw), dim3(256), 0, s2, a.setup, a.views, a.status, a.TS,
                       wide_list, a.bincnt, a.tile_cnt);
    if (depth_sorted) {
      // near first: the clipped triangles in depth-key order (one 8-bit
      // pass over keys[1] + cap_tri, the values the triangle ids themselves,
      // in chunk order within a key; results in vals[1] + cap_tri, the
      // sorted keys not written), then the bin offsets in that order
      uint32_t* const dkeys[2] = {a.keys[1] + a.cap_tri, a.keys[1] + 2ull * a.cap_tri};
      uint32_t* const dvals[2] = {nullptr, a.vals[1] + a.cap_tri};
      const int dc = enqueue_radix_sort(dkeys, dvals, &a.status->n_out, a.cap_tri, kDepthKeyBits, a.hist,
                                        a.partials, s, false, nullptr, 0u, false);
      join(s, a, 3);
      uint32_t* order = dvals[dc];
      // the scan of the counts in that order: its chunk sums here, the rest
      // inside the pair emit
      hipLaunchKernelGGL(k_scan_reduce, dim3(kScanG), dim3(256), 0, s, a.bincnt, &a.status->n_out, 0u, a.cap_tri,
                         a.partials, order);
      hipLaunchKernelGGL(k_bin_emit_near, dim3(kScanG), dim3(256), 0, s, a.status, a.cap_tri, wflag, a.binoff,
                         order, a.bincnt, a.partials, a.cap_pairs, a.keys[0], a.vals[0]);
    } else {
      scan_u32(a.bincnt, a.binoff, &a.status->n_out, 0, a.cap_tri, a.partials, &a.status->n_pairs, a.cap_pairs,
               &a.status->overflow, 2u, s);
      hipLaunchKernelGGL(k_bin_emit<true>, dim3(gb), dim3(256), 0, s, a.setup, a.views, a.status, a.cap_tri, a.TS,
                         wflag, a.binoff, a.tile_start, a.tile_cur, a.cap_pairs, a.keys[0], a.vals[0]);
    }
    hipLaunchKernelGGL(k_bin_emit_wide<true>, dim3(gw), dim3(256), 0, s, a.setup, a.views, a.status, wide_list,
                       a.TS, a.binoff, a.tile_start, a.tile_cur, a.cap_pairs, a.keys[0], a.vals[0]);
    // radix sort by tile key (one pass when the key has at most 11 bits, else
    // stable 8-bit passes), then the tile ranges: from the one pass's scanned
    // histogram, else from the sorted keys.  Near-first lists take the one
    // pass with free order within a chunk of the depth order;
    // submission-ordered lists stay stable.
    const bool one_pass = a.key_bits <= 11;
    const int cur = enqueue_radix_sort(a.keys, a.vals, &a.status->n_pairs, a.cap_pairs, a.key_bits, a.hist,
                                       a.partials, s, !one_pass, one_pass ? a.ranges : nullptr, a.n_tiles,
                                       !depth_sorted);
    b.vals[0] = a.vals[cur];
    if (!one_pass)
      hipLaunchKernelGGL(k_tile_ranges, dim3(std::min<uint32_t>(cdiv(a.cap_pairs, 256), 4096u)), dim3(256), 0, s,
                         a.keys[cur], &a.status->n_pairs, a.cap_pairs, a.ranges);
  } else {
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min<uint32_t>(cdiv(2ull * a.n_tiles, 256), 2048u)), dim3(256), 0, s,
                       a.tile_cnt, 0u, 2u * a.n_tiles);  // tile_cnt and tile_cur are adjacent
    hipLaunchKernelGGL(k_bin_count, dim3(gb), dim3(256), 0, s, a.setup, a.views, a.status, a.cap_tri, a.TS,
                       a.tile_cnt, wflag);
    hipLaunchKernelGGL(k_wide_list, dim3(gl), dim3(256), 0, s, wflag, a.status, a.cap_tri, wide_list, nullptr, 0u);
    hipLaunchKernelGGL(k_bin_count_wide<false>, dim3(gw), dim3(256), 0, s, a.setup, a.views, a.status, a.TS,
                       wide_list, a.bincnt, a.tile_cnt);
    scan_u32(a.tile_cnt, a.tile_start, nullptr, a.n_tiles, a.n_tiles, a.partials, &a.status->n_pairs, a.cap_pairs,
             &a.status->overflow, 2u, s);
    hipLaunchKernelGGL(k_bin_emit<false>, dim3(gb), dim3(256), 0, s, a.setup, a.views, a.status, a.cap_tri, a.TS,
                       wflag, a.binoff, a.tile_start, a.tile_cur, a.cap_pairs, a.keys[0], a.vals[0]);
    hipLaunchKernelGGL(k_bin_emit_wide<false>, dim3(gw), dim3(256), 0, s, a.setup, a.views, a.status, wide_list,
                       a.TS, a.binoff, a.tile_start, a.tile_cur, a.cap_pairs, a.keys[0], a.vals[0]);
    hipLaunchKernelGGL(k_ranges_from_counts, dim3(std::min<uint32_t>(cdiv(a.n_tiles, 256), 2048u)), dim3(256), 0, s,
                       a.tile_start, a.tile_cnt, a.n_tiles, a.cap_pairs, a.ranges);
  }
  if (a.ev[2]) (void)hipEventRecord(a.ev[2], s);
  return b;
}

// Before the transparent pass: object-group bases of the opaque ids (from
// the opaque pass's counts, still in cnt/blk), then each transparent mesh
// sorted back to front in place -- the mesh keeps the new order, as the
// reference's sort_by on the mesh's triangle vector does.
void enqueue_transparent_prep(const PipeArgs& opaque, const TransArgs& t, hipStream_t s) {
  if (t.n_groups)
    hipLaunchKernelGGL(k_group_base, dim3(cdiv(t.n_groups, 256)), dim3(256), 0, s, t.grp_first, t.n_groups,
                       opaque.n_tri_in, opaque.cnt, opaque.blk, opaque.status, t.grp_base);
  for (uint32_t m = 0; m < t.n_meshes; m++) {
    const TransMesh& tm = t.meshes[m];
    if (tm.n_tri < 2) continue;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, t.sort_n, tm.n_tri, 1u);
    hipLaunchKernelGGL(k_tsort_keys, dim3(cdiv(tm.n_tri, 256)), dim3(256), 0, s, tm.tri, tm.n_tri, tm.pos,
                       t.cam[0], t.cam[1], t.cam[2], opaque.keys[0], opaque.vals[0]);
    const int cur = enqueue_radix_sort(opaque.keys, opaque.vals, t.sort_n, opaque.cap_pairs, 32, opaque.hist,
                                       opaque.partials, s, false, nullptr, 0u);
    hipLaunchKernelGGL(k_permute_tris, dim3(cdiv(tm.n_tri, 256)), dim3(256), 0, s, tm.tri, opaque.vals[cur],
                       tm.n_tri, t.perm_tmp);
    (void)hipMemcpyAsync(tm.tri, t.perm_tmp, 36ull * tm.n_tri, hipMemcpyDeviceToDevice, s);
  }
}

// The frame's status words into coherent pinned host memory (one wave, in
// stream order after every kernel that writes them).
__global__ void k_status_out(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, uint32_t n_words) {
  for (uint32_t i = threadIdx.x; i < n_words; i += blockDim.x) dst[i] = src[i];
  __threadfence_system();
}

void enqueue_status_out(const Status* dev, Status* host_dev_ptr, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_status_out, dim3(1), dim3(64), 0, s, reinterpret_cast<const uint32_t*>(dev),
                     reinterpret_cast<uint32_t*>(host_dev_ptr), (uint32_t)(n * sizeof(Status) / 4));
}

// One shard of a sort-last frame: geometry, binning, raster, then the packed
// u64 visibility (no shading: dh_shade_shard shades after the merge).
void enqueue_visibility(const PipeArgs& a, uint32_t shard, uint64_t* out, hipStream_t s) {
  const PipeArgs b = enqueue_front(a, s);
  enqueue_raster(b, s);
  if (a.ev[3]) (void)hipEventRecord(a.ev[3], s);
  enqueue_vis_pack(a, shard, out, s);
  if (a.ev[4]) (void)hipEventRecord(a.ev[4], s);
}

void enqueue_frame(const PipeArgs& a, hipStream_t s) {
  PipeArgs b = enqueue_front(a, s);
  // raster, shade (dh_raster.hip): the kind-0 raster shades its own tiles
  b.shade_in_raster = a.raster_kind == 0;
  enqueue_raster(b, s);
  if (a.ev[3]) (void)hipEventRecord(a.ev[3], s);
  if (!b.shade_in_raster) enqueue_shade(a, s);
  if (a.ev[4]) (void)hipEventRecord(a.ev[4], s);
}

}  // namespace dh
