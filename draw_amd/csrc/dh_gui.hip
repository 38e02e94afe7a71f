// dh_gui.hip -- the GUI overlay rasterizer: Gui::render (app/gui.rs:382-478)
// over Canvas::draw_triangle (canvas.rs:435-575).  Per-vertex colour and
// alpha, the RGBA font texture through get_rgba_slice (floor(u * w), not
// w - 1; scene/mod.rs:137-152), a per-command clip rectangle, Pixel products
// truncated to u8 and sums wrapping, depth tested against 0 and never
// written.  Fragments blend over the current pixel, so each pixel applies
// the draw list in order: one workgroup per 16x16 tile compacts the
// triangles whose scan rectangle meets the tile (keeping their order), then
// every lane walks that list for its pixel.
#include "dh_common.hpp"
#include "dh_internal.hpp"

namespace dh {

struct alignas(16) GuiRec {
  float ea[4], eb[4], eg[4];  // edge functions bc, ca, ab: A, B, C, D
  float f[4];                 // f_alpha, f_beta, f_gama, -
  float uv[6];                // texture coordinates after the 1 - v flip
  float alpha[3];             // vertex alpha col[3] / 255
  uint32_t rgb[3];            // vertex colour r | g << 8 | b << 16
  uint32_t x0, x1, y0, y1;    // inclusive scan rectangle
  uint32_t ties, pad[2];
};
static_assert(sizeof(GuiRec) == kGuiRecBytes, "GuiRec layout");

// Edge function of (p, q) as A, B, C, D: ((A x + B y) + C) - D, canvas.rs:461-480.
__device__ inline void edge_of(float px, float py, float qx, float qy, float e[4]) {
  e[0] = py - qy;
  e[1] = qx - px;
  e[2] = px * qy;
  e[3] = qx * py;
}

// Rectangle from_coords / clip on (x0, y0, x1, y1) inclusive (canvas.rs:315-350).
struct Rect { uint64_t x0, y0, x1, y1; };
__device__ inline Rect rect_from(uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
  Rect r;
  r.x0 = a < c ? a : c; r.x1 = a < c ? c : a;
  r.y0 = b < d ? b : d; r.y1 = b < d ? d : b;
  return r;
}
__device__ inline Rect rect_clip(const Rect& a, const Rect& b) {
  uint64_t x0 = a.x0 > b.x0 ? a.x0 : b.x0, y0 = a.y0 > b.y0 ? a.y0 : b.y0;
  uint64_t x1 = a.x1 < b.x1 ? a.x1 : b.x1, y1 = a.y1 < b.y1 ? a.y1 : b.y1;
  if (x0 > x1) { x0 = 0; x1 = 0; }
  if (y0 > y1) { y0 = 0; y1 = 0; }
  return rect_from(x0, y0, x1, y1);
}

__global__ __launch_bounds__(256) void k_gui_setup(const dh_gui_vertex* __restrict__ vtx, const uint4* __restrict__ tris,
                                                   uint32_t n_tri, const uint4* __restrict__ clips, int W, int H,
                                                   GuiRec* __restrict__ recs) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= n_tri) return;
  const uint4 q = tris[t];
  const uint32_t vi[3] = {q.x, q.y, q.z};
  float cx[3], cy[3];
  GuiRec r;
  for (int k = 0; k < 3; k++) {  // gui.rs:420-466: invert y, flip v, alpha = col[3] / 255
    const dh_gui_vertex& v = vtx[vi[k]];
    const float sx = v.pos[0], sy = ((float)H - 1.0f) - v.pos[1];
    cx[k] = floorf(sx + 0.5f);  // pos_map_center, canvas.rs:896-904 (no canvas offset)
    cy[k] = floorf(sy + 0.5f);
    r.uv[2 * k] = v.uv[0];
    r.uv[2 * k + 1] = 1.0f - v.uv[1];
    r.alpha[k] = (float)v.col[3] / 255.0f;
    r.rgb[k] = (uint32_t)v.col[0] | ((uint32_t)v.col[1] << 8) | ((uint32_t)v.col[2] << 16);
  }
  edge_of(cx[1], cy[1], cx[2], cy[2], r.ea);  // f_bc
  edge_of(cx[2], cy[2], cx[0], cy[0], r.eb);  // f_ca
  edge_of(cx[0], cy[0], cx[1], cy[1], r.eg);  // f_ab
  r.f[0] = edge_eval(r.ea, cx[0], cy[0]);
  r.f[1] = edge_eval(r.eb, cx[1], cy[1]);
  r.f[2] = edge_eval(r.eg, cx[2], cy[2]);
  r.f[3] = 0.0f;
  r.ties = 0;
  if (r.f[0] * edge_eval(r.ea, -1.0f, -1.0f) > 0.0f) r.ties |= 1u;
  if (r.f[1] * edge_eval(r.eb, -1.0f, -1.0f) > 0.0f) r.ties |= 2u;
  if (r.f[2] * edge_eval(r.eg, -1.0f, -1.0f) > 0.0f) r.ties |= 4u;
  float mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;  // canvas.rs:491-510
  for (int k = 0; k < 3; k++) {
    if (cx[k] < mnx) mnx = cx[k];
    if (cy[k] < mny) mny = cy[k];
    if (cx[k] > mxx) mxx = cx[k];
    if (cy[k] > mxy) mxy = cy[k];
  }
  Rect drawable = rect_from(f2usize(mnx), f2usize(mny), f2usize(mxx), f2usize(mxy));
  drawable = rect_clip(drawable, rect_from(0, 0, (uint64_t)W - 1u, (uint64_t)H - 1u));
  const uint4 cl = clips[q.w];
  const Rect valid = rect_clip(rect_from(cl.x, cl.y, cl.z, cl.w), drawable);
  r.x0 = (uint32_t)valid.x0; r.x1 = (uint32_t)valid.x1;
  r.y0 = (uint32_t)valid.y0; r.y1 = (uint32_t)valid.y1;
  r.pad[0] = r.pad[1] = 0;
  recs[t] = r;
}

// get_rgba_slice (scene/mod.rs:137-152) in release semantics: false when the
// 4-byte slice leaves the image (the reference panics).
__device__ inline bool rgba_offset(int32_t w, int32_t h, int32_t comp, float u, float v, uint64_t* off) {
  const uint64_t ui = f2usize(floorf(u * (float)w));
  const uint64_t vi = ((uint64_t)h - 1u) - f2usize(floorf(v * (float)h));
  const uint64_t o = (vi * (uint64_t)w + ui) * (uint64_t)comp;
  const uint64_t end = o + 4u, len = (uint64_t)w * (uint64_t)h * (uint64_t)comp;
  *off = o;
  return !(o > end || end > len);
}

// Pixel * f32 then Pixel + Pixel per channel (canvas.rs:136-169).
__device__ inline uint32_t pmul(uint32_t p, float s) {
  return f2u8((float)(p & 255u) * s) | (f2u8((float)((p >> 8) & 255u) * s) << 8) |
         (f2u8((float)((p >> 16) & 255u) * s) << 16);
}
__device__ inline uint32_t padd(uint32_t a, uint32_t b) {
  return ((a + b) & 255u) | ((((a >> 8) + (b >> 8)) & 255u) << 8) | ((((a >> 16) + (b >> 16)) & 255u) << 16);
}

constexpr int kGuiTS = 16;

__global__ __launch_bounds__(256) void k_gui_raster(const GuiRec* __restrict__ recs, uint32_t n_tri, int W, int H,
                                                    uint32_t ntx, const uint8_t* __restrict__ font, int32_t fw,
                                                    int32_t fh, int32_t fcomp, uint32_t* __restrict__ bgra,
                                                    const float* __restrict__ depth, uint32_t* __restrict__ error) {
  constexpr int CH = 256;
  __shared__ uint32_t s_hit[CH];
  __shared__ uint32_t s_cnt[4];
  const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const int tx0 = (int)(tile % ntx) * kGuiTS, ty0 = (int)(tile / ntx) * kGuiTS;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int px = tx0 + tid % kGuiTS, py = ty0 + tid / kGuiTS;
  const bool ok = px < W && py < H;
  // the canvas keeps rgb in r, g, b order inside its B,G,R,A words
  uint32_t cur = 0;
  bool pass = false;
  if (ok) {
    const uint32_t w = bgra[(uint32_t)(H - 1 - py) * (uint32_t)W + (uint32_t)px];
    cur = ((w >> 16) & 255u) | (w & 0xFF00u) | ((w & 255u) << 16);
    pass = 0.0f < depth[(uint32_t)py * (uint32_t)W + (uint32_t)px];
  }
  bool wrote = false;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (uint32_t base = 0; base < n_tri; base += CH) {
    // compact the chunk's triangles that meet this tile, in order
    __syncthreads();
    const uint32_t t = base + (uint32_t)tid;
    bool hit = false;
    if (t < n_tri) {
      const GuiRec& r = recs[t];
      hit = !((int)r.x1 < tx0 || (int)r.x0 > tx0 + kGuiTS - 1 || (int)r.y1 < ty0 || (int)r.y0 > ty0 + kGuiTS - 1);
    }
    const uint64_t m = __ballot(hit);
    if (lane == 0) s_cnt[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int w = 0; w < 4; w++) {
      if (w < wave) before += s_cnt[w];
      total += s_cnt[w];
    }
    __syncthreads();
    if (hit) s_hit[before + (uint32_t)__popcll(m & lt)] = t;
    __syncthreads();
    for (uint32_t k = 0; k < total; k++) {
      const GuiRec& r = recs[s_hit[k]];
      if (!ok || px < (int)r.x0 || px > (int)r.x1 || py < (int)r.y0 || py > (int)r.y1) continue;
      const float xf = (float)px, yf = (float)py;
      const float al = edge_eval(r.ea, xf, yf) / r.f[0];
      const float be = edge_eval(r.eb, xf, yf) / r.f[1];
      const float ga = edge_eval(r.eg, xf, yf) / r.f[2];
      if (!(al >= 0.0f && be >= 0.0f && ga >= 0.0f)) continue;
      if (!((al > 0.0f || (r.ties & 1u)) && (be > 0.0f || (r.ties & 2u)) && (ga > 0.0f || (r.ties & 4u)))) continue;
      uint32_t col = padd(padd(pmul(r.rgb[0], al), pmul(r.rgb[1], be)), pmul(r.rgb[2], ga));
      const float calpha = (al * r.alpha[0] + be * r.alpha[1]) + ga * r.alpha[2];
      const float u = (r.uv[0] * al + r.uv[2] * be) + r.uv[4] * ga;
      const float v = (r.uv[1] * al + r.uv[3] * be) + r.uv[5] * ga;
      uint64_t off;
      if (!rgba_offset(fw, fh, fcomp, u, v, &off)) {
        atomicOr(error, 1u);
        continue;
      }
      const uint32_t tex = (uint32_t)font[off] | ((uint32_t)font[off + 1] << 8) | ((uint32_t)font[off + 2] << 16);
      const float ta = (float)font[off + 3] / 255.0f;
      col = padd(pmul(col, ta), pmul(tex, 1.0f - ta));
      const float fin = calpha * ta;
      const uint32_t out = fin < 1.0f ? padd(pmul(cur, 1.0f - fin), pmul(col, fin)) : col;
      if (pass) {  // draw_pixel_coord_with_depth(x, y, colour, alpha, 0.0); depth writes off
        cur = out;
        wrote = true;
      }
    }
  }
  if (ok && wrote)  // B, G, R and the pad byte 0 of a Pixel sum
    bgra[(uint32_t)(H - 1 - py) * (uint32_t)W + (uint32_t)px] =
        ((cur >> 16) & 255u) | (cur & 0xFF00u) | ((cur & 255u) << 16);
}

hipError_t enqueue_gui(const dh_gui_vertex* vtx, const uint4* tris, uint32_t n_tri, const uint4* clips, void* recs,
                       int W, int H, const uint8_t* font, int32_t fw, int32_t fh, int32_t fcomp, uint32_t* bgra,
                       const float* depth, uint32_t* error, hipStream_t s) {
  if (n_tri == 0) return hipSuccess;
  GuiRec* r = reinterpret_cast<GuiRec*>(recs);
  hipLaunchKernelGGL(k_gui_setup, dim3((n_tri + 255u) / 256u), dim3(256), 0, s, vtx, tris, n_tri, clips, W, H, r);
  const uint32_t ntx = (uint32_t)((W + kGuiTS - 1) / kGuiTS), nty = (uint32_t)((H + kGuiTS - 1) / kGuiTS);
  hipLaunchKernelGGL(k_gui_raster, dim3(ntx * nty), dim3(256), 0, s, r, n_tri, W, H, ntx, font, fw, fh, fcomp, bgra,
                     depth, error);
  return hipGetLastError();
}

}  // namespace dh
