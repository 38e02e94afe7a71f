// dh_api.cpp -- the C ABI of libdrawhip: contexts, device-resident resources
// (textures, materials, objects, meshes, canvases) and dh_render.
//
// A frame is enqueued without host synchronisation; the scratch buffers it
// needs are sized from the previous frame.  When a device-side count exceeds a
// capacity, the kernels stop writing past it and raise a flag, and dh_render
// grows the buffer to the measured size and renders the frame again.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "dh_common.hpp"
#include "dh_internal.hpp"
#include "dh_pipeline.hpp"

namespace {

enum Kind : uint32_t { kTex = 1, kMat = 2, kObj = 3, kMesh = 4, kCanvas = 5 };

inline dh_handle make_handle(Kind k, uint64_t id) { return ((uint64_t)k << 56) | id; }
inline Kind handle_kind(dh_handle h) { return (Kind)(h >> 56); }

struct TexRes {
  uint8_t* d = nullptr;
  int32_t w = 0, h = 0, comp = 0;
};
struct MatRes {
  uint32_t slot = 0;
  float alpha = 1.0f;
};
struct ObjRes {
  float* pos = nullptr;
  float* nrm = nullptr;
  float* uv = nullptr;
  uint32_t n_pos = 0, n_nrm = 0, n_uv = 0;
};
struct MeshRes {
  dh_handle obj = 0;
  uint32_t* tri = nullptr;
  float* soa = nullptr;  // per-triangle vertex streams and records (opaque material), or null
  uint32_t n_tri = 0;
  dh_handle mat = 0;
};
struct CanvasRes {
  int32_t W = 0, H = 0;
  uint32_t* bgra = nullptr;
  float* depth = nullptr;
  uint32_t* winner = nullptr;
};

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t ensure(size_t want) {
    if (want <= n && p) return hipSuccess;
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(want, 1) * sizeof(T));
    if (e == hipSuccess) n = std::max<size_t>(want, 1);
    return e;
  }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// Device-resident resources.  One pool may serve several contexts (streams)
// on the same device (dh_ctx_create_shared); every access holds mu.
struct ResPool {
  std::mutex mu;
  // held by a frame with a transparent pass from its launch until it has
  // completed on the device: that pass sorts each transparent mesh's index
  // buffer in place (scene/mod.rs:1100), so two contexts sharing the pool
  // must not sort one mesh at the same time
  std::mutex trans_mu;
  int device = 0;
  uint64_t next_id = 1;
  std::unordered_map<uint64_t, TexRes> tex;
  std::unordered_map<uint64_t, MatRes> mats;
  std::unordered_map<uint64_t, ObjRes> objs;
  std::unordered_map<uint64_t, MeshRes> meshes;
  std::unordered_map<uint64_t, CanvasRes> canvases;
  // material table (slot 0 = Texture::default)
  std::vector<dh::DevMaterial> mat_host;
  DevBuf<dh::DevMaterial> mat_dev;
  bool mat_dirty = true;
  uint8_t* white_rgb = nullptr;  // TextureMap::default
  ~ResPool();
};

}  // namespace

struct dh_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  // a second stream (default priority) for independent kernels of a frame,
  // forked from and joined back to `stream` with these events
  hipStream_t side = nullptr;
  hipEvent_t xev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::string err;
  std::shared_ptr<ResPool> res;
  // raster configuration
  int TS = 16, waves = 4;
  int raster_kind = 0;
  bool graph_frames = true;  // dh_render replays a captured launch sequence while the frame structure repeats
  bool profiling = false;
  hipEvent_t ev[5] = {};
  // scratch
  DevBuf<uint8_t> cnt;
  DevBuf<uint32_t> surv;  // (input triangle, first output id) of each input triangle that emits
  DevBuf<uint32_t> blk;
  DevBuf<dh::SetupRec> setup;
  DevBuf<dh::ShadeRec> shade;
  DevBuf<uint32_t> bincnt, binoff, big;
  DevBuf<uint32_t> keys0, keys1, vals0, vals1;
  DevBuf<uint32_t> hist, partials;
  DevBuf<uint2> ranges;
  DevBuf<uint32_t> tile_cnt, tile_start;
  int bin_mode = 2;  // dh_set_bin_order: 0 unordered, 1 submission order, 2 near first
  DevBuf<uint32_t> vis;
  DevBuf<dh::Status> status;
  // draw and view tables of a frame in one device buffer (one upload per
  // frame from the pinned h_tables): draws at 0, views at tables_view_off
  DevBuf<uint8_t> tables;
  size_t tables_view_off = 0;
  size_t tables_used = 0;  // bytes of the current frame's tables
  // sort-last shard frames: dh_render_shard sets shard_out for its call; the
  // frame it leaves (records, tables) is what dh_shade_shard shades
  uint64_t* shard_out = nullptr;
  uint32_t shard = 0;
  bool shard_frame = false;
  uint64_t shard_n_in = 0;
  uint32_t shard_n_tiles = 0, shard_n_pixels = 0;
  dh::DevDraw* draws = nullptr;
  dh::DevView* views = nullptr;
  uint8_t* h_tables = nullptr;  // coherent pinned host staging of the same layout
  uint32_t* d_tables_host = nullptr;  // its device address (read by the frame's first kernel)
  size_t h_tables_n = 0;
  dh::Status* d_status_host = nullptr;  // device address of h_status
  DevBuf<uint32_t> export_buf;  // RGBA staging of dh_canvas_read_rgba
  DevBuf<dh_gui_vertex> gui_vtx;  // GUI overlay draw list
  DevBuf<uint4> gui_tri, gui_clip;
  DevBuf<uint4> gui_rec;          // dh::kGuiRecBytes per triangle
  DevBuf<uint32_t> gui_err;
  uint64_t scratch_cap = 0;     // 0 = none; else frame scratch may not grow past it (DH_E_BUDGET)
  uint64_t scratch_need = 0;    // bytes the last frame refused by the cap asked for
  uint32_t cap_tri = 0, cap_pairs = 0;
  // pinned staging
  dh::Status* h_status = nullptr;
  std::vector<dh::DevDraw> h_draws;
  std::vector<dh::DevView> h_views;
  // transparent pass of a one-view frame
  std::vector<dh::DevDraw> h_tdraws;
  std::vector<dh::TransMesh> h_tmeshes;
  std::vector<uint32_t> h_grp_first;  // first opaque input triangle of each object group
  uint64_t n_tin = 0, max_tmesh = 0;
  DevBuf<dh::DevDraw> tdraws;
  DevBuf<uint32_t> grp;       // [2 * groups]: first input triangle, then first clipped id
  DevBuf<uint32_t> tperm;     // [9 * max_tmesh]
  DevBuf<uint32_t> sort_n;    // [1]
  dh_stats stats = {};
  // captured batch launch (DH_BATCH_GRAPH)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<uint64_t> graph_key;
};

namespace {

int fail(dh_ctx c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(dh_ctx c, hipError_t e, const char* what) {
  return fail(c, e == hipErrorOutOfMemory ? DH_E_NOMEM : DH_E_HIP,
              std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(c, expr)                                  \
  do {                                                   \
    hipError_t _e = (expr);                              \
    if (_e != hipSuccess) return hip_fail((c), _e, #expr); \
  } while (0)

int key_bits_for(uint32_t n) {
  int b = 1;
  while (b < 32 && (1ull << b) < (uint64_t)n) b++;
  return b;
}

int upload_materials(dh_ctx c) {
  if (!c->res->mat_dirty) return DH_OK;
  HIPCHK(c, c->res->mat_dev.ensure(c->res->mat_host.size()));
  HIPCHK(c, hipMemcpyAsync(c->res->mat_dev.p, c->res->mat_host.data(), c->res->mat_host.size() * sizeof(dh::DevMaterial),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->res->mat_dirty = false;
  return DH_OK;
}

dh::DevMaterial default_material(const uint8_t* white) {
  dh::DevMaterial m;
  memset(&m, 0, sizeof(m));
  for (int i = 0; i < 3; i++) {  // Texture::default, scene/mod.rs:237-252
    m.ka[i] = 0.9f;
    m.kd[i] = 0.4f;
    m.ks[i] = 0.5f;
  }
  m.alpha = 1.0f;
  m.power = 2;
  m.kd_img = white;
  m.ka_img = white;
  m.kd_w = m.kd_h = m.ka_w = m.ka_h = 1;
  m.kd_comp = m.ka_comp = 3;
  return m;
}

// The draw/view table layout: draws, then views at a 256-byte boundary
// (copied as 4-byte words by the frame's first kernel).
static_assert(sizeof(dh::DevView) % 4 == 0, "view table copied in words");
size_t draw_bytes(uint64_t n_draws) { return (std::max<uint64_t>(n_draws, 1) * sizeof(dh::DevDraw) + 255) & ~(size_t)255; }
size_t tables_bytes(uint64_t n_draws, uint64_t n_views) {
  return draw_bytes(n_draws) + std::max<uint64_t>(n_views, 1) * sizeof(dh::DevView);
}

// Bytes every scratch buffer would hold after ensure_scratch with these sizes.
uint64_t scratch_bytes_after(dh_ctx c, uint64_t n_in, uint64_t cap_tri, uint64_t cap_pairs, uint64_t n_tiles,
                             uint64_t n_pixels, uint64_t n_draws, uint64_t n_views) {
  auto mx = [](uint64_t have, uint64_t want) { return std::max<uint64_t>(have, std::max<uint64_t>(want, 1)); };
  uint64_t b = mx(c->cnt.n, n_in) + 4 * mx(c->blk.n, 3 * ((n_in + 255) / 256 + 1)) + 4 * mx(c->surv.n, 2 * n_in);
  b += (sizeof(dh::SetupRec) * mx(c->setup.n, cap_tri)) + sizeof(dh::ShadeRec) * mx(c->shade.n, cap_tri);
  b += 4 * (mx(c->bincnt.n, cap_tri) + mx(c->binoff.n, cap_tri) + mx(c->big.n, 3 * cap_tri + 4));
  b += 4 * (mx(c->keys0.n, cap_pairs) + mx(c->keys1.n, cap_pairs) + mx(c->vals0.n, cap_pairs) + mx(c->vals1.n, cap_pairs));
  b += 4 * (mx(c->hist.n, dh::kRadixHistEntries) + mx(c->partials.n, 1024)) + 8 * mx(c->ranges.n, n_tiles);
  b += 4 * (mx(c->tile_cnt.n, 2 * n_tiles) + mx(c->tile_start.n, n_tiles) + mx(c->vis.n, n_pixels));
  b += sizeof(dh::Status) * mx(c->status.n, 2) + mx(c->tables.n, tables_bytes(n_draws, n_views));
  return b;
}

// Grow every scratch buffer to the given capacities.
int ensure_scratch(dh_ctx c, uint32_t n_in, uint32_t cap_tri, uint32_t cap_pairs, uint32_t n_tiles,
                   uint32_t n_pixels, uint32_t n_draws, uint32_t n_views) {
  HIPCHK(c, c->cnt.ensure(n_in));
  HIPCHK(c, c->blk.ensure(3ull * ((n_in + 255) / 256 + 1)));  // output, unclipped and clipped emitter counts
  HIPCHK(c, c->surv.ensure(2ull * n_in));
  HIPCHK(c, c->setup.ensure(cap_tri));
  HIPCHK(c, c->shade.ensure(cap_tri));
  HIPCHK(c, c->bincnt.ensure(cap_tri));
  HIPCHK(c, c->binoff.ensure(cap_tri));
  HIPCHK(c, c->big.ensure(3ull * cap_tri + 4));  // flag bytes, then 8-byte bin boxes (16-byte aligned)
  HIPCHK(c, c->keys0.ensure(cap_pairs));
  HIPCHK(c, c->keys1.ensure(cap_pairs));
  HIPCHK(c, c->vals0.ensure(cap_pairs));
  HIPCHK(c, c->vals1.ensure(cap_pairs));
  HIPCHK(c, c->hist.ensure(dh::kRadixHistEntries));
  HIPCHK(c, c->partials.ensure(1024));
  HIPCHK(c, c->ranges.ensure(n_tiles));
  HIPCHK(c, c->tile_cnt.ensure(2ull * n_tiles));
  HIPCHK(c, c->tile_start.ensure(n_tiles));
  HIPCHK(c, c->vis.ensure(n_pixels));
  HIPCHK(c, c->status.ensure(2));  // opaque pass, transparent pass
  {
    const size_t need = tables_bytes(n_draws, n_views);
    if (c->tables.n < need || !c->tables.p) {
      HIPCHK(c, c->tables.ensure(need));
    }
    if (c->h_tables_n < need) {
      if (c->h_tables) hipHostFree(c->h_tables);
      c->h_tables = nullptr;
      c->h_tables_n = 0;
      HIPCHK(c, hipHostMalloc(reinterpret_cast<void**>(&c->h_tables), need, hipHostMallocCoherent));
      HIPCHK(c, hipHostGetDevicePointer(reinterpret_cast<void**>(&c->d_tables_host), c->h_tables, 0));
      c->h_tables_n = need;
    }
    c->tables_view_off = draw_bytes(n_draws);
    c->draws = reinterpret_cast<dh::DevDraw*>(c->tables.p);
    c->views = reinterpret_cast<dh::DevView*>(c->tables.p + c->tables_view_off);
  }
  c->cap_tri = (uint32_t)std::min<size_t>(c->setup.n, 0xFFFFFFF0u);
  c->cap_pairs = (uint32_t)std::min<size_t>(c->keys0.n, 0xFFFFFFF0u);
  // binning keeps the wide-triangle list (up to cap_tri entries) in keys[1]
  if (c->cap_pairs < c->cap_tri) return fail(c, DH_E_CAPACITY, "bin buffers smaller than the triangle records");
  return DH_OK;
}

}  // namespace

// ------------------------------------------------------------------ context
ResPool::~ResPool() {
  hipSetDevice(device);
  for (auto& kv : tex) hipFree(kv.second.d);
  for (auto& kv : objs) {
    hipFree(kv.second.pos);
    hipFree(kv.second.nrm);
    hipFree(kv.second.uv);
  }
  for (auto& kv : meshes) {
    hipFree(kv.second.tri);
    if (kv.second.soa) hipFree(kv.second.soa);
  }
  for (auto& kv : canvases) {
    hipFree(kv.second.bgra);
    hipFree(kv.second.depth);
    hipFree(kv.second.winner);
  }
  mat_dev.release();
  if (white_rgb) hipFree(white_rgb);
}

namespace {

// A context on `device` with its own stream, events and scratch; resources
// from `pool`, or a new pool when null.
int ctx_new(int32_t device, std::shared_ptr<ResPool> pool, dh_ctx* out) {
  dh_ctx c = new dh_ctx_s();
  c->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
  for (int i = 0; i < 4 && e == hipSuccess; i++) e = hipEventCreateWithFlags(&c->xev[i], hipEventDisableTiming);
  // the frame's last kernel writes the status words here (coherent: no copy, no cache flush)
  if (e == hipSuccess)
    e = hipHostMalloc(reinterpret_cast<void**>(&c->h_status), 2 * sizeof(dh::Status), hipHostMallocCoherent);
  if (e == hipSuccess) e = hipHostGetDevicePointer(reinterpret_cast<void**>(&c->d_status_host), c->h_status, 0);
  for (int i = 0; i < 5 && e == hipSuccess; i++) e = hipEventCreate(&c->ev[i]);
  // set-up of the fixed export staging on this device (its address, once per
  // device); a failure is reported by the entry point that needs it
  if (e == hipSuccess) (void)dh::rgba_scratch_init(device);
  if (e == hipSuccess && !pool) {
    pool = std::make_shared<ResPool>();
    pool->device = device;
    e = hipMalloc(reinterpret_cast<void**>(&pool->white_rgb), 4);
    if (e == hipSuccess) {
      const uint8_t w[4] = {255, 255, 255, 255};  // Color::White.as_slice(), scene/mod.rs:130
      e = hipMemcpy(pool->white_rgb, w, 4, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) pool->mat_host.push_back(default_material(pool->white_rgb));
  }
  c->res = pool;
  if (e != hipSuccess) {
    dh_ctx_destroy(c);
    return DH_E_HIP;
  }
  *out = c;
  return DH_OK;
}

}  // namespace

extern "C" int dh_ctx_create(int32_t device, dh_ctx* out) {
  if (!out) return DH_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return DH_E_UNSUPPORTED;
  if (device < 0 || device >= n) return DH_E_ARG;
  return ctx_new(device, nullptr, out);
}

extern "C" int dh_ctx_create_shared(dh_ctx parent, dh_ctx* out) {
  if (!parent || !out) return DH_E_ARG;
  *out = nullptr;
  const int rc = ctx_new(parent->device, parent->res, out);
  if (rc == DH_OK) {
    (*out)->TS = parent->TS;
    (*out)->waves = parent->waves;
    (*out)->raster_kind = parent->raster_kind;
    (*out)->bin_mode = parent->bin_mode;
    (*out)->graph_frames = parent->graph_frames;
  }
  return rc;
}

extern "C" int dh_ctx_destroy(dh_ctx c) {
  if (!c) return DH_E_ARG;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
  if (c->graph) hipGraphDestroy(c->graph);
  c->cnt.release(); c->blk.release(); c->surv.release(); c->setup.release(); c->shade.release();
  c->bincnt.release(); c->binoff.release(); c->big.release();
  c->keys0.release(); c->keys1.release(); c->vals0.release(); c->vals1.release();
  c->hist.release(); c->partials.release(); c->ranges.release(); c->vis.release();
  c->tile_cnt.release(); c->tile_start.release();
  c->status.release(); c->tables.release(); c->draws = nullptr; c->views = nullptr; c->export_buf.release();
  c->tdraws.release(); c->grp.release(); c->tperm.release(); c->sort_n.release();
  c->gui_vtx.release(); c->gui_tri.release(); c->gui_clip.release(); c->gui_rec.release(); c->gui_err.release();
  if (c->h_status) hipHostFree(c->h_status);
  if (c->h_tables) hipHostFree(c->h_tables);
  for (int i = 0; i < 5; i++)
    if (c->ev[i]) hipEventDestroy(c->ev[i]);
  for (int i = 0; i < 4; i++)
    if (c->xev[i]) hipEventDestroy(c->xev[i]);
  if (c->side) hipStreamDestroy(c->side);
  if (c->stream) hipStreamDestroy(c->stream);
  c->res.reset();  // the pool goes with its last context
  delete c;
  return DH_OK;
}

extern "C" const char* dh_last_error(dh_ctx c) { return c ? c->err.c_str() : "null context"; }

extern "C" int dh_set_tiling(dh_ctx c, int32_t tile, int32_t waves) {
  if (!c) return DH_E_ARG;
  if (!dh::raster_config_supported(tile, waves, 0)) return fail(c, DH_E_ARG, "unsupported tile/waves");
  c->raster_kind = 0;
  c->TS = tile;
  c->waves = waves;
  return DH_OK;
}

extern "C" int dh_set_raster_kernel(dh_ctx c, int32_t kind, int32_t tile, int32_t waves) {
  if (!c) return DH_E_ARG;
  if (!dh::raster_config_supported(tile, waves, kind)) return fail(c, DH_E_ARG, "unsupported raster kernel config");
  c->raster_kind = kind;
  c->TS = tile;
  c->waves = waves;
  return DH_OK;
}

namespace {
void drop_graph(dh_ctx c);
}

extern "C" int dh_set_graph(dh_ctx c, int32_t enable) {
  if (!c) return DH_E_ARG;
  c->graph_frames = enable != 0;
  if (!c->graph_frames) drop_graph(c);
  return DH_OK;
}

extern "C" int dh_set_bin_order(dh_ctx c, int32_t mode) {
  if (!c || mode < 0 || mode > 2) return DH_E_ARG;
  c->bin_mode = mode;
  return DH_OK;
}

extern "C" int dh_set_profiling(dh_ctx c, int32_t enable) {
  if (!c) return DH_E_ARG;
  c->profiling = enable != 0;
  return DH_OK;
}

// ---------------------------------------------------------------- resources
extern "C" int dh_texture_create(dh_ctx c, const uint8_t* img, int32_t w, int32_t h, int32_t comp, dh_handle* out) {
  if (!c || !img || !out || w <= 0 || h <= 0 || (comp != 3 && comp != 4)) return DH_E_ARG;
  std::lock_guard<std::mutex> lk(c->res->mu);
  HIPCHK(c, hipSetDevice(c->device));
  TexRes t;
  t.w = w;
  t.h = h;
  t.comp = comp;
  const size_t bytes = (size_t)w * (size_t)h * (size_t)comp;
  HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&t.d), bytes + 4));  // +4: 32-bit texel loads
  HIPCHK(c, hipMemcpy(t.d, img, bytes, hipMemcpyHostToDevice));
  const uint64_t id = c->res->next_id++;
  c->res->tex[id] = t;
  *out = make_handle(kTex, id);
  return DH_OK;
}

extern "C" int dh_material_create(dh_ctx c, const dh_material_desc* d, dh_handle* out) {
  if (!c || !d || !out || d->power < 0) return DH_E_ARG;
  std::lock_guard<std::mutex> lk(c->res->mu);
  dh::DevMaterial m = default_material(c->res->white_rgb);
  for (int i = 0; i < 3; i++) {
    m.ka[i] = d->ka[i];
    m.kd[i] = d->kd[i];
    m.ks[i] = d->ks[i];
  }
  m.alpha = d->alpha;
  m.power = d->power;
  const dh_handle maps[2] = {d->map_kd, d->map_ka};
  for (int k = 0; k < 2; k++) {
    if (!maps[k]) continue;
    if (handle_kind(maps[k]) != kTex) return fail(c, DH_E_HANDLE, "material map is not a texture");
    auto it = c->res->tex.find(maps[k] & ((1ull << 56) - 1));
    if (it == c->res->tex.end()) return fail(c, DH_E_HANDLE, "unknown texture");
    if (k == 0) {
      m.kd_img = it->second.d; m.kd_w = it->second.w; m.kd_h = it->second.h; m.kd_comp = it->second.comp;
    } else {
      m.ka_img = it->second.d; m.ka_w = it->second.w; m.ka_h = it->second.h; m.ka_comp = it->second.comp;
    }
  }
  MatRes r;
  r.slot = (uint32_t)c->res->mat_host.size();
  r.alpha = d->alpha;
  c->res->mat_host.push_back(m);
  c->res->mat_dirty = true;
  const uint64_t id = c->res->next_id++;
  c->res->mats[id] = r;
  *out = make_handle(kMat, id);
  return DH_OK;
}

extern "C" int dh_object_create(dh_ctx c, const float* pos, uint32_t n_pos, const float* nrm, uint32_t n_nrm,
                                const float* uv, uint32_t n_uv, dh_handle* out) {
  if (!c || !out || (n_pos && !pos) || (n_nrm && !nrm) || (n_uv && !uv)) return DH_E_ARG;
  std::lock_guard<std::mutex> lk(c->res->mu);
  HIPCHK(c, hipSetDevice(c->device));
  ObjRes o;
  o.n_pos = n_pos;
  o.n_nrm = n_nrm;
  o.n_uv = n_uv;
  HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&o.pos), std::max<size_t>(1, 3ull * n_pos) * 4));
  HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&o.nrm), std::max<size_t>(1, 3ull * n_nrm) * 4));
  HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&o.uv), std::max<size_t>(1, 2ull * n_uv) * 4));
  if (n_pos) HIPCHK(c, hipMemcpy(o.pos, pos, 12ull * n_pos, hipMemcpyHostToDevice));
  if (n_nrm) HIPCHK(c, hipMemcpy(o.nrm, nrm, 12ull * n_nrm, hipMemcpyHostToDevice));
  if (n_uv) HIPCHK(c, hipMemcpy(o.uv, uv, 8ull * n_uv, hipMemcpyHostToDevice));
  const uint64_t id = c->res->next_id++;
  c->res->objs[id] = o;
  *out = make_handle(kObj, id);
  return DH_OK;
}

extern "C" int dh_mesh_create(dh_ctx c, dh_handle object, const uint32_t* tri, uint32_t n_tri, dh_handle material,
                              dh_handle* out) {
  if (!c || !out || (n_tri && !tri)) return DH_E_ARG;
  std::lock_guard<std::mutex> lk(c->res->mu);
  if (handle_kind(object) != kObj) return fail(c, DH_E_HANDLE, "not an object handle");
  auto ot = c->res->objs.find(object & ((1ull << 56) - 1));
  if (ot == c->res->objs.end()) return fail(c, DH_E_HANDLE, "unknown object");
  if (material && (handle_kind(material) != kMat || !c->res->mats.count(material & ((1ull << 56) - 1))))
    return fail(c, DH_E_HANDLE, "unknown material");
  const ObjRes& o = ot->second;
  // IndexedMesh::vec3_list_from_indexed asserts (scene/mesh.rs:46-48)
  for (uint64_t t = 0; t < n_tri; t++) {
    const uint32_t* q = tri + 9 * t;
    for (int k = 0; k < 3; k++) {
      if (q[k] >= o.n_pos || q[3 + k] >= o.n_uv || q[6 + k] >= o.n_nrm) {
        char buf[128];
        snprintf(buf, sizeof buf, "triangle %llu: vertex index out of range", (unsigned long long)t);
        return fail(c, DH_E_INDEX, buf);
      }
    }
  }
  HIPCHK(c, hipSetDevice(c->device));
  MeshRes m;
  m.obj = object;
  m.n_tri = n_tri;
  m.mat = material;
  HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&m.tri), std::max<size_t>(1, 9ull * n_tri) * 4));
  if (n_tri) HIPCHK(c, hipMemcpy(m.tri, tri, 36ull * n_tri, hipMemcpyHostToDevice));
  // Opaque meshes (their triangle order never changes) get per-triangle vertex
  // streams and records for the geometry kernels; transparent ones are
  // re-sorted every frame and keep reading through their index triplets.
  const bool opaque = !material || c->res->mats.at(material & ((1ull << 56) - 1)).alpha >= 1.0f;
  if (opaque && n_tri) {
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(&m.soa), (size_t)dh::mesh_stream_floats(n_tri) * 4ull);
    if (e != hipSuccess) {
      hipFree(m.tri);
      return hip_fail(c, e, "hipMalloc(mesh streams)");
    }
    dh::enqueue_mesh_streams(m.tri, n_tri, o.pos, o.nrm, o.uv, m.soa, c->stream);
    const hipError_t e2 = hipStreamSynchronize(c->stream);
    if (e2 != hipSuccess) {
      hipFree(m.soa);
      hipFree(m.tri);
      return hip_fail(c, e2, "mesh streams");
    }
  }
  const uint64_t id = c->res->next_id++;
  c->res->meshes[id] = m;
  *out = make_handle(kMesh, id);
  return DH_OK;
}

extern "C" int dh_mesh_read(dh_ctx c, dh_handle mesh, uint32_t* tri) {
  if (!c || !tri) return DH_E_ARG;
  if (handle_kind(mesh) != kMesh) return fail(c, DH_E_HANDLE, "not a mesh handle");
  MeshRes m;
  {
    std::lock_guard<std::mutex> lk(c->res->mu);
    auto it = c->res->meshes.find(mesh & ((1ull << 56) - 1));
    if (it == c->res->meshes.end()) return fail(c, DH_E_HANDLE, "unknown mesh");
    m = it->second;
  }
  HIPCHK(c, hipSetDevice(c->device));
  if (m.n_tri) HIPCHK(c, hipMemcpyAsync(tri, m.tri, 36ull * m.n_tri, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DH_OK;
}

extern "C" int dh_canvas_create(dh_ctx c, int32_t W, int32_t H, dh_handle* out) {
  return dh_canvas_create_ex(c, W, H, 0u, out);
}

extern "C" int dh_canvas_create_ex(dh_ctx c, int32_t W, int32_t H, uint32_t flags, dh_handle* out) {
  if (!c || !out || W <= 0 || H <= 0 || W > 65535 || H > 65535 || (flags & ~(uint32_t)DH_CANVAS_NO_WINNER))
    return DH_E_ARG;
  std::lock_guard<std::mutex> lk(c->res->mu);
  HIPCHK(c, hipSetDevice(c->device));
  CanvasRes cv;
  cv.W = W;
  cv.H = H;
  const size_t n = (size_t)W * (size_t)H;
  HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&cv.bgra), n * 4));
  HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&cv.depth), n * 4));
  if (!(flags & DH_CANVAS_NO_WINNER)) HIPCHK(c, hipMalloc(reinterpret_cast<void**>(&cv.winner), n * 4));
  const uint64_t id = c->res->next_id++;
  c->res->canvases[id] = cv;
  *out = make_handle(kCanvas, id);
  return DH_OK;
}

extern "C" int dh_release(dh_ctx c, dh_handle h) {
  if (!c) return DH_E_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::lock_guard<std::mutex> lk(c->res->mu);
  const uint64_t id = h & ((1ull << 56) - 1);
  switch (handle_kind(h)) {
    case kTex: {
      auto it = c->res->tex.find(id);
      if (it == c->res->tex.end()) break;
      hipFree(it->second.d);
      c->res->tex.erase(it);
      return DH_OK;
    }
    case kMat: {
      auto it = c->res->mats.find(id);
      if (it == c->res->mats.end()) break;
      c->res->mats.erase(it);  // the slot stays in the table
      return DH_OK;
    }
    case kObj: {
      auto it = c->res->objs.find(id);
      if (it == c->res->objs.end()) break;
      hipFree(it->second.pos);
      hipFree(it->second.nrm);
      hipFree(it->second.uv);
      c->res->objs.erase(it);
      return DH_OK;
    }
    case kMesh: {
      auto it = c->res->meshes.find(id);
      if (it == c->res->meshes.end()) break;
      hipFree(it->second.tri);
      if (it->second.soa) hipFree(it->second.soa);
      c->res->meshes.erase(it);
      return DH_OK;
    }
    case kCanvas: {
      auto it = c->res->canvases.find(id);
      if (it == c->res->canvases.end()) break;
      hipFree(it->second.bgra);
      hipFree(it->second.depth);
      hipFree(it->second.winner);
      c->res->canvases.erase(it);
      return DH_OK;
    }
  }
  return fail(c, DH_E_HANDLE, "unknown handle");
}

// ---------------------------------------------------------------- rendering
namespace {

// Draw and view tables of a batch; key = everything a captured graph depends on.
int build_tables(dh_ctx c, const dh_view* views, uint32_t n_views, bool allow_transparent, uint64_t* n_in_out,
                 uint32_t* n_tiles_out, uint32_t* n_pixels_out, std::vector<uint64_t>* key) {
  c->h_draws.clear();
  c->h_views.clear();
  c->h_tdraws.clear();
  c->h_tmeshes.clear();
  c->h_grp_first.clear();
  c->n_tin = 0;
  c->max_tmesh = 0;
  dh_handle last_obj = 0;
  key->clear();
  key->push_back(n_views);
  key->push_back((uint64_t)c->TS << 32 | (uint64_t)c->raster_kind << 24 | (uint64_t)c->waves << 2 |
                 (uint64_t)c->bin_mode);
  uint64_t n_in = 0, n_tiles = 0, n_pixels = 0;
  for (uint32_t v = 0; v < n_views; v++) {
    const dh_view& vq = views[v];
    const dh_frame_params* fr = &vq.fr;
    if (fr->W <= 0 || fr->H <= 0 || (vq.n_meshes && !vq.meshes)) return fail(c, DH_E_ARG, "bad view");
    if (handle_kind(vq.canvas) != kCanvas) return fail(c, DH_E_HANDLE, "not a canvas handle");
    auto ct = c->res->canvases.find(vq.canvas & ((1ull << 56) - 1));
    if (ct == c->res->canvases.end()) return fail(c, DH_E_HANDLE, "unknown canvas");
    const CanvasRes& cv = ct->second;
    if (cv.W != fr->W || cv.H != fr->H) return fail(c, DH_E_ARG, "canvas size differs from frame size");
    dh::DevView dv;
    memset(&dv, 0, sizeof(dv));
    dv.fr = *fr;
    dv.ntx = (uint32_t)((fr->W + c->TS - 1) / c->TS);
    dv.nty = (uint32_t)((fr->H + c->TS - 1) / c->TS);
    dv.tile_base = (uint32_t)n_tiles;
    dv.px_base = (uint32_t)n_pixels;
    dv.out_bgra = cv.bgra;
    dv.out_depth = cv.depth;
    dv.out_winner = cv.winner;
    dv.tri_begin = (uint32_t)n_in;
    dv.id_base = 0;
    {
      // half the camera's distance to its near plane (plane_f = n.p + d)
      const float* pl = fr->planes[0];
      const double nn = std::sqrt((double)pl[0] * pl[0] + (double)pl[1] * pl[1] + (double)pl[2] * pl[2]);
      const double dn = std::fabs((double)pl[0] * fr->cam[0] + (double)pl[1] * fr->cam[1] +
                                  (double)pl[2] * fr->cam[2] + (double)pl[3]) / nn;
      const float h = (float)(0.5 * dn);
      uint32_t hb = 0;
      memcpy(&hb, &h, 4);
      dv.depth_base = (std::isfinite(h) && h > 0.0f) ? hb : 0u;
    }
    key->push_back((uint64_t)fr->W << 32 | (uint32_t)fr->H);
    key->push_back(vq.canvas);
    key->push_back(vq.n_meshes);
    // meshes in submission order (Scene::render's object/mesh loops)
    for (uint32_t i = 0; i < vq.n_meshes; i++) {
      key->push_back(vq.meshes[i]);
      if (handle_kind(vq.meshes[i]) != kMesh) return fail(c, DH_E_HANDLE, "not a mesh handle");
      auto mt = c->res->meshes.find(vq.meshes[i] & ((1ull << 56) - 1));
      if (mt == c->res->meshes.end()) return fail(c, DH_E_HANDLE, "unknown mesh");
      const MeshRes& m = mt->second;
      auto ot = c->res->objs.find(m.obj & ((1ull << 56) - 1));
      if (ot == c->res->objs.end()) return fail(c, DH_E_HANDLE, "mesh object was released");
      uint32_t slot = 0;
      float alpha = 1.0f;
      if (m.mat) {
        auto it = c->res->mats.find(m.mat & ((1ull << 56) - 1));
        if (it == c->res->mats.end()) return fail(c, DH_E_HANDLE, "mesh material was released");
        slot = it->second.slot;
        alpha = it->second.alpha;
      }
      // Object groups: consecutive meshes of one object, as Scene::render
      // visits an object's opaque meshes, then its transparent ones.
      if (i == 0 || m.obj != last_obj) {
        c->h_grp_first.push_back((uint32_t)n_in);
        last_obj = m.obj;
      }
      const uint32_t group = (uint32_t)c->h_grp_first.size() - 1u;
      if (alpha < 1.0f && !allow_transparent)
        return fail(c, DH_E_UNSUPPORTED, "transparent meshes (alpha < 1) need dh_render (one view)");
      if (m.n_tri == 0) continue;
      dh::DevDraw d;
      memset(&d, 0, sizeof(d));
      d.pos = ot->second.pos;
      d.nrm = ot->second.nrm;
      d.uv = ot->second.uv;
      d.tri = m.tri;
      d.soa = alpha < 1.0f ? nullptr : m.soa;
      d.n_tri = m.n_tri;
      d.material = slot;
      d.view = v;
      d.group = group;
      if (alpha < 1.0f) {  // transparent pass: no culling, own submission sequence
        d.tri_begin = (uint32_t)c->n_tin;
        d.flags = dh::kDrawNoCull;
        c->h_tdraws.push_back(d);
        c->h_tmeshes.push_back(dh::TransMesh{m.tri, m.n_tri, ot->second.pos});
        c->n_tin += m.n_tri;
        c->max_tmesh = std::max<uint64_t>(c->max_tmesh, m.n_tri);
        continue;
      }
      d.tri_begin = (uint32_t)n_in;
      c->h_draws.push_back(d);
      n_in += m.n_tri;
    }
    c->h_views.push_back(dv);
    n_tiles += (uint64_t)dv.ntx * dv.nty;
    n_pixels += (uint64_t)fr->W * (uint64_t)fr->H;
  }
  if (n_in > 0x7FFFFFFFull || c->n_tin > 0x7FFFFFFFull) return fail(c, DH_E_ARG, "too many triangles in one launch");
  if (n_tiles > 0x7FFFFFFFull || n_pixels > 0x7FFFFFFFull) return fail(c, DH_E_ARG, "too many pixels in one launch");
  if (c->h_draws.empty()) {  // keep the draw table non-empty
    dh::DevDraw d;
    memset(&d, 0, sizeof(d));
    c->h_draws.push_back(d);
  }
  *n_in_out = n_in;
  *n_tiles_out = (uint32_t)n_tiles;
  *n_pixels_out = (uint32_t)n_pixels;
  return DH_OK;
}

dh::PipeArgs make_args(dh_ctx c, const dh::DevMaterial* mats, uint32_t n_in, uint32_t n_tiles, uint32_t n_pixels,
                       bool events) {
  dh::PipeArgs a;
  memset(&a, 0, sizeof(a));
  a.draws = c->draws;
  a.n_draws = (uint32_t)c->h_draws.size();
  a.tables_src = c->d_tables_host;
  a.tables_dst = reinterpret_cast<uint32_t*>(c->tables.p);
  a.tables_words = (uint32_t)(c->tables_used / 4);
  a.n_tri_in = n_in;
  a.views = c->views;
  a.n_views = (uint32_t)c->h_views.size();
  a.n_tiles = n_tiles;
  a.n_pixels = n_pixels;
  a.key_bits = key_bits_for(n_tiles);
  a.mats = mats;
  a.TS = c->TS;
  a.waves = c->waves;
  a.raster_kind = c->raster_kind;
  a.cnt = c->cnt.p;
  a.blk = c->blk.p;
  a.surv = reinterpret_cast<uint2*>(c->surv.p);
  a.setup = c->setup.p;
  a.shade = c->shade.p;
  a.cap_tri = c->cap_tri;
  a.bincnt = c->bincnt.p;
  a.binoff = c->binoff.p;
  a.big = c->big.p;
  a.ordered_bins = c->bin_mode != 0;
  a.depth_order = c->bin_mode == 2;
  a.tile_cnt = c->tile_cnt.p;
  a.tile_cur = c->tile_cnt.p + n_tiles;
  a.tile_start = c->tile_start.p;
  a.keys[0] = c->keys0.p;
  a.keys[1] = c->keys1.p;
  a.vals[0] = c->vals0.p;
  a.vals[1] = c->vals1.p;
  a.cap_pairs = c->cap_pairs;
  a.hist = c->hist.p;
  a.partials = c->partials.p;
  a.ranges = c->ranges.p;
  a.vis = c->vis.p;
  a.status = c->status.p;
  for (int i = 0; i < 5; i++) a.ev[i] = events ? c->ev[i] : nullptr;
  a.side = c->side;
  for (int i = 0; i < 4; i++) a.xev[i] = c->xev[i];
  return a;
}

void drop_graph(dh_ctx c) {
  if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
  if (c->graph) hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_key.clear();
}

int render_views(dh_ctx c, const dh_view* views, uint32_t n_views, bool use_graph, bool allow_transparent) {
  HIPCHK(c, hipSetDevice(c->device));
  uint64_t n_in = 0;
  uint32_t n_tiles = 0, n_pixels = 0;
  std::vector<uint64_t> key;
  int rc;
  dh::DevMaterial* mats = nullptr;
  {
    std::lock_guard<std::mutex> lk(c->res->mu);
    rc = build_tables(c, views, n_views, allow_transparent, &n_in, &n_tiles, &n_pixels, &key);
    if (rc) return rc;
    rc = upload_materials(c);
    if (rc) return rc;
    mats = c->res->mat_dev.p;
  }
  const bool transparent = !c->h_tdraws.empty();
  std::unique_lock<std::mutex> trans_lock(c->res->trans_mu, std::defer_lock);
  if (transparent) {
    use_graph = false;
    trans_lock.lock();  // until this function returns, after the stream has synchronised
  }
  c->shard_frame = false;
  const uint64_t n_max = std::max<uint64_t>(n_in, c->n_tin);
  key.push_back((uint64_t)(uintptr_t)mats);
  uint64_t want_tri = std::max<uint64_t>(c->cap_tri, n_max + n_max / 4 + 1024);
  uint64_t want_pairs = std::max<uint64_t>(std::max<uint64_t>(c->cap_pairs, 4 * want_tri + n_tiles), c->max_tmesh);
  c->stats = dh_stats{};
  c->stats.tris_in = n_in + c->n_tin;
  for (int attempt = 0; attempt < 6; attempt++) {
    want_tri = std::min<uint64_t>(want_tri, 0xFFFFFFF0ull);
    want_pairs = std::min<uint64_t>(want_pairs, 0xFFFFFFF0ull);
    if (c->scratch_cap) {
      const uint64_t need = scratch_bytes_after(c, n_max, want_tri, want_pairs, n_tiles, n_pixels,
                                                c->h_draws.size(), c->h_views.size());
      if (need > c->scratch_cap) {
        c->scratch_need = need;
        return fail(c, DH_E_BUDGET, "frame scratch would exceed the context's cap");
      }
    }
    rc = ensure_scratch(c, (uint32_t)n_max, (uint32_t)want_tri, (uint32_t)want_pairs, n_tiles, n_pixels,
                        (uint32_t)c->h_draws.size(), (uint32_t)c->h_views.size());
    if (rc) return rc;
    // both tables into the pinned staging (the previous frame has completed);
    // the frame's first kernel copies them to the device
    memcpy(c->h_tables, c->h_draws.data(), c->h_draws.size() * sizeof(dh::DevDraw));
    memcpy(c->h_tables + c->tables_view_off, c->h_views.data(), c->h_views.size() * sizeof(dh::DevView));
    c->tables_used = c->tables_view_off + c->h_views.size() * sizeof(dh::DevView);
    const uint32_t n_groups = (uint32_t)c->h_grp_first.size();
    if (transparent) {
      HIPCHK(c, c->tdraws.ensure(c->h_tdraws.size()));
      HIPCHK(c, c->grp.ensure(2ull * n_groups));
      HIPCHK(c, c->tperm.ensure(9ull * c->max_tmesh));
      HIPCHK(c, c->sort_n.ensure(1));
      HIPCHK(c, hipMemcpyAsync(c->tdraws.p, c->h_tdraws.data(), c->h_tdraws.size() * sizeof(dh::DevDraw),
                               hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipMemcpyAsync(c->grp.p, c->h_grp_first.data(), n_groups * sizeof(uint32_t), hipMemcpyHostToDevice,
                               c->stream));
    }
    // everything a captured launch sequence bakes in
    std::vector<uint64_t> gkey = key;
    const void* ptrs[] = {c->cnt.p, c->blk.p, c->surv.p, c->setup.p, c->shade.p, c->bincnt.p, c->binoff.p, c->big.p,
                          c->keys0.p, c->keys1.p, c->vals0.p, c->vals1.p, c->hist.p, c->partials.p, c->ranges.p,
                          c->tile_cnt.p, c->tile_start.p, c->vis.p, c->status.p, c->draws, c->views,
                          c->d_tables_host};
    for (const void* p : ptrs) gkey.push_back((uint64_t)(uintptr_t)p);
    gkey.push_back((uint64_t)c->cap_tri << 32 | c->cap_pairs);
    gkey.push_back(n_in);
    gkey.push_back((uint64_t)(uintptr_t)c->shard_out);  // shard frames: visibility output and shard index
    gkey.push_back(c->shard);
    const bool replay = use_graph && c->graph_exec && gkey == c->graph_key;
    memset(c->h_status, 0, 2 * sizeof(dh::Status));
    if (replay) {
      if (c->profiling) HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
      HIPCHK(c, hipGraphLaunch(c->graph_exec, c->stream));
      if (c->profiling) HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
    } else {
      dh::PipeArgs oa = make_args(c, mats, (uint32_t)n_in, n_tiles, n_pixels, c->profiling);
      if (transparent) oa.ev[4] = nullptr;  // the frame ends after the transparent pass
      if (c->shard_out)
        dh::enqueue_visibility(oa, c->shard, c->shard_out, c->stream);
      else
        dh::enqueue_frame(oa, c->stream);
      if (transparent) {
        // transparent pass (scene/mod.rs:1088-1247): sort, geometry without
        // culling, ordered binning, per-tile blend
        dh::TransArgs ta;
        memset(&ta, 0, sizeof(ta));
        ta.meshes = c->h_tmeshes.data();
        ta.n_meshes = (uint32_t)c->h_tmeshes.size();
        for (int k = 0; k < 3; k++) ta.cam[k] = c->h_views[0].fr.cam[k];
        ta.sort_n = c->sort_n.p;
        ta.perm_tmp = c->tperm.p;
        ta.grp_first = c->grp.p;
        ta.grp_base = c->grp.p + n_groups;
        ta.n_groups = n_groups;
        ta.pass = make_args(c, mats, (uint32_t)c->n_tin, n_tiles, n_pixels, false);
        ta.pass.draws = c->tdraws.p;
        ta.pass.tables_src = nullptr;  // copied by the opaque pass
        ta.pass.n_draws = (uint32_t)c->h_tdraws.size();
        ta.pass.status = c->status.p + 1;
        ta.pass.ordered_bins = true;  // blending needs every tile list in submission order
        ta.pass.depth_order = false;
        dh::enqueue_transparent_prep(oa, ta, c->stream);
        const dh::PipeArgs tb = dh::enqueue_front(ta.pass, c->stream);
        dh::enqueue_blend(tb, ta, c->stream);
        if (c->profiling) HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
      }
      dh::enqueue_status_out(c->status.p, c->d_status_host, 2, c->stream);
      HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const dh::Status st = c->h_status[0];
    const dh::Status st2 = transparent ? c->h_status[1] : dh::Status{};
    c->stats.tris_out = (uint64_t)st.n_out + st2.n_out;
    c->stats.bin_entries = (uint64_t)st.n_pairs + st2.n_pairs;
    c->stats.frags = 0;
    for (uint32_t i = 0; i < dh::kFragWords; i++) c->stats.frags += st.frags[i];
    if (st.overflow || st2.overflow) {
      c->stats.retries++;
      const uint64_t n_out = std::max<uint32_t>(st.n_out, st2.n_out);
      const uint64_t n_pairs = std::max<uint32_t>(st.n_pairs, st2.n_pairs);
      if ((st.overflow | st2.overflow) & 1u) want_tri = n_out + n_out / 8 + 1024;
      if ((st.overflow | st2.overflow) & 2u) want_pairs = n_pairs + n_pairs / 8 + 4096;
      // bin entries are only exact once the triangle records fit
      if ((st.overflow | st2.overflow) & 1u) want_pairs = std::max<uint64_t>(want_pairs, 4 * want_tri + n_tiles);
      want_pairs = std::max<uint64_t>(want_pairs, c->max_tmesh);
      drop_graph(c);
      continue;
    }
    if (c->profiling) {
      if (replay) {
        c->stats.ms_geometry = c->stats.ms_binning = c->stats.ms_raster = c->stats.ms_shade = 0.0f;
      } else {
        float ms[4];
        for (int i = 0; i < 3; i++) hipEventElapsedTime(&ms[i], c->ev[i], c->ev[i + 1]);
        hipEventElapsedTime(&ms[3], c->ev[3], c->ev[4]);  // shade (+ the transparent pass)
        c->stats.ms_geometry = ms[0];
        c->stats.ms_binning = ms[1];
        c->stats.ms_raster = ms[2];
        c->stats.ms_shade = ms[3];
      }
      hipEventElapsedTime(&c->stats.ms_frame, c->ev[0], c->ev[4]);
    }
    if (use_graph && !replay) {
      // capture the same launch sequence (no stage events) for the next call
      drop_graph(c);
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      if (c->shard_out)
        dh::enqueue_visibility(make_args(c, mats, (uint32_t)n_in, n_tiles, n_pixels, false), c->shard, c->shard_out,
                               c->stream);
      else
        dh::enqueue_frame(make_args(c, mats, (uint32_t)n_in, n_tiles, n_pixels, false), c->stream);
      dh::enqueue_status_out(c->status.p, c->d_status_host, 2, c->stream);
      hipGraph_t g = nullptr;
      HIPCHK(c, hipStreamEndCapture(c->stream, &g));
      c->graph = g;
      HIPCHK(c, hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0));
      c->graph_key = gkey;
    }
    if (c->shard_out) {
      if (st.n_out >= (1u << dh::kShardShift))
        return fail(c, DH_E_CAPACITY, "a shard emits at most 2^29 clipped triangles");
      c->shard_frame = true;
      c->shard_n_in = n_in;
      c->shard_n_tiles = n_tiles;
      c->shard_n_pixels = n_pixels;
    }
    if ((st.error | st2.error) & 1u)
      return fail(c, DH_E_TEXCOORD, "texture coordinate out of range (get_rgb_slice would panic)");
    return DH_OK;
  }
  return fail(c, DH_E_CAPACITY, "scratch capacity did not converge");
}

}  // namespace

extern "C" int dh_render(dh_ctx c, const dh_frame_params* fr, const dh_handle* meshes, uint32_t n_meshes,
                         dh_handle canvas) {
  if (!c || !fr || (n_meshes && !meshes)) return DH_E_ARG;
  dh_view v;
  memset(&v, 0, sizeof(v));
  v.fr = *fr;
  v.canvas = canvas;
  v.meshes = meshes;
  v.n_meshes = n_meshes;
  return render_views(c, &v, 1, c->graph_frames, true);
}

extern "C" int dh_device_alloc(dh_ctx c, uint64_t bytes, void** out_dev) {
  if (!c || !out_dev) return DH_E_ARG;
  *out_dev = nullptr;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMalloc(out_dev, bytes ? bytes : 1));
  return DH_OK;
}

extern "C" int dh_device_free(dh_ctx c, void* dev) {
  if (!c) return DH_E_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  if (dev) HIPCHK(c, hipFree(dev));
  return DH_OK;
}

extern "C" int dh_device_read(dh_ctx c, void* host, const void* dev, uint64_t bytes) {
  if (!c || (bytes && (!host || !dev))) return DH_E_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  if (bytes) HIPCHK(c, hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DH_OK;
}

extern "C" int dh_render_shard(dh_ctx c, const dh_frame_params* fr, const dh_handle* meshes, uint32_t n_meshes,
                               uint32_t shard, dh_handle canvas, uint64_t* vis_dev) {
  if (!c || !fr || (n_meshes && !meshes) || !vis_dev) return DH_E_ARG;
  if (shard >= dh::kShardMax) return fail(c, DH_E_ARG, "shard index must be below 8");
  dh_view v;
  memset(&v, 0, sizeof(v));
  v.fr = *fr;
  v.canvas = canvas;
  v.meshes = meshes;
  v.n_meshes = n_meshes;
  c->shard_out = vis_dev;
  c->shard = shard;
  const int rc = render_views(c, &v, 1, c->graph_frames, false);
  c->shard_out = nullptr;
  return rc;
}

extern "C" int dh_vis_merge_min(dh_ctx c, uint64_t* acc_dev, const uint64_t* src_dev, uint64_t n) {
  if (!c || (n && (!acc_dev || !src_dev))) return DH_E_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  if (n) dh::enqueue_vis_min(acc_dev, src_dev, n, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DH_OK;
}

extern "C" int dh_shade_shard(dh_ctx c, const uint64_t* merged_dev, uint32_t shard) {
  if (!c || !merged_dev) return DH_E_ARG;
  if (!c->shard_frame || shard != c->shard)
    return fail(c, DH_E_ARG, "dh_shade_shard needs this shard's dh_render_shard frame on the same context");
  HIPCHK(c, hipSetDevice(c->device));
  dh::DevMaterial* mats = nullptr;
  {
    std::lock_guard<std::mutex> lk(c->res->mu);
    mats = c->res->mat_dev.p;
  }
  dh::PipeArgs a = make_args(c, mats, (uint32_t)c->shard_n_in, c->shard_n_tiles, c->shard_n_pixels, false);
  HIPCHK(c, hipMemsetAsync(c->status.p, 0, sizeof(dh::Status), c->stream));
  dh::enqueue_vis_unpack(merged_dev, c->shard_n_pixels, shard, c->vis.p, c->stream);
  dh::enqueue_shade(a, c->stream);
  dh::enqueue_status_out(c->status.p, c->d_status_host, 2, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->h_status[0].error & 1u)
    return fail(c, DH_E_TEXCOORD, "texture coordinate out of range (get_rgb_slice would panic)");
  return DH_OK;
}

extern "C" int dh_render_batch(dh_ctx c, const dh_view* views, uint32_t n_views, uint32_t flags) {
  if (!c || (n_views && !views)) return DH_E_ARG;
  if (n_views == 0) return DH_OK;
  return render_views(c, views, n_views, (flags & DH_BATCH_GRAPH) != 0, false);
}

namespace {
// The canvas behind a handle (copied out under the pool lock).
int find_canvas(dh_ctx c, dh_handle canvas, CanvasRes* out) {
  if (handle_kind(canvas) != kCanvas) return fail(c, DH_E_HANDLE, "not a canvas handle");
  std::lock_guard<std::mutex> lk(c->res->mu);
  auto ct = c->res->canvases.find(canvas & ((1ull << 56) - 1));
  if (ct == c->res->canvases.end()) return fail(c, DH_E_HANDLE, "unknown canvas");
  *out = ct->second;
  return DH_OK;
}
}  // namespace

extern "C" int dh_canvas_read(dh_ctx c, dh_handle canvas, uint8_t* bgra, float* depth, uint32_t* winner) {
  if (!c) return DH_E_ARG;
  CanvasRes cv;
  const int rc = find_canvas(c, canvas, &cv);
  if (rc) return rc;
  if (winner && !cv.winner) return fail(c, DH_E_ARG, "canvas created without a winner buffer");
  const size_t n = (size_t)cv.W * (size_t)cv.H;
  HIPCHK(c, hipSetDevice(c->device));
  if (bgra) HIPCHK(c, hipMemcpyAsync(bgra, cv.bgra, n * 4, hipMemcpyDeviceToHost, c->stream));
  if (depth) HIPCHK(c, hipMemcpyAsync(depth, cv.depth, n * 4, hipMemcpyDeviceToHost, c->stream));
  if (winner) HIPCHK(c, hipMemcpyAsync(winner, cv.winner, n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DH_OK;
}

extern "C" int dh_canvas_read_rgba(dh_ctx c, dh_handle canvas, uint8_t* rgba) {
  if (!c || !rgba) return DH_E_ARG;
  CanvasRes cv;
  const int rc = find_canvas(c, canvas, &cv);
  if (rc) return rc;
  const size_t n = (size_t)cv.W * (size_t)cv.H;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, c->export_buf.ensure(n));
  HIPCHK(c, dh::enqueue_bgra_to_rgba(cv.bgra, c->export_buf.p, (uint32_t)n, c->stream));
  HIPCHK(c, hipMemcpyAsync(rgba, c->export_buf.p, n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DH_OK;
}

extern "C" int dh_canvas_read_rgba_noalloc(dh_ctx c, dh_handle canvas, uint8_t* rgba) {
  if (!c || !rgba) return DH_E_ARG;
  CanvasRes cv;
  const int rc = find_canvas(c, canvas, &cv);
  if (rc) return rc;
  const uint64_t n = (uint64_t)cv.W * (uint64_t)cv.H;
  if (n > dh::kRgbaScratchPixels)
    return fail(c, DH_E_CAPACITY, "canvas larger than the fixed export staging (3840x2160 pixels)");
  uint32_t* staging = nullptr;
  std::mutex* mu = nullptr;
  const hipError_t e = dh::rgba_scratch(c->device, &staging, &mu);
  if (e != hipSuccess) return hip_fail(c, e, "hipGetSymbolAddress(export staging)");
  // every call on this device shares the array: one at a time
  std::lock_guard<std::mutex> lk(*mu);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, dh::enqueue_bgra_to_rgba(cv.bgra, staging, (uint32_t)n, c->stream));
  HIPCHK(c, hipMemcpyAsync(rgba, staging, n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DH_OK;
}

extern "C" int dh_canvas_export_png(dh_ctx c, dh_handle canvas, const char* path) {
  if (!c || !path) return DH_E_ARG;
  CanvasRes cv;
  int rc = find_canvas(c, canvas, &cv);
  if (rc) return rc;
  std::vector<uint8_t> rgba((size_t)cv.W * (size_t)cv.H * 4);
  rc = dh_canvas_read_rgba(c, canvas, rgba.data());
  if (rc) return rc;
  rc = dh_write_png(path, rgba.data(), cv.W, cv.H);
  return rc ? fail(c, rc, std::string("cannot write ") + path) : DH_OK;
}

extern "C" int dh_canvas_export_jpg(dh_ctx c, dh_handle canvas, const char* path, int32_t quality) {
  if (!c || !path) return DH_E_ARG;
  CanvasRes cv;
  int rc = find_canvas(c, canvas, &cv);
  if (rc) return rc;
  std::vector<uint8_t> rgba((size_t)cv.W * (size_t)cv.H * 4);
  rc = dh_canvas_read_rgba(c, canvas, rgba.data());
  if (rc) return rc;
  rc = dh_write_jpg(path, rgba.data(), cv.W, cv.H, quality);
  return rc ? fail(c, rc, std::string("cannot write ") + path) : DH_OK;
}

extern "C" int dh_draw_gui(dh_ctx c, dh_handle canvas, const dh_gui_vertex* vtx, uint32_t n_vtx, const uint32_t* idx,
                           uint32_t n_idx, const dh_gui_cmd* cmds, uint32_t n_cmds, dh_handle font) {
  if (!c || (n_vtx && !vtx) || (n_idx && !idx) || (n_cmds && !cmds)) return DH_E_ARG;
  CanvasRes cv;
  int rc = find_canvas(c, canvas, &cv);
  if (rc) return rc;
  TexRes ft;
  {
    std::lock_guard<std::mutex> lk(c->res->mu);
    auto it = handle_kind(font) == kTex ? c->res->tex.find(font & ((1ull << 56) - 1)) : c->res->tex.end();
    if (it == c->res->tex.end()) return fail(c, DH_E_HANDLE, "font is not a texture handle");
    ft = it->second;
  }
  // Gui::render (gui.rs:382-478): commands in order; an inverted clip rect
  // skips its command; the clip rect's y is flipped (H - y) and cast to usize.
  std::vector<uint4> tris, clips;
  const float Hf = (float)cv.H;
  for (uint32_t ci = 0; ci < n_cmds; ci++) {
    const dh_gui_cmd& cm = cmds[ci];
    const float* cr = cm.clip_rect;
    if (cr[0] > cr[2] || cr[1] > cr[3]) continue;
    auto cl = [](float v) { return (uint32_t)std::min<uint64_t>(dh::f2usize(v), 0x7FFFFFFFull); };  // order-preserving
    clips.push_back(make_uint4(cl(cr[0]), cl(Hf - cr[1]), cl(cr[2]), cl(Hf - cr[3])));
    if ((uint64_t)cm.idx_offset + cm.count > n_idx) return fail(c, DH_E_INDEX, "command indices past the index buffer");
    for (uint32_t t = 0; t + 3 <= cm.count; t += 3) {
      uint32_t v[3];
      for (int k = 0; k < 3; k++) {
        const uint64_t vi = (uint64_t)idx[cm.idx_offset + t + k] + cm.vtx_offset;
        if (vi >= n_vtx) return fail(c, DH_E_INDEX, "vertex index past the vertex buffer");
        v[k] = (uint32_t)vi;
      }
      tris.push_back(make_uint4(v[0], v[1], v[2], (uint32_t)clips.size() - 1u));
    }
  }
  if (tris.empty()) return DH_OK;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, c->gui_vtx.ensure(n_vtx));
  HIPCHK(c, c->gui_tri.ensure(tris.size()));
  HIPCHK(c, c->gui_clip.ensure(clips.size()));
  HIPCHK(c, c->gui_rec.ensure(tris.size() * (dh::kGuiRecBytes / sizeof(uint4))));
  HIPCHK(c, c->gui_err.ensure(1));
  HIPCHK(c, hipMemcpyAsync(c->gui_vtx.p, vtx, n_vtx * sizeof(dh_gui_vertex), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->gui_tri.p, tris.data(), tris.size() * sizeof(uint4), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->gui_clip.p, clips.data(), clips.size() * sizeof(uint4), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemsetAsync(c->gui_err.p, 0, 4, c->stream));
  HIPCHK(c, dh::enqueue_gui(c->gui_vtx.p, c->gui_tri.p, (uint32_t)tris.size(), c->gui_clip.p, c->gui_rec.p, cv.W,
                            cv.H, ft.d, ft.w, ft.h, ft.comp, cv.bgra, cv.depth, c->gui_err.p, c->stream));
  uint32_t err = 0;
  HIPCHK(c, hipMemcpyAsync(&err, c->gui_err.p, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (err) return fail(c, DH_E_TEXCOORD, "font sample outside the texture (get_rgba_slice would panic)");
  return DH_OK;
}

extern "C" int dh_canvas_device_ptrs(dh_ctx c, dh_handle canvas, void** bgra, void** depth, void** winner) {
  if (!c) return DH_E_ARG;
  CanvasRes cv;
  const int rc = find_canvas(c, canvas, &cv);
  if (rc) return rc;
  if (bgra) *bgra = cv.bgra;
  if (depth) *depth = cv.depth;
  if (winner) *winner = cv.winner;
  return DH_OK;
}

extern "C" int dh_get_stats(dh_ctx c, dh_stats* out) {
  if (!c || !out) return DH_E_ARG;
  *out = c->stats;
  return DH_OK;
}

// --------------------------------------------------------- internal helpers
namespace dh {

uint64_t ctx_scratch_bytes(dh_ctx c) {
  uint64_t b = 0;
  b += c->cnt.n + 4 * (c->blk.n + c->surv.n) + sizeof(SetupRec) * c->setup.n + sizeof(ShadeRec) * c->shade.n;
  b += 4 * (c->bincnt.n + c->binoff.n + c->big.n + c->keys0.n + c->keys1.n + c->vals0.n + c->vals1.n);
  b += 4 * (c->hist.n + c->partials.n + c->tile_cnt.n + c->tile_start.n + c->vis.n) + 8 * c->ranges.n;
  b += sizeof(Status) * c->status.n + c->tables.n;
  return b;
}

// Scratch a frame of n_in input triangles, n_tiles tiles and n_pixels pixels
// asks for on its first attempt (the sizing render_views starts from).
uint64_t scratch_estimate(uint64_t n_in, uint64_t n_tiles, uint64_t n_pixels, uint64_t n_draws,
                          uint64_t n_views) {
  const uint64_t tri = n_in + n_in / 4 + 1024, pairs = 4 * tri + n_tiles;
  return std::max<uint64_t>(n_in, 1) + 12 * ((n_in + 255) / 256 + 1) + 8 * n_in +
         (sizeof(SetupRec) + sizeof(ShadeRec) + 12) * tri + 16 * pairs +
         4ull * kRadixHistEntries + 4096 + 20 * n_tiles + 4 * n_pixels + sizeof(DevDraw) * n_draws +
         sizeof(DevView) * n_views + 2 * sizeof(Status);
}

int ctx_tile_size(dh_ctx c) { return c->TS; }
void ctx_set_scratch_cap(dh_ctx c, uint64_t cap) { c->scratch_cap = cap; }
uint64_t ctx_scratch_need(dh_ctx c) { return c->scratch_need; }

// Triangle count of a mesh handle (DH_E_HANDLE if unknown).
int mesh_triangles(dh_ctx c, dh_handle mesh, uint32_t* n_tri) {
  if (handle_kind(mesh) != kMesh) return fail(c, DH_E_HANDLE, "not a mesh handle");
  std::lock_guard<std::mutex> lk(c->res->mu);
  auto it = c->res->meshes.find(mesh & ((1ull << 56) - 1));
  if (it == c->res->meshes.end()) return fail(c, DH_E_HANDLE, "unknown mesh");
  *n_tri = it->second.n_tri;
  return DH_OK;
}

}  // namespace dh

namespace dh {
// Free every scratch buffer of a context (they are regrown by the next frame).
void ctx_release_scratch(dh_ctx c) {
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  drop_graph(c);
  c->cnt.release(); c->blk.release(); c->surv.release(); c->setup.release(); c->shade.release();
  c->bincnt.release(); c->binoff.release(); c->big.release();
  c->keys0.release(); c->keys1.release(); c->vals0.release(); c->vals1.release();
  c->hist.release(); c->partials.release(); c->ranges.release(); c->vis.release();
  c->tile_cnt.release(); c->tile_start.release();
  c->status.release(); c->tables.release(); c->draws = nullptr; c->views = nullptr; c->export_buf.release();
  c->tdraws.release(); c->grp.release(); c->tperm.release(); c->sort_n.release();
  c->gui_vtx.release(); c->gui_tri.release(); c->gui_clip.release(); c->gui_rec.release(); c->gui_err.release();
  c->cap_tri = c->cap_pairs = 0;
}
}  // namespace dh
