// dh_pipeline.hpp -- device buffers of one frame (or one batch of views) and
// the entry that enqueues the whole pipeline on a stream.
#pragma once
#include "dh_common.hpp"

namespace dh {

constexpr uint32_t kRadixHistEntries = 2048u * 1024u;  // (up to 11-bit) digits x radix-sort waves

struct PipeArgs {
  // inputs
  const DevDraw* draws;
  uint32_t n_draws;
  uint32_t n_tri_in;
  const DevView* views;
  uint32_t n_views;
  uint32_t n_tiles;    // over all views
  uint32_t n_pixels;   // over all views
  int key_bits;        // bits of (n_tiles - 1)
  const DevMaterial* mats;
  int TS;              // tile size in pixels
  int waves;           // waves per tile
  int raster_kind;     // 0: waves own 8x8 quadrants; 1: triangle-parallel rectangle sweeps
  // scratch
  uint8_t* cnt;        // [n_tri_in] clipped-triangle count per input triangle
  uint32_t* blk;       // [3 * (ceil(n_tri_in/256) + 1)] per-block counts -> offsets: clipped triangles,
                       // then (each ceil(n_tri_in/256) + 1 further) unclipped and clipped emitters
  uint2* surv;         // [n_tri_in] (input triangle, first output id) of the emitting triangles: the
                       // unclipped ones in order, then the clipped ones in order
  SetupRec* setup;     // [cap_tri]
  ShadeRec* shade;     // [cap_tri]
  uint32_t cap_tri;
  uint32_t* bincnt;    // [cap_tri]
  uint32_t* binoff;    // [cap_tri]
  uint32_t* big;       // [3 cap_tri + 4] words: a byte per triangle (1 = wide, binned by a wave; 2 =
                       // dropped), then 16-byte aligned an 8-byte bin box per triangle (ordered bins)
  bool ordered_bins;   // stable radix sort (per-tile lists in submission order)
  bool shade_in_raster;  // kind 0 rasters shade their own tiles (no visibility buffer, no k_shade)
  bool depth_order;    // ordered bins: per-tile lists near first (triangles sorted by depth key before binning)
  uint32_t* tile_cnt;  // [2 * n_tiles] per-tile counts, then per-tile cursors
  uint32_t* tile_cur;  // = tile_cnt + n_tiles
  uint32_t* tile_start;  // [n_tiles] exclusive scan of tile_cnt
  uint32_t* keys[2];   // [cap_pairs] each
  uint32_t* vals[2];
  uint32_t cap_pairs;
  uint32_t* hist;      // [256 * kRadixG]
  uint32_t* partials;  // [1024]
  uint2* ranges;       // [n_tiles]
  uint32_t* vis;       // [n_pixels] winner id per pixel
  Status* status;
  // the frame's draw/view tables, copied by the frame's first kernel from
  // coherent pinned host memory (tables_src: its device address; null: none)
  const uint32_t* tables_src;
  uint32_t* tables_dst;
  uint32_t tables_words;
  hipEvent_t ev[5];    // optional stage events (all null when not profiling)
  // independent kernels fork onto `side` (null: everything on the main
  // stream) and join back with xev[] (fork 0 / join 1: the two geometry
  // emits; fork 2 / join 3: the wide-triangle counts beside the depth sort)
  hipStream_t side;
  hipEvent_t xev[4];
};

// One transparent mesh (device pointers): its triangles, sorted in place.
struct TransMesh {
  uint32_t* tri;
  uint32_t n_tri;
  const float* pos;
};

// The transparent pass of a one-view frame (scene/mod.rs:1088-1247).
struct TransArgs {
  const TransMesh* meshes;     // host array
  uint32_t n_meshes;
  float cam[3];
  uint32_t* sort_n;            // [1] device scratch: item count of a sort
  uint32_t* perm_tmp;          // [9 * max transparent mesh triangles]
  const uint32_t* grp_first;   // [n_groups] first opaque input triangle of each object group
  uint32_t* grp_base;          // [n_groups] -> first opaque clipped-triangle id of each group
  uint32_t n_groups;
  PipeArgs pass;               // geometry + binning of the transparent draws (own Status)
};

PipeArgs enqueue_front(const PipeArgs& a, hipStream_t s);
void enqueue_frame(const PipeArgs& a, hipStream_t s);
void enqueue_visibility(const PipeArgs& a, uint32_t shard, uint64_t* out, hipStream_t s);
// Copies n Status words from device memory to coherent pinned host memory
// (host_dev_ptr: its device address) with a kernel, so that the copy can sit
// inside a captured frame graph.
void enqueue_status_out(const Status* dev, Status* host_dev_ptr, uint32_t n, hipStream_t s);
// A mesh's per-triangle vertex streams and records (mesh_stream_floats(n_tri) floats at soa).
void enqueue_mesh_streams(const uint32_t* tri, uint32_t n_tri, const float* pos, const float* nrm, const float* uv,
                          float* soa, hipStream_t s);
int enqueue_radix_sort(uint32_t* const keys[2], uint32_t* const vals[2], const uint32_t* n_ptr, uint32_t cap,
                       int key_bits, uint32_t* hist, uint32_t* partials, hipStream_t s, bool keys_out,
                       uint2* ranges, uint32_t n_ranges, bool stable = true);
void enqueue_transparent_prep(const PipeArgs& opaque, const TransArgs& t, hipStream_t s);
void enqueue_blend(const PipeArgs& b, const TransArgs& t, hipStream_t s);
bool raster_config_supported(int TS, int waves, int kind);
// Per-pixel stages (dh_raster.hip): a.vals[0] must hold the sorted bin values.
void enqueue_raster(const PipeArgs& a, hipStream_t s);
void enqueue_shade(const PipeArgs& a, hipStream_t s);
// Sort-last shards (dh_render_shard / dh_shade_shard / dh_vis_merge_min).
void enqueue_vis_pack(const PipeArgs& a, uint32_t shard, uint64_t* out, hipStream_t s);
void enqueue_vis_unpack(const uint64_t* merged, uint32_t n_pixels, uint32_t shard, uint32_t* vis, hipStream_t s);
void enqueue_vis_min(uint64_t* acc, const uint64_t* src, uint64_t n, hipStream_t s);

}  // namespace dh
