// dh_raster.hip -- the per-pixel stages of a frame on gfx950:
//   k_raster  one workgroup per screen tile: coverage and the strict depth test
//             of canvas.rs:668-682, 923 into an LDS visibility buffer of
//             (depth key << 32 | triangle id), with hierarchical early-Z
//   k_shade   deferred Phong + texture shading of each pixel's winner
//             (canvas.rs:684-745), BGRA8 + f32 depth + winner id writes
#include <math.h>

#include "dh_common.hpp"
#include "dh_pipeline.hpp"

namespace dh {

struct FragOut {
  float al, be, ga, d;
};

// Barycentrics and depth at a pixel, canvas.rs:673-682.
__device__ inline FragOut bary(const SetupRec& r, float x, float y) {
  FragOut o;
  o.al = edge_eval(r.ea, x, y) / r.f[0];
  o.be = edge_eval(r.eb, x, y) / r.f[1];
  o.ga = edge_eval(r.eg, x, y) / r.f[2];
  o.d = (o.al * r.d[0] + o.be * r.d[1]) + o.ga * r.d[2];
  return o;
}

// The reference samples both maps of every covered fragment before its depth
// test (canvas.rs:689-695), so an out-of-range texture coordinate fails the
// frame even for a hidden fragment.
__device__ inline bool uv_ok(const ShadeRec* __restrict__ shade, const DevMaterial* __restrict__ mats, uint32_t id,
                             const FragOut& o) {
  const ShadeRec& s = shade[id];
  const float u = (s.uv[0] * o.al + s.uv[2] * o.be) + s.uv[4] * o.ga;
  const float v = (s.uv[1] * o.al + s.uv[3] * o.be) + s.uv[5] * o.ga;
  const DevMaterial& m = mats[s.material];
  uint64_t off;
  return rgb_offset(m.kd_w, m.kd_h, m.kd_comp, u, v, &off) && rgb_offset(m.ka_w, m.ka_h, m.ka_comp, u, v, &off);
}


// popcount(m & lanes below this one) + base (v_mbcnt_lo / v_mbcnt_hi).
__device__ inline uint32_t lanes_below(uint64_t m, uint32_t base) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, base));
}

// Maximum over the wave with DPP moves (no LDS): quads, half rows and rows
// by permutes and mirrors, then rows 1 and 3 take lane 15 of the row before
// (row_bcast:15) and rows 2 and 3 lane 31 (row_bcast:31), so lane 63 holds
// the maximum.  Lanes a move does not write get 0, the identity of max.
template <int CTRL, int ROW_MASK>
__device__ inline uint32_t dpp_max(uint32_t v) {
  return max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xF, false));
}
__device__ inline uint32_t wave_max_u32(uint32_t v) {
  v = dpp_max<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
  v = dpp_max<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
  v = dpp_max<0x141, 0xF>(v);  // row_half_mirror
  v = dpp_max<0x140, 0xF>(v);  // row_mirror
  v = dpp_max<0x142, 0xA>(v);  // row_bcast:15
  v = dpp_max<0x143, 0xC>(v);  // row_bcast:31
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// ===================================================================== shade
__device__ inline void fetch_rgb(const uint8_t* img, int32_t comp, uint64_t off, uint32_t c[3]) {
  if (comp == 4) {
    const uint32_t w = *reinterpret_cast<const uint32_t*>(img + off);
    c[0] = w & 255u; c[1] = (w >> 8) & 255u; c[2] = (w >> 16) & 255u;
  } else {
    c[0] = img[off]; c[1] = img[off + 1]; c[2] = img[off + 2];
  }
}

// k / 255 for every byte k (Pixel::normalized_as_vec3, canvas.rs:81-87),
// divided once per workgroup into an LDS table that the shading reads, the
// block's nt threads filling it together.  The caller places a barrier
// between the fill and the first shade_pixel.
__device__ inline void fill_unit_table(float* unit, int tid, int nt) {
  for (int k = tid; k < 256; k += nt) unit[k] = (float)k / 255.0f;
}

// Shading of a covered pixel, canvas.rs:685-741; returns B | G<<8 | R<<16 | A<<24.
// unit: the k / 255 table (fill_unit_table).
__device__ inline uint32_t shade_pixel(const ShadeRec& s, const DevMaterial& m, const FragOut& o, Status* status,
                                       const float* unit) {
  const float u = (s.uv[0] * o.al + s.uv[2] * o.be) + s.uv[4] * o.ga;
  const float v = (s.uv[1] * o.al + s.uv[3] * o.be) + s.uv[5] * o.ga;
  uint64_t okd, oka;
  const bool ok1 = rgb_offset(m.kd_w, m.kd_h, m.kd_comp, u, v, &okd);
  const bool ok2 = rgb_offset(m.ka_w, m.ka_h, m.ka_comp, u, v, &oka);
  if (!(ok1 && ok2)) {
    atomicOr(&status->error, 1u);
    okd = 0;
    oka = 0;
  }
  uint32_t dk[3], ak[3];
  fetch_rgb(m.kd_img, m.kd_comp, okd, dk);
  fetch_rgb(m.ka_img, m.ka_comp, oka, ak);
  const V3 diffuse = v3(unit[dk[0]], unit[dk[1]], unit[dk[2]]);
  const V3 ambient = v3(unit[ak[0]], unit[ak[1]], unit[ak[2]]);
  const V3 pn = add(add(muls(v3(s.n[0], s.n[1], s.n[2]), o.al), muls(v3(s.n[3], s.n[4], s.n[5]), o.be)),
                    muls(v3(s.n[6], s.n[7], s.n[8]), o.ga));
  const V3 pl = add(add(muls(v3(s.l[0], s.l[1], s.l[2]), o.al), muls(v3(s.l[3], s.l[4], s.l[5]), o.be)),
                    muls(v3(s.l[6], s.l[7], s.l[8]), o.ga));
  const V3 ph = add(add(muls(v3(s.h[0], s.h[1], s.h[2]), o.al), muls(v3(s.h[3], s.h[4], s.h[5]), o.be)),
                    muls(v3(s.h[6], s.h[7], s.h[8]), o.ga));
  const V3 c_l = v3(m.ks[0], m.ks[1], m.ks[2]);
  const V3 c_r = v3(diffuse.x * m.kd[0], diffuse.y * m.kd[1], diffuse.z * m.kd[2]);
  const V3 c_a = v3(ambient.x * m.ka[0], ambient.y * m.ka[1], ambient.z * m.ka[2]);
  const float dln = dot(pl, pn);
  const float mx = (dln > 0.0f) ? dln : 0.0f;  // 0.0_f32.max(x); the sign of a zero is irrelevant below
  const float sfac = 1.0f - mx;
  const V3 inner = add(c_a, muls(c_l, sfac));
  const V3 t1 = v3(c_r.x * inner.x, c_r.y * inner.y, c_r.z * inner.z);
  const float h2 = powi_f(dot(ph, pn), m.power);
  const V3 col = add(t1, muls(c_l, h2));
  const V3 sc = muls(col, 255.0f);
  return f2u8(sc.z) | (f2u8(sc.y) << 8) | (f2u8(sc.x) << 16) | (255u << 24);
}

// Shading of up to 64 pixels by one wave; every lane calls it.  A lane's
// pixel (x, y) of view vw has winner id, or kNone (the clear colour); lanes
// with active false have no pixel.  Runs of equal winners along the wave load
// their records once into run[slot] (up to R runs; more take per-lane loads):
// the texture address units cost per requested byte per lane, and a run of
// pixels shares one triangle's 224 record bytes.  A material shared by the
// whole wave is read with scalar loads.
template <int R>
__device__ inline void shade_wave(uint32_t id, bool active, uint32_t x, uint32_t y, const DevView& vw,
                                  const SetupRec* __restrict__ setup, const ShadeRec* __restrict__ shade,
                                  const DevMaterial* __restrict__ mats, Status* __restrict__ status,
                                  uint4 (*run)[14], const float* unit) {
  const uint32_t lane = threadIdx.x & 63u;
  const bool covered = active && id != kNone;
  const uint32_t key = covered ? id : kSkip;
  // runs: a covered lane starts one unless the lane before holds the same id
  const uint32_t prev = (uint32_t)__shfl_up((int)key, 1, 64);
  const bool start = covered && (lane == 0u || prev != key);
  const uint64_t sm = __ballot(start);
  SetupRec r;
  ShadeRec sh;
  if (__popcll(sm) <= R) {
    const uint32_t slot = (uint32_t)__popcll(sm & (~0ull >> (63u - lane))) - 1u;  // the last start at or before
    uint4* const rs_run = run[covered ? slot : 0u];
    if (start) {
      const uint4* rs = reinterpret_cast<const uint4*>(setup + id);
      const uint4* ss = reinterpret_cast<const uint4*>(shade + id);
#pragma unroll
      for (int q = 0; q < 5; q++) rs_run[q] = rs[q];
#pragma unroll
      for (int q = 0; q < 9; q++) rs_run[5 + q] = ss[q];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (covered) {
      uint4* rd = reinterpret_cast<uint4*>(&r);
      uint4* sd = reinterpret_cast<uint4*>(&sh);
#pragma unroll
      for (int q = 0; q < 5; q++) rd[q] = rs_run[q];
#pragma unroll
      for (int q = 0; q < 9; q++) sd[q] = rs_run[5 + q];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  } else if (covered) {
    const uint4* rs = reinterpret_cast<const uint4*>(setup + id);
    uint4* rd = reinterpret_cast<uint4*>(&r);
#pragma unroll
    for (int q = 0; q < 5; q++) rd[q] = rs[q];
    sh = shade[id];
  }
  if (!active) return;
  const uint32_t W = (uint32_t)vw.fr.W, H = (uint32_t)vw.fr.H;
  uint32_t bgra;
  float depth;
  if (!covered) {
    const uint8_t* c = vw.fr.clear_bgra;
    bgra = (uint32_t)c[0] | ((uint32_t)c[1] << 8) | ((uint32_t)c[2] << 16) | ((uint32_t)c[3] << 24);
    depth = vw.fr.clear_depth;
  } else {
    const FragOut o = bary(r, (float)x, (float)y);
    // one material for the wave's covered pixels: scalar loads
    const uint32_t m0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)sh.material);
    if (__ballot(sh.material != m0) == 0ull)
      bgra = shade_pixel(sh, mats[m0], o, status, unit);
    else
      bgra = shade_pixel(sh, mats[sh.material], o, status, unit);
    depth = o.d;
  }
  const uint32_t p = y * W + x;
  vw.out_bgra[(H - 1u - y) * W + x] = bgra;  // draw_pixel_coord row flip, canvas.rs:955
  vw.out_depth[p] = depth;
  if (vw.out_winner) vw.out_winner[p] = covered ? id - vw.id_base : kNone;
}

// Per-tile visibility.  The tile's bin list is near first (or in submission
// order, or unordered: the result is the same); every
// covered fragment with depth < clear_depth contributes its key
// (depth_key(d) << 32 | id) to an LDS atomicMin, which leaves at each pixel the
// first fragment of minimum depth -- the winner of the reference's sequential
// strict `<` test.  Each wave owns QPW 8x8 quadrants of the tile:
//  1. records are staged in LDS in chunks of CH, each with the mask of the
//     tile's quadrants it can cover (its scan rectangle and, for an exact
//     record, exact per-quadrant edge tests), computed once for all waves;
//  2. lanes pre-filter 64 records at once: the wave's bits of that mask and
//     its largest recorded depth (hierarchical Z);
//  3. the wave visits the hits in order; lanes test coverage of their own
//     pixel without division (covered_fast) and against their pixel's recorded
//     depth, and append covered fragments to a per-wave LDS queue;
//  4. the queue is drained densely: divisions, depth, key, atomicMin.
// Early-Z only drops a fragment whose depth lower bound is strictly above a
// depth already recorded at its pixel, so it cannot be the minimum nor tie it.
// SHADE: the workgroup shades its tile's pixels itself once the tile's
// visibility is final (the tiles that finish first shade while others still
// rasterize), else it writes the winners to vis_out for k_shade.
template <int TS, int QPW, bool SHADE>
__global__ __launch_bounds__(64 * (TS / 8) * (TS / 8) / QPW) void k_raster(
    const DevView* __restrict__ views, uint32_t n_views, const uint2* __restrict__ ranges,
    const uint32_t* __restrict__ vals, const SetupRec* __restrict__ setup, const ShadeRec* __restrict__ shade,
    const DevMaterial* __restrict__ mats, uint32_t* __restrict__ vis_out, Status* __restrict__ status) {
  constexpr int QS = TS / 8, NQ = QS * QS, NW = NQ / QPW, NT = 64 * NW;
  constexpr int CH = 128;    // records staged per round
  // fragment queue per wave: drained once it holds more than kQFlush entries,
  // checked before each triangle, which appends at most 64 per quadrant
  constexpr int kQFlush = 224, QCAP = kQFlush + 64 * QPW;
  // the staged chunk: records, edge constants, depth-bound keys, ids,
  // quadrant masks (one LDS area, which the shading epilogue reuses for its
  // record runs).  A workgroup's LDS must fit in one of the CU's two 80 KB
  // halves, so <32, 2> (8 waves, 80 VGPRs) keeps below 40 KB to fit two per
  // half and run three per CU.
  constexpr int kChunkU4 = CH * (sizeof(SetupRec) + 48 + 4 + 4 + 4) / 16;
  __shared__ uint4 s_chunk[kChunkU4];
  SetupRec* const s_rec = reinterpret_cast<SetupRec*>(s_chunk);
  // per edge of an exact record: s*A, s*B and E, the edge's exact value
  // s*e - t at the tile origin (s = sign(f); t = 0 on a tie edge, else 1)
  float4 (*const s_pre)[3] = reinterpret_cast<float4 (*)[3]>(s_chunk + CH * sizeof(SetupRec) / 16);
  uint32_t* const s_zk = reinterpret_cast<uint32_t*>(s_chunk + CH * (sizeof(SetupRec) + 48) / 16);  // depth-bound key
  uint32_t* const s_id = s_zk + CH;
  // the tile's quadrants a record can cover (bit Q: quadrant row Q / QS,
  // column Q % QS), | kQExact for an exact record
  uint32_t* const s_qm = s_id + CH;
  __shared__ unsigned long long s_vis[TS * TS];
  __shared__ uint32_t s_q[NW][QCAP];
  __shared__ uint32_t s_nfrag;  // the tile's fragment count (one global atomic per tile)
  __shared__ uint32_t s_wmax[NW];  // each wave's largest recorded depth key, as of its last flush
  __shared__ float s_unit[SHADE ? 256 : 1];
  // How an exact hit's edge constants reach the wave.  kLdsHit: broadcast
  // reads of the staged s_pre.  Otherwise the pre-filtering lane reads them
  // and the wave takes them with readlanes: with one quadrant per wave a hit
  // has too little work to cover the LDS latency.
  constexpr bool kLdsHit = QPW >= 2;
  constexpr uint32_t kQExact = 1u << 16;

  const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const DevView& vw = views[wave_find_view_tile(views, n_views, tile)];
  const uint32_t lt = tile - vw.tile_base;
  const int tx = (int)(lt % vw.ntx), ty = (int)(lt / vw.ntx);
  const int W = vw.fr.W, H = vw.fr.H;
  const float clear_depth = vw.fr.clear_depth;
  const int tx0 = tx * TS, ty0 = ty * TS;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: keeps the wave's geometry scalar
  const uint2 rg = ranges[tile];

  // Pixels off the screen start at key 0, below every fragment's (depth
  // keys of recorded fragments and staged depth bounds are >= 1), so the
  // early-Z test alone keeps them uncovered.
  for (int p = tid; p < TS * TS; p += NT)
    s_vis[p] = (tx0 + p % TS < W && ty0 + p / TS < H) ? ~0ull : 0ull;
  if (tid == 0) s_nfrag = 0;
  if (tid < NW) s_wmax[tid] = 0xFFFFFFFFu;
  if constexpr (SHADE) fill_unit_table(s_unit, tid, NT);  // read after the barrier before the shading

  const int lx = lane & 7, ly = lane >> 3;
  const uint64_t lanemask_lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // The wave's quadrants: QPW consecutive ones in row-major order (part of a
  // quadrant row, or whole rows), bits wq0 .. wq0 + QPW - 1 of a record's
  // quadrant mask.
  const int wq0 = wave * QPW;
  int lp[QPW];        // local index of the lane's pixel in each of its quadrants
  uint32_t zk[QPW];   // recorded depth key at that pixel (0xFFFFFFFF: none yet; 0: off the screen)
  float dxf[QPW], dyf[QPW];  // the pixel's offset from the tile's origin
#pragma unroll
  for (int qq = 0; qq < QPW; qq++) {
    const int Q = wq0 + qq;
    lp[qq] = ((Q / QS) * 8 + ly) * TS + (Q % QS) * 8 + lx;
    const int px = tx0 + (Q % QS) * 8 + lx, py = ty0 + (Q / QS) * 8 + ly;
    dxf[qq] = (float)(px - tx0);
    dyf[qq] = (float)(py - ty0);
    zk[qq] = (px < W && py < H) ? 0xFFFFFFFFu : 0u;
  }
  uint32_t wmax = 0xFFFFFFFFu;  // largest recorded depth key over the wave's pixels
  // ...and over each of its quadrants (hierarchical Z per quadrant, with 4
  // quadrants per wave: a record's quadrant bits drop where it is behind;
  // with 2, the wave maximum per quadrant cost more than it saved)
  constexpr bool kQuadZ = QPW == 4;
  uint32_t qmax[QPW];
#pragma unroll
  for (int qq = 0; qq < QPW; qq++) qmax[qq] = 0xFFFFFFFFu;
  uint32_t qn = 0;
  uint32_t nfrag = 0;

  auto submit = [&](const SetupRec& r, uint32_t id, int px, int py) {
    const FragOut o = bary(r, (float)px, (float)py);
    if (!(r.flags & kUvSafe) && !uv_ok(shade, mats, id, o)) atomicOr(&status->error, 1u);
    if (o.d < clear_depth) {
      const unsigned long long key = ((unsigned long long)depth_key(o.d) << 32) | id;
      atomicMin(&s_vis[(py - ty0) * TS + (px - tx0)], key);
    }
  };
  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // queue writes by other lanes of this wave
    for (uint32_t j = (uint32_t)lane; j < qn; j += 64) {
      const uint32_t q = s_q[wave][j];
      const uint32_t k = q >> 16, l = q & 0xFFFFu;
      if (s_zk[k] > (uint32_t)(s_vis[l] >> 32)) continue;  // early-Z recheck
      submit(s_rec[k], s_id[k], tx0 + (int)(l % TS), ty0 + (int)(l / TS));
    }
    nfrag += qn;
    qn = 0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    uint32_t m = 0;
#pragma unroll
    for (int qq = 0; qq < QPW; qq++) {
      zk[qq] = (uint32_t)(s_vis[lp[qq]] >> 32);
      if constexpr (kQuadZ) {
        qmax[qq] = wave_max_u32(zk[qq]);
        m = max(m, qmax[qq]);
      } else {
        m = max(m, zk[qq]);
      }
    }
    wmax = kQuadZ ? m : wave_max_u32(m);
    s_wmax[wave] = wmax;  // read at the next staging, after a barrier
  };

  // Staging: P consecutive lanes of one wave per record of a chunk (P = NT /
  // CH, at least 1).  Lane part sp loads the record's 16-byte words sp,
  // sp + P, ... and stores them into LDS; the P lanes read back the words
  // they need (stores of their own wave: no barrier) and each tests the
  // quadrant rows sp, sp + P, ... of the tile; an OR over the P lanes joins
  // the rows.  With NT >= CH the loads run ahead of the raster: while a chunk
  // is rasterized, the next chunk's words (their ids loaded a chunk earlier)
  // and the ids of the one after are in flight in registers.
  constexpr int P = NT >= CH ? NT / CH : 1;
  constexpr int NPC = (6 + P - 1) / P;  // words per lane
  constexpr bool kPrefetch = NT >= CH;
  const int sr = NT >= CH ? tid / P : tid, sp = NT >= CH ? tid % P : 0;
  // This lane's words of a record: q_j = word sp + j P for j < NPC (named
  // values, not an array, which would live in scratch memory).  A lane with
  // fewer than NPC words loads word 5 again.
  auto load_words = [&](uint32_t id, uint4& q0, uint4& q1, uint4& q2, uint4& q3, uint4& q4, uint4& q5) {
    const uint4* src = reinterpret_cast<const uint4*>(setup + id);
    q0 = src[min(sp, 5)];
    if constexpr (NPC > 1) q1 = src[min(sp + P, 5)];
    if constexpr (NPC > 2) q2 = src[min(sp + 2 * P, 5)];
    if constexpr (NPC > 3) q3 = src[min(sp + 3 * P, 5)];
    if constexpr (NPC > 4) q4 = src[min(sp + 4 * P, 5)];
    if constexpr (NPC > 5) q5 = src[min(sp + 5 * P, 5)];
  };
  // OR of v over the P lanes of a record (all of them active)
  auto or_parts = [](uint32_t v) {
    if constexpr (P >= 2) v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);  // lane ^ 1
    if constexpr (P >= 4) v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);  // lane ^ 2
    if constexpr (P >= 8) v |= (uint32_t)__shfl_xor((int)v, 4, 64);
    return v;
  };
  uint32_t tile_wmax = 0xFFFFFFFFu;  // the tile's largest recorded depth key (s_wmax), for stage
  // Record i of a chunk into LDS, from this lane's words.
  auto stage = [&](int i, uint32_t id, uint4 q0, uint4 q1, uint4 q2, uint4 q3, uint4 q4, uint4 q5) {
    uint4* dst = reinterpret_cast<uint4*>(s_rec + i);
    if (sp < 6) dst[sp] = q0;
    if constexpr (NPC > 1) if (sp + P < 6) dst[sp + P] = q1;
    if constexpr (NPC > 2) if (sp + 2 * P < 6) dst[sp + 2 * P] = q2;
    if constexpr (NPC > 3) if (sp + 3 * P < 6) dst[sp + 3 * P] = q3;
    if constexpr (NPC > 4) if (sp + 4 * P < 6) dst[sp + 4 * P] = q4;
    if constexpr (NPC > 5) if (sp + 5 * P < 6) dst[sp + 5 * P] = q5;
    uint4 w0, w1, w2, w3, w5;
    uint32_t zb;  // the depth lower bound's bits
    if constexpr (P == 1) {
      w5 = q5;
      zb = q4.w;
    } else {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the other parts' stores
      w5 = dst[5];
      zb = __float_as_uint(s_rec[i].d[3]);
    }
    // depth bound key, at least 1: only off-screen pixels hold key 0 (a
    // fragment key of 0 is a NaN depth, which never passes the depth test)
    const uint32_t zkey = max(depth_key(__uint_as_float(zb)), 1u);
    if (sp == 0) {
      s_id[i] = id;
      s_zk[i] = zkey;
    }
    // A record whose depth bound is above every wave's largest recorded key
    // is hidden on the whole tile (each wave's pre-filter would drop it): no
    // quadrants, and no edge work.
    if (zkey > tile_wmax) {
      if (sp == 0) s_qm[i] = 0u;
      return;
    }
    if constexpr (P == 1) {
      w0 = q0; w1 = q1; w2 = q2; w3 = q3;
    } else {
      w0 = dst[0]; w1 = dst[1]; w2 = dst[2]; w3 = dst[3];
    }
    // The tile's quadrants the scan rectangle meets: quadrant column c meets
    // [x0, x1] iff floor((x0 - tx0) / 8) <= c <= floor((x1 - tx0) / 8), rows
    // likewise.
    const uint32_t flags = w5.z;
    const int c0 = max(((int)(w5.x & 0xFFFFu) - tx0) >> 3, 0), c1 = min(((int)(w5.x >> 16) - tx0) >> 3, QS - 1);
    const int r0 = max(((int)(w5.y & 0xFFFFu) - ty0) >> 3, 0), r1 = min(((int)(w5.y >> 16) - ty0) >> 3, QS - 1);
    uint32_t qmask = 0;
    if (c0 <= c1 && r0 <= r1) {
      const uint32_t cols = (2u << c1) - (1u << c0);
#pragma unroll
      for (int rr = 0; rr < QS; rr++)
        if (rr >= r0 && rr <= r1) qmask |= cols << (rr * QS);
    }
    // Exact records.  With S = |A| x1 + |B| y1 + |C| + |D| over the tile's
    // largest pixel coordinates (x1, y1) below 2^24 - 216 on every edge (the
    // float S is within 3 units of the exact one), every product and partial
    // sum of the edge function at the tile's pixels is an integer of
    // magnitude <= S, so s*e(p) - t = E + sA*(px - tx0) + sB*(py - ty0) is
    // exact in two fma steps, E = s*e(tx0, ty0) - t, and covered_fast on the
    // edge <=> s*e >= t <=> that value >= 0.  Pixels outside the triangle's
    // bounding box need no separate test: a covered pixel is inside the
    // snapped triangle, hence inside its bounding box unless it is off the
    // screen.
    bool exact = (flags & kFast) != 0u;
    float sA[3], sB[3], E[3];
    const uint4 edge[3] = {w0, w1, w2};
    const uint32_t fb[3] = {w3.x, w3.y, w3.z};
    const float tx0f = (float)tx0, ty0f = (float)ty0, tx1f = (float)(tx0 + TS - 1), ty1f = (float)(ty0 + TS - 1);
#pragma unroll
    for (int e = 0; e < 3; e++) {
      const float sg = __uint_as_float(fb[e]) > 0.0f ? 1.0f : -1.0f;
      const float A = __uint_as_float(edge[e].x), B = __uint_as_float(edge[e].y);
      const float C = __uint_as_float(edge[e].z), D = __uint_as_float(edge[e].w);
      sA[e] = sg * A;
      sB[e] = sg * B;
      const float S = (fabsf(sA[e]) * tx1f + fabsf(sB[e]) * ty1f) + (fabsf(C) + fabsf(D));
      if (!(S < 16777000.0f)) exact = false;
      E[e] = ((sA[e] * tx0f + sB[e] * ty0f) + sg * (C - D)) - ((flags & (kTieA << e)) ? 0.0f : 1.0f);
      if (e % P == sp) s_pre[i][e] = make_float4(sA[e], sB[e], E[e], 0.0f);
    }
    if (exact) {  // the same on the record's P lanes
      // ...less the quadrants an edge rules out: an edge's largest value over
      // quadrant (qx, qy) is at its pixel (8 qx + (sA > 0 ? 7 : 0), 8 qy +
      // (sB > 0 ? 7 : 0)) of the tile, computed exactly (a pixel of the
      // tile) as cv[qx] + sB * 8 qy; below zero there, no pixel of the
      // quadrant is covered.
      float cv[3][QS];
#pragma unroll
      for (int e = 0; e < 3; e++) {
        cv[e][0] = fmaf(sA[e], sA[e] > 0.0f ? 7.0f : 0.0f, fmaf(sB[e], sB[e] > 0.0f ? 7.0f : 0.0f, E[e]));
#pragma unroll
        for (int k = 1; k < QS; k++) cv[e][k] = fmaf(sA[e], (float)(8 * k), cv[e][0]);
      }
      uint32_t keep = 0;
#pragma unroll
      for (int j = 0; j < (QS >= P ? QS / P : 1); j++) {
        const int qy = sp + j * P;
        if (QS >= P || qy < QS) {
          float rv[3];
#pragma unroll
          for (int e = 0; e < 3; e++) rv[e] = sB[e] * (float)(8 * qy);
#pragma unroll
          for (int qx = 0; qx < QS; qx++) {
            const float v = fminf(fminf(cv[0][qx] + rv[0], cv[1][qx] + rv[1]), cv[2][qx] + rv[2]);
            keep |= (v >= 0.0f ? 1u : 0u) << (qy * QS + qx);
          }
        }
      }
      qmask = (qmask & or_parts(keep)) | kQExact;
    }
    if (sp == 0) s_qm[i] = qmask;
  };
  uint32_t pf_id = 0, pf_id2 = 0;  // ids of the next chunk and of the one after it
  uint4 p0{}, p1{}, p2{}, p3{}, p4{}, p5{};  // this lane's words of the next chunk's record
  if constexpr (kPrefetch) {
    if (rg.x + (uint32_t)sr < rg.y) {
      pf_id = vals[rg.x + sr];
      load_words(pf_id, p0, p1, p2, p3, p4, p5);
    }
    if (rg.x + CH + (uint32_t)sr < rg.y) pf_id2 = vals[rg.x + CH + sr];
  }

  for (uint32_t base = rg.x; base < rg.y; base += CH) {
    const int n = (int)min((uint32_t)CH, rg.y - base);
    __syncthreads();
    tile_wmax = 0u;
#pragma unroll
    for (int w = 0; w < NW; w++) tile_wmax = max(tile_wmax, s_wmax[w]);
    if constexpr (kPrefetch) {
      if (sr < n) stage(sr, pf_id, p0, p1, p2, p3, p4, p5);
    } else {
      for (int i = tid; i < n; i += NT) {
        const uint32_t id = vals[base + i];
        uint4 w0, w1, w2, w3, w4, w5;
        load_words(id, w0, w1, w2, w3, w4, w5);
        stage(i, id, w0, w1, w2, w3, w4, w5);
      }
    }
    __syncthreads();
    if constexpr (kPrefetch) {
      // the next chunk's records, the ids of the chunk after it
      const uint32_t b1 = base + CH, b2 = base + 2u * CH;
      pf_id = pf_id2;
      if (b1 + (uint32_t)sr < rg.y) load_words(pf_id, p0, p1, p2, p3, p4, p5);
      pf_id2 = 0;
      if (b2 + (uint32_t)sr < rg.y) pf_id2 = vals[b2 + sr];
    }
    for (int g0 = 0; g0 < n; g0 += 64) {
      bool hit = false, wexact = false;
      float pA[3] = {0.0f, 0.0f, 0.0f}, pB[3] = {0.0f, 0.0f, 0.0f}, pE[3] = {0.0f, 0.0f, 0.0f};
      uint32_t qm = 0;  // the wave's quadrants the record can cover
      uint32_t zb = 0;  // the record's depth-bound key
      if (g0 + lane < n) {
        zb = s_zk[g0 + lane];
        const uint32_t m = s_qm[g0 + lane];
        qm = (m >> wq0) & ((1u << QPW) - 1u);
        if constexpr (kQuadZ) {
#pragma unroll
          for (int qq = 0; qq < QPW; qq++)
            if (zb > qmax[qq]) qm &= ~(1u << qq);
        }
        hit = qm != 0u && !(zb > wmax);
        wexact = (m & kQExact) != 0u;
        if constexpr (!kLdsHit) {
          if (hit && wexact) {
#pragma unroll
            for (int e = 0; e < 3; e++) {
              const float4 q = s_pre[g0 + lane][e];
              pA[e] = q.x;
              pB[e] = q.y;
              pE[e] = q.z;
            }
          }
        }
      }
      // Exact two-FMA hits in a tight loop, then the rest (the visibility
      // minimum does not depend on the order of the visits).
      uint64_t hf = __ballot(hit && wexact);
      uint64_t hs = __ballot(hit) & ~hf;
      while (hf) {
        const int kl = (int)__builtin_ctzll(hf);
        const int k = g0 + kl;
        hf &= hf - 1;
        if (qn > kQFlush) flush();  // room for this triangle's visits
        // the record's quadrant mask and depth bound from the registers of the
        // lane that pre-filtered it, its edge constants as kLdsHit says
        const uint32_t qmk = (uint32_t)__builtin_amdgcn_readlane((int)qm, kl);
        const uint32_t zkey = (uint32_t)__builtin_amdgcn_readlane((int)zb, kl);
        float A[3], B[3], E[3];
        if constexpr (kLdsHit) {
#pragma unroll
          for (int e = 0; e < 3; e++) {
            const float4 q = s_pre[k][e];
            A[e] = q.x;
            B[e] = q.y;
            E[e] = q.z;
          }
        } else {
#pragma unroll
          for (int e = 0; e < 3; e++) {
            A[e] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pA[e]), kl));
            B[e] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pB[e]), kl));
            const float Es = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pE[e]), kl));
            // once per triangle into a vector register: a VALU instruction
            // reads at most one scalar register, and each fma below already
            // takes A or B
            asm("v_mov_b32 %0, %1" : "=v"(E[e]) : "s"(Es));
          }
        }
#pragma unroll
        for (int qq = 0; qq < QPW; qq++) {
          if (!((qmk >> qq) & 1u)) continue;
          const float w0 = fmaf(A[0], dxf[qq], fmaf(B[0], dyf[qq], E[0]));
          const float w1 = fmaf(A[1], dxf[qq], fmaf(B[1], dyf[qq], E[1]));
          const float w2 = fmaf(A[2], dxf[qq], fmaf(B[2], dyf[qq], E[2]));
          // bitwise, not short-circuit: straight-line code, no exec-mask
          // branches.  The ballot of each comparison is its lane mask; the
          // masks are ANDed on the scalar unit.
          const bool inside = fminf(fminf(w0, w1), w2) >= 0.0f, nearer = zkey <= zk[qq];
          const bool cov = inside & nearer;
          const uint64_t m = __ballot(inside) & __ballot(nearer);
          uint32_t* const qb = &s_q[wave][qn];
          if (cov) qb[lanes_below(m, 0u)] = ((uint32_t)k << 16) | (uint32_t)lp[qq];
          qn += (uint32_t)__popcll(m);
        }
      }
      while (hs) {
        const int kl = (int)__builtin_ctzll(hs);
        const int k = g0 + kl;
        hs &= hs - 1;
        if (qn > kQFlush) flush();
        const SetupRec& r = s_rec[k];
        const uint32_t hx = r.rx, hy = r.ry;
        const int rx0 = (int)(hx & 0xFFFFu), rx1 = (int)(hx >> 16);
        const int ry0 = (int)(hy & 0xFFFFu), ry1 = (int)(hy >> 16);
        const uint32_t flags = r.flags;
        const uint32_t zkey = (uint32_t)__builtin_amdgcn_readlane((int)zb, kl);
#pragma unroll
        for (int qq = 0; qq < QPW; qq++) {
          const int Q = wq0 + qq;
          const int qx0 = tx0 + (Q % QS) * 8, qy0 = ty0 + (Q / QS) * 8;
          if (rx1 < qx0 || rx0 > qx0 + 7 || ry1 < qy0 || ry0 > qy0 + 7) continue;
          const int px = qx0 + lx, py = qy0 + ly;
          const bool in = px >= rx0 && px <= rx1 && py >= ry0 && py <= ry1;
          const float xf = (float)px, yf = (float)py;
          if (flags & kFast) {
            bool cov = false;
            if (in && !(zkey > zk[qq])) {
              const float e0 = edge_eval(r.ea, xf, yf);
              const float e1 = edge_eval(r.eb, xf, yf);
              const float e2 = edge_eval(r.eg, xf, yf);
              cov = covered_fast(e0, r.f[0], flags & kTieA) && covered_fast(e1, r.f[1], flags & kTieB) &&
                    covered_fast(e2, r.f[2], flags & kTieG);
            }
            const uint64_t m = __ballot(cov);
            if (m) {
              if (cov) {
                const uint32_t pos = qn + (uint32_t)__popcll(m & lanemask_lt);
                s_q[wave][pos] = ((uint32_t)k << 16) | (uint32_t)lp[qq];
              }
              qn += (uint32_t)__popcll(m);
            }
          } else {
            // Any other triangle: the reference's predicate on the divided values.
            bool cov = false;
            if (in) {
              const FragOut o = bary(r, xf, yf);
              cov = o.al >= 0.0f && o.be >= 0.0f && o.ga >= 0.0f && (o.al > 0.0f || (flags & kTieA)) &&
                    (o.be > 0.0f || (flags & kTieB)) && (o.ga > 0.0f || (flags & kTieG));
            }
            if (cov) submit(r, s_id[k], px, py);
            nfrag += (uint32_t)__popcll(__ballot(cov));
          }
        }
      }
      if (qn) flush();
    }
  }
  // A wave with fragments ran the chunk loop, whose barriers order this
  // after the zeroing above.  Same-address atomics from every wave of a large
  // batch serialize in memory: one per tile, spread over kFragWords words.
  if (nfrag && lane == 0) atomicAdd(&s_nfrag, nfrag);  // nfrag is wave-uniform
  __syncthreads();
  if (tid == 0 && s_nfrag) atomicAdd(&status->frags[tile % kFragWords], s_nfrag);
  if constexpr (SHADE) {
    // each wave shades 64 consecutive pixels of the tile at a time, its
    // record runs in its share of the chunk area
    constexpr int R = (int)(kChunkU4 / (NW * 14)) < 16 ? (int)(kChunkU4 / (NW * 14)) : 16;
    uint4 (*const run)[14] = reinterpret_cast<uint4 (*)[14]>(s_chunk) + wave * R;
    for (int p0 = wave * 64; p0 < TS * TS; p0 += NT) {
      const int p = p0 + lane;
      const int px = tx0 + p % TS, py = ty0 + p / TS;
      const bool on = px < W && py < H;
      const unsigned long long v = s_vis[p];
      shade_wave<R>((v == ~0ull) ? kNone : (uint32_t)v, on, (uint32_t)px, (uint32_t)py, vw, setup, shade, mats,
                    status, run, s_unit);
    }
  } else {
    for (int p = tid; p < TS * TS; p += NT) {
      const int px = tx0 + p % TS, py = ty0 + p / TS;
      if (px < W && py < H) {
        const unsigned long long v = s_vis[p];
        vis_out[vw.px_base + (uint32_t)py * (uint32_t)W + (uint32_t)px] = (v == ~0ull) ? kNone : (uint32_t)v;
      }
    }
  }
}

// Triangle-parallel variant of the per-tile visibility pass.  The waves of a
// tile take the staged triangles one at a time from an LDS counter; a
// triangle's scan rectangle clipped to the tile is covered row-major, 64
// pixels per visit (a 9x9 box takes 2 visits where 8x8-aligned quadrants take
// 4 or more).  Coverage is the same exact test as k_raster: two fma steps per
// edge when every edge value over the rectangle is an integer below 2^24,
// covered_fast otherwise, the divided predicate for other triangles.
// Early-Z reads the pixel's recorded key from LDS; covered fragments go to a
// per-wave queue drained through the exact divisions into the LDS atomicMin.
template <int TS, int NW>
__global__ __launch_bounds__(64 * NW) void k_raster_tp(
    const DevView* __restrict__ views, uint32_t n_views, const uint2* __restrict__ ranges,
    const uint32_t* __restrict__ vals, const SetupRec* __restrict__ setup, const ShadeRec* __restrict__ shade,
    const DevMaterial* __restrict__ mats, uint32_t* __restrict__ vis_out, Status* __restrict__ status) {
  constexpr int NT = 64 * NW;
  constexpr int CH = 128;
  constexpr int QCAP = 256;
  __shared__ SetupRec s_rec[CH];
  __shared__ uint4 s_hdr[CH];
  __shared__ float4 s_pre[CH][3];
  __shared__ uint32_t s_id[CH];
  __shared__ unsigned long long s_vis[TS * TS];
  __shared__ uint32_t s_q[NW][QCAP];
  __shared__ uint32_t s_next;
  __shared__ uint32_t s_nfrag;  // the tile's fragment count (one global atomic per tile)

  const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const DevView& vw = views[wave_find_view_tile(views, n_views, tile)];
  const uint32_t lt = tile - vw.tile_base;
  const int tx0 = (int)(lt % vw.ntx) * TS, ty0 = (int)(lt / vw.ntx) * TS;
  const int W = vw.fr.W, H = vw.fr.H;
  const float clear_depth = vw.fr.clear_depth;
  const int tx1 = min(tx0 + TS, W) - 1, ty1 = min(ty0 + TS, H) - 1;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const uint2 rg = ranges[tile];
  const uint64_t lanemask_lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int p = tid; p < TS * TS; p += NT) s_vis[p] = ~0ull;
  if (tid == 0) s_nfrag = 0;
  uint32_t qn = 0, nfrag = 0;

  auto submit = [&](const SetupRec& r, uint32_t id, int px, int py) {
    const FragOut o = bary(r, (float)px, (float)py);
    if (!(r.flags & kUvSafe) && !uv_ok(shade, mats, id, o)) atomicOr(&status->error, 1u);
    if (o.d < clear_depth) {
      const unsigned long long key = ((unsigned long long)depth_key(o.d) << 32) | id;
      atomicMin(&s_vis[(py - ty0) * TS + (px - tx0)], key);
    }
  };
  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (uint32_t j = (uint32_t)lane; j < qn; j += 64) {
      const uint32_t q = s_q[wave][j];
      const uint32_t k = q >> 16, l = q & 0xFFFFu;
      if (s_hdr[k].w > (uint32_t)(s_vis[l] >> 32)) continue;  // early-Z recheck
      submit(s_rec[k], s_id[k], tx0 + (int)(l % TS), ty0 + (int)(l / TS));
    }
    nfrag += qn;
    qn = 0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  };
  const volatile uint32_t* vis_hi = reinterpret_cast<const volatile uint32_t*>(s_vis);

  for (uint32_t base = rg.x; base < rg.y; base += CH) {
    const int n = (int)min((uint32_t)CH, rg.y - base);
    __syncthreads();
    for (int i = tid; i < n; i += NT) {
      const uint32_t id = vals[base + i];
      s_id[i] = id;
      const uint4* src = reinterpret_cast<const uint4*>(setup + id);
      uint4* dst = reinterpret_cast<uint4*>(s_rec + i);
      uint4 w[6];
#pragma unroll
      for (int q = 0; q < 6; q++) w[q] = src[q];
#pragma unroll
      for (int q = 0; q < 6; q++) dst[q] = w[q];
      s_hdr[i] = make_uint4(w[5].x, w[5].y, w[5].z, depth_key(__uint_as_float(w[4].w)));
      const float fs[3] = {__uint_as_float(w[3].x), __uint_as_float(w[3].y), __uint_as_float(w[3].z)};
#pragma unroll
      for (int e = 0; e < 3; e++) {
        const float sg = fs[e] > 0.0f ? 1.0f : -1.0f;
        const float A = __uint_as_float(w[e].x), B = __uint_as_float(w[e].y);
        const float C = __uint_as_float(w[e].z), D = __uint_as_float(w[e].w);
        s_pre[i][e] = make_float4(sg * A, sg * B, sg * (C - D), fabsf(C) + fabsf(D));
      }
    }
    if (tid == 0) s_next = 0;
    __syncthreads();
    for (;;) {
      uint32_t kk = 0;
      if (lane == 0) kk = atomicAdd(&s_next, 1u);
      const int k = (int)__builtin_amdgcn_readfirstlane(kk);
      if (k >= n) break;
      const uint4 h = s_hdr[k];
      const int cx0 = max((int)(__builtin_amdgcn_readfirstlane(h.x) & 0xFFFFu), tx0);
      const int cx1 = min((int)(__builtin_amdgcn_readfirstlane(h.x) >> 16), tx1);
      const int cy0 = max((int)(__builtin_amdgcn_readfirstlane(h.y) & 0xFFFFu), ty0);
      const int cy1 = min((int)(__builtin_amdgcn_readfirstlane(h.y) >> 16), ty1);
      if (cx0 > cx1 || cy0 > cy1) continue;
      const uint32_t flags = __builtin_amdgcn_readfirstlane(h.z);
      const uint32_t zkey = __builtin_amdgcn_readfirstlane(h.w);
      const int bw = cx1 - cx0 + 1, area = bw * (cy1 - cy0 + 1);
      const float inv_w = 1.0f / (float)bw;
      const SetupRec& r = s_rec[k];
      // Exact edge values over the rectangle (see k_raster): E at its corner.
      bool exact = (flags & kFast) != 0;
      float A[3], B[3], E[3];
#pragma unroll
      for (int e = 0; e < 3; e++) {
        const float4 q = s_pre[k][e];
        const float S = (fabsf(q.x) * (float)cx1 + fabsf(q.y) * (float)cy1) + q.w;
        if (!(S < 16777000.0f)) exact = false;
        A[e] = q.x;
        B[e] = q.y;
        E[e] = ((q.x * (float)cx0 + q.y * (float)cy0) + q.z) - ((flags & (kTieA << e)) ? 0.0f : 1.0f);
      }
      for (int v0 = 0; v0 < area; v0 += 64) {
        const int idx = v0 + lane;
        const int dy = (int)(((float)idx + 0.5f) * inv_w);
        const int dx = idx - dy * bw;
        const int px = cx0 + dx, py = cy0 + dy;
        const int l = (py - ty0) * TS + (px - tx0);
        bool cov = false;
        if (idx < area) {
          if (exact) {
            const float fx = (float)dx, fy = (float)dy;
            const float w0 = fmaf(A[0], fx, fmaf(B[0], fy, E[0]));
            const float w1 = fmaf(A[1], fx, fmaf(B[1], fy, E[1]));
            const float w2 = fmaf(A[2], fx, fmaf(B[2], fy, E[2]));
            cov = fminf(fminf(w0, w1), w2) >= 0.0f;
          } else if (flags & kFast) {
            const float xf = (float)px, yf = (float)py;
            cov = covered_fast(edge_eval(r.ea, xf, yf), r.f[0], flags & kTieA) &&
                  covered_fast(edge_eval(r.eb, xf, yf), r.f[1], flags & kTieB) &&
                  covered_fast(edge_eval(r.eg, xf, yf), r.f[2], flags & kTieG);
          } else {
            const FragOut o = bary(r, (float)px, (float)py);
            cov = o.al >= 0.0f && o.be >= 0.0f && o.ga >= 0.0f && (o.al > 0.0f || (flags & kTieA)) &&
                  (o.be > 0.0f || (flags & kTieB)) && (o.ga > 0.0f || (flags & kTieG));
          }
          if (cov) cov = !(zkey > vis_hi[2 * l + 1]);  // early-Z on the recorded depth key
        }
        if (!(flags & kFast)) {  // any other triangle: submitted directly
          if (cov) submit(r, s_id[k], px, py);
          nfrag += (uint32_t)__popcll(__ballot(cov));
          continue;
        }
        const uint64_t m = __ballot(cov);
        if (m) {
          if (cov) s_q[wave][qn + (uint32_t)__popcll(m & lanemask_lt)] = ((uint32_t)k << 16) | (uint32_t)l;
          qn += (uint32_t)__popcll(m);
          if (qn > QCAP - 64) flush();
        }
      }
    }
    if (qn) flush();
  }
  if (nfrag && lane == 0) atomicAdd(&s_nfrag, nfrag);  // after the loop's barriers, as in k_raster
  __syncthreads();
  if (tid == 0 && s_nfrag) atomicAdd(&status->frags[tile % kFragWords], s_nfrag);
  for (int p = tid; p < TS * TS; p += NT) {
    const int px = tx0 + p % TS, py = ty0 + p / TS;
    if (px < W && py < H) {
      const unsigned long long v = s_vis[p];
      vis_out[vw.px_base + (uint32_t)py * (uint32_t)W + (uint32_t)px] = (v == ~0ull) ? kNone : (uint32_t)v;
    }
  }
}

// One thread per pixel of the visibility buffer (see shade_wave).
constexpr int kShadeRuns = 16;

__global__ __launch_bounds__(256) void k_shade(const DevView* __restrict__ views, uint32_t n_views, uint32_t n_pixels,
                                               const uint32_t* __restrict__ vis, const SetupRec* __restrict__ setup,
                                               const ShadeRec* __restrict__ shade,
                                               const DevMaterial* __restrict__ mats, Status* __restrict__ status) {
  __shared__ uint4 s_run[4][kShadeRuns][14];  // setup bytes 0-79, then the 144 shade bytes
  __shared__ float s_unit[256];
  fill_unit_table(s_unit, (int)threadIdx.x, 256);
  __syncthreads();
  const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint32_t g0 = xcd_remap(blockIdx.x, gridDim.x) * 256u + wave * 64u;  // the wave's first pixel
  const uint32_t g = g0 + (threadIdx.x & 63u);
  const uint32_t id = g < n_pixels ? vis[g] : kSkip;
  const bool active = id != kSkip;  // (kSkip: shaded by another shard)
  const uint32_t v = wave_find_view_px(views, n_views, g0, g);
  const DevView& vw = views[active ? v : 0u];
  const uint32_t p = active ? g - vw.px_base : 0u;
  const uint32_t W = (uint32_t)vw.fr.W;
  const uint32_t y = p / W, x = p - y * W;
  shade_wave<kShadeRuns>(id, active, x, y, vw, setup, shade, mats, status, s_run[wave], s_unit);
}


// ========================================================= transparent blend
// Transparent fragments (Canvas::draw_pixel_coord_with_depth with opacity < 1,
// canvas.rs:906-930; Pixel Mul/Add, canvas.rs:136-169): per channel
// bg*(1 - a) and c*a each truncated to u8, summed with u8 wrap-around, pad 0.
__device__ inline uint32_t blend_px(uint32_t bg, uint32_t c, float a) {
  const float ia = 1.0f - a;
  uint32_t out = 0;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const uint32_t p = f2u8((float)((bg >> (8 * i)) & 255u) * ia);
    const uint32_t q = f2u8((float)((c >> (8 * i)) & 255u) * a);
    out |= ((p + q) & 255u) << (8 * i);
  }
  return out;  // pad byte 0
}

// One workgroup per tile of a one-view frame.  The opaque pass has left each
// pixel's final colour, depth D and winner.  Every later opaque fragment of
// the reference's sequence loses to that winner, so from the winner on the
// depth buffer holds D: a transparent fragment matters iff its object group
// is not before the winner's (transparent meshes are drawn after their own
// object's opaque meshes, scene/mod.rs:1088) and its depth is < D.  Lanes own
// pixels and apply the tile's transparent triangles in submission order.
template <int TS>
__global__ __launch_bounds__(256) void k_blend(const DevView* __restrict__ views, const uint2* __restrict__ ranges,
                                               const uint32_t* __restrict__ vals, const SetupRec* __restrict__ setup,
                                               const ShadeRec* __restrict__ shade,
                                               const DevMaterial* __restrict__ mats,
                                               const uint32_t* __restrict__ grp_base, uint32_t n_groups,
                                               Status* __restrict__ status) {
  constexpr int PPT = (TS * TS + 255) / 256;
  constexpr int CH = 64;
  __shared__ SetupRec s_rec[CH];
  __shared__ uint32_t s_id[CH];
  __shared__ float s_unit[256];
  const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const DevView& vw = views[0];
  const int W = vw.fr.W, H = vw.fr.H;
  const int tx0 = (int)(tile % vw.ntx) * TS, ty0 = (int)(tile / vw.ntx) * TS;
  const int tid = threadIdx.x;
  fill_unit_table(s_unit, tid, 256);  // read after the chunk loop's first barrier
  uint32_t col[PPT];
  float D[PPT];
  int gw[PPT];
  bool ok[PPT];
#pragma unroll
  for (int q = 0; q < PPT; q++) {
    const int p = tid + 256 * q;
    const int px = tx0 + p % TS, py = ty0 + p / TS;
    ok[q] = p < TS * TS && px < W && py < H;
    col[q] = 0;
    D[q] = 0.0f;
    gw[q] = -1;
    if (ok[q]) {
      col[q] = vw.out_bgra[(uint32_t)(H - 1 - py) * (uint32_t)W + (uint32_t)px];
      D[q] = vw.out_depth[(uint32_t)py * (uint32_t)W + (uint32_t)px];
      const uint32_t w = vw.out_winner[(uint32_t)py * (uint32_t)W + (uint32_t)px];
      if (w != kNone) {  // last group whose first id is <= w
        uint32_t lo = 0, hi = n_groups;
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (grp_base[mid] <= w) lo = mid;
          else hi = mid;
        }
        gw[q] = (int)lo;
      }
    }
  }
  const uint2 rg = ranges[tile];
  for (uint32_t base = rg.x; base < rg.y; base += CH) {
    const int n = (int)min((uint32_t)CH, rg.y - base);
    __syncthreads();
    for (int i = tid; i < n * 6; i += 256) {
      const int k = i / 6, w = i % 6;
      const uint32_t id = vals[base + k];
      if (w == 0) s_id[k] = id;
      reinterpret_cast<uint4*>(s_rec + k)[w] = reinterpret_cast<const uint4*>(setup + id)[w];
    }
    __syncthreads();
    for (int k = 0; k < n; k++) {
      const SetupRec& r = s_rec[k];
      if (r.flags & kDrop) continue;
      const int rx0 = (int)(r.rx & 0xFFFFu), rx1 = (int)(r.rx >> 16);
      const int ry0 = (int)(r.ry & 0xFFFFu), ry1 = (int)(r.ry >> 16);
#pragma unroll
      for (int q = 0; q < PPT; q++) {
        const int p = tid + 256 * q;
        const int px = tx0 + p % TS, py = ty0 + p / TS;
        if (!ok[q] || px < rx0 || px > rx1 || py < ry0 || py > ry1) continue;
        const float xf = (float)px, yf = (float)py;
        bool cov;
        FragOut o;
        if (r.flags & kFast) {
          cov = covered_fast(edge_eval(r.ea, xf, yf), r.f[0], r.flags & kTieA) &&
                covered_fast(edge_eval(r.eb, xf, yf), r.f[1], r.flags & kTieB) &&
                covered_fast(edge_eval(r.eg, xf, yf), r.f[2], r.flags & kTieG);
          if (cov) o = bary(r, xf, yf);
        } else {
          o = bary(r, xf, yf);
          cov = o.al >= 0.0f && o.be >= 0.0f && o.ga >= 0.0f && (o.al > 0.0f || (r.flags & kTieA)) &&
                (o.be > 0.0f || (r.flags & kTieB)) && (o.ga > 0.0f || (r.flags & kTieG));
        }
        if (!cov) continue;
        // both maps are sampled before the depth test (canvas.rs:689-695)
        const ShadeRec& sr = shade[s_id[k]];
        const DevMaterial& m = mats[sr.material];
        const uint32_t c = shade_pixel(sr, m, o, status, s_unit);
        if (o.d < D[q] && (int)sr.group >= gw[q]) col[q] = blend_px(col[q], c, m.alpha);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < PPT; q++) {
    const int p = tid + 256 * q;
    const int px = tx0 + p % TS, py = ty0 + p / TS;
    if (ok[q]) vw.out_bgra[(uint32_t)(H - 1 - py) * (uint32_t)W + (uint32_t)px] = col[q];
  }
}

// ============================================================ launch helpers
static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

template <int TS, int QPW>
static void launch_raster(const PipeArgs& a, hipStream_t s) {
  constexpr int NT = 64 * (TS / 8) * (TS / 8) / QPW;
  if (a.shade_in_raster)
    hipLaunchKernelGGL((k_raster<TS, QPW, true>), dim3(a.n_tiles), dim3(NT), 0, s, a.views, a.n_views, a.ranges,
                       a.vals[0], a.setup, a.shade, a.mats, a.vis, a.status);
  else
    hipLaunchKernelGGL((k_raster<TS, QPW, false>), dim3(a.n_tiles), dim3(NT), 0, s, a.views, a.n_views, a.ranges,
                       a.vals[0], a.setup, a.shade, a.mats, a.vis, a.status);
}

template <int TS, int NW>
static void launch_raster_tp(const PipeArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((k_raster_tp<TS, NW>), dim3(a.n_tiles), dim3(64 * NW), 0, s, a.views, a.n_views, a.ranges,
                     a.vals[0], a.setup, a.shade, a.mats, a.vis, a.status);
}

bool raster_config_supported(int TS, int waves, int kind) {
  if (kind == 1) return (TS == 16 || TS == 32) && (waves == 2 || waves == 4 || waves == 8);
  if (kind != 0) return false;
  if (TS != 8 && TS != 16 && TS != 32) return false;
  const int nq = (TS / 8) * (TS / 8);
  if (waves < 1 || nq % waves) return false;
  const int qpw = nq / waves;
  return qpw == 1 || qpw == 2 || qpw == 4 || qpw == 8 || qpw == 16;
}

void enqueue_raster(const PipeArgs& a, hipStream_t s) {
  if (a.raster_kind == 1) {
    if (a.TS == 16) {
      if (a.waves == 2) launch_raster_tp<16, 2>(a, s);
      else if (a.waves == 4) launch_raster_tp<16, 4>(a, s);
      else launch_raster_tp<16, 8>(a, s);
    } else {
      if (a.waves == 2) launch_raster_tp<32, 2>(a, s);
      else if (a.waves == 4) launch_raster_tp<32, 4>(a, s);
      else launch_raster_tp<32, 8>(a, s);
    }
    return;
  }
  const int qpw = (a.TS / 8) * (a.TS / 8) / a.waves;
  if (a.TS == 8) launch_raster<8, 1>(a, s);
  else if (a.TS == 16) {
    if (qpw == 1) launch_raster<16, 1>(a, s);
    else if (qpw == 2) launch_raster<16, 2>(a, s);
    else launch_raster<16, 4>(a, s);
  } else {
    if (qpw == 1) launch_raster<32, 1>(a, s);
    else if (qpw == 2) launch_raster<32, 2>(a, s);
    else if (qpw == 4) launch_raster<32, 4>(a, s);
    else if (qpw == 8) launch_raster<32, 8>(a, s);
    else launch_raster<32, 16>(a, s);
  }
}

void enqueue_blend(const PipeArgs& b, const TransArgs& t, hipStream_t s) {
  const dim3 g(b.n_tiles), blk(256);
  if (b.TS == 8)
    hipLaunchKernelGGL(k_blend<8>, g, blk, 0, s, b.views, b.ranges, b.vals[0], b.setup, b.shade, b.mats,
                       t.grp_base, t.n_groups, b.status);
  else if (b.TS == 16)
    hipLaunchKernelGGL(k_blend<16>, g, blk, 0, s, b.views, b.ranges, b.vals[0], b.setup, b.shade, b.mats,
                       t.grp_base, t.n_groups, b.status);
  else
    hipLaunchKernelGGL(k_blend<32>, g, blk, 0, s, b.views, b.ranges, b.vals[0], b.setup, b.shade, b.mats,
                       t.grp_base, t.n_groups, b.status);
}

// Sort-last shards.  k_vis_pack writes a shard's full-screen visibility as
// u64 (depth key << 32 | shard << 29 | clipped-triangle id): the same key the
// raster's atomicMin kept (bary and depth_key at the winner's pixel), with
// ids ordered by (shard, submission) -- so the u64 minimum over the shards
// of a frame split by submission range is the single-GPU winner, ties
// included.  No winner: ~0.
__global__ __launch_bounds__(256) void k_vis_pack(const DevView* __restrict__ views, uint32_t n_pixels,
                                                  const uint32_t* __restrict__ vis, const SetupRec* __restrict__ setup,
                                                  uint32_t tag, unsigned long long* __restrict__ out) {
  const uint32_t g = blockIdx.x * 256u + threadIdx.x;
  if (g >= n_pixels) return;
  const DevView& vw = views[0];
  const uint32_t W = (uint32_t)vw.fr.W;
  const uint32_t y = g / W, x = g - y * W;
  const uint32_t id = vis[g];
  unsigned long long k = ~0ull;
  if (id != kNone) {
    const FragOut o = bary(setup[id], (float)x, (float)y);
    k = ((unsigned long long)depth_key(o.d) << 32) | (tag | id);
  }
  out[g] = k;
}

// The merged visibility back to this shard's winner ids: its own pixels keep
// their id, pixels without a winner go to shard 0 (kNone: the clear colour),
// the rest are kSkip.
__global__ __launch_bounds__(256) void k_vis_unpack(const unsigned long long* __restrict__ merged, uint32_t n_pixels,
                                                    uint32_t shard, uint32_t* __restrict__ vis) {
  const uint32_t g = blockIdx.x * 256u + threadIdx.x;
  if (g >= n_pixels) return;
  const unsigned long long k = merged[g];
  uint32_t v;
  if (k == ~0ull) {
    v = shard == 0 ? kNone : kSkip;
  } else {
    const uint32_t lo = (uint32_t)k;
    v = (lo >> kShardShift) == shard ? (lo & ((1u << kShardShift) - 1u)) : kSkip;
  }
  vis[g] = v;
}

__global__ __launch_bounds__(256) void k_vis_min(unsigned long long* __restrict__ acc,
                                                 const unsigned long long* __restrict__ src, uint64_t n) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256ull)
    acc[i] = min(acc[i], src[i]);
}

void enqueue_vis_pack(const PipeArgs& a, uint32_t shard, uint64_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_vis_pack, dim3(cdiv(a.n_pixels, 256)), dim3(256), 0, s, a.views, a.n_pixels, a.vis, a.setup,
                     shard << kShardShift, reinterpret_cast<unsigned long long*>(out));
}

void enqueue_vis_unpack(const uint64_t* merged, uint32_t n_pixels, uint32_t shard, uint32_t* vis, hipStream_t s) {
  hipLaunchKernelGGL(k_vis_unpack, dim3(cdiv(n_pixels, 256)), dim3(256), 0, s,
                     reinterpret_cast<const unsigned long long*>(merged), n_pixels, shard, vis);
}

void enqueue_vis_min(uint64_t* acc, const uint64_t* src, uint64_t n, hipStream_t s) {
  const uint64_t nb = (n + 255) / 256;
  const uint32_t blocks = nb > 8192u ? 8192u : (nb ? (uint32_t)nb : 1u);
  hipLaunchKernelGGL(k_vis_min, dim3(blocks), dim3(256), 0, s,
                     reinterpret_cast<unsigned long long*>(acc), reinterpret_cast<const unsigned long long*>(src), n);
}

void enqueue_shade(const PipeArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_shade, dim3(cdiv(a.n_pixels, 256)), dim3(256), 0, s, a.views, a.n_views, a.n_pixels, a.vis,
                     a.setup, a.shade, a.mats, a.status);
}

}  // namespace dh
