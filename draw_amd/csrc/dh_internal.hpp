// dh_internal.hpp -- library-internal helpers shared by dh_api.cpp and the
// render server (dh_server.cpp).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "../../include/drawhip.h"

namespace dh {
uint64_t ctx_scratch_bytes(dh_ctx c);
uint64_t scratch_estimate(uint64_t n_in, uint64_t n_tiles, uint64_t n_pixels, uint64_t n_draws, uint64_t n_views);
int ctx_tile_size(dh_ctx c);
// Frame scratch of the context may not grow past cap bytes (0 = no cap); a
// frame that would returns DH_E_BUDGET and ctx_scratch_need gives its need.
void ctx_set_scratch_cap(dh_ctx c, uint64_t cap);
uint64_t ctx_scratch_need(dh_ctx c);
int mesh_triangles(dh_ctx c, dh_handle mesh, uint32_t* n_tri);
void ctx_release_scratch(dh_ctx c);
hipError_t enqueue_gui(const dh_gui_vertex* vtx, const uint4* tris, uint32_t n_tri, const uint4* clips, void* recs,
                       int W, int H, const uint8_t* font, int32_t fw, int32_t fh, int32_t fcomp, uint32_t* bgra,
                       const float* depth, uint32_t* error, hipStream_t s);
constexpr size_t kGuiRecBytes = 144;  // sizeof(GuiRec)
hipError_t enqueue_bgra_to_rgba(const uint32_t* src, uint32_t* dst, uint32_t n, hipStream_t s);
// The fixed RGBA staging array of dh_canvas_read_rgba_noalloc: at most
// kRgbaScratchPixels pixels (one 3840x2160 frame, 31.6 MiB per device).
constexpr uint32_t kRgbaScratchPixels = 3840u * 2160u;
// Looks up the array's device address on `device` once (set-up: called by
// dh_ctx_create); later calls return the stored status.
hipError_t rgba_scratch_init(int device);
// The array's device address on `device` and the lock its users share.
hipError_t rgba_scratch(int device, uint32_t** ptr, std::mutex** mu);
}  // namespace dh
