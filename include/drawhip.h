/* drawhip.h -- C ABI of libdrawhip, the MI355X (gfx950) raster path of
 * mororo18/draw.
 *
 * The reference has no FFI: its seam is two Rust methods, Scene::render
 * (src/renderer/scene/mod.rs:901) and Canvas::draw_triangle_with_attributes
 * (src/renderer/canvas.rs:577), called once per frame and once per clipped
 * triangle.  A per-triangle FFI would be launch-bound, so this ABI replaces
 * the body of Scene::render (scene/mod.rs:901-1249) for whole frames, while
 * the caller keeps its own scene, camera and draw-call API.
 *
 * Conventions
 *  - Every call returns int: DH_OK (0) or a negative DH_E_* code.  Nothing
 *    unwinds across the ABI.  The reference panics where this returns an error.
 *  - Host buffers are caller-owned and only borrowed for the duration of a call.
 *  - Device objects (textures, materials, objects, meshes, canvases) live in
 *    HBM, owned by the context, behind opaque 64-bit handles.
 *  - One context = one HIP device + one HIP stream + its scratch memory.  A
 *    context may move between threads but must not be used by two at once.
 */
#ifndef DRAWHIP_H
#define DRAWHIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    DH_OK = 0,
    DH_E_ARG = -1,        /* invalid argument                                           */
    DH_E_INDEX = -2,      /* vertex index out of range: IndexedMesh::vec3_list_from_indexed
                             assert! (scene/mesh.rs:46-48)                               */
    DH_E_TEXCOORD = -3,   /* texture slice out of range: TextureMap::get_rgb_slice
                             (scene/mod.rs:167) panics on that fragment                   */
    DH_E_CAMERA = -4,     /* Camera::new asserts (scene/mod.rs:327-331)                   */
    DH_E_PLANE = -5,      /* ViewPlane::new assert (scene/mod.rs:624)                      */
    DH_E_HIP = -6,        /* HIP runtime error                                           */
    DH_E_NOMEM = -7,      /* device allocation failed                                    */
    DH_E_HANDLE = -8,     /* unknown or wrong-kind handle                                */
    DH_E_UNSUPPORTED = -9,/* not available in this build (e.g. no GPU)                    */
    DH_E_BUDGET = -10,    /* request larger than the server's memory budget              */
    DH_E_CAPACITY = -11,  /* scratch capacity could not be grown                         */
};

typedef struct dh_ctx_s* dh_ctx;
typedef uint64_t dh_handle;

/* Per-frame uniforms, computed on the host (reference a1 row): the product
 * M = viewport*orth*persp*view of Scene::gen_transformation_matrix
 * (scene/mod.rs:817-899), row-major; the six world-space planes of
 * Camera::gen_view_planes (scene/mod.rs:481-593) in the order near, far,
 * right, left, top, bottom, each (normal.xyz, k); camera position; light
 * position; canvas size, offset (Canvas::apply_offset, canvas.rs:382-385),
 * clear depth (Canvas::init_depth, canvas.rs:403) and clear colour
 * (Pixel::azul_bb, canvas.rs:131-133) as B,G,R,A. */
typedef struct {
    float M[16];
    float cam[3];
    float light[3];
    float planes[6][4];
    int32_t W, H;
    float off_x, off_y;
    float clear_depth;
    uint8_t clear_bgra[4];
} dh_frame_params;

/* Camera::new(pos, dir, ratio) (scene/mod.rs:297-357) with the three values
 * it hard-codes made explicit: near (-10), far (near-500) and fov_x (135 deg).
 * near/far keep the reference's sign convention: far < near < 0. */
typedef struct {
    float pos[3];
    float dir[3];
    float fov_x_deg;
    float near_;
    float far_;
} dh_camera;

/* Texture (material) of scene/mod.rs:206-216.  map_kd/map_ka: texture
 * handles, 0 = TextureMap::default (1x1 white RGB, scene/mod.rs:128-135).
 * power: the specular exponent canvas.rs:724 fixes at 2 (must be >= 0). */
typedef struct {
    float ka[3], kd[3], ks[3];
    float alpha;
    int32_t power;
    dh_handle map_kd;
    dh_handle map_ka;
} dh_material_desc;

/* Frame counters of the last render on a context. */
typedef struct {
    uint64_t tris_in;        /* input triangles submitted                       */
    uint64_t tris_out;       /* clipped triangles emitted (opaque)              */
    uint64_t bin_entries;    /* (tile, triangle) bin entries                    */
    uint64_t frags;          /* covered fragments that reached the depth test   */
    uint64_t retries;        /* capacity regrowths during the last render       */
    /* Device time of the last render's final attempt (HIP events on the
     * context stream; zero unless dh_set_profiling(ctx, 1)). */
    float ms_geometry;       /* cull + clip + setup                             */
    float ms_binning;        /* tile counts, pairs, radix sort, tile ranges     */
    float ms_raster;         /* per-tile coverage + depth test                  */
    float ms_shade;          /* deferred shading + framebuffer writes           */
    float ms_frame;          /* whole frame                                     */
} dh_stats;

/* ---- host-side uniforms ---------------------------------------------- */
/* Camera::new + Scene::gen_transformation_matrix + Camera::gen_view_planes
 * (scene/mod.rs:297-357, 817-899, 481-593).  Pure host code; light may be
 * NULL (then (0,300,300), scene/mod.rs:766).  Sets clear_depth 100000
 * (app/mod.rs:80), clear colour azul_bb and offset 0. */
int dh_frame_uniforms(const dh_camera* cam, int32_t W, int32_t H, const float light[3],
                      dh_frame_params* out);
/* The reference application's configuration: Scene::new(screen_w, screen_h)
 * (scene/mod.rs:760-786) sets the camera's aspect ratio and the viewport
 * matrix from the screen size, and the frame is drawn into a separate
 * Canvas::new(canvas_w, canvas_h) (canvas.rs:366) moved by
 * Canvas::apply_offset(off_x, off_y) (canvas.rs:382-385); the app uses an
 * 800x600 canvas at ((screen_w-800)/2, (screen_h-600)/2) (app/mod.rs:59-63,
 * 81-84).  W/H of the result are the canvas size, and every snapped vertex is
 * screen_coord - offset (canvas.rs:585-587); the raster's screen rectangle is
 * the canvas (canvas.rs:647).  dh_frame_uniforms is this with canvas = screen
 * and offset 0. */
int dh_frame_uniforms_canvas(const dh_camera* cam, int32_t screen_w, int32_t screen_h, int32_t canvas_w,
                             int32_t canvas_h, int32_t off_x, int32_t off_y, const float light[3],
                             dh_frame_params* out);

/* ---- context --------------------------------------------------------- */
int dh_ctx_create(int32_t device, dh_ctx* out);
/* A second context (own HIP stream and scratch) on the parent's device that
 * shares the parent's resources: handles created through either are valid in
 * both, so frames can render concurrently from one resident mesh pool.  The
 * resources live until the last context sharing them is destroyed.  It
 * starts with the parent's raster kernel, tiling, bin order and graph
 * settings.  No counterpart in the reference (one canvas, one thread). */
int dh_ctx_create_shared(dh_ctx parent, dh_ctx* out);
int dh_ctx_destroy(dh_ctx ctx);
/* Human-readable text of the last error on this context (never NULL). */
const char* dh_last_error(dh_ctx ctx);
/* Raster tuning: tile size 8/16/32 px and waves per tile (1..16, divides
 * (tile/8)^2) for raster kernel kind 0, which this selects.  Defaults 16 and 4. */
int dh_set_tiling(dh_ctx ctx, int32_t tile, int32_t waves_per_tile);
/* Raster kernel and its tiling together: kind 0 = waves own 8x8 quadrants of
 * the tile (the dh_set_tiling configurations); kind 1 = triangle-parallel:
 * the waves of a tile take triangles from a shared counter and sweep each
 * one's rectangle (tile 16 or 32, waves 2, 4 or 8).  Output is identical. */
int dh_set_raster_kernel(dh_ctx ctx, int32_t kind, int32_t tile, int32_t waves_per_tile);
/* dh_render captures its launch sequence in a hipGraph after a frame and
 * replays it while the next frames keep the same structure (meshes, canvas,
 * sizes, tiling); uniforms may change.  1 (default) enables this, 0 launches
 * every kernel eagerly (per-stage profiling times need eager frames). */
int dh_set_graph(dh_ctx ctx, int32_t enable);
/* Bin order of the per-tile triangle lists.  2 (default): near first -- the
 * clipped triangles are ordered by their nearest vertex depth (11-bit key,
 * stable) before binning, so that the raster's hierarchical depth test
 * rejects most occluded triangles; 1: submission order (stable radix sort
 * by tile); 0: built with atomics, in no particular order.  The frame is the
 * same in every mode: each pixel keeps the minimum of (depth, submission
 * index).  The transparent pass always bins in submission order. */
int dh_set_bin_order(dh_ctx ctx, int32_t mode);

/* ---- resources (Scene::add_obj, scene/mod.rs:788-801; Object::new,
 *      object.rs:34-71; TextureMap::new, scene/mod.rs:113-126) ---------- */
/* img: h rows of w pixels, row 0 = top, comp = 3 or 4 bytes per pixel. */
int dh_texture_create(dh_ctx ctx, const uint8_t* img, int32_t w, int32_t h, int32_t comp,
                      dh_handle* out);
int dh_material_create(dh_ctx ctx, const dh_material_desc* desc, dh_handle* out);
/* Object vertex arrays: pos[3*n_pos], nrm[3*n_nrm], uv[2*n_uv]. */
int dh_object_create(dh_ctx ctx, const float* pos, uint32_t n_pos, const float* nrm, uint32_t n_nrm,
                     const float* uv, uint32_t n_uv, dh_handle* out);
/* IndexedMesh of an object (scene/mesh.rs:30-35): tri[9*n_tri] = pos a,b,c,
 * uv a,b,c, normal a,b,c.  Indices are validated here (DH_E_INDEX).
 * material 0 = Texture::default (scene/mod.rs:1072-1075). */
int dh_mesh_create(dh_ctx ctx, dh_handle object, const uint32_t* tri, uint32_t n_tri, dh_handle material,
                   dh_handle* out);
/* Copy a mesh's triangle triplets back (9 * n_tri u32).  The transparent
 * pass sorts a transparent mesh's triangles in place every frame, as the
 * reference's sort_by on the mesh's triangle vector does (scene/mod.rs:1100),
 * so a caller that mirrors its meshes on the host reads the new order here. */
int dh_mesh_read(dh_ctx ctx, dh_handle mesh, uint32_t* tri);
/* Device-resident Canvas (canvas.rs:353-433): BGRA8 frame + f32 depth +
 * u32 winner id per pixel. */
int dh_canvas_create(dh_ctx ctx, int32_t W, int32_t H, dh_handle* out);
/* The same with flags.  DH_CANVAS_NO_WINNER: no winner buffer, as the
 * reference's Canvas (canvas.rs:353-433) has none; frames render and shade
 * the same, without the 4 B per pixel of winner writes, and
 * dh_canvas_read / dh_canvas_device_ptrs report no winners (DH_E_ARG if a
 * winner array is asked of dh_canvas_read).  Other flags: DH_E_ARG. */
enum { DH_CANVAS_NO_WINNER = 1u };
int dh_canvas_create_ex(dh_ctx ctx, int32_t W, int32_t H, uint32_t flags, dh_handle* out);
int dh_release(dh_ctx ctx, dh_handle h);

/* ---- rendering -------------------------------------------------------- */
/* Scene::render (scene/mod.rs:901-1249) of the given meshes, in submission
 * order, into a device canvas; returns when the frame is complete on the
 * device.  Consecutive meshes of one object form an object group: as the
 * reference's object loop does, a group's opaque meshes are drawn first, in
 * list order, then its transparent meshes (material alpha < 1), each sorted
 * back to front in place (scene/mod.rs:1088-1247), without back-face culling
 * and without depth writes. */
int dh_render(dh_ctx ctx, const dh_frame_params* fr, const dh_handle* meshes, uint32_t n_meshes,
              dh_handle canvas);
/* One view of a batch: uniforms, target canvas (same size as fr), meshes in
 * submission order. */
typedef struct {
    dh_frame_params fr;
    dh_handle canvas;
    const dh_handle* meshes;
    uint32_t n_meshes;
    uint32_t pad;
} dh_view;

enum { DH_BATCH_GRAPH = 1u };

/* Batched Scene::render: n_views independent frames (each exactly what
 * dh_render would produce for it alone, winner ids included) through one
 * pipeline pass -- every kernel launches once over all views' triangles, tiles
 * and pixels.  DH_BATCH_GRAPH captures that launch sequence in a hipGraph and
 * replays it while the batch keeps the same structure (views, sizes, canvases,
 * meshes); uniforms may change between replays.  No counterpart in the
 * reference, which renders one frame per call (app/mod.rs:96-235).
 * Transparent meshes are not accepted here (DH_E_UNSUPPORTED): their in-place
 * sort is a per-frame sequence. */
int dh_render_batch(dh_ctx ctx, const dh_view* views, uint32_t n_views, uint32_t flags);

/* ---- sort-last shards: one frame split by submission range over GPUs ----
 * No counterpart in the reference (its render is one sequential loop); this
 * is the per-GPU part of a multi-GPU frame.  A caller splits the frame's mesh
 * list into consecutive submission ranges, one per GPU (shard < 8), and:
 *  1. dh_render_shard renders its range into vis_dev, W*H u64 on its device:
 *     (depth key << 32 | shard << 29 | clipped-triangle id), ~0 where no
 *     triangle covers the pixel (at most 2^29 clipped triangles per shard);
 *  2. the callers reduce the shards' buffers with an element-wise u64
 *     minimum (an RCCL reduce-scatter or all-reduce with ncclMin on
 *     ncclUint64; dh_vis_merge_min does the same on one device) -- the result
 *     is the single-GPU winner of every pixel, submission-order ties included;
 *  3. dh_shade_shard, on the context that rendered the shard and with no
 *     other frame in between, shades the pixels whose winner is its own into
 *     the canvas given to dh_render_shard (shard 0 also writes the clear
 *     colour of pixels without a winner).  The winner ids it writes are the
 *     shard's clipped-triangle ids.
 * Opaque meshes only (alpha < 1: DH_E_UNSUPPORTED). */
int dh_render_shard(dh_ctx ctx, const dh_frame_params* fr, const dh_handle* meshes, uint32_t n_meshes,
                    uint32_t shard, dh_handle canvas, uint64_t* vis_dev);
int dh_vis_merge_min(dh_ctx ctx, uint64_t* acc_dev, const uint64_t* src_dev, uint64_t n);
int dh_shade_shard(dh_ctx ctx, const uint64_t* merged_dev, uint32_t shard);

/* Device memory for callers without their own allocator (visibility
 * buffers above): hipMalloc/hipFree on the context's device, and a copy to
 * host memory in the context's stream order. */
int dh_device_alloc(dh_ctx ctx, uint64_t bytes, void** out_dev);
int dh_device_free(dh_ctx ctx, void* dev);
int dh_device_read(dh_ctx ctx, void* host, const void* dev, uint64_t bytes);

/* Copy a canvas to host memory: bgra W*H*4 bytes, rows top-down
 * (Canvas::as_bytes_slice, canvas.rs:974-978); depth W*H floats, rows
 * bottom-up (canvas.rs:415); winner W*H ids, rows bottom-up, 0xFFFFFFFF =
 * clear.  Any pointer may be NULL. */
int dh_canvas_read(dh_ctx ctx, dh_handle canvas, uint8_t* bgra, float* depth, uint32_t* winner);
/* Frame export (App::export_frame_as, app/mod.rs:316-379): the canvas frame
 * with bytes 0 and 2 of every pixel swapped (BGRA -> RGBA, app/mod.rs:347-351)
 * into rgba, W*H*4 bytes, rows top-down; the swap runs on the device. */
int dh_canvas_read_rgba(dh_ctx ctx, dh_handle canvas, uint8_t* rgba);
/* dh_canvas_read_rgba without allocating: no counterpart in the reference.
 * For callers that cannot allocate device memory (tight serving loops, code
 * that must not call hipMalloc).  The BGRA -> RGBA staging is a fixed array
 * of 3840*2160 pixels (31.6 MiB) that every GPU loading the library holds
 * until the process ends; a larger canvas returns DH_E_CAPACITY.  Calls on
 * one GPU from one process share that array and take turns on it (a lock);
 * each process has its own copy.  Same output as dh_canvas_read_rgba. */
int dh_canvas_read_rgba_noalloc(dh_ctx ctx, dh_handle canvas, uint8_t* rgba);
/* That RGBA frame written as an 8-bit RGBA PNG file (the reference writes it
 * with stb_image_write; decoded pixels are the same, the compressed bytes are
 * not). */
int dh_canvas_export_png(dh_ctx ctx, dh_handle canvas, const char* path);
/* Host-only PNG writer behind dh_canvas_export_png: W*H RGBA8 pixels, rows
 * top-down. */
int dh_write_png(const char* path, const uint8_t* rgba, int32_t W, int32_t H);
/* The same RGBA frame as a baseline JPEG file (app/mod.rs:316-379 with the
 * JPEG format).  quality as stbi_write_jpg takes it: 0 means 90, anything
 * else is clamped to 1..100; the reference passes width*4, i.e. 100.  A
 * baseline sequential encoder (ITU-T T.81, JFIF): YCbCr without chroma
 * subsampling, the Annex K quantisation tables scaled by quality (every
 * step 1 at 100) and Annex K Huffman tables.  The compressed bytes are not
 * stb's. */
int dh_canvas_export_jpg(dh_ctx ctx, dh_handle canvas, const char* path, int32_t quality);
/* Host-only JPEG writer behind dh_canvas_export_jpg: W*H RGBA8 pixels (alpha
 * ignored), rows top-down, W and H at most 65535. */
int dh_write_jpg(const char* path, const uint8_t* rgba, int32_t W, int32_t H, int32_t quality);
/* Device pointers of a canvas (for zero-copy consumers on the same device). */
int dh_canvas_device_ptrs(dh_ctx ctx, dh_handle canvas, void** bgra, void** depth, void** winner);
int dh_get_stats(dh_ctx ctx, dh_stats* out);
/* Enable (1) or disable (0) per-stage HIP event timing of dh_render. */
int dh_set_profiling(dh_ctx ctx, int32_t enable);

/* ---- GUI overlay ------------------------------------------------------ */
/* Gui::render (app/gui.rs:382-478) over Canvas::draw_triangle
 * (canvas.rs:435-575): an imgui draw list drawn onto a canvas after the
 * scene.  Vertices as ImDrawVert (screen position with y down, uv, RGBA
 * colour), u32 indices, commands (clip rect x0,y0,x1,y1 with y down, index
 * offset and count, vertex offset) applied in order; font: an RGBA texture
 * handle sampled with get_rgba_slice.  Fragments blend over the canvas, are
 * depth-tested against 0 and write no depth.  DH_E_INDEX for indices past
 * the buffers, DH_E_TEXCOORD when a sample leaves the font image (the
 * reference panics on both). */
typedef struct { float pos[2]; float uv[2]; uint8_t col[4]; } dh_gui_vertex;
typedef struct { float clip_rect[4]; uint32_t idx_offset, count, vtx_offset, pad; } dh_gui_cmd;
int dh_draw_gui(dh_ctx ctx, dh_handle canvas, const dh_gui_vertex* vtx, uint32_t n_vtx, const uint32_t* idx,
                uint32_t n_idx, const dh_gui_cmd* cmds, uint32_t n_cmds, dh_handle font);

/* ---- texture files ---------------------------------------------------- */
/* TextureMap::load_from_file (scene/mod.rs:174-202), whose stb_image calls
 * are stbi_info_from_reader then stbi_load_from_reader: a PNG or JPEG file's
 * size and component count (3 or 4; every colour JPEG reports 3), then its
 * 8-bit pixels, rows top-down, into out (cap bytes, at least W*H*comp, else
 * DH_E_CAPACITY).  PNG: 8/16-bit RGB, RGBA and palette images, not
 * interlaced.  JPEG: baseline and progressive Huffman, 8-bit, YCbCr, RGB,
 * CMYK or YCCK, any sampling factors that divide the largest.  Grey images
 * (the reference reaches unreachable!()) and other forms return
 * DH_E_UNSUPPORTED; corrupt files DH_E_ARG.  On failure err (if given)
 * receives a message.  The pixels go to dh_texture_create. */
int dh_image_info(const char* path, int32_t* W, int32_t* H, int32_t* comp, char* err, uint32_t err_len);
int dh_image_load(const char* path, uint8_t* out, uint64_t cap, int32_t* W, int32_t* H, int32_t* comp,
                  char* err, uint32_t err_len);

/* ---- OBJ / MTL ingest ------------------------------------------------- */
/* Object::load_from_file (object.rs:106-419) on the host: OBJ + MTL parsing
 * (the reference uses the `obj` crate), PNG and JPEG textures decoded as
 * dh_image_load decodes them, normals normalised,
 * positions rescaled to a largest norm of 100, quads split (a,b,c)+(c,d,a),
 * a dummy texture coordinate per group lacking them, generated normals per
 * group lacking them.  Textures: index 0 is Texture::default, then every
 * material of every library in file order; a mesh's texture is the first
 * with its group's material name. */
typedef struct dh_obj_s* dh_obj;

typedef struct {
    uint32_t n_pos, n_nrm, n_uv, n_meshes, n_textures;
    uint32_t pad;
    const float* pos;   /* xyz per vertex, after rescaling    */
    const float* nrm;   /* xyz per normal                     */
    const float* uv;    /* uv per texture coordinate          */
} dh_obj_summary;

typedef struct {
    const uint32_t* tri;  /* 9 per triangle: pos a,b,c; uv a,b,c; nrm a,b,c */
    uint32_t n_tri;
    uint32_t texture_idx;
    const char* name;     /* the OBJ group name                                */
} dh_obj_mesh_desc;

typedef struct {
    float ka[3], kd[3], ks[3];
    float alpha;               /* MTL d (default 1)                           */
    const char* name;
    const uint8_t* map_kd;     /* NULL = TextureMap::default                  */
    int32_t kd_w, kd_h, kd_comp;
    const uint8_t* map_ka;
    int32_t ka_w, ka_h, ka_comp;
} dh_obj_texture_desc;

/* Parse path (and the material libraries and textures it names, relative to
 * its directory).  On failure err (if given) receives a message. */
int dh_obj_load(const char* path, dh_obj* out, char* err, uint32_t err_len);
int dh_obj_info(dh_obj obj, dh_obj_summary* out);
int dh_obj_mesh(dh_obj obj, uint32_t i, dh_obj_mesh_desc* out);
int dh_obj_texture(dh_obj obj, uint32_t i, dh_obj_texture_desc* out);
/* Upload into ctx's resident pool: one object, its textures and materials,
 * and one mesh per OBJ group (meshes[cap], in order). */
int dh_obj_upload(dh_ctx ctx, dh_obj obj, dh_handle* object, dh_handle* meshes, uint32_t cap);
int dh_obj_free(dh_obj obj);

/* ---- serving ---------------------------------------------------------- */
/* A render server with continuous batching.  No counterpart in the reference,
 * whose app loop renders one frame per iteration (app/mod.rs:96-235); each job
 * is one Scene::render of the given meshes and uniforms, and its result is
 * exactly what dh_render produces for it alone.
 *
 * Worker threads, each with its own stream and scratch (contexts sharing the
 * server context's resources), take queued jobs in arrival order and render
 * as many as fit in one dh_render_batch launch.  A job is admitted only when
 * its canvas and the scratch of the launch fit the HBM budget together with
 * every canvas and scratch buffer already held; otherwise it waits until
 * memory is released.  Jobs that arrive while launches are on the device join
 * the next launch of whichever worker frees up first, so launches of
 * different workers overlap on the device. */
typedef struct dh_server_s* dh_server;

typedef struct {
    uint64_t budget_bytes;  /* HBM cap for job canvases + worker scratch; 0 = no cap   */
    uint32_t workers;       /* concurrent launches (streams), >= 1; 0 = 2               */
    uint32_t max_views;     /* jobs per launch; 0 = 64                                  */
    uint64_t max_tris;      /* input triangles per launch; 0 = 16M                      */
} dh_server_config;

enum { DH_JOB_AUTORELEASE = 1u };  /* recycle the canvas as soon as the job is done */

typedef struct {
    dh_frame_params fr;
    const dh_handle* meshes;  /* copied by dh_srv_submit                              */
    uint32_t n_meshes;
    uint32_t flags;           /* DH_JOB_*                                             */
} dh_job;

enum { DH_JOB_QUEUED = 1, DH_JOB_RUNNING = 2, DH_JOB_DONE = 0 };

typedef struct {
    int32_t status;           /* DH_JOB_QUEUED/RUNNING/DONE or a negative DH_E_* code  */
    uint32_t batch_views;     /* jobs in the launch this job ran in                   */
    dh_handle canvas;         /* result canvas (resource of the server context) when
                                 DONE; valid until dh_srv_release                      */
    double t_submit_ms;       /* server clock: ms since dh_srv_create                  */
    double t_admit_ms;
    double t_done_ms;         /* the job's launch completed on the device              */
} dh_job_info;

typedef struct {
    uint64_t submitted, completed, failed, launches;
    uint64_t bytes_in_use;    /* canvases (held and pooled) + worker scratch now        */
    uint64_t bytes_peak;
    uint64_t queued;          /* jobs waiting for admission now                        */
} dh_server_stats;

/* ctx: the context whose resources (meshes, textures, materials) jobs use; it
 * must outlive the server.  Its tiling applies to every worker. */
int dh_srv_create(dh_ctx ctx, const dh_server_config* cfg, dh_server* out);
/* Queue a job.  DH_E_BUDGET if its canvas plus the scratch of a launch of it
 * alone exceed the budget; DH_E_HANDLE for an unknown mesh. */
int dh_srv_submit(dh_server srv, const dh_job* job, uint64_t* ticket);
/* State of a job; wait != 0 blocks until it is done or failed. */
int dh_srv_poll(dh_server srv, uint64_t ticket, int32_t wait, dh_job_info* out);
/* Forget a finished job and return its canvas to the server's pool. */
int dh_srv_release(dh_server srv, uint64_t ticket);
int dh_srv_stats(dh_server srv, dh_server_stats* out);
/* Finish every queued and running job, then stop the workers. */
int dh_srv_destroy(dh_server srv);

#ifdef __cplusplus
}
#endif

#endif
