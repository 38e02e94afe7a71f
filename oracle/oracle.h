/* oracle.h -- CPU restatement of mororo18/draw's per-frame raster path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path in draw_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in draw_amd/ links or calls it.
 *
 * Every function follows the reference's IEEE-754 binary32 operation order
 * (no FMA contraction, correctly rounded / and sqrt, Rust release-mode
 * integer semantics: saturating float->int casts, wrapping usize arithmetic).
 * Build flags: gcc -O2 -ffp-contract=off -fno-fast-math (x86-64 SSE scalar).
 */
#ifndef DRAW_ORACLE_H
#define DRAW_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OR_OK = 0,
    OR_E_ARG = -1,        /* bad argument (null pointer, zero size)               */
    OR_E_INDEX = -2,      /* vertex index out of range: mesh.rs:46-48 assert!      */
    OR_E_TEXCOORD = -3,   /* texture slice out of range: scene/mod.rs:167 panic    */
    OR_E_CAMERA = -4,     /* Camera::new asserts, scene/mod.rs:327-331             */
    OR_E_PLANE = -5,      /* ViewPlane::new assert, scene/mod.rs:624               */
};

/* Per-frame uniforms (a1).  planes: near, far, right, left, top, bottom;
 * each (normal.x, normal.y, normal.z, k). */
typedef struct {
    float M[16];          /* row-major viewport*orth*persp*view                   */
    float cam[3];
    float light[3];
    float planes[6][4];
    int32_t W, H;
    float off_x, off_y;   /* Canvas::offset (canvas.rs:357)                        */
    float clear_depth;    /* Canvas::depth_max (canvas.rs:362, app/mod.rs:80)      */
    uint8_t clear_bgra[4];
} or_frame;

typedef struct {
    float pos[3], dir[3];
    float fov_x_deg, near_, far_;   /* reference: 135, -10, -510 (scene/mod.rs:308-310) */
} or_camera_desc;

typedef struct {
    float ka[3], kd[3], ks[3];
    float alpha;
    int32_t power;                  /* canvas.rs:724 hard-codes 2                    */
    const uint8_t* map_kd; int32_t kd_w, kd_h, kd_comp;
    const uint8_t* map_ka; int32_t ka_w, ka_h, ka_comp;
} or_material;

typedef struct {
    const uint32_t* tri;            /* 9 per triangle: pos a,b,c; uv a,b,c; nrm a,b,c */
    uint32_t n_tri;
    int32_t material;               /* index into materials; <0 or >= n -> default   */
} or_mesh;

typedef struct {
    const float* pos; uint32_t n_pos;   /* xyz */
    const float* nrm; uint32_t n_nrm;   /* xyz */
    const float* uv;  uint32_t n_uv;    /* uv (the reference's Vec3 z is never read) */
    const or_mesh* meshes; uint32_t n_meshes;
} or_object;

typedef struct {
    uint64_t tris_in;         /* triangles submitted (opaque + transparent)         */
    uint64_t tris_culled;     /* back-face culled                                   */
    uint64_t tris_emitted;    /* clipped triangles handed to the rasterizer (opaque) */
    uint64_t tris_emitted_tr; /* same, transparent meshes                           */
    uint64_t px_tested;       /* bbox pixels evaluated                              */
    uint64_t frags_covered;   /* covered fragments (shaded before the depth test)   */
    uint64_t frags_written;   /* fragments that passed the depth test               */
} or_stats;

/* Vertex attributes as VertexAttributes (canvas.rs:193-203), uv as 2 floats. */
typedef struct {
    float sx, sy;             /* screen_coord */
    float depth;
    float normal[3], light[3], halfway[3];
    float uv[2];
} or_attr;

/* a1: Camera::new + gen_transformation_matrix + gen_view_planes. */
int or_frame_uniforms(const or_camera_desc* cam, int32_t W, int32_t H,
                      const float light[3], or_frame* out);
/* a1 in the reference app's configuration: scene (screen) size W x H for the
 * camera and viewport, a canvas_w x canvas_h canvas at offset (off_x, off_y)
 * (app/mod.rs:59-63, 81-84). */
int or_frame_uniforms_canvas(const or_camera_desc* cam, int32_t screen_w, int32_t screen_h, int32_t canvas_w,
                             int32_t canvas_h, int32_t off_x, int32_t off_y, const float light[3],
                             or_frame* out);

/* a2: per-vertex visual info (scene/mod.rs:920-925): out = light[3], eye[3],
 * halfway[3], depth. */
void or_vertex_visual(const float v[3], const float cam[3], const float light[3],
                      float out[10]);

/* a4: Triangle::clip_against_planes.  in: 3 world positions (9 floats) + 3
 * attrs.  out: up to 12 triangles (positions 9 floats each, attrs 3 each).
 * Returns the count. */
int or_clip(const or_frame* fr, const float pos[9], const or_attr attr[3],
            float out_pos[12 * 9], or_attr out_attr[12 * 3]);

/* Canvas-level entry: draw_triangle_with_attributes onto caller buffers.
 * bgra: W*H*4 top-down; depth: W*H bottom-up; winner: W*H (may be NULL).
 * depth_update: Canvas::depth_update_enabled.  Returns OR_OK or an error. */
int or_draw_triangle(const or_frame* fr, const or_attr* a, const or_attr* b,
                     const or_attr* c, const or_material* mat, int depth_update,
                     uint32_t id, uint8_t* bgra, float* depth, uint32_t* winner,
                     or_stats* st);

/* GUI overlay (f4): one imgui-style draw list through Gui::render
 * (app/gui.rs:382-478) and Canvas::draw_triangle (canvas.rs:435-575):
 * vertices (pos, uv, RGBA col), u32 indices, commands (clip rect x0,y0,x1,y1
 * in imgui's top-down convention, idx_offset, count, vtx_offset), an RGBA
 * font texture (get_rgba_slice).  Depth is tested (`0 < depth`) but never
 * written.  Returns OR_OK, OR_E_INDEX or OR_E_TEXCOORD (a panic). */
typedef struct { float pos[2], uv[2]; uint8_t col[4]; } or_gui_vertex;
typedef struct { float clip_rect[4]; uint32_t idx_offset, count, vtx_offset, pad; } or_gui_cmd;
int or_gui_draw(int32_t W, int32_t H, const or_gui_vertex* vtx, uint32_t n_vtx, const uint32_t* idx,
                uint32_t n_idx, const or_gui_cmd* cmds, uint32_t n_cmds, const uint8_t* font, int32_t fw,
                int32_t fh, int32_t fcomp, uint8_t* bgra, const float* depth);

/* Canvas::clear (canvas.rs:425-433). */
void or_clear(const or_frame* fr, uint8_t* bgra, float* depth, uint32_t* winner);

/* Scene::render (scene/mod.rs:901-1249): clear, opaque pass, transparent pass.
 * winner: id of the opaque clipped triangle that last wrote each pixel
 * (0xFFFFFFFF = none), ids counted in emission order over the whole frame. */
int or_render(const or_frame* fr, const or_object* objs, uint32_t n_obj,
              const or_material* mats, uint32_t n_mat,
              uint8_t* bgra, float* depth, uint32_t* winner, or_stats* st);

/* Same as or_render but only draws opaque triangles whose input index (over the
 * frame's opaque submission stream) lies in [tri_begin, tri_end); used to time a
 * bounded sample of a large workload. */
int or_render_range(const or_frame* fr, const or_object* objs, uint32_t n_obj,
                    const or_material* mats, uint32_t n_mat,
                    uint64_t tri_begin, uint64_t tri_end,
                    uint8_t* bgra, float* depth, uint32_t* winner, or_stats* st);

/* or_render restricted to the pixels with x % sx == ox and y % sy == oy
 * (other pixels keep the clear values).  Exact on the sampled pixels. */
int or_render_grid(const or_frame* fr, const or_object* objs, uint32_t n_obj,
                   const or_material* mats, uint32_t n_mat,
                   uint32_t sx, uint32_t sy, uint32_t ox, uint32_t oy,
                   uint8_t* bgra, float* depth, uint32_t* winner, or_stats* st);

#ifdef __cplusplus
}
#endif
#endif
