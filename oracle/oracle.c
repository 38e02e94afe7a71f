/* oracle.c -- CPU restatement of mororo18/draw's per-frame raster path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker and the CPU baseline); see
 * oracle.h.  Each function cites the reference file:line it restates.  The
 * reference is Rust; all arithmetic below is IEEE binary32 in the same order
 * as the Rust source (rustc never contracts a*b+c into an FMA, and `a+b+c`
 * is (a+b)+c), so this file must be compiled with -ffp-contract=off.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_EPS 0.0000001f /* linalg.rs:6 */

typedef struct { float x, y; } v2;
typedef struct { float x, y, z; } v3;
typedef struct { float m[4][4]; } m4;

/* ---------------------------------------------------------------- linalg.rs */
static v3 mk3(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
/* Vec3 Add/Sub/Mul<f32>/Div<f32>: linalg.rs:213-259 (component-wise). */
static v3 add3(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 sub3(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 mul3(v3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
static v3 div3(v3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
/* Vec3::norm linalg.rs:167-171: powi(2) is x*x; sum is ((x2+y2)+z2). */
static float norm3(v3 a) {
    float s = a.x * a.x + a.y * a.y;
    s = s + a.z * a.z;
    return sqrtf(s);
}
/* Vec3::normalized linalg.rs:173-175 */
static v3 normalized3(v3 a) { return div3(a, norm3(a)); }
/* Vec3::dot linalg.rs:182-184 */
static float dot3(v3 a, v3 b) {
    float s = a.x * b.x + a.y * b.y;
    return s + a.z * b.z;
}
/* Vec3::cross linalg.rs:186-200 */
static v3 cross3(v3 a, v3 b) {
    return mk3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
/* Matrix4 * Matrix4 linalg.rs:328-344: c[i][j] += a[i][k]*b[k][j], from 0. */
static m4 mulm(m4 a, m4 b) {
    m4 c;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float s = 0.0f;
            for (int k = 0; k < 4; k++) s = s + a.m[i][k] * b.m[k][j];
            c.m[i][j] = s;
        }
    return c;
}
/* Matrix4 * Vec4 linalg.rs:346-360 */
static void mulv(const m4* a, const float v[4], float out[4]) {
    for (int i = 0; i < 4; i++) {
        float s = 0.0f;
        for (int j = 0; j < 4; j++) s = s + a->m[i][j] * v[j];
        out[i] = s;
    }
}
/* Matrix4::transposed linalg.rs:275-287 */
static m4 transposed(m4 a) {
    m4 c = a;
    for (int i = 0; i < 4; i++)
        for (int j = i + 1; j < 4; j++) { float t = c.m[i][j]; c.m[i][j] = c.m[j][i]; c.m[j][i] = t; }
    return c;
}
static m4 mkm(const float* r) {
    m4 a;
    memcpy(a.m, r, sizeof(a.m));
    return a;
}

/* Rust `f as usize` (64-bit): NaN -> 0, saturating, truncating toward zero. */
static uint64_t f2usize(float v) {
    if (!(v > 0.0f)) return 0; /* NaN, negative, -0, +0 */
    if (v >= 18446744073709551616.0f) return UINT64_MAX;
    return (uint64_t)v;
}
/* Rust `f as u8`: NaN -> 0, saturating, truncating. */
static uint8_t f2u8(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 255.0f) return 255;
    return (uint8_t)v;
}
/* f32::powi with a non-negative exponent: binary exponentiation, low bit first
 * (LLVM's constant-exponent expansion and compiler-rt __powisf2 both multiply
 * in this order; power 2 is x*x as at canvas.rs:739). */
static float powi_f(float x, int p) {
    float r = 1.0f;
    int first = 1;
    while (p) {
        if (p & 1) { r = first ? x : r * x; first = 0; }
        p >>= 1;
        if (p) x = x * x;
    }
    return r;
}

/* ------------------------------------------------------------ camera (a1) */
typedef struct {
    v3 position, direction, up, u, v, w;
    float top, bottom, right, left, near_, far_;
} camera;

/* Camera::update_basis scene/mod.rs:438-451 */
static void update_basis(camera* c) {
    v3 g = c->direction;
    v3 w = mul3(div3(g, norm3(g)), -1.0f);
    v3 t_x_w = cross3(c->up, w);
    v3 u = div3(t_x_w, norm3(t_x_w));
    v3 v = cross3(w, u);
    c->u = normalized3(u);
    c->v = normalized3(v);
    c->w = normalized3(w);
}
/* Camera::get_basis_matrix scene/mod.rs:415-428 */
static m4 basis_matrix(const camera* c) {
    float r[16] = {c->u.x, c->v.x, c->w.x, 0.0f, c->u.y, c->v.y, c->w.y, 0.0f,
                   c->u.z, c->v.z, c->w.z, 0.0f, 0.0f,   0.0f,   0.0f,   1.0f};
    return mkm(r);
}
/* Camera::gen_matrix scene/mod.rs:453-479 */
static m4 gen_matrix(camera* c) {
    v3 p = c->position;
    float r[16] = {1, 0, 0, -p.x, 0, 1, 0, -p.y, 0, 0, 1, -p.z, 0, 0, 0, 1};
    m4 matrix_pos = mkm(r);
    update_basis(c);
    m4 bt = transposed(basis_matrix(c));
    return mulm(bt, matrix_pos);
}
/* (M*p.as_vec4()).vec3_over_w() + camera_pos, scene/mod.rs:504-527 */
static v3 cam_to_world(const m4* mb, v3 p, v3 campos) {
    float in[4] = {p.x, p.y, p.z, 1.0f}, o[4];
    mulv(mb, in, o);
    v3 q = mk3(o[0] / o[3], o[1] / o[3], o[2] / o[3]); /* Vec4::vec3_over_w linalg.rs:89-91 */
    return add3(q, campos);
}
/* ViewPlane::new scene/mod.rs:603-632 */
static int view_plane(v3 a, v3 b, v3 c, v3 visible, float out[4]) {
    v3 p = sub3(b, a), q = sub3(c, b);
    v3 n = cross3(p, q);
    float k = -dot3(n, a);
    float test = dot3(n, visible) + k;
    if (test < 0.0f) {
        n = cross3(q, p);
        k = -dot3(n, a);
        test = dot3(n, visible) + k;
        if (!(test > 0.0f)) return OR_E_PLANE;
    }
    out[0] = n.x; out[1] = n.y; out[2] = n.z; out[3] = k;
    return OR_OK;
}

/* Scene::new(W, H) (scene/mod.rs:760-786) drawn into Canvas::new(cw, ch)
 * (canvas.rs:366) with Canvas::apply_offset(ox, oy) (canvas.rs:382-385): the
 * aspect ratio and viewport use the scene (screen) size, the raster the
 * canvas size and offset (canvas.rs:585-587, 647). */
static int frame_uniforms(const or_camera_desc* cd, int32_t W, int32_t H, int32_t cw, int32_t ch,
                          int32_t ox, int32_t oy, const float light[3], or_frame* out) {
    if (!cd || !out || W <= 0 || H <= 0 || cw <= 0 || ch <= 0) return OR_E_ARG;
    memset(out, 0, sizeof(*out));
    camera c;
    memset(&c, 0, sizeof(c));
    /* Camera::new scene/mod.rs:297-357 (near/far/fov_x are parameters here) */
    float ratio = (float)W / (float)H; /* Scene::new scene/mod.rs:768 */
    float near_ = cd->near_, far_ = cd->far_, fov_x = cd->fov_x_deg;
    const float pi = 3.14159265358979323846f;
    float fov_x_rad = fov_x * (pi / 180.0f); /* f32::to_radians */
    float right = fabsf(near_) * tanf(fov_x_rad / 2.0f);
    float left = -right;
    float top = (1.0f / ratio) * right; /* ratio.recip() * right */
    float bottom = -top;
    if (!(fov_x < 180.0f) || !(near_ < 0.0f) || !(near_ > far_) || !(right > left) || !(top > bottom))
        return OR_E_CAMERA;
    c.position = mk3(cd->pos[0], cd->pos[1], cd->pos[2]);
    c.direction = normalized3(mk3(cd->dir[0], cd->dir[1], cd->dir[2]));
    c.up = mk3(0.0f, 1.0f, 0.0f);
    c.top = top; c.bottom = bottom; c.right = right; c.left = left;
    c.near_ = near_; c.far_ = far_;

    /* Scene::gen_transformation_matrix scene/mod.rs:817-899 */
    float n_x = (float)W, n_y = (float)H;
    float n = c.near_, f = c.far_, r = c.right, l = c.left, t = c.top, b = c.bottom;
    m4 matrix_cam = gen_matrix(&c);
    float pr[16] = {n, 0, 0, 0, 0, n, 0, 0, 0, 0, (n + f), -(n * f), 0, 0, 1, 0};
    float orr[16] = {2.0f / (r - l), 0, 0, -(r + l) / (r - l),
                     0, 2.0f / (t - b), 0, -(t + b) / (t - b),
                     0, 0, 2.0f / (n - f), -(n + f) / (n - f),
                     0, 0, 0, 1};
    float vp[16] = {n_x / 2.0f, 0, 0, (n_x - 1.0f) / 2.0f, 0, n_y / 2.0f, 0, (n_y - 1.0f) / 2.0f,
                    0, 0, 1, 0, 0, 0, 0, 1};
    m4 M = mulm(mulm(mulm(mkm(vp), mkm(orr)), mkm(pr)), matrix_cam);
    memcpy(out->M, M.m, sizeof(out->M));

    /* Camera::gen_view_planes scene/mod.rs:481-593 */
    update_basis(&c);
    m4 mb = basis_matrix(&c);
    v3 cp = c.position;
    v3 urn = cam_to_world(&mb, mk3(r, t, n), cp);
    v3 uln = cam_to_world(&mb, mk3(l, t, n), cp);
    v3 lrn = cam_to_world(&mb, mk3(r, b, n), cp);
    v3 lln = cam_to_world(&mb, mk3(l, b, n), cp);
    float xc = (l + r) / 2.0f, yc = (b + t) / 2.0f;
    v3 ucf = cam_to_world(&mb, mk3(xc, (f * t) / n, f), cp);
    v3 lcf = cam_to_world(&mb, mk3(xc, (f * b) / n, f), cp);
    v3 rcf = cam_to_world(&mb, mk3((f * r) / n, yc, f), cp);
    v3 lfc = cam_to_world(&mb, mk3((f * l) / n, yc, f), cp);
    v3 visible = div3(add3(urn, lcf), 2.0f);
    int e = 0;
    e |= view_plane(urn, lrn, lln, visible, out->planes[0]); /* near   */
    e |= view_plane(lfc, rcf, ucf, visible, out->planes[1]); /* far    */
    e |= view_plane(rcf, urn, lrn, visible, out->planes[2]); /* right  */
    e |= view_plane(lfc, lln, uln, visible, out->planes[3]); /* left   */
    e |= view_plane(ucf, uln, urn, visible, out->planes[4]); /* top    */
    e |= view_plane(lcf, lln, lrn, visible, out->planes[5]); /* bottom */
    if (e) return OR_E_PLANE;

    out->cam[0] = cp.x; out->cam[1] = cp.y; out->cam[2] = cp.z;
    if (light) { out->light[0] = light[0]; out->light[1] = light[1]; out->light[2] = light[2]; }
    out->W = cw; out->H = ch;
    out->off_x = (float)ox; out->off_y = (float)oy;   /* apply_offset: x as f32 */
    out->clear_depth = 100000.0f;                /* app/mod.rs:80            */
    out->clear_bgra[0] = 255; out->clear_bgra[1] = 186; /* Pixel::azul_bb canvas.rs:131-133 */
    out->clear_bgra[2] = 155; out->clear_bgra[3] = 255;
    return OR_OK;
}

int or_frame_uniforms(const or_camera_desc* cd, int32_t W, int32_t H, const float light[3],
                      or_frame* out) {
    return frame_uniforms(cd, W, H, W, H, 0, 0, light, out);
}

int or_frame_uniforms_canvas(const or_camera_desc* cd, int32_t screen_w, int32_t screen_h, int32_t canvas_w,
                             int32_t canvas_h, int32_t off_x, int32_t off_y, const float light[3],
                             or_frame* out) {
    return frame_uniforms(cd, screen_w, screen_h, canvas_w, canvas_h, off_x, off_y, light, out);
}

/* ---------------------------------------------------------- per-vertex (a2) */
void or_vertex_visual(const float vp[3], const float camp[3], const float lightp[3], float out[10]) {
    /* scene/mod.rs:920-925 */
    v3 v = mk3(vp[0], vp[1], vp[2]), cam = mk3(camp[0], camp[1], camp[2]);
    v3 L = mk3(lightp[0], lightp[1], lightp[2]);
    v3 eye_dir = sub3(v, cam);
    v3 light = normalized3(sub3(v, L));
    v3 eye = normalized3(eye_dir);
    float depth = norm3(eye_dir);
    v3 halfway = normalized3(add3(light, eye));
    out[0] = light.x; out[1] = light.y; out[2] = light.z;
    out[3] = eye.x; out[4] = eye.y; out[5] = eye.z;
    out[6] = halfway.x; out[7] = halfway.y; out[8] = halfway.z;
    out[9] = depth;
}

/* ------------------------------------------------------------ clipping (a4) */
typedef struct { v3 p[3]; or_attr a[3]; } tri_t;

/* VertexAttributes a + (c - a) * t (canvas.rs:242-291), per attribute. */
static or_attr lerp_attr(const or_attr* a, const or_attr* c, float t) {
    or_attr r;
    r.sx = a->sx + ((c->sx + (-a->sx)) * t); /* Vec2 Sub is self + (-rhs) linalg.rs:37-43 */
    r.sy = a->sy + ((c->sy + (-a->sy)) * t);
    r.depth = a->depth + ((c->depth - a->depth) * t);
    for (int i = 0; i < 3; i++) {
        r.normal[i] = a->normal[i] + ((c->normal[i] - a->normal[i]) * t);
        r.light[i] = a->light[i] + ((c->light[i] - a->light[i]) * t);
        r.halfway[i] = a->halfway[i] + ((c->halfway[i] - a->halfway[i]) * t);
    }
    for (int i = 0; i < 2; i++) r.uv[i] = a->uv[i] + ((c->uv[i] - a->uv[i]) * t);
    return r;
}
static float plane_func(const float pl[4], v3 p) { /* ViewPlane::func scene/mod.rs:634-636 */
    return dot3(mk3(pl[0], pl[1], pl[2]), p) + pl[3];
}
/* ViewPlane::at_least_partially_visible scene/mod.rs:641-660 */
static int partially_visible(const float pl[4], const tri_t* t) {
    float fa = plane_func(pl, t->p[0]), fb = plane_func(pl, t->p[1]), fc = plane_func(pl, t->p[2]);
    if (fa > 0.0f && fb > 0.0f && fc > 0.0f) return 1;
    if (fa <= 0.0f && fb <= 0.0f && fc <= 0.0f) return 0;
    return 1;
}
/* ViewPlane::clip scene/mod.rs:662-746 */
static int plane_clip(const float pl[4], const tri_t* tri, tri_t* out) {
    v3 a = tri->p[0], b = tri->p[1], c = tri->p[2];
    or_attr aa = tri->a[0], ba = tri->a[1], ca = tri->a[2];
    float fa = plane_func(pl, a), fb = plane_func(pl, b), fc = plane_func(pl, c);
    if (fa > 0.0f && fb > 0.0f && fc > 0.0f) { out[0] = *tri; return 1; }
    if (fa <= 0.0f && fb <= 0.0f && fc <= 0.0f) return 0;
    if (fa * fc >= 0.0f) { /* (a,b,c) <- (c,a,b) */
        v3 tp = c; c = b; b = a; a = tp;
        or_attr ta = ca; ca = ba; ba = aa; aa = ta;
        float tf = fc; fc = fb; fb = fa; fa = tf;
    } else if (fb * fc >= 0.0f) { /* (a,b,c) <- (b,c,a) */
        v3 tp = a; a = b; b = c; c = tp;
        or_attr ta = aa; aa = ba; ba = ca; ca = ta;
        float tf = fa; fa = fb; fb = fc; fc = tf;
    }
    (void)fa; (void)fb;
    v3 nrm = mk3(pl[0], pl[1], pl[2]);
    float t_a = plane_func(pl, a) / dot3(nrm, sub3(a, c)) - OR_EPS;
    v3 na = add3(a, mul3(sub3(c, a), t_a));
    or_attr naa = lerp_attr(&aa, &ca, t_a);
    float t_b = plane_func(pl, b) / dot3(nrm, sub3(b, c)) - OR_EPS;
    v3 nb = add3(b, mul3(sub3(c, b), t_b));
    or_attr nba = lerp_attr(&ba, &ca, t_b);
    if (fc <= 0.0f) {
        out[0].p[0] = a; out[0].p[1] = na; out[0].p[2] = nb;
        out[0].a[0] = aa; out[0].a[1] = naa; out[0].a[2] = nba;
        out[1].p[0] = a; out[1].p[1] = b; out[1].p[2] = nb;
        out[1].a[0] = aa; out[1].a[1] = ba; out[1].a[2] = nba;
        return 2;
    }
    out[0].p[0] = c; out[0].p[1] = na; out[0].p[2] = nb;
    out[0].a[0] = ca; out[0].a[1] = naa; out[0].a[2] = nba;
    return 1;
}
/* Triangle::clip_against_planes scene/mod.rs:43-90 */
static int clip_against_planes(const or_frame* fr, const tri_t* tri, tri_t pool[12]) {
    int n = 0;
    if (partially_visible(fr->planes[2], tri) && partially_visible(fr->planes[3], tri) &&
        partially_visible(fr->planes[4], tri) && partially_visible(fr->planes[5], tri)) {
        pool[0] = *tri;
        n = 1;
    }
    tri_t tmp[12];
    tri_t* cur = pool;
    tri_t* nxt = tmp;
    for (int p = 0; p < 2; p++) { /* depth planes: near, far */
        int m = 0;
        for (int i = 0; i < n; i++) m += plane_clip(fr->planes[p], &cur[i], nxt + m);
        tri_t* s = cur; cur = nxt; nxt = s;
        n = m;
    }
    if (cur != pool) memcpy(pool, cur, sizeof(tri_t) * (size_t)n);
    return n;
}

int or_clip(const or_frame* fr, const float pos[9], const or_attr attr[3], float out_pos[12 * 9],
            or_attr out_attr[12 * 3]) {
    tri_t t, pool[12];
    for (int i = 0; i < 3; i++) { t.p[i] = mk3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]); t.a[i] = attr[i]; }
    int n = clip_against_planes(fr, &t, pool);
    for (int k = 0; k < n; k++)
        for (int i = 0; i < 3; i++) {
            out_pos[9 * k + 3 * i] = pool[k].p[i].x;
            out_pos[9 * k + 3 * i + 1] = pool[k].p[i].y;
            out_pos[9 * k + 3 * i + 2] = pool[k].p[i].z;
            out_attr[3 * k + i] = pool[k].a[i];
        }
    return n;
}

/* ------------------------------------------------------- texture + shading */
static const uint8_t k_white_rgb[3] = {255, 255, 255}; /* TextureMap::default scene/mod.rs:128-135 */

/* TextureMap::get_rgb_slice scene/mod.rs:154-168 in release semantics:
 * saturating casts, wrapping usize arithmetic, bounds-checked slice. */
static int rgb_slice(const uint8_t* img, int32_t w, int32_t h, int32_t comp, float u, float v,
                     uint8_t out[3]) {
    float fw = (float)w, fh = (float)h;
    uint64_t u_idx = f2usize(floorf(u * (fw - 1.0f)));
    uint64_t v_idx = ((uint64_t)h - 1u) - f2usize(floorf(v * (fh - 1.0f)));
    uint64_t off = (v_idx * (uint64_t)w + u_idx) * (uint64_t)comp;
    uint64_t end = off + 3u;
    uint64_t len = (uint64_t)w * (uint64_t)h * (uint64_t)comp;
    if (off > end || end > len) return OR_E_TEXCOORD;
    out[0] = img[off]; out[1] = img[off + 1]; out[2] = img[off + 2];
    return OR_OK;
}

/* Texture::default scene/mod.rs:237-252 */
static or_material default_material(void) {
    or_material m;
    memset(&m, 0, sizeof(m));
    for (int i = 0; i < 3; i++) { m.ka[i] = 0.9f; m.kd[i] = 0.4f; m.ks[i] = 0.5f; }
    m.alpha = 1.0f;
    m.power = 2;
    m.map_kd = k_white_rgb; m.kd_w = 1; m.kd_h = 1; m.kd_comp = 3;
    m.map_ka = k_white_rgb; m.ka_w = 1; m.ka_h = 1; m.ka_comp = 3;
    return m;
}

/* Shading at one covered pixel, canvas.rs:685-741.  bgra out: B,G,R,255. */
static int shade(const or_attr* a, const or_attr* b, const or_attr* c, const or_material* mat,
                 float al, float be, float ga, uint8_t out[4]) {
    float u = (a->uv[0] * al + b->uv[0] * be) + c->uv[0] * ga;
    float v = (a->uv[1] * al + b->uv[1] * be) + c->uv[1] * ga;
    uint8_t dk[3], ak[3];
    int e = rgb_slice(mat->map_kd, mat->kd_w, mat->kd_h, mat->kd_comp, u, v, dk);
    if (e) return e;
    e = rgb_slice(mat->map_ka, mat->ka_w, mat->ka_h, mat->ka_comp, u, v, ak);
    if (e) return e;
    /* Pixel::normalized_as_vec3 canvas.rs:81-87 */
    v3 diffuse = mk3((float)dk[0] / 255.0f, (float)dk[1] / 255.0f, (float)dk[2] / 255.0f);
    v3 ambient = mk3((float)ak[0] / 255.0f, (float)ak[1] / 255.0f, (float)ak[2] / 255.0f);
    v3 an = mk3(a->normal[0], a->normal[1], a->normal[2]), bn = mk3(b->normal[0], b->normal[1], b->normal[2]),
       cn = mk3(c->normal[0], c->normal[1], c->normal[2]);
    v3 pn = add3(add3(mul3(an, al), mul3(bn, be)), mul3(cn, ga));
    v3 al_ = mk3(a->light[0], a->light[1], a->light[2]), bl = mk3(b->light[0], b->light[1], b->light[2]),
       cl = mk3(c->light[0], c->light[1], c->light[2]);
    v3 pl = add3(add3(mul3(al_, al), mul3(bl, be)), mul3(cl, ga));
    v3 ah = mk3(a->halfway[0], a->halfway[1], a->halfway[2]), bh = mk3(b->halfway[0], b->halfway[1], b->halfway[2]),
       ch = mk3(c->halfway[0], c->halfway[1], c->halfway[2]);
    v3 ph = add3(add3(mul3(ah, al), mul3(bh, be)), mul3(ch, ga));
    v3 c_l = mk3(mat->ks[0], mat->ks[1], mat->ks[2]);
    v3 c_r = mk3(diffuse.x * mat->kd[0], diffuse.y * mat->kd[1], diffuse.z * mat->kd[2]);
    v3 c_a = mk3(ambient.x * mat->ka[0], ambient.y * mat->ka[1], ambient.z * mat->ka[2]);
    float dln = dot3(pl, pn);
    float mx = (dln > 0.0f) ? dln : 0.0f; /* 0.0_f32.max(x): NaN -> 0; +-0 irrelevant below */
    float s = 1.0f - mx;
    v3 inner = add3(c_a, mul3(c_l, s));
    v3 t1 = mk3(c_r.x * inner.x, c_r.y * inner.y, c_r.z * inner.z);
    float h2 = powi_f(dot3(ph, pn), mat->power);
    v3 col = add3(t1, mul3(c_l, h2));
    v3 sc = mul3(col, 255.0f); /* Pixel::from_normalized_vec3 canvas.rs:89-92 */
    out[0] = f2u8(sc.z); out[1] = f2u8(sc.y); out[2] = f2u8(sc.x); out[3] = 255;
    return OR_OK;
}

/* ----------------------------------------------------- rasterizer (a6-a9) */
static float edge_fn(v2 p, v2 q, float x, float y) { /* canvas.rs:597-616 */
    return (p.y - q.y) * x + (q.x - p.x) * y + (p.x * q.y) - (q.x * p.y);
}
static v2 map_center(const or_frame* fr, float sx, float sy) {
    /* pos_map_center(screen_coord - offset), canvas.rs:585-587, 896-904 */
    v2 r;
    r.x = floorf((sx + (-fr->off_x)) + 0.5f);
    r.y = floorf((sy + (-fr->off_y)) + 0.5f);
    return r;
}
static float min3f(float a, float b, float c) { /* canvas.rs:618-627 */
    float r = INFINITY;
    if (a < r) r = a;
    if (b < r) r = b;
    if (c < r) r = c;
    return r;
}
static float max3f(float a, float b, float c) { /* canvas.rs:629-638 */
    float r = -INFINITY;
    if (a > r) r = a;
    if (b > r) r = b;
    if (c > r) r = c;
    return r;
}
/* Rectangle::from_coords + Rectangle::clip, canvas.rs:315-350, 646-658,
 * along one axis: returns the inclusive scan range. */
static void rect_axis(uint64_t lo_in, uint64_t hi_in, uint64_t screen_max, uint64_t* lo, uint64_t* hi) {
    uint64_t a = lo_in < hi_in ? lo_in : hi_in; /* from_coords */
    uint64_t b = lo_in < hi_in ? hi_in : lo_in;
    /* clip(drawable, screen): max of mins (screen min 0), min of maxes */
    uint64_t mn = a, mx = b < screen_max ? b : screen_max;
    if (mn > mx) { mn = 0; mx = 0; }
    /* clip(screen, drawable) leaves [mn, mx] unchanged (it lies inside the screen) */
    *lo = mn; *hi = mx;
}

/* Canvas::draw_pixel_coord_with_depth canvas.rs:906-930 */
static void put_pixel(const or_frame* fr, uint64_t x, uint64_t y, const uint8_t col[4], float opacity,
                      float d, int depth_update, uint32_t id, uint8_t* bgra, float* depth,
                      uint32_t* winner, or_stats* st) {
    uint64_t W = (uint64_t)fr->W, H = (uint64_t)fr->H;
    uint8_t* px = bgra + 4u * (W * (H - y - 1u) + x); /* draw_pixel_coord canvas.rs:955-956 */
    uint8_t nc[4];
    if (opacity < 1.0f) {
        /* bg * (1 - a) + c * a with Pixel Mul<f32>/Add (canvas.rs:136-169); pad 0 */
        float ia = 1.0f - opacity;
        for (int i = 0; i < 3; i++) {
            uint8_t p = f2u8((float)px[i] * ia);
            uint8_t q = f2u8((float)col[i] * opacity);
            nc[i] = (uint8_t)(p + q); /* u8 add wraps in release builds */
        }
        nc[3] = 0;
    } else {
        memcpy(nc, col, 4);
    }
    if (d < depth[W * y + x]) {
        memcpy(px, nc, 4);
        if (depth_update) {
            depth[W * y + x] = d;
            if (winner) winner[W * y + x] = id;
        }
        if (st) st->frags_written++;
    }
}

/* Pixel sampling grid: only pixels with x % sx == ox and y % sy == oy are
 * evaluated (1,1,0,0 = every pixel).  Each pixel's result depends only on the
 * fragments at that pixel, so a sampled render is exact on its samples. */
static uint32_t g_sx = 1, g_sy = 1, g_ox = 0, g_oy = 0;

static uint64_t first_on_grid(uint64_t lo, uint32_t step, uint32_t off) {
    uint64_t r = lo - (lo % step) + off;
    return r < lo ? r + step : r;
}

int or_draw_triangle(const or_frame* fr, const or_attr* aa, const or_attr* ba, const or_attr* ca,
                     const or_material* mat, int depth_update, uint32_t id, uint8_t* bgra, float* depth,
                     uint32_t* winner, or_stats* st) {
    /* Canvas::draw_triangle_with_attributes canvas.rs:577-750 */
    v2 a = map_center(fr, aa->sx, aa->sy), b = map_center(fr, ba->sx, ba->sy), c = map_center(fr, ca->sx, ca->sy);
    float ad = aa->depth, bd = ba->depth, cd = ca->depth;
    uint64_t x0 = f2usize(min3f(a.x, b.x, c.x)), y0 = f2usize(min3f(a.y, b.y, c.y));
    uint64_t x1 = f2usize(max3f(a.x, b.x, c.x)), y1 = f2usize(max3f(a.y, b.y, c.y));
    uint64_t xl, xh, yl, yh;
    rect_axis(x0, x1, (uint64_t)fr->W - 1u, &xl, &xh);
    rect_axis(y0, y1, (uint64_t)fr->H - 1u, &yl, &yh);
    float f_alpha = edge_fn(b, c, a.x, a.y);
    float f_beta = edge_fn(c, a, b.x, b.y);
    float f_gama = edge_fn(a, b, c.x, c.y);
    float f_alpha_out = edge_fn(b, c, -1.0f, -1.0f);
    float f_beta_out = edge_fn(c, a, -1.0f, -1.0f);
    float f_gama_out = edge_fn(a, b, -1.0f, -1.0f);
    int tie_a = f_alpha * f_alpha_out > 0.0f, tie_b = f_beta * f_beta_out > 0.0f,
        tie_g = f_gama * f_gama_out > 0.0f;
    for (uint64_t y = first_on_grid(yl, g_sy, g_oy); y <= yh; y += g_sy) {
        float yf = (float)y;
        for (uint64_t x = first_on_grid(xl, g_sx, g_ox); x <= xh; x += g_sx) {
            float xf = (float)x;
            if (st) st->px_tested++;
            float al = edge_fn(b, c, xf, yf) / f_alpha;
            float be = edge_fn(c, a, xf, yf) / f_beta;
            float ga = edge_fn(a, b, xf, yf) / f_gama;
            if (al >= 0.0f && be >= 0.0f && ga >= 0.0f) {
                if ((al > 0.0f || tie_a) && (be > 0.0f || tie_b) && (ga > 0.0f || tie_g)) {
                    float d = (al * ad + be * bd) + ga * cd;
                    uint8_t col[4];
                    int e = shade(aa, ba, ca, mat, al, be, ga, col);
                    if (e) return e;
                    if (st) st->frags_covered++;
                    put_pixel(fr, x, y, col, mat->alpha, d, depth_update, id, bgra, depth, winner, st);
                }
            }
        }
    }
    return OR_OK;
}

void or_clear(const or_frame* fr, uint8_t* bgra, float* depth, uint32_t* winner) {
    size_t n = (size_t)fr->W * (size_t)fr->H;
    for (size_t i = 0; i < n; i++) {
        memcpy(bgra + 4 * i, fr->clear_bgra, 4);
        depth[i] = fr->clear_depth;
        if (winner) winner[i] = 0xFFFFFFFFu;
    }
}

/* ------------------------------------------------------ Scene::render */
typedef struct { float light[3], halfway[3], depth; } visual_t;

/* Gather + attribute build for one indexed triangle, scene/mod.rs:938-1014. */
static int assemble(const or_object* o, const visual_t* vis, const uint32_t* t9, tri_t* tri) {
    for (int k = 0; k < 3; k++) {
        uint32_t pi = t9[k], ti = t9[3 + k], ni = t9[6 + k];
        if (pi >= o->n_pos || ti >= o->n_uv || ni >= o->n_nrm) return OR_E_INDEX; /* mesh.rs:46-48 */
        tri->p[k] = mk3(o->pos[3 * pi], o->pos[3 * pi + 1], o->pos[3 * pi + 2]);
        or_attr* a = &tri->a[k];
        a->sx = 0.0f; a->sy = 0.0f;
        a->depth = vis[pi].depth;
        for (int i = 0; i < 3; i++) {
            a->normal[i] = o->nrm[3 * ni + i];
            a->light[i] = vis[pi].light[i];
            a->halfway[i] = vis[pi].halfway[i];
        }
        a->uv[0] = o->uv[2 * ti];
        a->uv[1] = o->uv[2 * ti + 1];
    }
    return OR_OK;
}

/* Project a clipped triangle and draw it, scene/mod.rs:1047-1083. */
static int project_draw(const or_frame* fr, tri_t* ct, const or_material* mat, int depth_update, uint32_t id,
                        uint8_t* bgra, float* depth, uint32_t* winner, or_stats* st) {
    const m4* M = (const m4*)fr->M;
    for (int k = 0; k < 3; k++) {
        float in[4] = {ct->p[k].x, ct->p[k].y, ct->p[k].z, 1.0f}, o[4];
        mulv(M, in, o);
        ct->a[k].sx = o[0] / o[3];
        ct->a[k].sy = o[1] / o[3];
    }
    return or_draw_triangle(fr, &ct->a[0], &ct->a[1], &ct->a[2], mat, depth_update, id, bgra, depth, winner, st);
}

static const or_material* pick_material(const or_material* mats, uint32_t n_mat, int32_t idx,
                                        const or_material* dflt) {
    if (idx < 0 || (uint32_t)idx >= n_mat) return dflt; /* scene/mod.rs:1072-1075 */
    return &mats[idx];
}

typedef struct { float d; uint32_t i; } sortkey_t;
static int total_cmp_desc(const void* pa, const void* pb) {
    /* f32::total_cmp(...).reverse(); ties broken by the previous position so
     * that qsort reproduces Rust's stable sort_by. */
    const sortkey_t* a = (const sortkey_t*)pa;
    const sortkey_t* b = (const sortkey_t*)pb;
    int32_t ia, ib;
    memcpy(&ia, &a->d, 4);
    memcpy(&ib, &b->d, 4);
    ia ^= (int32_t)(((uint32_t)(ia >> 31)) >> 1);
    ib ^= (int32_t)(((uint32_t)(ib >> 31)) >> 1);
    if (ia != ib) return ia < ib ? 1 : -1;
    return a->i < b->i ? -1 : (a->i > b->i ? 1 : 0);
}

static int render_impl(const or_frame* fr, const or_object* objs, uint32_t n_obj, const or_material* mats,
                       uint32_t n_mat, uint64_t tri_begin, uint64_t tri_end, uint8_t* bgra, float* depth,
                       uint32_t* winner, or_stats* st) {
    if (!fr || !bgra || !depth || fr->W <= 0 || fr->H <= 0) return OR_E_ARG;
    or_stats local;
    if (!st) st = &local;
    memset(st, 0, sizeof(*st));
    or_clear(fr, bgra, depth, winner);
    or_material dflt = default_material();
    v3 cam = mk3(fr->cam[0], fr->cam[1], fr->cam[2]);
    v3 L = mk3(fr->light[0], fr->light[1], fr->light[2]);
    uint32_t next_id = 0;
    uint64_t opaque_index = 0;
    for (uint32_t oi = 0; oi < n_obj; oi++) {
        const or_object* o = &objs[oi];
        visual_t* vis = (visual_t*)malloc(sizeof(visual_t) * (o->n_pos ? o->n_pos : 1));
        if (!vis) return OR_E_ARG;
        for (uint32_t i = 0; i < o->n_pos; i++) { /* scene/mod.rs:917-926 */
            v3 v = mk3(o->pos[3 * i], o->pos[3 * i + 1], o->pos[3 * i + 2]);
            v3 eye_dir = sub3(v, cam);
            v3 light = normalized3(sub3(v, L));
            v3 eye = normalized3(eye_dir);
            vis[i].light[0] = light.x; vis[i].light[1] = light.y; vis[i].light[2] = light.z;
            vis[i].depth = norm3(eye_dir);
            v3 h = normalized3(add3(light, eye));
            vis[i].halfway[0] = h.x; vis[i].halfway[1] = h.y; vis[i].halfway[2] = h.z;
        }
        /* opaque meshes (Object::new splits by alpha < 1, object.rs:45-53) */
        for (uint32_t mi = 0; mi < o->n_meshes; mi++) {
            const or_mesh* mesh = &o->meshes[mi];
            const or_material* mat = pick_material(mats, n_mat, mesh->material, &dflt);
            if (mat->alpha < 1.0f) continue;
            for (uint32_t t = 0; t < mesh->n_tri; t++, opaque_index++) {
                if (opaque_index < tri_begin || opaque_index >= tri_end) continue;
                st->tris_in++;
                tri_t tri;
                int e = assemble(o, vis, mesh->tri + 9u * t, &tri);
                if (e) { free(vis); return e; }
                /* back-face culling scene/mod.rs:1016-1027 */
                v3 nrm = cross3(sub3(tri.p[1], tri.p[0]), sub3(tri.p[2], tri.p[1]));
                v3 center = add3(add3(add3(mk3(0.0f, 0.0f, 0.0f), tri.p[0]), tri.p[1]), tri.p[2]);
                center = div3(center, 3.0f); /* Triangle::get_center scene/mod.rs:92-99 */
                if (dot3(sub3(cam, center), nrm) <= 0.0f) { st->tris_culled++; continue; }
                tri_t pool[12];
                int n = clip_against_planes(fr, &tri, pool);
                for (int k = 0; k < n; k++) {
                    st->tris_emitted++;
                    e = project_draw(fr, &pool[k], mat, 1, next_id++, bgra, depth, winner, st);
                    if (e) { free(vis); return e; }
                }
            }
        }
        /* transparent meshes scene/mod.rs:1090-1247: depth update disabled,
         * per-mesh back-to-front stable sort, no back-face culling. */
        for (uint32_t mi = 0; mi < o->n_meshes; mi++) {
            const or_mesh* mesh = &o->meshes[mi];
            const or_material* mat = pick_material(mats, n_mat, mesh->material, &dflt);
            if (!(mat->alpha < 1.0f)) continue;
            if (tri_begin != 0 || tri_end != UINT64_MAX) continue; /* sampled runs: opaque only */
            sortkey_t* keys = (sortkey_t*)malloc(sizeof(sortkey_t) * (mesh->n_tri ? mesh->n_tri : 1));
            if (!keys) { free(vis); return OR_E_ARG; }
            for (uint32_t t = 0; t < mesh->n_tri; t++) {
                const uint32_t* t9 = mesh->tri + 9u * t;
                for (int k = 0; k < 3; k++)
                    if (t9[k] >= o->n_pos) { free(keys); free(vis); return OR_E_INDEX; }
                v3 a = mk3(o->pos[3 * t9[0]], o->pos[3 * t9[0] + 1], o->pos[3 * t9[0] + 2]);
                v3 b = mk3(o->pos[3 * t9[1]], o->pos[3 * t9[1] + 1], o->pos[3 * t9[1] + 2]);
                v3 c = mk3(o->pos[3 * t9[2]], o->pos[3 * t9[2] + 1], o->pos[3 * t9[2] + 2]);
                v3 ctr = div3(add3(add3(a, b), c), 3.0f); /* scene/mod.rs:1108 */
                keys[t].d = norm3(sub3(ctr, cam));        /* Vec3::dist linalg.rs:177-180 */
                keys[t].i = t;
            }
            qsort(keys, mesh->n_tri, sizeof(sortkey_t), total_cmp_desc);
            for (uint32_t s = 0; s < mesh->n_tri; s++) {
                st->tris_in++;
                tri_t tri;
                int e = assemble(o, vis, mesh->tri + 9u * keys[s].i, &tri);
                if (e) { free(keys); free(vis); return e; }
                tri_t pool[12];
                int n = clip_against_planes(fr, &tri, pool);
                for (int k = 0; k < n; k++) {
                    st->tris_emitted_tr++;
                    e = project_draw(fr, &pool[k], mat, 0, 0xFFFFFFFFu, bgra, depth, winner, st);
                    if (e) { free(keys); free(vis); return e; }
                }
            }
            free(keys);
        }
        free(vis);
    }
    return OR_OK;
}

int or_render(const or_frame* fr, const or_object* objs, uint32_t n_obj, const or_material* mats, uint32_t n_mat,
              uint8_t* bgra, float* depth, uint32_t* winner, or_stats* st) {
    return render_impl(fr, objs, n_obj, mats, n_mat, 0, UINT64_MAX, bgra, depth, winner, st);
}

int or_render_grid(const or_frame* fr, const or_object* objs, uint32_t n_obj, const or_material* mats,
                   uint32_t n_mat, uint32_t sx, uint32_t sy, uint32_t ox, uint32_t oy, uint8_t* bgra, float* depth,
                   uint32_t* winner, or_stats* st) {
    if (sx == 0 || sy == 0 || ox >= sx || oy >= sy) return OR_E_ARG;
    g_sx = sx; g_sy = sy; g_ox = ox; g_oy = oy;
    int rc = render_impl(fr, objs, n_obj, mats, n_mat, 0, UINT64_MAX, bgra, depth, winner, st);
    g_sx = 1; g_sy = 1; g_ox = 0; g_oy = 0;
    return rc;
}

int or_render_range(const or_frame* fr, const or_object* objs, uint32_t n_obj, const or_material* mats,
                    uint32_t n_mat, uint64_t tri_begin, uint64_t tri_end, uint8_t* bgra, float* depth,
                    uint32_t* winner, or_stats* st) {
    return render_impl(fr, objs, n_obj, mats, n_mat, tri_begin, tri_end, bgra, depth, winner, st);
}

/* ------------------------------------------------------------ GUI overlay */
typedef struct { uint64_t x, y, w, h; } rect_t; /* Rectangle, canvas.rs:293-313 */
static rect_t rect_from_coords(uint64_t x0, uint64_t y0, uint64_t x1, uint64_t y1) { /* canvas.rs:315-330 */
    rect_t r;
    uint64_t xa = x0 < x1 ? x0 : x1, xb = x0 < x1 ? x1 : x0;
    uint64_t ya = y0 < y1 ? y0 : y1, yb = y0 < y1 ? y1 : y0;
    r.x = xa; r.y = ya; r.w = xb - xa; r.h = yb - ya;
    return r;
}
static rect_t rect_clip(rect_t a, rect_t b) { /* canvas.rs:332-350 */
    uint64_t x0 = a.x > b.x ? a.x : b.x, y0 = a.y > b.y ? a.y : b.y;
    uint64_t ax = a.x + a.w, bx = b.x + b.w, ay = a.y + a.h, by = b.y + b.h;
    uint64_t x1 = ax < bx ? ax : bx, y1 = ay < by ? ay : by;
    if (x0 > x1) { x0 = 0; x1 = 0; }
    if (y0 > y1) { y0 = 0; y1 = 0; }
    return rect_from_coords(x0, y0, x1, y1);
}
/* TextureMap::get_rgba_slice scene/mod.rs:137-152, release semantics: the
 * slice [offset, offset + 4) must lie in the image, else the frame panics. */
static int rgba_slice(const uint8_t* img, int32_t w, int32_t h, int32_t comp, float u, float v, uint8_t out[4]) {
    uint64_t ui = f2usize(floorf(u * (float)w));
    uint64_t vi = ((uint64_t)h - 1u) - f2usize(floorf(v * (float)h));
    uint64_t off = (vi * (uint64_t)w + ui) * (uint64_t)comp;
    uint64_t end = off + 4u, len = (uint64_t)w * (uint64_t)h * (uint64_t)comp;
    if (off > end || end > len) return OR_E_TEXCOORD;
    memcpy(out, img + off, 4);
    return OR_OK;
}
/* Pixel * f32 (canvas.rs:155-169): each channel truncated to u8. */
static void px_mul(const uint8_t p[3], float s, uint8_t o[3]) {
    for (int i = 0; i < 3; i++) o[i] = f2u8((float)p[i] * s);
}
static void px_add(const uint8_t a[3], const uint8_t b[3], uint8_t o[3]) { /* canvas.rs:136-153, wrapping */
    for (int i = 0; i < 3; i++) o[i] = (uint8_t)(a[i] + b[i]);
}

typedef struct { v2 s, uv; uint8_t rgb[3]; float alpha; } gvert_t; /* VertexSimpleAttributes canvas.rs:186-191 */

static int gui_triangle(int32_t W, int32_t H, const gvert_t* A, const gvert_t* B, const gvert_t* Cv, rect_t clip,
                        const uint8_t* font, int32_t fw, int32_t fh, int32_t fcomp, uint8_t* bgra,
                        const float* depth) {
    /* Canvas::draw_triangle canvas.rs:435-575 (no canvas offset) */
    v2 a, b, c;
    a.x = floorf(A->s.x + 0.5f); a.y = floorf(A->s.y + 0.5f); /* pos_map_center canvas.rs:896-904 */
    b.x = floorf(B->s.x + 0.5f); b.y = floorf(B->s.y + 0.5f);
    c.x = floorf(Cv->s.x + 0.5f); c.y = floorf(Cv->s.y + 0.5f);
    rect_t drawable = rect_from_coords(f2usize(min3f(a.x, b.x, c.x)), f2usize(min3f(a.y, b.y, c.y)),
                                       f2usize(max3f(a.x, b.x, c.x)), f2usize(max3f(a.y, b.y, c.y)));
    rect_t screen = rect_from_coords(0, 0, (uint64_t)W - 1u, (uint64_t)H - 1u);
    drawable = rect_clip(drawable, screen);
    rect_t valid = rect_clip(clip, drawable);
    float fa = edge_fn(b, c, a.x, a.y), fb = edge_fn(c, a, b.x, b.y), fg = edge_fn(a, b, c.x, c.y);
    float oa = edge_fn(b, c, -1.0f, -1.0f), ob = edge_fn(c, a, -1.0f, -1.0f), og = edge_fn(a, b, -1.0f, -1.0f);
    for (uint64_t y = valid.y; y <= valid.y + valid.h; y++) {
        float yf = (float)y;
        for (uint64_t x = valid.x; x <= valid.x + valid.w; x++) {
            float xf = (float)x;
            float al = edge_fn(b, c, xf, yf) / fa;
            float be = edge_fn(c, a, xf, yf) / fb;
            float ga = edge_fn(a, b, xf, yf) / fg;
            if (!(al >= 0.0f && be >= 0.0f && ga >= 0.0f)) continue;
            if (!((al > 0.0f || fa * oa > 0.0f) && (be > 0.0f || fb * ob > 0.0f) && (ga > 0.0f || fg * og > 0.0f)))
                continue;
            uint8_t t0[3], t1[3], t2[3], col[3];
            px_mul(A->rgb, al, t0);
            px_mul(B->rgb, be, t1);
            px_mul(Cv->rgb, ga, t2);
            px_add(t0, t1, col);
            px_add(col, t2, col);
            float calpha = (al * A->alpha + be * B->alpha) + ga * Cv->alpha;
            float u = (A->uv.x * al + B->uv.x * be) + Cv->uv.x * ga;
            float v = (A->uv.y * al + B->uv.y * be) + Cv->uv.y * ga;
            uint8_t tex[4];
            if (rgba_slice(font, fw, fh, fcomp, u, v, tex)) return OR_E_TEXCOORD;
            float ta = (float)tex[3] / 255.0f;
            uint8_t ct[3] = {tex[0], tex[1], tex[2]}; /* Color::Custom -> Pixel::new(r, g, b) */
            px_mul(col, ta, t0);
            px_mul(ct, 1.0f - ta, t1);
            px_add(t0, t1, col);
            float fin = calpha * ta;
            uint8_t out[3];
            uint64_t W64 = (uint64_t)W, H64 = (uint64_t)H;
            uint8_t* px = bgra + 4u * (W64 * (H64 - y - 1u) + x);
            if (fin < 1.0f) { /* draw_pixel_coord_with_depth canvas.rs:906-930 */
                uint8_t bg[3] = {px[2], px[1], px[0]};
                px_mul(bg, 1.0f - fin, t0);
                px_mul(col, fin, t1);
                px_add(t0, t1, out);
            } else {
                memcpy(out, col, 3);
            }
            if (0.0f < depth[W64 * y + x]) {
                px[0] = out[2]; px[1] = out[1]; px[2] = out[0]; px[3] = 0; /* Pixel sums carry pad 0 */
            }
        }
    }
    return OR_OK;
}

int or_gui_draw(int32_t W, int32_t H, const or_gui_vertex* vtx, uint32_t n_vtx, const uint32_t* idx,
                uint32_t n_idx, const or_gui_cmd* cmds, uint32_t n_cmds, const uint8_t* font, int32_t fw,
                int32_t fh, int32_t fcomp, uint8_t* bgra, const float* depth) {
    if (W <= 0 || H <= 0 || !bgra || !depth || !font || fw <= 0 || fh <= 0) return OR_E_ARG;
    for (uint32_t ci = 0; ci < n_cmds; ci++) { /* Gui::render gui.rs:382-478 */
        const or_gui_cmd* cm = &cmds[ci];
        float cr[4] = {cm->clip_rect[0], cm->clip_rect[1], cm->clip_rect[2], cm->clip_rect[3]};
        if (cr[0] > cr[2] || cr[1] > cr[3]) continue;
        cr[1] = (float)H - cr[1];
        cr[3] = (float)H - cr[3];
        rect_t clip = rect_from_coords(f2usize(cr[0]), f2usize(cr[1]), f2usize(cr[2]), f2usize(cr[3]));
        if ((uint64_t)cm->idx_offset + cm->count > n_idx) return OR_E_INDEX;
        for (uint32_t t = 0; t + 3 <= cm->count; t += 3) {
            gvert_t gv[3];
            for (int k = 0; k < 3; k++) {
                uint64_t vi = (uint64_t)idx[cm->idx_offset + t + k] + cm->vtx_offset;
                if (vi >= n_vtx) return OR_E_INDEX;
                const or_gui_vertex* q = &vtx[vi];
                gv[k].s.x = q->pos[0];
                gv[k].s.y = ((float)H - 1.0f) - q->pos[1]; /* invert_y */
                gv[k].uv.x = q->uv[0];
                gv[k].uv.y = 1.0f - q->uv[1];
                memcpy(gv[k].rgb, q->col, 3);
                gv[k].alpha = (float)q->col[3] / 255.0f;
            }
            int e = gui_triangle(W, H, &gv[0], &gv[1], &gv[2], clip, font, fw, fh, fcomp, bgra, depth);
            if (e) return e;
        }
    }
    return OR_OK;
}
