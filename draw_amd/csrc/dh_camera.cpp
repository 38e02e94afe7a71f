// dh_camera.cpp -- per-frame uniforms on the host (reference row a1).
//
// Camera::new (scene/mod.rs:297-357), Camera::update_basis (438-451),
// Camera::gen_matrix (453-479), Camera::gen_view_planes (481-593),
// ViewPlane::new (603-632) and Scene::gen_transformation_matrix (817-899),
// in binary32 and the reference's operation order.  Pure host code: it runs
// once per frame and its 16+24 floats are passed to the kernels.
#include <math.h>
#include <string.h>

#include "dh_common.hpp"

namespace {

using dh::V3;
using dh::v3;

struct Mat4 {
  float a[4][4];
};

// Matrix4 * Matrix4, linalg.rs:328-344: c[i][j] accumulates a[i][k]*b[k][j]
// from 0 in k order.
Mat4 mul(const Mat4& x, const Mat4& y) {
  Mat4 c;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      float s = 0.0f;
      for (int k = 0; k < 4; k++) s = s + x.a[i][k] * y.a[k][j];
      c.a[i][j] = s;
    }
  return c;
}

// Matrix4 * Vec4, linalg.rs:346-360.
void mul_vec(const Mat4& m, const float v[4], float o[4]) {
  for (int i = 0; i < 4; i++) {
    float s = 0.0f;
    for (int j = 0; j < 4; j++) s = s + m.a[i][j] * v[j];
    o[i] = s;
  }
}

Mat4 rows(float a00, float a01, float a02, float a03, float a10, float a11, float a12, float a13, float a20,
          float a21, float a22, float a23, float a30, float a31, float a32, float a33) {
  Mat4 m = {{{a00, a01, a02, a03}, {a10, a11, a12, a13}, {a20, a21, a22, a23}, {a30, a31, a32, a33}}};
  return m;
}

struct Basis {
  V3 u, v, w;
};

// Camera::update_basis, scene/mod.rs:438-451 (up = (0,1,0)).
Basis update_basis(V3 direction) {
  const V3 up = v3(0.0f, 1.0f, 0.0f);
  V3 w = dh::muls(dh::divs(direction, dh::norm(direction)), -1.0f);
  V3 t_x_w = dh::cross(up, w);
  V3 u = dh::divs(t_x_w, dh::norm(t_x_w));
  V3 v = dh::cross(w, u);
  Basis b;
  b.u = dh::normalized(u);
  b.v = dh::normalized(v);
  b.w = dh::normalized(w);
  return b;
}

// Camera::get_basis_matrix, scene/mod.rs:415-428.
Mat4 basis_matrix(const Basis& b) {
  return rows(b.u.x, b.v.x, b.w.x, 0.0f, b.u.y, b.v.y, b.w.y, 0.0f, b.u.z, b.v.z, b.w.z, 0.0f, 0.0f, 0.0f, 0.0f,
              1.0f);
}

Mat4 transposed(const Mat4& m) {  // linalg.rs:275-287
  Mat4 t = m;
  for (int i = 0; i < 4; i++)
    for (int j = i + 1; j < 4; j++) {
      float s = t.a[i][j];
      t.a[i][j] = t.a[j][i];
      t.a[j][i] = s;
    }
  return t;
}

// (basis * p.as_vec4()).vec3_over_w() + camera_pos, scene/mod.rs:504-527.
V3 to_world(const Mat4& basis, V3 p, V3 cam) {
  float in[4] = {p.x, p.y, p.z, 1.0f}, o[4];
  mul_vec(basis, in, o);
  return dh::add(v3(o[0] / o[3], o[1] / o[3], o[2] / o[3]), cam);
}

// ViewPlane::new, scene/mod.rs:603-632: the interior point must evaluate
// positive; the flipped orientation is asserted.
int view_plane(V3 a, V3 b, V3 c, V3 visible, float out[4]) {
  V3 p = dh::sub(b, a), q = dh::sub(c, b);
  V3 n = dh::cross(p, q);
  float k = -dh::dot(n, a);
  float test = dh::dot(n, visible) + k;
  if (test < 0.0f) {
    n = dh::cross(q, p);
    k = -dh::dot(n, a);
    test = dh::dot(n, visible) + k;
    if (!(test > 0.0f)) return DH_E_PLANE;
  }
  out[0] = n.x;
  out[1] = n.y;
  out[2] = n.z;
  out[3] = k;
  return DH_OK;
}

// The uniforms of a Scene of screen size W x H (Scene::new, scene/mod.rs:760-786:
// the camera's aspect ratio and the viewport matrix use the screen size)
// drawn into a canvas of cw x ch pixels whose origin sits at (ox, oy) on that
// screen (Canvas::new + Canvas::apply_offset, canvas.rs:366, 382-385): the
// raster's screen rectangle is the canvas (canvas.rs:647) and every snapped
// vertex is screen_coord - offset (canvas.rs:585-587).
int uniforms(const dh_camera* cam, int32_t W, int32_t H, int32_t cw, int32_t ch, int32_t ox, int32_t oy,
             const float light[3], dh_frame_params* out) {
  if (!cam || !out || W <= 0 || H <= 0 || cw <= 0 || ch <= 0) return DH_E_ARG;
  memset(out, 0, sizeof(*out));
  // Camera::new: the window half-extents from |near| and fov_x.
  const float ratio = (float)W / (float)H;  // Scene::new, scene/mod.rs:768
  const float near_ = cam->near_, far_ = cam->far_, fov_x = cam->fov_x_deg;
  const float pi = 3.14159265358979323846f;
  const float fov_x_rad = fov_x * (pi / 180.0f);  // f32::to_radians
  const float right = fabsf(near_) * tanf(fov_x_rad / 2.0f);
  const float left = -right;
  const float top = (1.0f / ratio) * right;  // ratio.recip() * right
  const float bottom = -top;
  if (!(fov_x < 180.0f) || !(near_ < 0.0f) || !(near_ > far_) || !(right > left) || !(top > bottom))
    return DH_E_CAMERA;
  const V3 pos = v3(cam->pos[0], cam->pos[1], cam->pos[2]);
  const V3 direction = dh::normalized(v3(cam->dir[0], cam->dir[1], cam->dir[2]));

  // Scene::gen_transformation_matrix: viewport * orth * persp * view.
  const float n_x = (float)W, n_y = (float)H;
  const float n = near_, f = far_, r = right, l = left, t = top, b = bottom;
  const Basis basis = update_basis(direction);
  const Mat4 matrix_pos = rows(1, 0, 0, -pos.x, 0, 1, 0, -pos.y, 0, 0, 1, -pos.z, 0, 0, 0, 1);
  const Mat4 view = mul(transposed(basis_matrix(basis)), matrix_pos);
  const Mat4 persp = rows(n, 0, 0, 0, 0, n, 0, 0, 0, 0, (n + f), -(n * f), 0, 0, 1, 0);
  const Mat4 orth = rows(2.0f / (r - l), 0, 0, -(r + l) / (r - l), 0, 2.0f / (t - b), 0, -(t + b) / (t - b), 0, 0,
                         2.0f / (n - f), -(n + f) / (n - f), 0, 0, 0, 1);
  const Mat4 viewport =
      rows(n_x / 2.0f, 0, 0, (n_x - 1.0f) / 2.0f, 0, n_y / 2.0f, 0, (n_y - 1.0f) / 2.0f, 0, 0, 1, 0, 0, 0, 0, 1);
  const Mat4 M = mul(mul(mul(viewport, orth), persp), view);
  memcpy(out->M, M.a, sizeof(out->M));

  // Camera::gen_view_planes: near-plane corners, far-plane edge centres.
  const Mat4 mb = basis_matrix(update_basis(direction));
  const V3 urn = to_world(mb, v3(r, t, n), pos);
  const V3 uln = to_world(mb, v3(l, t, n), pos);
  const V3 lrn = to_world(mb, v3(r, b, n), pos);
  const V3 lln = to_world(mb, v3(l, b, n), pos);
  const float xc = (l + r) / 2.0f, yc = (b + t) / 2.0f;
  const V3 ucf = to_world(mb, v3(xc, (f * t) / n, f), pos);
  const V3 lcf = to_world(mb, v3(xc, (f * b) / n, f), pos);
  const V3 rcf = to_world(mb, v3((f * r) / n, yc, f), pos);
  const V3 lfc = to_world(mb, v3((f * l) / n, yc, f), pos);
  const V3 visible = dh::divs(dh::add(urn, lcf), 2.0f);
  int e = DH_OK;
  if ((e = view_plane(urn, lrn, lln, visible, out->planes[0]))) return e;  // near
  if ((e = view_plane(lfc, rcf, ucf, visible, out->planes[1]))) return e;  // far
  if ((e = view_plane(rcf, urn, lrn, visible, out->planes[2]))) return e;  // right
  if ((e = view_plane(lfc, lln, uln, visible, out->planes[3]))) return e;  // left
  if ((e = view_plane(ucf, uln, urn, visible, out->planes[4]))) return e;  // top
  if ((e = view_plane(lcf, lln, lrn, visible, out->planes[5]))) return e;  // bottom

  out->cam[0] = pos.x;
  out->cam[1] = pos.y;
  out->cam[2] = pos.z;
  const float dflt_light[3] = {0.0f, 300.0f, 300.0f};  // Scene::new, scene/mod.rs:766
  const float* L = light ? light : dflt_light;
  out->light[0] = L[0];
  out->light[1] = L[1];
  out->light[2] = L[2];
  out->W = cw;
  out->H = ch;
  out->off_x = (float)ox;  // Canvas::apply_offset: x as f32 (canvas.rs:382-385)
  out->off_y = (float)oy;
  out->clear_depth = 100000.0f;  // app/mod.rs:80
  out->clear_bgra[0] = 255;      // Pixel::azul_bb = (r 155, g 186, b 255), canvas.rs:131-133
  out->clear_bgra[1] = 186;
  out->clear_bgra[2] = 155;
  out->clear_bgra[3] = 255;
  return DH_OK;
}

}  // namespace

extern "C" int dh_frame_uniforms(const dh_camera* cam, int32_t W, int32_t H, const float light[3],
                                 dh_frame_params* out) {
  return uniforms(cam, W, H, W, H, 0, 0, light, out);
}

extern "C" int dh_frame_uniforms_canvas(const dh_camera* cam, int32_t screen_w, int32_t screen_h, int32_t canvas_w,
                                        int32_t canvas_h, int32_t off_x, int32_t off_y, const float light[3],
                                        dh_frame_params* out) {
  return uniforms(cam, screen_w, screen_h, canvas_w, canvas_h, off_x, off_y, light, out);
}
