// dh_obj.cpp -- OBJ/MTL ingest on the host (Object::load_from_file,
// object.rs:106-419; TextureMap::load_from_file, scene/mod.rs:174-202), then
// upload into a context's resident pool.
//
// The reference parses with the third-party `obj` crate (git HEAD, not in the
// reference tree) and decodes textures with stb_image.  This restates the
// behaviour those calls have on the reference's inputs:
//   OBJ  v / vt / vn / f (v, v/vt, v//vn, v/vt/vn; 1-based, negative =
//        relative), o, g, usemtl (a group that already has faces is split),
//        mtllib; other statements ignored.  Floats parse correctly rounded
//        (strtof), as Rust's str::parse::<f32> does.
//   MTL  newmtl, Ka, Kd, Ks (required: the reference unwraps them), d
//        (default 1), map_Ka, map_Kd.
//   maps PNG or JPEG files, decoded as dh_image.cpp describes.
// and then the reference's own post-processing, in its f32 operation order:
// normals normalised, positions rescaled so the largest norm is 100, faces of
// 3 or 4 vertices (quads split a,b,c + c,d,a), a dummy (0,0) texture
// coordinate per group with missing coordinates, generated per-vertex normals
// per group with missing normals.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include <algorithm>
#include <exception>
#include <new>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "dh_common.hpp"
#include "dh_image.hpp"

using dh::V3;
using dh::add;
using dh::cross;
using dh::norm;
using dh::sub;
using dh::v3;

namespace {

struct ObjTexture {
  std::string name;
  float ka[3], kd[3], ks[3];
  float alpha;
  std::vector<uint8_t> map_kd, map_ka;  // empty = TextureMap::default
  int32_t kd_w = 0, kd_h = 0, kd_comp = 0, ka_w = 0, ka_h = 0, ka_comp = 0;
};

struct ObjMesh {
  std::string name;
  std::vector<uint32_t> tri;  // 9 per triangle
  uint32_t texture_idx = 0;
};

}  // namespace

struct dh_obj_s {
  std::string name;
  std::vector<float> pos, nrm, uv;  // xyz, xyz, uv
  std::vector<ObjMesh> meshes;
  std::vector<ObjTexture> textures;
  std::string err;
};

namespace {

std::string dir_of(const std::string& p) {
  const size_t k = p.find_last_of('/');
  return k == std::string::npos ? std::string() : p.substr(0, k + 1);
}

bool read_file(const std::string& path, std::vector<uint8_t>* out) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return false;
  fseek(f, 0, SEEK_END);
  const long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  out->resize(n > 0 ? (size_t)n : 0);
  const size_t got = n > 0 ? fread(out->data(), 1, (size_t)n, f) : 0;
  fclose(f);
  return got == out->size();
}

std::vector<std::string> split_ws(const std::string& s) {
  std::vector<std::string> out;
  std::istringstream is(s);
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}

std::string rest_of(const std::string& line, const std::string& key) {
  size_t k = line.find(key);
  if (k == std::string::npos) return {};
  k += key.size();
  while (k < line.size() && (line[k] == ' ' || line[k] == '\t')) k++;
  size_t e = line.size();
  while (e > k && (line[e - 1] == ' ' || line[e - 1] == '\t' || line[e - 1] == '\r')) e--;
  return line.substr(k, e - k);
}

bool parse_f(const std::string& t, float* v) {
  char* end = nullptr;
  *v = strtof(t.c_str(), &end);
  return end && *end == 0 && end != t.c_str();
}

int load_mtl(dh_obj o, const std::string& path, const std::string& dir) {
  std::ifstream in(path);
  if (!in) {
    o->err = "cannot open material library " + path;
    return DH_E_ARG;
  }
  std::string line;
  ObjTexture* cur = nullptr;
  bool has_ka = false, has_kd = false, has_ks = false;
  std::string map_ka, map_kd;
  auto finish = [&]() -> int {
    if (!cur) return DH_OK;
    if (!(has_ka && has_kd && has_ks)) {  // material.ka/kd/ks.as_ref().unwrap(), object.rs:180-182
      o->err = "material " + cur->name + " lacks Ka, Kd or Ks";
      return DH_E_ARG;
    }
    for (int k = 0; k < 2; k++) {
      const std::string& fn = k == 0 ? map_kd : map_ka;
      if (fn.empty()) continue;
      std::vector<uint8_t> file;
      if (!read_file(dir + fn, &file)) {
        o->err = "cannot open texture " + dir + fn;
        return DH_E_ARG;
      }
      std::vector<uint8_t> img;
      int32_t w, h, c;
      std::string e;
      const int rc = dh::image_decode(file.data(), file.size(), &img, &w, &h, &c, &e);
      if (rc) {
        o->err = "texture " + fn + ": " + e;
        return rc;
      }
      if (k == 0) { cur->map_kd.swap(img); cur->kd_w = w; cur->kd_h = h; cur->kd_comp = c; }
      else { cur->map_ka.swap(img); cur->ka_w = w; cur->ka_h = h; cur->ka_comp = c; }
    }
    return DH_OK;
  };
  while (std::getline(in, line)) {
    const std::vector<std::string> t = split_ws(line);
    if (t.empty() || t[0][0] == '#') continue;
    if (t[0] == "newmtl") {
      int rc = finish();
      if (rc) return rc;
      o->textures.emplace_back();
      cur = &o->textures.back();
      cur->name = rest_of(line, "newmtl");
      cur->alpha = 1.0f;  // material.d.unwrap_or(1.0), object.rs:183
      has_ka = has_kd = has_ks = false;
      map_ka.clear();
      map_kd.clear();
    } else if (!cur) {
      continue;
    } else if (t[0] == "Ka" || t[0] == "Kd" || t[0] == "Ks") {
      float v[3];
      if (t.size() < 4 || !parse_f(t[1], &v[0]) || !parse_f(t[2], &v[1]) || !parse_f(t[3], &v[2])) {
        o->err = "bad " + t[0] + " in " + path;
        return DH_E_ARG;
      }
      float* dst = t[0] == "Ka" ? cur->ka : (t[0] == "Kd" ? cur->kd : cur->ks);
      memcpy(dst, v, sizeof v);
      (t[0] == "Ka" ? has_ka : (t[0] == "Kd" ? has_kd : has_ks)) = true;
    } else if (t[0] == "d" && t.size() >= 2) {
      if (!parse_f(t[1], &cur->alpha)) {
        o->err = "bad d in " + path;
        return DH_E_ARG;
      }
    } else if (t[0] == "map_Kd") {
      map_kd = rest_of(line, "map_Kd");
    } else if (t[0] == "map_Ka") {
      map_ka = rest_of(line, "map_Ka");
    }
  }
  return finish();
}

struct Corner {
  int64_t v, t, n;  // resolved 0-based indices, -1 = absent
};

struct Group {
  std::string name, material;
  std::vector<std::vector<Corner>> faces;
};

V3 vat(const std::vector<float>& a, size_t i) { return v3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }

int load_obj(dh_obj o, const std::string& path) {
  std::ifstream in(path);
  if (!in) {
    o->err = "cannot open " + path;
    return DH_E_ARG;
  }
  const std::string dir = dir_of(path);
  o->name = path.substr(dir.size());
  std::vector<std::string> mtllibs;
  std::vector<Group> groups;  // over all objects, in file order
  groups.push_back(Group{"default", "", {}});
  std::string line;
  size_t n_v = 0, n_vt = 0, n_vn = 0;
  auto resolve = [](const std::string& s, size_t n, int64_t* out) -> bool {
    if (s.empty()) { *out = -1; return true; }
    char* end = nullptr;
    const long long i = strtoll(s.c_str(), &end, 10);
    if (!end || *end || i == 0) return false;
    const long long r = i > 0 ? i - 1 : (long long)n + i;
    if (r < 0 || (size_t)r >= n) return false;
    *out = r;
    return true;
  };
  int line_no = 0;
  while (std::getline(in, line)) {
    line_no++;
    const std::vector<std::string> t = split_ws(line);
    if (t.empty() || t[0][0] == '#') continue;
    const std::string& k = t[0];
    if (k == "v" || k == "vn") {
      float v[3];
      if (t.size() < 4 || !parse_f(t[1], &v[0]) || !parse_f(t[2], &v[1]) || !parse_f(t[3], &v[2])) {
        o->err = "line " + std::to_string(line_no) + ": bad " + k;
        return DH_E_ARG;
      }
      std::vector<float>& dst = k == "v" ? o->pos : o->nrm;
      dst.insert(dst.end(), v, v + 3);
      (k == "v" ? n_v : n_vn)++;
    } else if (k == "vt") {
      float u = 0.0f, v = 0.0f;
      if (t.size() < 2 || !parse_f(t[1], &u) || (t.size() >= 3 && !parse_f(t[2], &v))) {
        o->err = "line " + std::to_string(line_no) + ": bad vt";
        return DH_E_ARG;
      }
      o->uv.push_back(u);
      o->uv.push_back(v);
      n_vt++;
    } else if (k == "f") {
      std::vector<Corner> face;
      for (size_t i = 1; i < t.size(); i++) {
        const std::string& c = t[i];
        const size_t s1 = c.find('/');
        const size_t s2 = s1 == std::string::npos ? std::string::npos : c.find('/', s1 + 1);
        Corner q;
        const std::string a = c.substr(0, s1);
        const std::string b = s1 == std::string::npos ? "" : c.substr(s1 + 1, s2 == std::string::npos ? std::string::npos : s2 - s1 - 1);
        const std::string d = s2 == std::string::npos ? "" : c.substr(s2 + 1);
        if (a.empty() || !resolve(a, n_v, &q.v) || !resolve(b, n_vt, &q.t) || !resolve(d, n_vn, &q.n)) {
          o->err = "line " + std::to_string(line_no) + ": bad face index " + c;
          return DH_E_INDEX;
        }
        face.push_back(q);
      }
      groups.back().faces.push_back(std::move(face));
    } else if (k == "o") {
      groups.push_back(Group{"default", groups.back().material, {}});
    } else if (k == "g") {
      groups.push_back(Group{t.size() > 1 ? rest_of(line, "g") : std::string("default"), groups.back().material, {}});
    } else if (k == "usemtl") {
      const std::string m = rest_of(line, "usemtl");
      if (!groups.back().faces.empty()) groups.push_back(Group{groups.back().name, m, {}});
      else groups.back().material = m;
    } else if (k == "mtllib") {
      for (size_t i = 1; i < t.size(); i++) mtllibs.push_back(t[i]);
    }
  }
  // Object::load_from_file post-processing, object.rs:140-419
  if (n_v == 0) {
    o->err = "no vertices";
    return DH_E_ARG;
  }
  for (size_t i = 0; i < n_vn; i++) {  // normals normalised
    const V3 n = vat(o->nrm, i);
    const float l = norm(n);
    o->nrm[3 * i] = n.x / l;
    o->nrm[3 * i + 1] = n.y / l;
    o->nrm[3 * i + 2] = n.z / l;
  }
  float vmax = -INFINITY;  // the largest norm (sort_by partial_cmp: NaN panics)
  for (size_t i = 0; i < n_v; i++) {
    const float l = norm(vat(o->pos, i));
    if (isnan(l)) {
      o->err = "NaN vertex";
      return DH_E_ARG;
    }
    vmax = std::max(vmax, l);
  }
  const float factor = 100.0f / vmax;
  for (float& x : o->pos) x = x * factor;
  // materials: Texture::default first, then every material of every library
  o->textures.clear();
  {
    ObjTexture d;
    d.name = "default";
    for (int k = 0; k < 3; k++) { d.ka[k] = 0.9f; d.kd[k] = 0.4f; d.ks[k] = 0.5f; }
    d.alpha = 1.0f;
    o->textures.push_back(d);
  }
  for (const std::string& lib : mtllibs) {
    const int rc = load_mtl(o, dir + lib, dir);
    if (rc) return rc;
  }
  for (const Group& g : groups) {
    if (g.faces.empty()) continue;  // "Group doesnt have faces"
    ObjMesh m;
    m.name = g.name;
    bool miss_t = false, miss_n = false;
    struct T9 { int64_t v[3], t[3], n[3]; bool has_t, has_n; };
    std::vector<T9> tris;
    for (const auto& f : g.faces) {
      if (f.size() < 3 || f.size() > 4) {
        o->err = "face with " + std::to_string(f.size()) + " vertices (3 or 4 supported)";
        return DH_E_ARG;
      }
      bool ft = true, fn = true;
      for (const Corner& c : f) { ft &= c.t >= 0; fn &= c.n >= 0; }
      miss_t |= !ft;
      miss_n |= !fn;
      const int order[2][3] = {{0, 1, 2}, {2, 3, 0}};  // quads split (a, b, c) + (c, d, a)
      for (int h = 0; h < (f.size() == 4 ? 2 : 1); h++) {
        T9 q;
        for (int k = 0; k < 3; k++) {
          const Corner& c = f[order[h][k]];
          q.v[k] = c.v; q.t[k] = c.t; q.n[k] = c.n;
        }
        q.has_t = ft;
        q.has_n = fn;
        tris.push_back(q);
      }
    }
    if (miss_t) {  // one dummy (0, 0) coordinate for the group
      const int64_t dummy = (int64_t)(o->uv.size() / 2);
      o->uv.push_back(0.0f);
      o->uv.push_back(0.0f);
      for (T9& q : tris)
        if (!q.has_t) { q.t[0] = q.t[1] = q.t[2] = dummy; q.has_t = true; }
    }
    m.texture_idx = 0;
    for (size_t i = 0; i < o->textures.size(); i++)
      if (o->textures[i].name == (g.material.empty() ? std::string("default") : g.material)) {
        m.texture_idx = (uint32_t)i;
        break;
      }
    if (miss_n) {  // per-vertex normals from this group's triangles, appended
      std::vector<V3> gen(n_v, v3(0.0f, 0.0f, 0.0f));
      for (const T9& q : tris) {
        const V3 a = vat(o->pos, q.v[0]), b = vat(o->pos, q.v[1]), c = vat(o->pos, q.v[2]);
        const V3 nn = cross(sub(b, a), sub(c, b));  // IndexedTriangle::calc_normal, mesh.rs:16-27
        for (int k = 0; k < 3; k++) gen[q.v[k]] = add(gen[q.v[k]], nn);
      }
      const int64_t base = (int64_t)(o->nrm.size() / 3);
      for (const V3& g3 : gen) {
        const float l = norm(g3);
        o->nrm.push_back(g3.x / l);
        o->nrm.push_back(g3.y / l);
        o->nrm.push_back(g3.z / l);
      }
      for (T9& q : tris)
        if (!q.has_n) { for (int k = 0; k < 3; k++) q.n[k] = q.v[k] + base; q.has_n = true; }
    }
    for (const T9& q : tris) {
      for (int k = 0; k < 3; k++) m.tri.push_back((uint32_t)q.v[k]);
      for (int k = 0; k < 3; k++) m.tri.push_back((uint32_t)q.t[k]);
      for (int k = 0; k < 3; k++) m.tri.push_back((uint32_t)q.n[k]);
    }
    o->meshes.push_back(std::move(m));
  }
  return DH_OK;
}

}  // namespace

extern "C" int dh_obj_load(const char* path, dh_obj* out, char* err, uint32_t err_len) {
  if (!path || !out) return DH_E_ARG;
  *out = nullptr;
  dh_obj o = nullptr;
  int rc;
  try {
    o = new dh_obj_s();
    rc = load_obj(o, path);
  } catch (const std::bad_alloc&) {
    rc = DH_E_NOMEM;
    if (o) o->err = "out of host memory";
  } catch (const std::exception& e) {
    rc = DH_E_ARG;
    if (o) o->err = e.what();
  }
  if (rc) {
    if (err && err_len) snprintf(err, err_len, "%s", o ? o->err.c_str() : "out of host memory");
    delete o;
    return rc;
  }
  *out = o;
  return DH_OK;
}

extern "C" int dh_obj_info(dh_obj o, dh_obj_summary* out) {
  if (!o || !out) return DH_E_ARG;
  out->n_pos = (uint32_t)(o->pos.size() / 3);
  out->n_nrm = (uint32_t)(o->nrm.size() / 3);
  out->n_uv = (uint32_t)(o->uv.size() / 2);
  out->n_meshes = (uint32_t)o->meshes.size();
  out->n_textures = (uint32_t)o->textures.size();
  out->pos = o->pos.data();
  out->nrm = o->nrm.data();
  out->uv = o->uv.data();
  return DH_OK;
}

extern "C" int dh_obj_mesh(dh_obj o, uint32_t i, dh_obj_mesh_desc* out) {
  if (!o || !out || i >= o->meshes.size()) return DH_E_ARG;
  const ObjMesh& m = o->meshes[i];
  out->tri = m.tri.data();
  out->n_tri = (uint32_t)(m.tri.size() / 9);
  out->texture_idx = m.texture_idx;
  out->name = m.name.c_str();
  return DH_OK;
}

extern "C" int dh_obj_texture(dh_obj o, uint32_t i, dh_obj_texture_desc* out) {
  if (!o || !out || i >= o->textures.size()) return DH_E_ARG;
  const ObjTexture& t = o->textures[i];
  memcpy(out->ka, t.ka, sizeof t.ka);
  memcpy(out->kd, t.kd, sizeof t.kd);
  memcpy(out->ks, t.ks, sizeof t.ks);
  out->alpha = t.alpha;
  out->name = t.name.c_str();
  out->map_kd = t.map_kd.empty() ? nullptr : t.map_kd.data();
  out->kd_w = t.kd_w; out->kd_h = t.kd_h; out->kd_comp = t.kd_comp;
  out->map_ka = t.map_ka.empty() ? nullptr : t.map_ka.data();
  out->ka_w = t.ka_w; out->ka_h = t.ka_h; out->ka_comp = t.ka_comp;
  return DH_OK;
}

namespace {
int obj_upload(dh_ctx ctx, dh_obj o, dh_handle* object, dh_handle* meshes);
}

extern "C" int dh_obj_upload(dh_ctx ctx, dh_obj o, dh_handle* object, dh_handle* meshes, uint32_t cap) {
  if (!ctx || !o || !object || (o->meshes.size() && !meshes) || cap < o->meshes.size()) return DH_E_ARG;
  try {
    return obj_upload(ctx, o, object, meshes);
  } catch (const std::bad_alloc&) {
    return DH_E_NOMEM;
  } catch (const std::exception&) {
    return DH_E_ARG;
  }
}

namespace {
int obj_upload(dh_ctx ctx, dh_obj o, dh_handle* object, dh_handle* meshes) {
  int rc = dh_object_create(ctx, o->pos.data(), (uint32_t)(o->pos.size() / 3), o->nrm.data(),
                            (uint32_t)(o->nrm.size() / 3), o->uv.data(), (uint32_t)(o->uv.size() / 2), object);
  if (rc) return rc;
  std::vector<dh_handle> mats(o->textures.size(), 0);
  for (size_t i = 0; i < o->textures.size(); i++) {
    const ObjTexture& t = o->textures[i];
    dh_material_desc d;
    memset(&d, 0, sizeof d);
    memcpy(d.ka, t.ka, sizeof t.ka);
    memcpy(d.kd, t.kd, sizeof t.kd);
    memcpy(d.ks, t.ks, sizeof t.ks);
    d.alpha = t.alpha;
    d.power = 2;  // canvas.rs:724
    if (!t.map_kd.empty() && (rc = dh_texture_create(ctx, t.map_kd.data(), t.kd_w, t.kd_h, t.kd_comp, &d.map_kd)))
      return rc;
    if (!t.map_ka.empty() && (rc = dh_texture_create(ctx, t.map_ka.data(), t.ka_w, t.ka_h, t.ka_comp, &d.map_ka)))
      return rc;
    if ((rc = dh_material_create(ctx, &d, &mats[i]))) return rc;
  }
  for (size_t i = 0; i < o->meshes.size(); i++) {
    const ObjMesh& m = o->meshes[i];
    rc = dh_mesh_create(ctx, *object, m.tri.data(), (uint32_t)(m.tri.size() / 9), mats[m.texture_idx], &meshes[i]);
    if (rc) return rc;
  }
  return DH_OK;
}
}  // namespace

extern "C" int dh_obj_free(dh_obj o) {
  if (!o) return DH_E_ARG;
  delete o;
  return DH_OK;
}
