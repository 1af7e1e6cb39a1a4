// th_pack.cpp -- host side of libtoaster_hip: scene packing, the median-split BVH
// builder, camera, light list, environment CDF and the host display conversion.
//
// Reference: crates/toaster-gpu/src/scene_upload.rs:60-475 (scene_to_gpu),
// crates/toaster-bvh (median split + depth-first flatten), crates/toaster-core/src/
// {camera.rs, light.rs, color.rs}, crates/toaster-scene/src/environment.rs:245-278.
// The packed arrays are byte-identical to what scene_to_gpu produces: the BVH
// node/primitive order decides the traversal order and therefore which of two
// equal-distance primitives the reference reports, so the builder reproduces
// toaster-bvh's result exactly (stable sort on the split axis, split at n/2, left
// subtree first) -- it only runs the two halves of large nodes on separate threads.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <future>
#include <thread>
#include <memory>
#include <string>
#include <vector>

#include "th_common.h"

namespace {
thread_local std::string g_last_error = "";
}

int32_t th_set_error(int32_t code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

extern "C" const char* th_last_error(void) { return g_last_error.c_str(); }

namespace {

using th::v3;

v3 ld3(const float* p) { return th::mk(p[0], p[1], p[2]); }
void st4(float* d, v3 v, float w) {
  d[0] = v.x;
  d[1] = v.y;
  d[2] = v.z;
  d[3] = w;
}
// glam Vec3::length / normalize (v * (1 / length)) as used by the reference host code.
float vlen(v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
v3 vnormalize(v3 a) {
  float r = 1.0f / vlen(a);
  return th::mk(a.x * r, a.y * r, a.z * r);
}
float lmin(float a, float b) { return b < a ? b : a; }
float lmax(float a, float b) { return a < b ? b : a; }

// ------------------------------------------------------------------ BVH builder
struct Prim {
  float mn[3], mx[3], c[3];
  uint32_t key[3];  // f32::total_cmp order of c[axis]
  TH_GpuPrimitiveRef ref;
};

uint32_t order_key(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

struct Flat {  // flat node with indices local to its vector
  float mn[3], mx[3];
  uint32_t first_or_right, count;
};

class BvhBuilder {
 public:
  explicit BvhBuilder(std::vector<Prim>& p) : p_(p) {}

  // Builds [lo, hi) into a pre-order node vector whose indices start at 0.
  std::vector<Flat> build(size_t lo, size_t hi, int par_depth) {
    std::vector<Flat> out;
    if (hi <= lo) return out;
    if (par_depth > 0 && hi - lo >= kParallelMin) {
      Flat self;
      size_t mid;
      if (!split(lo, hi, &self, &mid)) {
        out.push_back(self);
        return out;
      }
      auto left_f = std::async(std::launch::async, [&] { return build(lo, mid, par_depth - 1); });
      std::vector<Flat> right = build(mid, hi, par_depth - 1);
      std::vector<Flat> left = left_f.get();
      out.reserve(1 + left.size() + right.size());
      self.first_or_right = (uint32_t)(1 + left.size());
      self.count = 0;
      out.push_back(self);
      for (Flat f : left) {
        if (f.count == 0) f.first_or_right += 1;
        out.push_back(f);
      }
      uint32_t off = (uint32_t)(1 + left.size());
      for (Flat f : right) {
        if (f.count == 0) f.first_or_right += off;
        out.push_back(f);
      }
      return out;
    }
    serial(lo, hi, out);
    return out;
  }

 private:
  static constexpr size_t kParallelMin = 16384;
  static constexpr size_t kMaxLeaf = 4;  // bvh.rs:7
  std::vector<Prim>& p_;

  void bounds_of(size_t lo, size_t hi, Flat* f) const {  // union of primitive bounds
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = lo; i < hi; ++i)
      for (int a = 0; a < 3; ++a) {
        mn[a] = lmin(mn[a], p_[i].mn[a]);
        mx[a] = lmax(mx[a], p_[i].mx[a]);
      }
    for (int a = 0; a < 3; ++a) {
      f->mn[a] = mn[a];
      f->mx[a] = mx[a];
    }
  }

  // Fills the node for [lo, hi); returns false for a leaf (first = lo: leaf ranges are
  // the final primitive order), true with the split position for an interior node.
  bool split(size_t lo, size_t hi, Flat* node, size_t* mid) {
    bounds_of(lo, hi, node);
    size_t n = hi - lo;
    if (n <= kMaxLeaf) {  // bvh.rs:107-109
      node->first_or_right = (uint32_t)lo;
      node->count = (uint32_t)n;
      return false;
    }
    float cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = lo; i < hi; ++i)
      for (int a = 0; a < 3; ++a) {
        cmn[a] = lmin(cmn[a], p_[i].c[a]);
        cmx[a] = lmax(cmx[a], p_[i].c[a]);
      }
    float e[3] = {cmx[0] - cmn[0], cmx[1] - cmn[1], cmx[2] - cmn[2]};
    int axis = (e[0] >= e[1] && e[0] >= e[2]) ? 0 : (e[1] >= e[2] ? 1 : 2);  // aabb.rs:62-71
    if (e[axis] <= 1.1920929e-7f) {  // bvh.rs:119-121: degenerate centroids -> one leaf
      node->first_or_right = (uint32_t)lo;
      node->count = (uint32_t)n;
      return false;
    }
    std::stable_sort(p_.begin() + lo, p_.begin() + hi,
                     [axis](const Prim& a, const Prim& b) { return a.key[axis] < b.key[axis]; });
    *mid = lo + n / 2;
    return true;
  }

  // Pre-order emission: a node's index is reserved before its left subtree is built.
  void serial(size_t lo, size_t hi, std::vector<Flat>& out) {
    size_t me = out.size();
    out.push_back(Flat{});
    Flat node;
    size_t mid;
    if (!split(lo, hi, &node, &mid)) {
      out[me] = node;
      return;
    }
    serial(lo, mid, out);
    node.first_or_right = (uint32_t)out.size();
    node.count = 0;
    serial(mid, hi, out);
    out[me] = node;
  }
};

// Primitive records of build_gpu_bvh (scene_upload.rs:296-346): spheres first, then
// triangles with bounds padded by 1e-5.
void collect_prims(const TH_GpuSphere* sph, uint32_t ns, const TH_GpuTriangle* tri, uint32_t nt,
                   std::vector<Prim>& out) {
  out.resize((size_t)ns + nt);
  for (uint32_t i = 0; i < ns; ++i) {
    Prim& q = out[i];
    const float* cr = sph[i].center_radius;
    for (int a = 0; a < 3; ++a) {
      q.mn[a] = cr[a] - cr[3];
      q.mx[a] = cr[a] + cr[3];
    }
    q.ref = TH_GpuPrimitiveRef{0u, i};
  }
  for (uint32_t i = 0; i < nt; ++i) {
    Prim& q = out[(size_t)ns + i];
    const TH_GpuTriangle& t = tri[i];
    for (int a = 0; a < 3; ++a) {
      q.mn[a] = lmin(lmin(t.v0[a], t.v1[a]), t.v2[a]) - 1e-5f;
      q.mx[a] = lmax(lmax(t.v0[a], t.v1[a]), t.v2[a]) + 1e-5f;
    }
    q.ref = TH_GpuPrimitiveRef{1u, i};
  }
  for (Prim& q : out)
    for (int a = 0; a < 3; ++a) {
      q.c[a] = 0.5f * (q.mn[a] + q.mx[a]);
      q.key[a] = order_key(q.c[a]);
    }
}

void build_bvh(const TH_GpuSphere* sph, uint32_t ns, const TH_GpuTriangle* tri, uint32_t nt,
               std::vector<TH_GpuBvhNode>& nodes, std::vector<TH_GpuPrimitiveRef>& prims) {
  std::vector<Prim> p;
  collect_prims(sph, ns, tri, nt, p);
  unsigned hw = std::thread::hardware_concurrency();
  int par = hw >= 8 ? 3 : (hw >= 4 ? 2 : (hw >= 2 ? 1 : 0));
  BvhBuilder b(p);
  std::vector<Flat> flat = b.build(0, p.size(), par);
  nodes.resize(flat.size());
  for (size_t i = 0; i < flat.size(); ++i) {
    TH_GpuBvhNode& g = nodes[i];
    memset(&g, 0, sizeof g);
    for (int a = 0; a < 3; ++a) {
      g.min[a] = flat[i].mn[a];
      g.max[a] = flat[i].mx[a];
    }
    g.first_or_right = flat[i].first_or_right;
    g.primitive_count = flat[i].count;
  }
  prims.resize(p.size());
  for (size_t i = 0; i < p.size(); ++i) prims[i] = p[i].ref;
}

// ------------------------------------------------------------------ environment CDF
// environment.rs:245-278: weight = Y(rgb) * sin(theta_row) * intensity in f64, one
// flat row-major CDF, cast to f32, last entry forced to 1.
void build_env_cdf(uint32_t w, uint32_t h, const float* rgb, float intensity, float* cdf_out,
                   size_t stride) {
  size_t n = (size_t)w * h;
  std::vector<double> wt(n);
  double total = 0.0;
  const double pi32 = (double)3.14159274f;  // std::f32::consts::PI widened
  for (size_t y = 0; y < h; ++y) {
    double st = sin(pi32 * ((double)y + 0.5) / (double)h);
    for (size_t x = 0; x < w; ++x) {
      size_t i = y * w + x;
      double lum = 0.2126 * (double)rgb[3 * i] + 0.7152 * (double)rgb[3 * i + 1] +
                   0.0722 * (double)rgb[3 * i + 2];
      wt[i] = lum * st * (double)intensity;
      total += wt[i];
    }
  }
  if (!(total > 0.0) || !std::isfinite(total)) {
    for (size_t i = 0; i < n; ++i) cdf_out[i * stride] = 0.0f;
    return;
  }
  double acc = 0.0;
  for (size_t i = 0; i < n; ++i) {
    acc += wt[i] / total;
    cdf_out[i * stride] = (float)acc;
  }
  if (n) cdf_out[(n - 1) * stride] = 1.0f;
}

}  // namespace

struct th_packed_scene {
  th_scene_desc desc;
  std::vector<TH_GpuSphere> spheres;
  std::vector<TH_GpuTriangle> triangles;
  std::vector<TH_GpuTriangleAttributes> attributes;
  std::vector<TH_GpuMaterial> materials;
  std::vector<TH_GpuLight> lights;
  std::vector<TH_GpuBvhNode> nodes;
  std::vector<TH_GpuPrimitiveRef> prims;
  std::vector<uint32_t> texels;
  std::vector<float> env;
};

void th_srgb_table(float out[256]) {
  for (int b = 0; b < 256; ++b) {
    float v = (float)b * (1.0f / 255.0f);
    out[b] = v <= 0.04045f ? v / 12.92f : powf((v + 0.055f) / 1.055f, 2.4f);
  }
}

static float host_linear_to_srgb(float c) {  // color.rs:32-39
  if (c <= 0.0031308f) return 12.92f * c;
  return 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
}

void th_srgb_thresholds(float out[256]) {
  out[0] = 0.0f;
  for (int k = 1; k < 256; ++k) {
    // smallest v in [0, 1] with th_u8(linear_to_srgb(v)) >= k, by bisection on the
    // binary32 bit pattern (monotone for non-negative floats)
    uint32_t lo = 0, hi = 0x3f800000u;  // [0, 1.0]
    float one = 1.0f;
    if (th_u8(host_linear_to_srgb(one)) < k) {
      out[k] = 2.0f;  // unreachable byte
      continue;
    }
    while (lo < hi) {
      uint32_t mid = lo + (hi - lo) / 2;
      float v;
      memcpy(&v, &mid, 4);
      if (th_u8(host_linear_to_srgb(v)) >= k)
        hi = mid;
      else
        lo = mid + 1;
    }
    memcpy(&out[k], &lo, 4);
  }
}

extern "C" {

int32_t th_make_camera(const float position[3], const float look_at[3], const float up[3],
                       float fov, float aspect, TH_GpuCamera* out) {
  if (!position || !look_at || !up || !out) return th_set_error(TH_E_INVALID, "null argument");
  v3 pos = ld3(position), look = ld3(look_at), u = ld3(up);
  // scene_upload.rs:352-368 (== camera.rs:61-98)
  v3 backward = vnormalize(th::sub(pos, look));
  v3 right = vnormalize(th::cross(u, backward));
  v3 true_up = th::cross(backward, right);
  float vh = 2.0f * tanf(0.5f * (fov * (3.14159274f / 180.0f)));
  float vw = aspect * vh;
  v3 hor = th::smul(vw, right), ver = th::smul(vh, true_up);
  v3 ll = th::sub(th::sub(th::sub(pos, th::muls(hor, 0.5f)), th::muls(ver, 0.5f)), backward);
  st4(out->origin, pos, 0.0f);
  st4(out->lower_left_corner, ll, 0.0f);
  st4(out->horizontal, hor, 0.0f);
  st4(out->vertical, ver, 0.0f);
  return TH_OK;
}

int32_t th_build_bvh(const TH_GpuSphere* spheres, uint32_t ns, const TH_GpuTriangle* tris,
                     uint32_t nt, TH_GpuBvhNode* nodes_out, uint32_t node_capacity,
                     uint32_t* node_count, TH_GpuPrimitiveRef* prims_out, uint32_t prim_capacity) {
  if ((ns && !spheres) || (nt && !tris) || !nodes_out || !node_count || !prims_out)
    return th_set_error(TH_E_INVALID, "null argument");
  if (prim_capacity < ns + nt) return th_set_error(TH_E_CAPACITY, "primitive capacity too small");
  std::vector<TH_GpuBvhNode> nodes;
  std::vector<TH_GpuPrimitiveRef> prims;
  build_bvh(spheres, ns, tris, nt, nodes, prims);
  if (nodes.size() > node_capacity) return th_set_error(TH_E_CAPACITY, "node capacity too small");
  memcpy(nodes_out, nodes.data(), nodes.size() * sizeof(TH_GpuBvhNode));
  memcpy(prims_out, prims.data(), prims.size() * sizeof(TH_GpuPrimitiveRef));
  *node_count = (uint32_t)nodes.size();
  return TH_OK;
}

int32_t th_pack_scene(const th_scene_input* in, int32_t flags, th_packed_scene** out) {
  if (!in || !out) return th_set_error(TH_E_INVALID, "null argument");
  *out = nullptr;
  if (flags & ~(1 | TH_PACK_NO_BVH)) return th_set_error(TH_E_INVALID, "unknown packing flags");
  const bool include_static = (flags & 1) != 0, with_bvh = (flags & TH_PACK_NO_BVH) == 0;
  if (in->sphere_count == 0 && in->triangle_count == 0)
    return th_set_error(TH_E_INVALID, "GPU renderer requires at least one object");
  if (in->material_count == 0 || !in->materials)
    return th_set_error(TH_E_INVALID, "GPU renderer requires at least one material");
  if ((in->sphere_count && !in->spheres) || (in->triangle_count && !in->triangles))
    return th_set_error(TH_E_INVALID, "null geometry array");
  if (in->width == 0 || in->height == 0)
    return th_set_error(TH_E_INVALID, "render dimensions must be nonzero");
  if (in->background > TH_BG_ENVIRONMENT) return th_set_error(TH_E_INVALID, "unknown background");
  if (in->background == TH_BG_ENVIRONMENT && !in->environment)
    return th_set_error(TH_E_INVALID, "environment background requires an environment map");
  if (in->texture_count && !in->textures) return th_set_error(TH_E_INVALID, "null texture array");
  std::unique_ptr<th_packed_scene> s(new th_packed_scene());
  const uint32_t ns = in->sphere_count, nt = in->triangle_count, nm = in->material_count;

  s->spheres.resize(ns);
  for (uint32_t i = 0; i < ns; ++i) {
    const th_sphere& q = in->spheres[i];
    if (q.material_index >= nm) return th_set_error(TH_E_INVALID, "sphere material index out of range");
    TH_GpuSphere& g = s->spheres[i];
    memset(&g, 0, sizeof g);
    g.center_radius[0] = q.center[0];
    g.center_radius[1] = q.center[1];
    g.center_radius[2] = q.center[2];
    g.center_radius[3] = q.radius;
    g.material_index = q.material_index;
  }
  s->triangles.resize(nt);
  s->attributes.resize(nt);
  for (uint32_t i = 0; i < nt; ++i) {
    const th_triangle& q = in->triangles[i];
    if (q.material_index >= nm) return th_set_error(TH_E_INVALID, "triangle material index out of range");
    TH_GpuTriangle& g = s->triangles[i];
    memset(&g, 0, sizeof g);
    for (int a = 0; a < 3; ++a) {
      g.v0[a] = q.v[0][a];
      g.v1[a] = q.v[1][a];
      g.v2[a] = q.v[2][a];
    }
    g.material_index = q.material_index;
    TH_GpuTriangleAttributes& at = s->attributes[i];
    memset(&at, 0, sizeof at);
    if (in->triangle_attributes) {
      const th_triangle_attributes& qa = in->triangle_attributes[i];
      if (qa.has_normals) {
        at.flags |= 1u;
        for (int a = 0; a < 3; ++a) {
          at.n0[a] = qa.normals[0][a];
          at.n1[a] = qa.normals[1][a];
          at.n2[a] = qa.normals[2][a];
        }
      }
      if (qa.has_tex_coords) {
        at.flags |= 2u;
        for (int a = 0; a < 2; ++a) {
          at.uv0[a] = qa.tex_coords[0][a];
          at.uv1[a] = qa.tex_coords[1][a];
          at.uv2[a] = qa.tex_coords[2][a];
        }
      }
    }
  }
  // texture atlas metadata and texels (scene_upload.rs:168-188)
  std::vector<uint32_t> toff(in->texture_count);
  uint64_t texel_total = 0;
  for (uint32_t t = 0; t < in->texture_count; ++t) {
    const th_texture& tx = in->textures[t];
    uint64_t n = (uint64_t)tx.width * tx.height;
    if (n && !tx.rgba8) return th_set_error(TH_E_INVALID, "null texture data");
    if (texel_total + n > 0xffffffffull) return th_set_error(TH_E_INVALID, "texture atlas is too large");
    toff[t] = (uint32_t)texel_total;
    if (include_static) {
      for (uint64_t i = 0; i < n; ++i) {
        const uint8_t* b = tx.rgba8 + 4 * i;
        s->texels.push_back((uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) |
                            ((uint32_t)b[3] << 24));
      }
    }
    texel_total += n;
  }
  // materials (scene_upload.rs:371-404)
  s->materials.resize(nm);
  for (uint32_t i = 0; i < nm; ++i) {
    const th_material& m = in->materials[i];
    TH_GpuMaterial& g = s->materials[i];
    memset(&g, 0, sizeof g);
    v3 albedo = ld3(m.color);
    if (m.kind == TH_MAT_DIFFUSE) {
      g.kind = 0;
    } else if (m.kind == TH_MAT_TEXTURED_DIFFUSE) {
      if (m.texture_index >= in->texture_count)
        return th_set_error(TH_E_INVALID, "material references an unknown texture");
      g.kind = 0;
      g.texture_offset = toff[m.texture_index];
      g.texture_width = in->textures[m.texture_index].width;
      g.texture_height = in->textures[m.texture_index].height;
    } else if (m.kind == TH_MAT_METAL) {
      g.kind = 1;
      g.params[0] = m.roughness;
    } else if (m.kind == TH_MAT_DIELECTRIC) {
      g.kind = 2;
      albedo = th::mk(1.0f, 1.0f, 1.0f);
      g.params[1] = m.ior;
    } else if (m.kind == TH_MAT_EMISSIVE) {
      g.kind = 3;
      g.params[2] = m.strength;
    } else {
      return th_set_error(TH_E_INVALID, "unknown material kind");
    }
    st4(g.albedo, albedo, 1.0f);
  }
  // power-weighted emitters (scene_upload.rs:412-475): triangles, then spheres
  float cumulative = 0.0f;
  auto luminance = [](const float* c) { return 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2]; };
  for (uint32_t i = 0; i < nt; ++i) {
    const th_triangle& q = in->triangles[i];
    const th_material& m = in->materials[q.material_index];
    if (m.kind != TH_MAT_EMISSIVE) continue;
    v3 a = ld3(q.v[0]), b = ld3(q.v[1]), c = ld3(q.v[2]);
    float area = 0.5f * vlen(th::cross(th::sub(b, a), th::sub(c, a)));
    float weight = area * luminance(m.color) * m.strength;  // light.rs:11-13
    if (weight <= 0.0f) continue;
    cumulative += weight;
    TH_GpuLight L;
    memset(&L, 0, sizeof L);
    L.kind = 0;
    L.material_index = q.material_index;
    st4(L.v0, a, 0.0f);
    st4(L.v1, b, 0.0f);
    st4(L.v2, c, 0.0f);
    L.area_weight_cumulative[0] = area;
    L.area_weight_cumulative[1] = weight;
    L.area_weight_cumulative[2] = cumulative;
    s->lights.push_back(L);
  }
  for (uint32_t i = 0; i < ns; ++i) {
    const th_sphere& q = in->spheres[i];
    const th_material& m = in->materials[q.material_index];
    if (m.kind != TH_MAT_EMISSIVE) continue;
    float area = 4.0f * 3.14159274f * q.radius * q.radius;
    float weight = area * luminance(m.color) * m.strength;
    if (weight <= 0.0f) continue;
    cumulative += weight;
    TH_GpuLight L;
    memset(&L, 0, sizeof L);
    L.kind = 1;
    L.material_index = q.material_index;
    L.center_radius[0] = q.center[0];
    L.center_radius[1] = q.center[1];
    L.center_radius[2] = q.center[2];
    L.center_radius[3] = q.radius;
    L.area_weight_cumulative[0] = area;
    L.area_weight_cumulative[1] = weight;
    L.area_weight_cumulative[2] = cumulative;
    s->lights.push_back(L);
  }
  if (with_bvh) build_bvh(s->spheres.data(), ns, s->triangles.data(), nt, s->nodes, s->prims);
  // environment (scene_upload.rs:214-246, environment.rs:42-88)
  uint32_t ew = 0, eh = 0;
  float eint = 0.0f, erot = 0.0f;
  if (in->environment) {
    const th_environment& e = *in->environment;
    if (e.width == 0 || e.height == 0 || !e.rgb)
      return th_set_error(TH_E_INVALID, "environment dimensions must be nonzero");
    if (!(e.intensity >= 0.0f) || !std::isfinite(e.intensity))
      return th_set_error(TH_E_INVALID, "environment intensity must be finite and non-negative");
    if (!std::isfinite(e.rotation_degrees))
      return th_set_error(TH_E_INVALID, "environment rotation_degrees must be finite");
    size_t n = (size_t)e.width * e.height;
    for (size_t i = 0; i < 3 * n; ++i)
      if (!(e.rgb[i] >= 0.0f) || !std::isfinite(e.rgb[i]))
        return th_set_error(TH_E_INVALID, "environment pixels must be finite and non-negative");
    float rot = fmodf(e.rotation_degrees, 360.0f);  // rem_euclid
    if (rot < 0.0f) rot += 360.0f;
    if (in->background == TH_BG_ENVIRONMENT) {
      ew = e.width;
      eh = e.height;
      eint = e.intensity;
      erot = rot;
    }
    if (include_static) {
      s->env.resize(4 * n);
      for (size_t i = 0; i < n; ++i) {
        s->env[4 * i] = e.rgb[3 * i];
        s->env[4 * i + 1] = e.rgb[3 * i + 1];
        s->env[4 * i + 2] = e.rgb[3 * i + 2];
      }
      build_env_cdf(e.width, e.height, e.rgb, e.intensity, s->env.data() + 3, 4);
    }
  }
  th_scene_desc& d = s->desc;
  memset(&d, 0, sizeof d);
  TH_GpuRenderParams& P = d.params;
  P.width = in->width;
  P.height = in->height;
  P.samples = in->samples;
  P.max_bounces = in->max_bounces;
  P.sphere_count = ns;
  P.triangle_count = nt;
  P.material_count = nm;
  P.background_kind = in->background;
  P.light_count = (uint32_t)s->lights.size();
  P.total_light_weight = cumulative;
  P.environment_width = ew;
  P.environment_height = eh;
  P.environment_intensity = eint;
  P.environment_rotation_degrees = erot;
  P.bvh_node_count = (uint32_t)s->nodes.size();
  P.bvh_primitive_count = (uint32_t)s->prims.size();
  float aspect = (float)in->width / (float)in->height;
  th_make_camera(in->camera_position, in->camera_look_at, in->camera_up, in->fov_degrees, aspect,
                 &d.camera);
  *out = s.release();
  return TH_OK;
}

int32_t th_packed_view(const th_packed_scene* s, th_scene_desc* d) {
  if (!s || !d) return th_set_error(TH_E_INVALID, "null argument");
  *d = s->desc;
  d->spheres = s->spheres.data();
  d->triangles = s->triangles.data();
  d->triangle_attributes = s->attributes.data();
  d->materials = s->materials.data();
  d->lights = s->lights.data();
  d->bvh_nodes = s->nodes.data();
  d->bvh_primitives = s->prims.data();
  d->texture_pixels = s->texels.data();
  d->environment_pixels = s->env.data();
  d->sphere_count = (uint32_t)s->spheres.size();
  d->triangle_count = (uint32_t)s->triangles.size();
  d->material_count = (uint32_t)s->materials.size();
  d->light_count = (uint32_t)s->lights.size();
  d->bvh_node_count = (uint32_t)s->nodes.size();
  d->bvh_primitive_count = (uint32_t)s->prims.size();
  d->texture_pixel_count = (uint32_t)s->texels.size();
  d->environment_pixel_count = (uint32_t)(s->env.size() / 4);
  return TH_OK;
}

void th_packed_free(th_packed_scene* s) { delete s; }

int32_t th_display_rgba8(const float* rgba, uint32_t width, uint32_t height, float exposure_stops,
                         uint8_t* out, size_t out_bytes) {
  if (!rgba || !out) return th_set_error(TH_E_INVALID, "null argument");
  size_t n = (size_t)width * height;
  if (out_bytes != 4 * n) return th_set_error(TH_E_INVALID, "output size must be width*height*4");
  float scale = exp2f(exposure_stops);  // color.rs:23-25
  for (size_t i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) {
      float v = th_aces(rgba[4 * i + c] * scale);
      v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
      out[4 * i + c] = th_u8(host_linear_to_srgb(v));
    }
    out[4 * i + 3] = th_u8(rgba[4 * i + 3]);
  }
  return TH_OK;
}

}  // extern "C"
