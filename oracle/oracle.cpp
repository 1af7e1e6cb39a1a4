// oracle.cpp -- CPU restatement of the reference path-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY: the checker of the HIP library.  Only tests/, the smoke
// check and the CPU-baseline leg of bench.py load liboracle.so; the shipped library
// never links it.
//
// Each function names the reference lines it restates.  The estimator is the
// reference GPU kernel's (shaders/pathtrace_spheres.wgsl): its per-pixel PCG stream,
// its draw order and its formulas, written as plain sequential C++ in the structure of
// toaster-cpu (crates/toaster-cpu/src/{integrator,intersect,light}.rs).  Scalar math
// (sqrt, sin, cos, atan2, acos, the PCG hash) comes from th_math.h so that the same
// binary32 operation sequence runs here and on the GPU.
//
// Pinning (tests/test_oracle_*.py): every known-answer test the reference holds for
// this path is restated there against these functions; closed forms (analytic sphere
// and plane hits, white furnace, CDF/pdf invariants, brute-force BVH equivalence)
// cover what the reference's tests leave open.
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../toaster_amd/csrc/th_math.h"

using th::v3;

namespace {

thread_local std::string g_error;
int32_t fail(const std::string& msg) {
  g_error = msg;
  return TH_E_INVALID;
}

// ============================================================ glam-style host ops
// Host-side packing runs in the reference on glam (Vec3 scalar path); normalize there
// is v * (1 / length) (glam Vec3::normalize -> length_recip).
v3 glam_normalize(v3 a) {
  float len = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
  float r = 1.0f / len;
  return th::mk(a.x * r, a.y * r, a.z * r);
}
float glam_length(v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
v3 v3_of(const float* p) { return th::mk(p[0], p[1], p[2]); }
void store4(float* dst, v3 a, float w) {
  dst[0] = a.x;
  dst[1] = a.y;
  dst[2] = a.z;
  dst[3] = w;
}

// ============================================================ BVH (toaster-bvh)
struct Aabb {  // aabb.rs:7-12
  v3 mn, mx;
};
Aabb aabb_empty() {  // aabb.rs:21-26
  const float inf = INFINITY;
  return Aabb{th::mk(inf, inf, inf), th::mk(-inf, -inf, -inf)};
}
float rmin(float a, float b) { return b < a ? b : a; }  // f32::min on finite values
float rmax(float a, float b) { return a < b ? b : a; }
Aabb aabb_union(Aabb a, Aabb b) {  // aabb.rs:29-34
  return Aabb{th::mk(rmin(a.mn.x, b.mn.x), rmin(a.mn.y, b.mn.y), rmin(a.mn.z, b.mn.z)),
              th::mk(rmax(a.mx.x, b.mx.x), rmax(a.mx.y, b.mx.y), rmax(a.mx.z, b.mx.z))};
}
v3 aabb_centroid(Aabb a) {  // aabb.rs:46-48: 0.5 * (min + max)
  return th::mk(0.5f * (a.mn.x + a.mx.x), 0.5f * (a.mn.y + a.mx.y), 0.5f * (a.mn.z + a.mx.z));
}
int longest_axis(Aabb a) {  // aabb.rs:62-71
  v3 e = th::sub(a.mx, a.mn);
  if (e.x >= e.y && e.x >= e.z) return 0;
  if (e.y >= e.z) return 1;
  return 2;
}

struct PrimInfo {  // bvh.rs:27-35
  Aabb bounds;
  v3 centroid;
  uint32_t kind;  // 0 sphere, 1 triangle (gpu encoding)
  uint32_t index;
};

// f32::total_cmp ordering as an unsigned key.
uint32_t total_key(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

struct TreeNode {  // bvh.rs:39-58
  Aabb bounds;
  int left = -1, right = -1;
  uint32_t first = 0, count = 0;  // leaf when count > 0
};

struct TreeBuilder {
  std::vector<TreeNode> nodes;
  std::vector<TH_GpuPrimitiveRef> ordered;
  size_t max_leaf = 4;  // bvh.rs:7 DEFAULT_MAX_LEAF_PRIMITIVES

  int create_leaf(Aabb bounds, PrimInfo* p, size_t n) {  // bvh.rs:141-154
    TreeNode node;
    node.bounds = bounds;
    node.first = (uint32_t)ordered.size();
    node.count = (uint32_t)n;
    for (size_t i = 0; i < n; ++i) ordered.push_back(TH_GpuPrimitiveRef{p[i].kind, p[i].index});
    nodes.push_back(node);
    return (int)nodes.size() - 1;
  }

  // bvh.rs:92-139 build_recursive
  int build(PrimInfo* p, size_t n) {
    if (n == 0) return -1;
    Aabb bounds = aabb_empty();
    for (size_t i = 0; i < n; ++i) bounds = aabb_union(bounds, p[i].bounds);
    if (n <= max_leaf) return create_leaf(bounds, p, n);
    Aabb cb = aabb_empty();
    for (size_t i = 0; i < n; ++i) cb = aabb_union(cb, Aabb{p[i].centroid, p[i].centroid});
    int axis = longest_axis(cb);
    float extent = th::get(th::sub(cb.mx, cb.mn), axis);
    if (extent <= 1.1920929e-7f) return create_leaf(bounds, p, n);  // f32::EPSILON
    // Rust sort_by is stable; ties keep the slice order.
    std::stable_sort(p, p + n, [axis](const PrimInfo& a, const PrimInfo& b) {
      return total_key(th::get(a.centroid, axis)) < total_key(th::get(b.centroid, axis));
    });
    size_t split = n / 2;
    int left = build(p, split);
    int right = build(p + split, n - split);
    TreeNode node;
    node.bounds = bounds;
    node.left = left;
    node.right = right;
    nodes.push_back(node);
    return (int)nodes.size() - 1;
  }

  // flatten.rs:72-97 depth-first flatten: left child follows its parent.
  uint32_t flatten(int idx, std::vector<TH_GpuBvhNode>& out) {
    uint32_t node_index = (uint32_t)out.size();
    TH_GpuBvhNode g;
    memset(&g, 0, sizeof(g));
    const TreeNode& t = nodes[idx];
    store4(g.min, t.bounds.mn, 0.0f);
    store4(g.max, t.bounds.mx, 0.0f);
    out.push_back(g);
    if (t.count > 0) {
      out[node_index].first_or_right = t.first;
      out[node_index].primitive_count = t.count;
    } else {
      flatten(t.left, out);
      uint32_t right = flatten(t.right, out);
      out[node_index].first_or_right = right;
    }
    return node_index;
  }
};

// ============================================================ packed scene
}  // namespace

struct or_scene {
  TH_GpuRenderParams params;
  TH_GpuCamera camera;
  std::vector<TH_GpuSphere> spheres;
  std::vector<TH_GpuTriangle> triangles;
  std::vector<TH_GpuTriangleAttributes> attributes;
  std::vector<TH_GpuMaterial> materials;
  std::vector<TH_GpuLight> lights;
  std::vector<TH_GpuBvhNode> nodes;
  std::vector<TH_GpuPrimitiveRef> prims;
  std::vector<uint32_t> texture_pixels;
  std::vector<float> env;  // float4 per texel
};

namespace {

float rec709(v3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }  // light.rs:6-8

// scene_upload.rs:352-368 make_camera (identical to camera.rs:61-98).
TH_GpuCamera make_camera(v3 position, v3 look_at, v3 up, float fov, float aspect) {
  v3 backward = glam_normalize(th::sub(position, look_at));
  v3 right = glam_normalize(th::cross(up, backward));
  v3 true_up = th::cross(backward, right);
  float to_rad = fov * (3.14159274f / 180.0f);  // f32::to_radians
  float viewport_height = 2.0f * tanf(0.5f * to_rad);
  float viewport_width = aspect * viewport_height;
  v3 horizontal = th::smul(viewport_width, right);
  v3 vertical = th::smul(viewport_height, true_up);
  v3 lower_left = th::sub(th::sub(th::sub(position, th::muls(horizontal, 0.5f)),
                                  th::muls(vertical, 0.5f)),
                          backward);
  TH_GpuCamera cam;
  store4(cam.origin, position, 0.0f);
  store4(cam.lower_left_corner, lower_left, 0.0f);
  store4(cam.horizontal, horizontal, 0.0f);
  store4(cam.vertical, vertical, 0.0f);
  return cam;
}

// environment.rs:245-278 build_importance_distribution (f64 accumulation).
std::vector<float> env_cdf(uint32_t w, uint32_t h, const float* rgb, float intensity) {
  size_t n = (size_t)w * h;
  std::vector<double> weights(n);
  double total = 0.0;
  const double pi = (double)3.14159274f;  // std::f32::consts::PI as f64
  for (size_t i = 0; i < n; ++i) {
    size_t y = i / w;
    double theta = pi * ((double)y + 0.5) / (double)h;
    double lum = 0.2126 * (double)rgb[3 * i] + 0.7152 * (double)rgb[3 * i + 1] +
                 0.0722 * (double)rgb[3 * i + 2];
    double weight = lum * sin(theta) * (double)intensity;
    weights[i] = weight;
    total += weight;
  }
  std::vector<float> cdf(n, 0.0f);
  if (!(total > 0.0) || !isfinite(total)) return cdf;
  double acc = 0.0;
  for (size_t i = 0; i < n; ++i) {
    acc += weights[i] / total;
    cdf[i] = (float)acc;
  }
  if (n) cdf[n - 1] = 1.0f;
  return cdf;
}

// ============================================================ intersection (WGSL)
struct Ray {
  v3 o, d;
};
struct Hit {
  float t = TH_FLT_MAX;
  v3 p{0, 0, 0}, n{0, 0, 0};
  bool front = false;
  uint32_t mat = 0;
  float tu = 0, tv = 0;
  uint32_t slot = 0xffffffffu;
};

// wgsl:551-596 hit_sphere
Hit hit_sphere(const Ray& r, const TH_GpuSphere& s) {
  Hit no;
  v3 c = v3_of(s.center_radius);
  float rad = s.center_radius[3];
  v3 offset = th::sub(r.o, c);
  float a = th::dot(r.d, r.d);
  float half_b = th::dot(offset, r.d);
  float cc = th::dot(offset, offset) - rad * rad;
  float disc = half_b * half_b - a * cc;
  if (disc < 0.0f) return no;
  float sq = th::fsqrt(disc);
  float t = (-half_b - sq) / a;
  if (t < TH_MIN_DISTANCE) {
    t = (-1.0f * half_b + sq) / a;
    if (t < TH_MIN_DISTANCE) return no;
  }
  Hit h;
  h.t = t;
  h.p = th::add(r.o, th::smul(t, r.d));
  v3 tn = th::divs(th::sub(h.p, c), rad);
  h.front = th::dot(r.d, tn) < 0.0f;
  h.n = h.front ? tn : th::neg(tn);
  h.mat = s.material_index;
  return h;
}

// wgsl:599-664 hit_triangle
Hit hit_triangle(const Ray& r, const TH_GpuTriangle& tri, const TH_GpuTriangleAttributes& at) {
  Hit no;
  v3 v0 = v3_of(tri.v0), v1 = v3_of(tri.v1), v2 = v3_of(tri.v2);
  v3 e1 = th::sub(v1, v0), e2 = th::sub(v2, v0);
  v3 p = th::cross(r.d, e2);
  float det = th::dot(p, e1);
  if (th::fabs_(det) < TH_EPSILON) return no;
  float inv = 1.0f / det;
  v3 oo = th::sub(r.o, v0);
  float u = th::dot(oo, p) * inv;
  if (u < 0.0f || u > 1.0f) return no;
  v3 q = th::cross(oo, e1);
  float v = th::dot(r.d, q) * inv;
  if (v < 0.0f || (u + v) > 1.0f) return no;
  float t = th::dot(e2, q) * inv;
  if (t < TH_MIN_DISTANCE || t > TH_MAX_DISTANCE) return no;
  v3 gn = th::normalize(th::cross(e1, e2));
  bool front = th::dot(r.d, gn) < 0.0f;
  float w0 = 1.0f - u - v;
  v3 outward = gn;
  if (at.flags & 1u) {
    outward = th::normalize(th::add(th::add(th::smul(w0, v3_of(at.n0)), th::smul(u, v3_of(at.n1))),
                                    th::smul(v, v3_of(at.n2))));
    if (th::dot(outward, gn) < 0.0f) outward = th::neg(outward);
  }
  Hit h;
  h.t = t;
  h.p = th::add(r.o, th::smul(t, r.d));
  h.n = front ? outward : th::smul(-1.0f, outward);
  h.front = front;
  h.mat = tri.material_index;
  if (at.flags & 2u) {
    h.tu = (w0 * at.uv0[0] + u * at.uv1[0]) + v * at.uv2[0];
    h.tv = (w0 * at.uv0[1] + u * at.uv1[1]) + v * at.uv2[1];
  }
  return h;
}

// wgsl:700-730 intersects_aabb
bool intersects_aabb(const TH_GpuBvhNode& node, v3 o, v3 d, v3 inv, float tmin, float tmax) {
  for (int axis = 0; axis < 3; ++axis) {
    float da = th::get(d, axis), oa = th::get(o, axis);
    if (th::fabs_(da) < TH_EPSILON) {
      if (oa < node.min[axis] || oa > node.max[axis]) return false;
      continue;
    }
    float ia = th::get(inv, axis);
    float nearv = (node.min[axis] - oa) * ia;
    float farv = (node.max[axis] - oa) * ia;
    tmin = th::fmax_(tmin, th::fmin_(nearv, farv));
    tmax = th::fmin_(tmax, th::fmax_(nearv, farv));
    if (tmax < tmin) return false;
  }
  return true;
}

struct RayCounters {
  uint64_t nodes = 0, primrefs = 0, tris = 0, spheres = 0;
};

// wgsl:667-698 linear fallback
Hit closest_linear(const or_scene& s, const Ray& r, RayCounters* c) {
  Hit rec;
  for (uint32_t i = 0; i < s.params.sphere_count; ++i) {
    Hit h = hit_sphere(r, s.spheres[i]);
    if (c) c->spheres++;
    if (rec.t > h.t) {
      rec = h;
      rec.slot = i;
    }
  }
  for (uint32_t i = 0; i < s.params.triangle_count; ++i) {
    Hit h = hit_triangle(r, s.triangles[i], s.attributes[i]);
    if (c) c->tris++;
    if (rec.t > h.t) {
      rec = h;
      rec.slot = 0x80000000u | i;
    }
  }
  return rec;
}

// wgsl:732-795 stack traversal: pop, test box against [MIN, best], leaf loop with
// strict <, interior pushes left then right (right is popped first), 64 entries.
Hit closest_bvh(const or_scene& s, const Ray& r, RayCounters* c) {
  Hit rec;
  v3 inv = th::mk(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
  uint32_t stack[64];
  uint32_t sp = 1;
  stack[0] = 0;
  while (sp > 0) {
    sp -= 1;
    uint32_t ni = stack[sp];
    const TH_GpuBvhNode& node = s.nodes[ni];
    if (c) c->nodes++;
    if (!intersects_aabb(node, r.o, r.d, inv, TH_MIN_DISTANCE, rec.t)) continue;
    if (node.primitive_count > 0) {
      for (uint32_t i = 0; i < node.primitive_count; ++i) {
        const TH_GpuPrimitiveRef& pr = s.prims[node.first_or_right + i];
        Hit cand;
        if (c) c->primrefs++;
        if (pr.kind == 0u) {
          cand = hit_sphere(r, s.spheres[pr.index]);
          if (c) c->spheres++;
        } else {
          cand = hit_triangle(r, s.triangles[pr.index], s.attributes[pr.index]);
          if (c) c->tris++;
        }
        if (cand.t < rec.t) {
          rec = cand;
          rec.slot = node.first_or_right + i;
        }
      }
    } else if (sp + 2u <= 64u) {
      stack[sp++] = ni + 1u;
      stack[sp++] = node.first_or_right;
    }
  }
  return rec;
}

Hit closest(const or_scene& s, const Ray& r, RayCounters* c) {  // wgsl:797-802
  if (s.params.bvh_node_count == 0u) return closest_linear(s, r, c);
  return closest_bvh(s, r, c);
}

// ============================================================ environment (WGSL)
struct Env {
  const or_scene& s;
  uint32_t w, h;
  float intensity, rot;
};

v3 env_texel(const or_scene& s, uint32_t w, uint32_t h, int x, int y) {  // wgsl:273-277
  int sw = (int)w;
  uint32_t wx = (uint32_t)(((x % sw) + sw) % sw);
  int cy = y < 0 ? 0 : (y > (int)h - 1 ? (int)h - 1 : y);
  const float* p = &s.env[4 * ((size_t)cy * w + wx)];
  return th::mk(p[0], p[1], p[2]);
}

v3 env_lookup(const or_scene& s, const TH_GpuRenderParams& P, v3 dir) {  // wgsl:280-305
  v3 ud = th::normalize(dir);
  float u = th::fract(th::fatan2(ud.z, ud.x) / (2.0f * TH_PI) + 0.5f +
                      P.environment_rotation_degrees / 360.0f);
  float v = th::facos(th::clampf(ud.y, -1.0f, 1.0f)) / TH_PI;
  float px = u * (float)P.environment_width - 0.5f;
  float py = v * (float)P.environment_height - 0.5f;
  int bx = (int)th::ffloor(px), by = (int)th::ffloor(py);
  float ax = th::fract(px), ay = th::fract(py);
  uint32_t W = P.environment_width, H = P.environment_height;
  v3 top = th::mix(env_texel(s, W, H, bx, by), env_texel(s, W, H, bx + 1, by), ax);
  v3 bottom = th::mix(env_texel(s, W, H, bx, by + 1), env_texel(s, W, H, bx + 1, by + 1), ax);
  return th::muls(th::mix(top, bottom, ay), P.environment_intensity);
}

float env_probability(const or_scene& s, uint32_t index) {  // wgsl:308-314
  float prev = 0.0f;
  if (index > 0u) prev = s.env[4 * (size_t)(index - 1u) + 3];
  return th::fmax_(0.0f, s.env[4 * (size_t)index + 3] - prev);
}

float env_pdf(const or_scene& s, const TH_GpuRenderParams& P, v3 dir) {  // wgsl:317-335
  v3 ud = th::normalize(dir);
  float u = th::fract(th::fatan2(ud.z, ud.x) / (2.0f * TH_PI) + 0.5f +
                      P.environment_rotation_degrees / 360.0f);
  float theta = th::facos(th::clampf(ud.y, -1.0f, 1.0f));
  float v = theta / TH_PI;
  uint32_t W = P.environment_width, H = P.environment_height;
  uint32_t x = std::min((uint32_t)th::ffloor(u * (float)W), W - 1u);
  uint32_t y = std::min((uint32_t)th::ffloor(v * (float)H), H - 1u);
  float prob = env_probability(s, y * W + x);
  float st = th::fsin(theta);
  if (prob <= 0.0f || st <= 0.0f) return 0.0f;
  return prob * (float)(W * H) / (2.0f * TH_PI * TH_PI * st);
}

struct EnvSample {
  v3 dir{0, 0, 0}, rad{0, 0, 0};
  float pdf = 0.0f;
};

// wgsl:345-381 with the three draws supplied (selection, jitter x, jitter y).
EnvSample env_sample_with(const or_scene& s, const TH_GpuRenderParams& P, float sel_in, float jx_in,
                          float jy_in, uint64_t* search_steps) {
  EnvSample out;
  uint32_t W = P.environment_width, H = P.environment_height;
  uint32_t count = W * H;
  float sel = th::fmin_(sel_in, 0.99999994f);
  uint32_t lo = 0, hi = count;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2u;
    if (search_steps) (*search_steps)++;
    if (s.env[4 * (size_t)mid + 3] <= sel)
      lo = mid + 1u;
    else
      hi = mid;
  }
  uint32_t index = std::min(lo, count - 1u);
  uint32_t x = index % W, y = index / W;
  float jx = th::fmin_(jx_in, 0.99999994f), jy = th::fmin_(jy_in, 0.99999994f);
  float u = ((float)x + jx) / (float)W;
  float v = ((float)y + jy) / (float)H;
  float theta = TH_PI * v;
  float phi = 2.0f * TH_PI * (u - 0.5f - P.environment_rotation_degrees / 360.0f);
  float st = th::fsin(theta);
  out.dir = th::mk(st * th::fcos(phi), th::fcos(theta), st * th::fsin(phi));
  float prob = env_probability(s, index);
  if (prob <= 0.0f || st <= 0.0f) return out;
  out.pdf = prob * (float)count / (2.0f * TH_PI * TH_PI * st);
  out.rad = env_lookup(s, P, out.dir);
  return out;
}

bool env_has_mass(const or_scene& s, const TH_GpuRenderParams& P) {
  uint32_t count = P.environment_width * P.environment_height;
  return !(count == 0u || s.env[4 * (size_t)(count - 1u) + 3] <= 0.0f);
}

float power_heuristic(float a, float b) {  // wgsl:475-483
  float a2 = a * a, b2 = b * b;
  float den = a2 + b2;
  if (den <= 0.0f) return 0.0f;
  return a2 / den;
}

v3 reflect_vec(v3 d, v3 n) { return th::sub(d, th::muls(n, th::dot(n, d) * 2.0f)); }  // wgsl:919

v3 refract_vec(v3 d, v3 n, float ratio) {  // wgsl:924-929
  float ct = th::fmin_(th::dot(th::neg(d), n), 1.0f);
  v3 perp = th::smul(ratio, th::add(d, th::smul(ct, n)));
  v3 par = th::smul(-th::fsqrt(th::fabs_(1.0f - th::dot(perp, perp))), n);
  return th::add(perp, par);
}

float reflectance(float cosine, float ratio) {  // wgsl:932-937
  float r0 = (1.0f - ratio) / (1.0f + ratio);
  r0 = r0 * r0;
  float t = 1.0f - cosine;
  return r0 + (1.0f - r0) * t * t * t * t * t;
}

v3 random_unit_vector(uint32_t* st) {  // wgsl:202-207
  float z = th::rand01(st) * 2.0f - 1.0f;
  float a = th::rand01(st) * 6.2831853f;
  float r = th::fsqrt(th::fmax_(0.0f, 1.0f - z * z));
  return th::mk(r * th::fcos(a), r * th::fsin(a), z);
}

struct Tracer {
  const or_scene& s;
  const TH_GpuRenderParams& P;
  or_counters* cnt;

  Hit trace(const Ray& r, bool shadow) {
    RayCounters rc;
    Hit h = closest(s, r, cnt ? &rc : nullptr);
    if (cnt) {
      cnt->nodes_visited += rc.nodes;
      cnt->primref_tests += rc.primrefs;
      cnt->triangle_tests += rc.tris;
      cnt->sphere_tests += rc.spheres;
      if (shadow) {
        cnt->shadow_rays++;
      } else {
        cnt->closest_rays++;
        if (h.t <= TH_MAX_DISTANCE) {
          cnt->closest_hits++;
          if (h.slot != 0xffffffffu) {
            bool tri = s.params.bvh_node_count ? (s.prims[h.slot].kind == 1u)
                                               : ((h.slot & 0x80000000u) != 0u);
            if (tri) cnt->triangle_hits++;
          }
        }
      }
    }
    return h;
  }

  // wgsl:390-403 sample_light: first light whose cumulative weight >= r*W (else last).
  const TH_GpuLight& sample_light(uint32_t* st) {
    float target = th::rand01(st) * P.total_light_weight;
    const TH_GpuLight* sel = &s.lights[0];
    for (uint32_t i = 0; i < P.light_count; ++i) {
      if (cnt) cnt->lights_scanned++;
      sel = &s.lights[i];
      if (s.lights[i].area_weight_cumulative[2] >= target) break;
    }
    return *sel;
  }

  // wgsl:406-472 direct_light
  v3 direct_light(const Hit& hit, v3 albedo, uint32_t* st) {
    if (P.light_count == 0u || P.total_light_weight <= 0.0f) return th::mk(0, 0, 0);
    const TH_GpuLight& L = sample_light(st);
    v3 pos, ln;
    if (L.kind == 0u) {
      float u = th::fsqrt(th::rand01(st));
      float v = th::rand01(st);
      float w0 = 1.0f - u, w1 = u * (1.0f - v), w2 = u * v;
      v3 a = v3_of(L.v0), b = v3_of(L.v1), c = v3_of(L.v2);
      pos = th::add(th::add(th::smul(w0, a), th::smul(w1, b)), th::smul(w2, c));
      ln = th::normalize(th::cross(th::sub(b, a), th::sub(c, a)));
    } else {
      ln = random_unit_vector(st);
      pos = th::add(v3_of(L.center_radius), th::smul(L.center_radius[3], ln));
    }
    const TH_GpuMaterial& em = s.materials[L.material_index];
    v3 so = th::add(hit.p, th::muls(hit.n, TH_MIN_DISTANCE));
    v3 tl = th::sub(pos, so);
    float d2 = th::dot(tl, tl);
    if (d2 <= TH_MIN_DISTANCE * TH_MIN_DISTANCE) return th::mk(0, 0, 0);
    float dist = th::fsqrt(d2);
    v3 dir = th::divs(tl, dist);
    float sc = th::dot(dir, hit.n);
    float lc = th::dot(th::neg(dir), ln);
    if (sc <= 0.0f || lc <= 0.0f) return th::mk(0, 0, 0);
    Hit sh = trace(Ray{so, dir}, true);
    if (sh.t < dist - 2.0f * TH_MIN_DISTANCE) return th::mk(0, 0, 0);
    v3 brdf = th::divs(albedo, TH_PI);
    float area = L.area_weight_cumulative[0];
    float weight = L.area_weight_cumulative[1];
    float sel_pdf = weight / P.total_light_weight;
    float surf_pdf = 1.0f / area;
    float pdf_area = sel_pdf * surf_pdf;
    v3 r = th::mul(brdf, v3_of(em.albedo));
    r = th::muls(r, em.params[2]);
    r = th::muls(r, sc);
    r = th::muls(r, lc);
    return th::divs(r, d2 * pdf_area);
  }

  // wgsl:486-505 direct_environment
  v3 direct_environment(const Hit& hit, v3 albedo, uint32_t* st) {
    if (P.background_kind != 2u) return th::mk(0, 0, 0);
    EnvSample es;
    if (env_has_mass(s, P)) {
      float sel = th::rand01(st);
      float jx = th::rand01(st);
      float jy = th::rand01(st);
      uint64_t steps = 0;
      es = env_sample_with(s, P, sel, jx, jy, &steps);
      if (cnt) {
        cnt->env_samples++;
        if (es.pdf > 0.0f) cnt->env_lookups++;
      }
    }
    float sc = th::fmax_(0.0f, th::dot(hit.n, es.dir));
    if (sc <= 0.0f || es.pdf <= 0.0f) return th::mk(0, 0, 0);
    Hit sh = trace(Ray{th::add(hit.p, th::muls(hit.n, TH_MIN_DISTANCE)), es.dir}, true);
    if (sh.t <= TH_MAX_DISTANCE) return th::mk(0, 0, 0);
    float bsdf_pdf = sc / TH_PI;
    float mis = power_heuristic(es.pdf, bsdf_pdf);
    v3 brdf = th::divs(albedo, TH_PI);
    v3 r = th::mul(brdf, es.rad);
    r = th::muls(r, sc);
    r = th::muls(r, mis);
    return th::divs(r, es.pdf);
  }

  // wgsl:805-916 ray_color
  v3 ray_color(Ray ray, uint32_t* st) {
    v3 throughput = th::mk(1, 1, 1);
    v3 radiance = th::mk(0, 0, 0);
    bool include_emissive = true, prev_diffuse = false;
    float prev_pdf = 0.0f;
    for (uint32_t i = 0; i < P.max_bounces; ++i) {
      Hit rec = trace(ray, false);
      if (rec.t > TH_MAX_DISTANCE) {
        if (P.background_kind == 0u) {
          float lt = (th::normalize(ray.d).y + 1.0f) * 0.5f;
          radiance = th::add(radiance, th::mul(throughput, th::mix(th::mk(1.0f, 1.0f, 1.0f),
                                                                    th::mk(0.35f, 0.65f, 1.0f), lt)));
        } else if (P.background_kind == 2u) {
          float mis = 1.0f;
          if (prev_diffuse) mis = power_heuristic(prev_pdf, env_pdf(s, P, ray.d));
          if (cnt) cnt->env_lookups++;
          radiance = th::add(radiance, th::muls(th::mul(throughput, env_lookup(s, P, ray.d)), mis));
        }
        break;
      }
      const TH_GpuMaterial& m = s.materials[rec.mat];
      v3 att = th::mk(0, 0, 0);
      Ray next = ray;
      bool brk = false;
      if (m.kind == 0u) {
        v3 albedo = v3_of(m.albedo);
        if (m.texture_width != 0u && m.texture_height != 0u) albedo = texture_albedo(m, rec);
        radiance = th::add(radiance, th::mul(throughput, direct_light(rec, albedo, st)));
        radiance = th::add(radiance, th::mul(throughput, direct_environment(rec, albedo, st)));
        v3 dir = th::add(rec.n, random_unit_vector(st));
        if (th::dot(dir, dir) < TH_EPSILON) dir = rec.n;
        next = Ray{rec.p, th::normalize(dir)};
        att = albedo;
        include_emissive = false;
        prev_diffuse = true;
        prev_pdf = th::fmax_(0.0f, th::dot(rec.n, next.d)) / TH_PI;
      } else if (m.kind == 1u) {
        v3 refl = reflect_vec(ray.d, rec.n);
        float rough = th::clampf(m.params[0], 0.0f, 1.0f);
        v3 dir = th::add(refl, th::smul(rough, random_unit_vector(st)));
        if (th::dot(dir, rec.n) <= 0.0f) brk = true;
        next = Ray{rec.p, th::normalize(dir)};
        att = v3_of(m.albedo);
        include_emissive = true;
        prev_diffuse = false;
      } else if (m.kind == 2u) {
        float ratio = rec.front ? 1.0f / m.params[1] : m.params[1];
        v3 ud = th::normalize(ray.d);
        float ct = th::fmin_(1.0f, th::dot(th::neg(ud), rec.n));
        float stt = th::fsqrt(1.0f - ct * ct);
        bool cannot = (ratio * stt) > 1.0f;
        v3 dir;
        if (cannot || reflectance(ct, ratio) > th::rand01(st))
          dir = reflect_vec(ud, rec.n);
        else
          dir = refract_vec(ud, rec.n, ratio);
        next = Ray{rec.p, th::normalize(dir)};
        att = th::mk(1, 1, 1);
        include_emissive = true;
        prev_diffuse = false;
      } else if (m.kind == 3u) {
        if (include_emissive)
          radiance = th::add(radiance, th::muls(th::mul(throughput, v3_of(m.albedo)), m.params[2]));
        brk = true;
      } else {
        brk = true;
      }
      if (brk) break;
      throughput = th::mul(throughput, att);
      ray = next;
    }
    return radiance;
  }

  // wgsl:210-270 textured base colour (sRGB atlas, repeat, bilinear).
  static float srgb_to_linear(float v) {
    if (v <= 0.04045f) return v / 12.92f;
    return powf((v + 0.055f) / 1.055f, 2.4f);
  }
  v3 tex_texel(const TH_GpuMaterial& m, int x, int y) {
    int sw = (int)m.texture_width, sh = (int)m.texture_height;
    uint32_t wx = (uint32_t)(((x % sw) + sw) % sw), wy = (uint32_t)(((y % sh) + sh) % sh);
    uint32_t packed = s.texture_pixels[m.texture_offset + wy * m.texture_width + wx];
    const float scale = 1.0f / 255.0f;
    return th::mk(srgb_to_linear((float)(packed & 255u) * scale),
                  srgb_to_linear((float)((packed >> 8u) & 255u) * scale),
                  srgb_to_linear((float)((packed >> 16u) & 255u) * scale));
  }
  v3 texture_albedo(const TH_GpuMaterial& m, const Hit& h) {
    float ru = th::fract(h.tu), rv = th::fract(h.tv);
    float px = ru * (float)m.texture_width - 0.5f, py = rv * (float)m.texture_height - 0.5f;
    int bx = (int)th::ffloor(px), by = (int)th::ffloor(py);
    float ax = th::fract(px), ay = th::fract(py);
    v3 top = th::mix(tex_texel(m, bx, by), tex_texel(m, bx + 1, by), ax);
    v3 bottom = th::mix(tex_texel(m, bx, by + 1), tex_texel(m, bx + 1, by + 1), ax);
    return th::mul(v3_of(m.albedo), th::mix(top, bottom, ay));
  }
};

Ray camera_ray(const TH_GpuCamera& cam, const TH_GpuRenderParams& P, uint32_t x, uint32_t y,
               uint32_t* st) {  // wgsl:525-534
  float u = ((float)x + th::rand01(st)) / (float)P.width;
  float v = 1.0f - (((float)y + th::rand01(st)) / (float)P.height);
  v3 o = v3_of(cam.origin);
  v3 d = th::add(th::add(v3_of(cam.lower_left_corner), th::smul(u, v3_of(cam.horizontal))),
                 th::smul(v, v3_of(cam.vertical)));
  d = th::sub(d, o);
  return Ray{o, th::normalize(d)};
}

// color.rs:61-74 / 32-39 / 42-55
float aces(float c) {
  if (isnan(c)) return 0.0f;
  if (!isfinite(c)) return c > 0 ? 1.0f : 0.0f;
  c = c < 0.0f ? 0.0f : c;
  return (c * (2.51f * c + 0.03f)) / (c * (2.43f * c + 0.59f) + 0.14f);
}
float lin_to_srgb(float c) {
  if (c <= 0.0031308f) return 12.92f * c;
  return 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
}
uint8_t to_u8(float c) {  // (c.clamp(0,1)*255).round() as u8
  if (!(c >= 0.0f)) c = 0.0f;  // NaN.clamp stays NaN in Rust; NaN as u8 saturates to 0
  if (c > 1.0f) c = 1.0f;
  return (uint8_t)roundf(c * 255.0f);
}

// ============================================================ toaster-cpu mode
// The CPU renderer's estimator, for converged-image comparison with the GPU kernels:
// crates/toaster-cpu/src/integrator.rs:47-63 (pixel loop), 99-165 (recursive
// ray_color_inner: scatter, then the subpath, then the light and environment samples),
// 168-244 (direct light and environment), 270-389 (scatter, reflect/refract/Schlick,
// rejection-sampled unit vectors: unit-ball metal fuzz); intersect.rs:28-148 (sphere,
// triangle), 172-263 (front-to-back traversal, infinite far bound), 266-289 (linear),
// 322-338 (entry distance); toaster-bvh/src/aabb.rs:74-111 (slab test with infinite
// inverse components); light.rs:126-171 (strict `t < cumulative` pick, sphere points
// from z and azimuth ranges); toaster-scene/src/environment.rs:132-242 (sample,
// importance sample with the 1 - f32::EPSILON clamp, solid-angle pdf);
// toaster-core/src/camera.rs:70-77.  It shares no sampling routine, traversal,
// transcendental or generator with the "pcg" mode above or with the kernels: libm
// sin/cos/atan2/acos, and a SplitMix64 stream per pixel (Steele, Lea & Flood 2014) in
// place of rand's StdRng, whose stream is not restated -- images agree with toaster-cpu
// in distribution, not per pixel.
namespace cpu {

const float EPS = 0.001f;  // integrator.rs:11 RAY_EPSILON
const float PI_F = 3.14159265358979323846f;
const float TAU_F = 6.28318530717958647692f;

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t s;
  uint64_t next64() { return mix64(s += 0x9E3779B97F4A7C15ull); }
  float f() { return (float)(next64() >> 40) * (1.0f / 16777216.0f); }  // [0, 1)
  float range(float lo, float hi) {
    float v = lo + (hi - lo) * f();
    return v < hi ? v : lo;  // half-open as rand's random_range
  }
};

v3 lerp(v3 a, v3 b, float t) { return th::add(a, th::muls(th::sub(b, a), t)); }  // glam Vec3::lerp
float rem1(float x) {  // f32::rem_euclid(1.0)
  float r = fmodf(x, 1.0f);
  return r < 0.0f ? r + 1.0f : r;
}

struct CHit {
  float t = 0.0f;
  v3 p{0, 0, 0}, n{0, 0, 0};
  bool front = false;
  uint32_t mat = 0;
  float tu = 0.0f, tv = 0.0f;
};

bool sphere(const Ray& r, const TH_GpuSphere& s, float tmin, float tmax, CHit* h) {  // intersect.rs:28-68
  v3 c = v3_of(s.center_radius);
  float rad = s.center_radius[3];
  v3 oc = th::sub(r.o, c);
  float a = th::dot(r.d, r.d);
  float hb = th::dot(oc, r.d);
  float cc = th::dot(oc, oc) - rad * rad;
  float disc = hb * hb - a * cc;
  if (disc < 0.0f) return false;
  float sq = sqrtf(disc);
  float t = (-hb - sq) / a;
  if (t < tmin || t > tmax) {
    t = (-hb + sq) / a;
    if (t < tmin || t > tmax) return false;
  }
  v3 p = th::add(r.o, th::smul(t, r.d));
  v3 out = th::divs(th::sub(p, c), rad);
  h->front = th::dot(r.d, out) < 0.0f;
  h->t = t;
  h->p = p;
  h->n = h->front ? out : th::neg(out);
  h->mat = s.material_index;
  h->tu = h->tv = 0.0f;
  return true;
}

bool triangle(const Ray& r, const TH_GpuTriangle& tr, const TH_GpuTriangleAttributes& at, float tmin, float tmax,
              CHit* h) {  // intersect.rs:87-148
  v3 v0 = v3_of(tr.v0), v1 = v3_of(tr.v1), v2 = v3_of(tr.v2);
  v3 e1 = th::sub(v1, v0), e2 = th::sub(v2, v0);
  v3 pv = th::cross(r.d, e2);
  float det = th::dot(e1, pv);
  if (fabsf(det) < 1e-8f) return false;
  float inv = 1.0f / det;
  v3 ov = th::sub(r.o, v0);
  float u = th::dot(ov, pv) * inv;
  if (!(u >= 0.0f && u <= 1.0f)) return false;
  v3 qv = th::cross(ov, e1);
  float v = th::dot(r.d, qv) * inv;
  if (v < 0.0f || u + v > 1.0f) return false;
  float t = th::dot(e2, qv) * inv;
  if (t < tmin || t > tmax) return false;
  v3 gn = glam_normalize(th::cross(e1, e2));
  bool front = th::dot(r.d, gn) < 0.0f;
  float w0 = 1.0f - u - v;
  v3 out = gn;
  if (at.flags & 1u)
    out = glam_normalize(th::add(th::add(th::smul(w0, v3_of(at.n0)), th::smul(u, v3_of(at.n1))),
                                 th::smul(v, v3_of(at.n2))));
  if (th::dot(out, gn) < 0.0f) out = th::neg(out);
  h->tu = h->tv = 0.0f;
  if (at.flags & 2u) {
    h->tu = w0 * at.uv0[0] + u * at.uv1[0] + v * at.uv2[0];
    h->tv = w0 * at.uv0[1] + u * at.uv1[1] + v * at.uv2[1];
  }
  h->t = t;
  h->p = th::add(r.o, th::smul(t, r.d));
  h->n = front ? out : th::neg(out);
  h->front = front;
  h->mat = tr.material_index;
  return true;
}

bool aabb_hit(const TH_GpuBvhNode& b, v3 o, v3 inv, float tmin, float tmax) {  // aabb.rs:74-111
  const float oa[3] = {o.x, o.y, o.z}, ia[3] = {inv.x, inv.y, inv.z};
  for (int k = 0; k < 3; ++k) {
    if (isinf(ia[k])) {
      if (oa[k] < b.min[k] || oa[k] > b.max[k]) return false;
      continue;
    }
    float n = (b.min[k] - oa[k]) * ia[k], f = (b.max[k] - oa[k]) * ia[k];
    if (n > f) std::swap(n, f);
    tmin = std::fmax(tmin, n);  // Rust f32::max / min ignore a NaN operand
    tmax = std::fmin(tmax, f);
    if (tmax < tmin) return false;
  }
  return true;
}

float entry_distance(const TH_GpuBvhNode& b, v3 o, v3 inv) {  // intersect.rs:322-338
  const float oa[3] = {o.x, o.y, o.z}, ia[3] = {inv.x, inv.y, inv.z};
  float e = -INFINITY;
  for (int k = 0; k < 3; ++k) {
    if (isinf(ia[k])) continue;
    float n = (b.min[k] - oa[k]) * ia[k], f = (b.max[k] - oa[k]) * ia[k];
    if (n > f) std::swap(n, f);
    e = std::fmax(e, n);
  }
  return e;
}

bool prim_hit(const or_scene& s, uint32_t kind, uint32_t index, const Ray& r, float tmin, float tmax, CHit* h) {
  if (kind == 0u) return sphere(r, s.spheres[index], tmin, tmax, h);
  return triangle(r, s.triangles[index], s.attributes[index], tmin, tmax, h);
}

// intersect_scene (intersect.rs:172-263); a scene without a tree takes the linear scan
// (intersect_scene_linear, 266-289).
bool scene_hit(const or_scene& s, const Ray& r, float tmin, CHit* out) {
  float closest = INFINITY;
  bool hit = false;
  CHit c;
  if (s.nodes.empty()) {
    for (uint32_t i = 0; i < (uint32_t)s.spheres.size(); ++i)
      if (sphere(r, s.spheres[i], tmin, closest, &c)) {
        closest = c.t;
        *out = c;
        hit = true;
      }
    for (uint32_t i = 0; i < (uint32_t)s.triangles.size(); ++i)
      if (triangle(r, s.triangles[i], s.attributes[i], tmin, closest, &c)) {
        closest = c.t;
        *out = c;
        hit = true;
      }
    return hit;
  }
  v3 inv = th::mk(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
  uint32_t stack[512];
  int sp = 0;
  stack[sp++] = 0u;
  while (sp > 0) {
    uint32_t ni = stack[--sp];
    const TH_GpuBvhNode& node = s.nodes[ni];
    if (!aabb_hit(node, r.o, inv, tmin, closest)) continue;
    if (node.primitive_count > 0u) {
      for (uint32_t k = node.first_or_right; k < node.first_or_right + node.primitive_count; ++k)
        if (prim_hit(s, s.prims[k].kind, s.prims[k].index, r, tmin, closest, &c)) {
          closest = c.t;
          *out = c;
          hit = true;
        }
    } else {
      uint32_t l = ni + 1u, rr = node.first_or_right;
      bool lh = aabb_hit(s.nodes[l], r.o, inv, tmin, closest);
      bool rh = aabb_hit(s.nodes[rr], r.o, inv, tmin, closest);
      if (sp + 2 > 512) abort();  // a median-split tree is far shallower
      if (lh && rh) {
        if (entry_distance(s.nodes[l], r.o, inv) < entry_distance(s.nodes[rr], r.o, inv)) {
          stack[sp++] = rr;
          stack[sp++] = l;
        } else {
          stack[sp++] = l;
          stack[sp++] = rr;
        }
      } else if (lh) {
        stack[sp++] = l;
      } else if (rh) {
        stack[sp++] = rr;
      }
    }
  }
  return hit;
}

v3 random_in_unit_sphere(Rng& g) {  // integrator.rs:377-389
  for (;;) {
    v3 v = th::mk(g.range(-1.0f, 1.0f), g.range(-1.0f, 1.0f), g.range(-1.0f, 1.0f));
    float l2 = th::dot(v, v);
    if (l2 >= 1e-8f && l2 <= 1.0f) return v;
  }
}
v3 random_unit_vector(Rng& g) { return glam_normalize(random_in_unit_sphere(g)); }  // 372-374

v3 reflect(v3 d, v3 n) { return th::sub(d, th::muls(n, 2.0f * th::dot(d, n))); }  // 349-351
v3 refract(v3 d, v3 n, float ratio) {                                             // 354-359
  float ct = std::fmin(th::dot(th::neg(d), n), 1.0f);
  v3 perp = th::smul(ratio, th::add(d, th::smul(ct, n)));
  v3 par = th::smul(-sqrtf(fabsf(1.0f - th::dot(perp, perp))), n);
  return th::add(perp, par);
}
float schlick(float c, float ratio) {  // 362-365
  float r0 = (1.0f - ratio) / (1.0f + ratio);
  r0 = r0 * r0;
  return r0 + (1.0f - r0) * powf(1.0f - c, 5.0f);
}
float power_heuristic(float a, float b) {  // 246-256
  float a2 = a * a, b2 = b * b, den = a2 + b2;
  return den > 0.0f ? a2 / den : 0.0f;
}

struct Estimator {
  const or_scene& s;
  const TH_GpuRenderParams& P;

  // ---- environment.rs:132-242
  v3 env_texel(int x, int y) const {
    int W = (int)P.environment_width, H = (int)P.environment_height;
    int xx = ((x % W) + W) % W, yy = y < 0 ? 0 : (y > H - 1 ? H - 1 : y);
    const float* p = &s.env[4 * ((size_t)yy * W + xx)];
    return th::mk(p[0], p[1], p[2]);
  }
  bool dir_uv(v3 d, float* u, float* v, float* sin_t) const {
    float l = glam_length(d);
    if (!(l > 0.0f) || isinf(l)) return false;  // normalize_or_zero
    v3 n = glam_normalize(d);
    *u = rem1(atan2f(n.z, n.x) / (2.0f * PI_F) + 0.5f + P.environment_rotation_degrees / 360.0f);
    float th = acosf(std::fmin(std::fmax(n.y, -1.0f), 1.0f));
    *v = th / PI_F;
    *sin_t = sinf(th);
    return true;
  }
  v3 env_sample(v3 d) const {
    float u, v, st;
    if (!dir_uv(d, &u, &v, &st)) return th::mk(0, 0, 0);
    float x = u * (float)P.environment_width - 0.5f, y = v * (float)P.environment_height - 0.5f;
    int x0 = (int)floorf(x), y0 = (int)floorf(y);
    float ax = x - floorf(x), ay = y - floorf(y);
    v3 top = lerp(env_texel(x0, y0), env_texel(x0 + 1, y0), ax);
    v3 bottom = lerp(env_texel(x0, y0 + 1), env_texel(x0 + 1, y0 + 1), ax);
    return th::muls(lerp(top, bottom, ay), P.environment_intensity);
  }
  float cdf(size_t i) const { return s.env[4 * i + 3]; }
  float texel_pdf(size_t index, float sin_t) const {
    float prev = index > 0 ? cdf(index - 1) : 0.0f;
    float p = std::fmax(cdf(index) - prev, 0.0f);
    if (p <= 0.0f || sin_t <= 0.0f) return 0.0f;
    return p * (float)P.environment_width * (float)P.environment_height / (2.0f * PI_F * PI_F * sin_t);
  }
  float env_pdf(v3 d) const {
    float u, v, st;
    if (!dir_uv(d, &u, &v, &st)) return 0.0f;
    size_t W = P.environment_width, H = P.environment_height;
    size_t x = (size_t)floorf(u * (float)W) % W;
    size_t y = std::min((size_t)floorf(v * (float)H), H - 1);
    return texel_pdf(y * W + x, st);
  }
  static float unit_interval(float x) {
    return isfinite(x) ? std::fmin(std::fmax(x, 0.0f), 1.0f - 1.1920929e-7f) : 0.0f;
  }
  bool env_importance(float sel, float ju, float jv, v3* dir, v3* rad, float* pdf) const {
    size_t n = (size_t)P.environment_width * P.environment_height;
    if (n == 0 || cdf(n - 1) <= 0.0f) return false;
    float x_sel = unit_interval(sel);
    size_t lo = 0, hi = n;  // partition_point(c <= x_sel)
    while (lo < hi) {
      size_t mid = lo + (hi - lo) / 2;
      if (cdf(mid) <= x_sel)
        lo = mid + 1;
      else
        hi = mid;
    }
    size_t index = std::min(lo, n - 1);
    size_t W = P.environment_width;
    float u = ((float)(index % W) + unit_interval(ju)) / (float)W;
    float v = ((float)(index / W) + unit_interval(jv)) / (float)P.environment_height;
    float theta = PI_F * v;
    float phi = 2.0f * PI_F * (u - 0.5f - P.environment_rotation_degrees / 360.0f);
    float st = sinf(theta);
    *dir = th::mk(st * cosf(phi), cosf(theta), st * sinf(phi));
    *pdf = texel_pdf(index, st);
    *rad = env_sample(*dir);
    return true;
  }

  // ---- materials (toaster-scene Material via the packed records; textured diffuse
  // resolved to its hit albedo as material_at_hit, integrator.rs:340-346)
  v3 albedo_at(const TH_GpuMaterial& m, const CHit& h) const {
    if (m.texture_width != 0u && m.texture_height != 0u) {
      Tracer tr{s, P, nullptr};
      Hit wh;
      wh.tu = h.tu;
      wh.tv = h.tv;
      return tr.texture_albedo(m, wh);
    }
    return v3_of(m.albedo);
  }

  // ---- light.rs:126-171 + integrator.rs:168-205
  v3 direct_area_light(const CHit& h, v3 albedo, Rng& g) const {
    if (P.light_count == 0u || P.total_light_weight <= 0.0f) return th::mk(0, 0, 0);
    float target = g.f() * P.total_light_weight;
    const TH_GpuLight* L = &s.lights.back();
    for (const TH_GpuLight& c : s.lights)
      if (target < c.area_weight_cumulative[2]) {
        L = &c;
        break;
      }
    v3 pos, ln;
    if (L->kind == 0u) {
      float su = sqrtf(g.f());
      float v = g.f();
      v3 a = v3_of(L->v0), b = v3_of(L->v1), c = v3_of(L->v2);
      pos = th::add(th::add(th::smul(1.0f - su, a), th::smul(su * (1.0f - v), b)), th::smul(su * v, c));
      ln = glam_normalize(th::cross(th::sub(b, a), th::sub(c, a)));
    } else {
      float z = g.range(-1.0f, 1.0f);
      float az = g.range(0.0f, TAU_F);
      float radial = sqrtf(std::fmax(1.0f - z * z, 0.0f));
      ln = th::mk(radial * cosf(az), radial * sinf(az), z);
      pos = th::add(v3_of(L->center_radius), th::smul(L->center_radius[3], ln));
    }
    const TH_GpuMaterial& em = s.materials[L->material_index];
    v3 emission = th::muls(v3_of(em.albedo), em.params[2]);
    float sel_pdf = L->area_weight_cumulative[1] / P.total_light_weight;
    float pdf_area = sel_pdf / L->area_weight_cumulative[0];
    v3 so = th::add(h.p, th::muls(h.n, EPS));
    v3 tl = th::sub(pos, so);
    float d2 = th::dot(tl, tl);
    if (d2 <= EPS * EPS) return th::mk(0, 0, 0);
    float dist = sqrtf(d2);
    v3 dir = th::divs(tl, dist);
    float sc = std::fmax(th::dot(h.n, dir), 0.0f);
    float lc = std::fmax(th::dot(ln, th::neg(dir)), 0.0f);
    if (sc <= 0.0f || lc <= 0.0f) return th::mk(0, 0, 0);
    CHit b;
    if (scene_hit(s, Ray{so, dir}, EPS, &b) && b.t < dist - 2.0f * EPS) return th::mk(0, 0, 0);
    v3 r = th::divs(albedo, PI_F);
    r = th::mul(r, emission);
    r = th::muls(r, sc);
    r = th::muls(r, lc);
    return th::divs(r, d2 * pdf_area);
  }

  // ---- integrator.rs:208-244
  v3 direct_environment(const CHit& h, v3 albedo, Rng& g) const {
    if (P.background_kind != 2u) return th::mk(0, 0, 0);
    float r1 = g.f(), r2 = g.f(), r3 = g.f();
    v3 dir, rad;
    float pdf;
    if (!env_importance(r1, r2, r3, &dir, &rad, &pdf)) return th::mk(0, 0, 0);
    float sc = std::fmax(th::dot(h.n, dir), 0.0f);
    if (sc <= 0.0f || pdf <= 0.0f) return th::mk(0, 0, 0);
    CHit b;
    if (scene_hit(s, Ray{th::add(h.p, th::muls(h.n, EPS)), dir}, EPS, &b)) return th::mk(0, 0, 0);
    float mis = power_heuristic(pdf, sc / PI_F);
    v3 r = th::divs(albedo, PI_F);
    r = th::mul(r, rad);
    r = th::muls(r, sc);
    r = th::muls(r, mis);
    return th::divs(r, pdf);
  }

  // ---- integrator.rs:99-165
  v3 ray_color(const Ray& ray, uint32_t depth, bool include_emission, bool has_pdf, float bsdf_pdf,
               Rng& g) const {
    if (depth == 0u) return th::mk(0, 0, 0);
    CHit h;
    if (scene_hit(s, ray, EPS, &h)) {
      const TH_GpuMaterial& m = s.materials[h.mat];
      if (m.kind == 3u) return include_emission ? th::muls(v3_of(m.albedo), m.params[2]) : th::mk(0, 0, 0);
      Ray next;
      v3 att;
      bool next_has_pdf = false;
      float next_pdf = 0.0f;
      v3 albedo = th::mk(0, 0, 0);
      if (m.kind == 0u) {  // diffuse, textured or not
        albedo = albedo_at(m, h);
        v3 d = th::add(h.n, random_unit_vector(g));
        if (th::dot(d, d) < 1e-8f) d = h.n;
        v3 nd = glam_normalize(d);
        next = Ray{h.p, nd};
        att = albedo;
        next_has_pdf = true;
        next_pdf = std::fmax(th::dot(h.n, nd), 0.0f) / PI_F;
      } else if (m.kind == 1u) {  // metal: unit-ball fuzz
        v3 refl = reflect(glam_normalize(ray.d), h.n);
        float rough = std::fmin(std::fmax(m.params[0], 0.0f), 1.0f);
        v3 d = th::add(refl, th::smul(rough, random_in_unit_sphere(g)));
        if (th::dot(d, h.n) <= 0.0f) return th::mk(0, 0, 0);
        next = Ray{h.p, glam_normalize(d)};
        att = v3_of(m.albedo);
      } else if (m.kind == 2u) {  // dielectric
        float ratio = h.front ? 1.0f / m.params[1] : m.params[1];
        v3 ud = glam_normalize(ray.d);
        float ct = std::fmin(th::dot(th::neg(ud), h.n), 1.0f);
        float st = sqrtf(1.0f - ct * ct);
        bool cannot = ratio * st > 1.0f;
        v3 d = (cannot || schlick(ct, ratio) > g.f()) ? reflect(ud, h.n) : refract(ud, h.n, ratio);
        next = Ray{h.p, glam_normalize(d)};
        att = th::mk(1, 1, 1);
      } else {
        return th::mk(0, 0, 0);
      }
      v3 indirect = th::mul(att, ray_color(next, depth - 1u, m.kind != 0u, next_has_pdf, next_pdf, g));
      v3 direct = th::mk(0, 0, 0);
      if (m.kind == 0u)
        direct = th::add(direct_area_light(h, albedo, g), direct_environment(h, albedo, g));
      return th::add(direct, indirect);
    }
    if (P.background_kind == 0u) {
      v3 d = glam_normalize(ray.d);
      return lerp(th::mk(1, 1, 1), th::mk(0.35f, 0.65f, 1.0f), 0.5f * (d.y + 1.0f));
    }
    if (P.background_kind == 2u) {
      v3 rad = env_sample(ray.d);
      float w = has_pdf ? power_heuristic(bsdf_pdf, env_pdf(ray.d)) : 1.0f;
      return th::muls(rad, w);
    }
    return th::mk(0, 0, 0);
  }
};

}  // namespace cpu

}  // namespace

// ============================================================ C ABI
extern "C" {

const char* or_last_error(void) { return g_error.c_str(); }

int32_t or_pack(const th_scene_input* in, or_scene** out) {
  if (!in || !out) return fail("null argument");
  *out = nullptr;
  // scene_upload.rs:84-92 validation
  if (in->sphere_count == 0 && in->triangle_count == 0)
    return fail("GPU renderer requires at least one object");
  if (in->material_count == 0 || !in->materials)
    return fail("GPU renderer requires at least one material");
  if ((in->sphere_count && !in->spheres) || (in->triangle_count && !in->triangles))
    return fail("null geometry array");
  if (in->width == 0 || in->height == 0) return fail("render dimensions must be nonzero");
  for (uint32_t i = 0; i < in->sphere_count; ++i)
    if (in->spheres[i].material_index >= in->material_count) return fail("material index out of range");
  for (uint32_t i = 0; i < in->triangle_count; ++i)
    if (in->triangles[i].material_index >= in->material_count) return fail("material index out of range");
  std::unique_ptr<or_scene> s(new or_scene());
  memset(&s->params, 0, sizeof(s->params));
  float aspect = (float)in->width / (float)in->height;
  s->camera = make_camera(v3_of(in->camera_position), v3_of(in->camera_look_at), v3_of(in->camera_up),
                          in->fov_degrees, aspect);
  for (uint32_t i = 0; i < in->sphere_count; ++i) {
    TH_GpuSphere g;
    memset(&g, 0, sizeof g);
    const th_sphere& sp = in->spheres[i];
    g.center_radius[0] = sp.center[0];
    g.center_radius[1] = sp.center[1];
    g.center_radius[2] = sp.center[2];
    g.center_radius[3] = sp.radius;
    g.material_index = sp.material_index;
    s->spheres.push_back(g);
  }
  for (uint32_t i = 0; i < in->triangle_count; ++i) {
    TH_GpuTriangle g;
    memset(&g, 0, sizeof g);
    const th_triangle& t = in->triangles[i];
    for (int k = 0; k < 3; ++k) {
      g.v0[k] = t.v[0][k];
      g.v1[k] = t.v[1][k];
      g.v2[k] = t.v[2][k];
    }
    g.material_index = t.material_index;
    s->triangles.push_back(g);
    TH_GpuTriangleAttributes a;
    memset(&a, 0, sizeof a);
    if (in->triangle_attributes) {
      const th_triangle_attributes& ta = in->triangle_attributes[i];
      if (ta.has_normals) {
        a.flags |= 1u;
        for (int k = 0; k < 3; ++k) {
          a.n0[k] = ta.normals[0][k];
          a.n1[k] = ta.normals[1][k];
          a.n2[k] = ta.normals[2][k];
        }
      }
      if (ta.has_tex_coords) {
        a.flags |= 2u;
        for (int k = 0; k < 2; ++k) {
          a.uv0[k] = ta.tex_coords[0][k];
          a.uv1[k] = ta.tex_coords[1][k];
          a.uv2[k] = ta.tex_coords[2][k];
        }
      }
    }
    s->attributes.push_back(a);
  }
  // texture atlas (scene_upload.rs:168-188)
  std::vector<uint32_t> tex_off, tex_w, tex_h;
  uint64_t texel_total = 0;
  for (uint32_t t = 0; t < in->texture_count; ++t) {
    const th_texture& tx = in->textures[t];
    tex_off.push_back((uint32_t)texel_total);
    tex_w.push_back(tx.width);
    tex_h.push_back(tx.height);
    uint64_t n = (uint64_t)tx.width * tx.height;
    for (uint64_t i = 0; i < n; ++i) {
      const uint8_t* p = tx.rgba8 + 4 * i;
      s->texture_pixels.push_back((uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
                                  ((uint32_t)p[3] << 24));
    }
    texel_total += n;
  }
  // materials (scene_upload.rs:371-404)
  for (uint32_t i = 0; i < in->material_count; ++i) {
    const th_material& m = in->materials[i];
    TH_GpuMaterial g;
    memset(&g, 0, sizeof g);
    v3 albedo = v3_of(m.color);
    switch (m.kind) {
      case TH_MAT_DIFFUSE: g.kind = 0; break;
      case TH_MAT_TEXTURED_DIFFUSE:
        if (m.texture_index >= in->texture_count) return fail("material references an unknown texture");
        g.kind = 0;
        g.texture_offset = tex_off[m.texture_index];
        g.texture_width = tex_w[m.texture_index];
        g.texture_height = tex_h[m.texture_index];
        break;
      case TH_MAT_METAL: g.kind = 1; g.params[0] = m.roughness; break;
      case TH_MAT_DIELECTRIC: g.kind = 2; albedo = th::mk(1, 1, 1); g.params[1] = m.ior; break;
      case TH_MAT_EMISSIVE: g.kind = 3; g.params[2] = m.strength; break;
      default: return fail("unknown material kind");
    }
    store4(g.albedo, albedo, 1.0f);
    s->materials.push_back(g);
  }
  // lights (scene_upload.rs:412-475): triangles first, then spheres
  float cumulative = 0.0f;
  for (uint32_t i = 0; i < in->triangle_count; ++i) {
    const th_triangle& t = in->triangles[i];
    const th_material& m = in->materials[t.material_index];
    if (m.kind != TH_MAT_EMISSIVE) continue;
    v3 a = v3_of(t.v[0]), b = v3_of(t.v[1]), c = v3_of(t.v[2]);
    float area = 0.5f * glam_length(th::cross(th::sub(b, a), th::sub(c, a)));
    float weight = area * rec709(v3_of(m.color)) * m.strength;
    if (weight <= 0.0f) continue;
    cumulative += weight;
    TH_GpuLight L;
    memset(&L, 0, sizeof L);
    L.kind = 0;
    L.material_index = t.material_index;
    store4(L.v0, a, 0);
    store4(L.v1, b, 0);
    store4(L.v2, c, 0);
    L.area_weight_cumulative[0] = area;
    L.area_weight_cumulative[1] = weight;
    L.area_weight_cumulative[2] = cumulative;
    s->lights.push_back(L);
  }
  for (uint32_t i = 0; i < in->sphere_count; ++i) {
    const th_sphere& sp = in->spheres[i];
    const th_material& m = in->materials[sp.material_index];
    if (m.kind != TH_MAT_EMISSIVE) continue;
    float area = 4.0f * 3.14159274f * sp.radius * sp.radius;
    float weight = area * rec709(v3_of(m.color)) * m.strength;
    if (weight <= 0.0f) continue;
    cumulative += weight;
    TH_GpuLight L;
    memset(&L, 0, sizeof L);
    L.kind = 1;
    L.material_index = sp.material_index;
    for (int k = 0; k < 3; ++k) L.center_radius[k] = sp.center[k];
    L.center_radius[3] = sp.radius;
    L.area_weight_cumulative[0] = area;
    L.area_weight_cumulative[1] = weight;
    L.area_weight_cumulative[2] = cumulative;
    s->lights.push_back(L);
  }
  // BVH (scene_upload.rs:296-346): spheres first, then triangles (padded 1e-5)
  std::vector<PrimInfo> infos;
  for (uint32_t i = 0; i < in->sphere_count; ++i) {
    const th_sphere& sp = in->spheres[i];
    v3 c = v3_of(sp.center), r = th::mk(sp.radius, sp.radius, sp.radius);
    Aabb b{th::sub(c, r), th::add(c, r)};
    infos.push_back(PrimInfo{b, aabb_centroid(b), 0u, i});
  }
  for (uint32_t i = 0; i < in->triangle_count; ++i) {
    const th_triangle& t = in->triangles[i];
    v3 a = v3_of(t.v[0]), b = v3_of(t.v[1]), c = v3_of(t.v[2]);
    v3 mn = th::mk(rmin(rmin(a.x, b.x), c.x), rmin(rmin(a.y, b.y), c.y), rmin(rmin(a.z, b.z), c.z));
    v3 mx = th::mk(rmax(rmax(a.x, b.x), c.x), rmax(rmax(a.y, b.y), c.y), rmax(rmax(a.z, b.z), c.z));
    const float pad = 1e-5f;
    Aabb bb{th::mk(mn.x - pad, mn.y - pad, mn.z - pad), th::mk(mx.x + pad, mx.y + pad, mx.z + pad)};
    infos.push_back(PrimInfo{bb, aabb_centroid(bb), 1u, i});
  }
  TreeBuilder tb;
  int root = tb.build(infos.data(), infos.size());
  if (root >= 0) tb.flatten(root, s->nodes);
  s->prims = tb.ordered;
  // environment (scene_upload.rs:214-246)
  uint32_t bg = in->background;
  if (bg > 2) return fail("unknown background");
  float env_int = 0.0f, env_rot = 0.0f;
  uint32_t ew = 0, eh = 0;
  if (bg == TH_BG_ENVIRONMENT && !in->environment)
    return fail("environment background requires an environment map");
  if (in->environment) {
    const th_environment& e = *in->environment;
    if (e.width == 0 || e.height == 0 || !e.rgb) return fail("environment dimensions must be nonzero");
    float rot = fmodf(e.rotation_degrees, 360.0f);  // f32::rem_euclid(360)
    if (rot < 0.0f) rot += 360.0f;
    std::vector<float> cdf = env_cdf(e.width, e.height, e.rgb, e.intensity);
    size_t n = (size_t)e.width * e.height;
    s->env.resize(4 * n);
    for (size_t i = 0; i < n; ++i) {
      s->env[4 * i] = e.rgb[3 * i];
      s->env[4 * i + 1] = e.rgb[3 * i + 1];
      s->env[4 * i + 2] = e.rgb[3 * i + 2];
      s->env[4 * i + 3] = cdf[i];
    }
    if (bg == TH_BG_ENVIRONMENT) {
      ew = e.width;
      eh = e.height;
      env_int = e.intensity;
      env_rot = rot;
    }
  }
  TH_GpuRenderParams& P = s->params;
  P.width = in->width;
  P.height = in->height;
  P.samples = in->samples;
  P.max_bounces = in->max_bounces;
  P.sphere_count = in->sphere_count;
  P.triangle_count = in->triangle_count;
  P.material_count = in->material_count;
  P.background_kind = bg;
  P.light_count = (uint32_t)s->lights.size();
  P.total_light_weight = cumulative;
  P.environment_width = ew;
  P.environment_height = eh;
  P.environment_intensity = env_int;
  P.environment_rotation_degrees = env_rot;
  P.bvh_node_count = (uint32_t)s->nodes.size();
  P.bvh_primitive_count = (uint32_t)s->prims.size();
  *out = s.release();
  return TH_OK;
}

// Drops the packed BVH so renders take the linear scan (wgsl:797-802), as a scene
// uploaded with bvh_node_count == 0 does.
int32_t or_drop_bvh(or_scene* s) {
  if (!s) return fail("null argument");
  s->nodes.clear();
  s->prims.clear();
  s->params.bvh_node_count = 0;
  s->params.bvh_primitive_count = 0;
  return TH_OK;
}

int32_t or_view(const or_scene* s, th_scene_desc* d) {
  if (!s || !d) return fail("null argument");
  memset(d, 0, sizeof *d);
  d->params = s->params;
  d->camera = s->camera;
  d->spheres = s->spheres.data();
  d->triangles = s->triangles.data();
  d->triangle_attributes = s->attributes.data();
  d->materials = s->materials.data();
  d->lights = s->lights.data();
  d->bvh_nodes = s->nodes.data();
  d->bvh_primitives = s->prims.data();
  d->texture_pixels = s->texture_pixels.data();
  d->environment_pixels = s->env.data();
  d->sphere_count = (uint32_t)s->spheres.size();
  d->triangle_count = (uint32_t)s->triangles.size();
  d->material_count = (uint32_t)s->materials.size();
  d->light_count = (uint32_t)s->lights.size();
  d->bvh_node_count = (uint32_t)s->nodes.size();
  d->bvh_primitive_count = (uint32_t)s->prims.size();
  d->texture_pixel_count = (uint32_t)s->texture_pixels.size();
  d->environment_pixel_count = (uint32_t)(s->env.size() / 4);
  return TH_OK;
}

void or_free(or_scene* s) { delete s; }

int32_t or_render(const or_scene* s, const TH_GpuRenderParams* params, const TH_GpuCamera* camera,
                  uint32_t row_begin, uint32_t row_end, int32_t threads, float* rgba,
                  or_counters* counters) {
  return or_render_ex(s, params, camera, row_begin, row_end, threads, 0u, 1u, rgba, counters);
}

int32_t or_render_ex(const or_scene* s, const TH_GpuRenderParams* params, const TH_GpuCamera* camera,
                     uint32_t row_begin, uint32_t row_end, int32_t threads, uint32_t rng_mode,
                     uint32_t group, float* rgba, or_counters* counters) {
  if (group == 0u || group > 64u || (group & (group - 1u)) != 0u) return fail("group must be a power of two <= 64");
  if (!s || !params || !camera || !rgba) return fail("null argument");
  const TH_GpuRenderParams P = *params;
  if (P.samples == 0) return fail("samples must be nonzero");
  if (row_end > P.height) row_end = P.height;
  or_counters total;
  memset(&total, 0, sizeof total);
  const TH_GpuCamera cam = *camera;
  const int64_t nrows = row_end > row_begin ? (int64_t)(row_end - row_begin) : 0;
  const int64_t npix = nrows * (int64_t)P.width;
  std::vector<or_counters> per_thread(threads > 1 ? threads : 1);
  for (auto& c : per_thread) memset(&c, 0, sizeof c);
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 1 ? threads : 1)
  for (int64_t k = 0; k < npix; ++k) {
#ifdef _OPENMP
    int tid = omp_get_thread_num();
#else
    int tid = 0;
#endif
    or_counters* c = counters ? &per_thread[tid] : nullptr;
    Tracer tr{*s, P, c};
    uint32_t y = row_begin + (uint32_t)(k / P.width), x = (uint32_t)(k % P.width);
    uint32_t idx = y * P.width + x;
    uint32_t base = idx ^ (P.frame_index * 9781u);
    uint32_t st = th::pcg_hash(base);  // wgsl:519
    float cr = 0, cg = 0, cb = 0, ca = 0;
    if (rng_mode == 0u) {
      for (uint32_t i = 0; i < P.samples; ++i) {
        Ray r = camera_ray(cam, P, x, y, &st);
        v3 L = tr.ray_color(r, &st);
        cr = cr + L.x;
        cg = cg + L.y;
        cb = cb + L.z;
        ca = ca + 1.0f;
        if (c) c->paths++;
      }
    } else {
      // per-sample streams pcg_hash(base + s * 0x9E3779B9); group j sums samples
      // j, j+G, ... and the G partial sums combine as a butterfly (xor 1, 2, 4, ...)
      float part[64][4];
      for (uint32_t j = 0; j < group; ++j) {
        part[j][0] = part[j][1] = part[j][2] = part[j][3] = 0.0f;
        for (uint32_t i = j; i < P.samples; i += group) {
          uint32_t ss = th::pcg_hash(base + i * 0x9E3779B9u);
          Ray r = camera_ray(cam, P, x, y, &ss);
          v3 L = tr.ray_color(r, &ss);
          part[j][0] = part[j][0] + L.x;
          part[j][1] = part[j][1] + L.y;
          part[j][2] = part[j][2] + L.z;
          part[j][3] = part[j][3] + 1.0f;
          if (c) c->paths++;
        }
      }
      for (uint32_t off = 1; off < group; off <<= 1)
        for (uint32_t j = 0; j < group; ++j)
          if ((j & off) == 0u)
            for (int k = 0; k < 4; ++k) {
              float a = part[j][k] + part[j ^ off][k];
              part[j][k] = a;
              part[j ^ off][k] = a;
            }
      cr = part[0][0];
      cg = part[0][1];
      cb = part[0][2];
      ca = part[0][3];
    }
    float ns = (float)P.samples;
    float b0 = cr / ns, b1 = cg / ns, b2 = cb / ns, b3 = ca / ns;
    float* o = rgba + 4 * (size_t)idx;
    if (P.accumulated_samples == 0u) {  // wgsl:538-547
      o[0] = b0;
      o[1] = b1;
      o[2] = b2;
      o[3] = b3;
    } else {
      float prev = (float)P.accumulated_samples;
      float tot = prev + ns;
      o[0] = (o[0] * prev + b0 * ns) / tot;
      o[1] = (o[1] * prev + b1 * ns) / tot;
      o[2] = (o[2] * prev + b2 * ns) / tot;
      o[3] = (o[3] * prev + b3 * ns) / tot;
    }
  }
  if (counters) {
    for (auto& c : per_thread) {
      total.paths += c.paths;
      total.closest_rays += c.closest_rays;
      total.shadow_rays += c.shadow_rays;
      total.nodes_visited += c.nodes_visited;
      total.primref_tests += c.primref_tests;
      total.triangle_tests += c.triangle_tests;
      total.sphere_tests += c.sphere_tests;
      total.closest_hits += c.closest_hits;
      total.triangle_hits += c.triangle_hits;
      total.lights_scanned += c.lights_scanned;
      total.env_lookups += c.env_lookups;
      total.env_samples += c.env_samples;
    }
    *counters = total;
  }
  return TH_OK;
}

// toaster-cpu mode render of rows [row_begin, row_end): per pixel `samples` camera rays
// (integrator.rs:47-63) on a SplitMix64 stream seeded from the pixel index, the frame
// index and the CPU renderer's seed 0x00544F4153544552 (integrator.rs:10), averaged, then
// folded into `rgba` as the progressive contract does (accumulated_samples).
int32_t or_render_cpu(const or_scene* s, const TH_GpuRenderParams* params, const TH_GpuCamera* camera,
                      uint32_t row_begin, uint32_t row_end, int32_t threads, float* rgba) {
  if (!s || !params || !camera || !rgba) return fail("null argument");
  const TH_GpuRenderParams P = *params;
  if (P.samples == 0) return fail("samples must be nonzero");
  if (row_end > P.height) row_end = P.height;
  const TH_GpuCamera cam = *camera;
  const int64_t nrows = row_end > row_begin ? (int64_t)(row_end - row_begin) : 0;
  const int64_t npix = nrows * (int64_t)P.width;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 1 ? threads : 1)
  for (int64_t k = 0; k < npix; ++k) {
    cpu::Estimator est{*s, P};
    uint32_t y = row_begin + (uint32_t)(k / P.width), x = (uint32_t)(k % P.width);
    uint32_t idx = y * P.width + x;
    cpu::Rng g{cpu::mix64(0x00544F4153544552ull ^ ((uint64_t)P.frame_index << 32) ^ (uint64_t)idx)};
    v3 color = th::mk(0, 0, 0);
    for (uint32_t i = 0; i < P.samples; ++i) {
      float u = ((float)x + g.f()) / (float)P.width;
      float v = 1.0f - ((float)y + g.f()) / (float)P.height;
      v3 d = th::sub(th::add(th::add(v3_of(cam.lower_left_corner), th::smul(u, v3_of(cam.horizontal))),
                             th::smul(v, v3_of(cam.vertical))),
                     v3_of(cam.origin));  // camera.rs:70-77
      Ray r{v3_of(cam.origin), glam_normalize(d)};
      color = th::add(color, est.ray_color(r, P.max_bounces, true, false, 0.0f, g));
    }
    float ns = (float)P.samples;
    float b0 = color.x / ns, b1 = color.y / ns, b2 = color.z / ns, b3 = 1.0f;
    float* o = rgba + 4 * (size_t)idx;
    if (P.accumulated_samples == 0u) {
      o[0] = b0;
      o[1] = b1;
      o[2] = b2;
      o[3] = b3;
    } else {
      float prev = (float)P.accumulated_samples, tot = prev + ns;
      o[0] = (o[0] * prev + b0 * ns) / tot;
      o[1] = (o[1] * prev + b1 * ns) / tot;
      o[2] = (o[2] * prev + b2 * ns) / tot;
      o[3] = (o[3] * prev + b3 * ns) / tot;
    }
  }
  return TH_OK;
}

int32_t or_primary_hits(const or_scene* s, const TH_GpuRenderParams* params,
                        const TH_GpuCamera* camera, uint32_t* slot_out, float* distance_out) {
  if (!s || !params || !camera || !slot_out || !distance_out) return fail("null argument");
  const TH_GpuRenderParams P = *params;
  for (uint32_t y = 0; y < P.height; ++y)
    for (uint32_t x = 0; x < P.width; ++x) {
      uint32_t idx = y * P.width + x;
      uint32_t st = th::pcg_hash(idx ^ (P.frame_index * 9781u));
      Ray r = camera_ray(*camera, P, x, y, &st);
      Hit h = closest(*s, r, nullptr);
      slot_out[idx] = h.slot;
      distance_out[idx] = h.t;
    }
  return TH_OK;
}

static void fill_hit(const Hit& h, or_hit* out) {
  out->distance = h.t;
  out->point[0] = h.p.x;
  out->point[1] = h.p.y;
  out->point[2] = h.p.z;
  out->normal[0] = h.n.x;
  out->normal[1] = h.n.y;
  out->normal[2] = h.n.z;
  out->front_face = h.front ? 1u : 0u;
  out->material_index = h.mat;
  out->tex_coord[0] = h.tu;
  out->tex_coord[1] = h.tv;
  out->slot = h.slot;
}

int32_t or_closest(const or_scene* s, const float origin[3], const float direction[3],
                   int32_t use_bvh, or_hit* out) {
  if (!s || !origin || !direction || !out) return fail("null argument");
  Ray r{v3_of(origin), v3_of(direction)};
  Hit h = use_bvh ? (s->nodes.empty() ? Hit() : closest_bvh(*s, r, nullptr)) : closest_linear(*s, r, nullptr);
  fill_hit(h, out);
  return h.t <= TH_MAX_DISTANCE ? 1 : 0;
}

// intersect.rs:28-68 (CPU form: inclusive [tmin, tmax] on both roots)
int32_t or_intersect_sphere(const float origin[3], const float direction[3], const float center[3],
                            float radius, float tmin, float tmax, or_hit* out) {
  v3 o = v3_of(origin), d = v3_of(direction), c = v3_of(center);
  v3 off = th::sub(o, c);
  float a = th::dot(d, d);
  float hb = th::dot(off, d);
  float cc = th::dot(off, off) - radius * radius;
  float disc = hb * hb - a * cc;
  if (disc < 0.0f) return 0;
  float sq = th::fsqrt(disc);
  float t = (-hb - sq) / a;
  if (t < tmin || t > tmax) {
    t = (-hb + sq) / a;
    if (t < tmin || t > tmax) return 0;
  }
  Hit h;
  h.t = t;
  h.p = th::add(o, th::smul(t, d));
  v3 on = th::divs(th::sub(h.p, c), radius);
  h.front = th::dot(d, on) < 0.0f;
  h.n = h.front ? on : th::neg(on);
  if (out) fill_hit(h, out);
  return 1;
}

// intersect.rs:87-148
int32_t or_intersect_triangle(const float origin[3], const float direction[3], const float v[9],
                              const TH_GpuTriangleAttributes* attributes, float tmin, float tmax,
                              or_hit* out) {
  v3 o = v3_of(origin), d = v3_of(direction);
  v3 v0 = v3_of(v), v1 = v3_of(v + 3), v2 = v3_of(v + 6);
  v3 e1 = th::sub(v1, v0), e2 = th::sub(v2, v0);
  v3 p = th::cross(d, e2);
  float det = th::dot(e1, p);
  if (th::fabs_(det) < 1e-8f) return 0;
  float inv = 1.0f / det;
  v3 oo = th::sub(o, v0);
  float u = th::dot(oo, p) * inv;
  if (!(u >= 0.0f && u <= 1.0f)) return 0;
  v3 q = th::cross(oo, e1);
  float vv = th::dot(d, q) * inv;
  if (vv < 0.0f || u + vv > 1.0f) return 0;
  float t = th::dot(e2, q) * inv;
  if (t < tmin || t > tmax) return 0;
  v3 gn = glam_normalize(th::cross(e1, e2));
  bool front = th::dot(d, gn) < 0.0f;
  float w0 = 1.0f - u - vv;
  v3 outward = gn;
  TH_GpuTriangleAttributes at;
  memset(&at, 0, sizeof at);
  if (attributes) at = *attributes;
  if (at.flags & 1u)
    outward = glam_normalize(th::add(th::add(th::smul(w0, v3_of(at.n0)), th::smul(u, v3_of(at.n1))),
                                     th::smul(vv, v3_of(at.n2))));
  if (th::dot(outward, gn) < 0.0f) outward = th::neg(outward);
  Hit h;
  h.t = t;
  h.p = th::add(o, th::smul(t, d));
  h.front = front;
  h.n = front ? outward : th::neg(outward);
  if (at.flags & 2u) {
    h.tu = (w0 * at.uv0[0] + u * at.uv1[0]) + vv * at.uv2[0];
    h.tv = (w0 * at.uv0[1] + u * at.uv1[1]) + vv * at.uv2[1];
  }
  if (out) fill_hit(h, out);
  return 1;
}

// aabb.rs:74-111 (infinite inverse component = parallel axis)
int32_t or_aabb_intersects(const float mn[3], const float mx[3], const float origin[3],
                           const float inv[3], float tmin, float tmax) {
  for (int a = 0; a < 3; ++a) {
    if (isinf(inv[a])) {
      if (origin[a] < mn[a] || origin[a] > mx[a]) return 0;
      continue;
    }
    float n = (mn[a] - origin[a]) * inv[a];
    float f = (mx[a] - origin[a]) * inv[a];
    if (n > f) {
      float t = n;
      n = f;
      f = t;
    }
    tmin = rmax(tmin, n);
    tmax = rmin(tmax, f);
    if (tmax < tmin) return 0;
  }
  return 1;
}

int32_t or_env_lookup(const or_scene* s, const float direction[3], float rgb[3]) {
  if (!s || s->params.environment_width == 0) return fail("scene has no environment background");
  v3 r = env_lookup(*s, s->params, v3_of(direction));
  rgb[0] = r.x;
  rgb[1] = r.y;
  rgb[2] = r.z;
  return TH_OK;
}

float or_env_pdf(const or_scene* s, const float direction[3]) {
  if (!s || s->params.environment_width == 0) return 0.0f;
  return env_pdf(*s, s->params, v3_of(direction));
}

int32_t or_env_sample(const or_scene* s, float selection, float jitter_u, float jitter_v,
                      float direction[3], float rgb[3], float* pdf) {
  if (!s || s->params.environment_width == 0) return fail("scene has no environment background");
  if (!env_has_mass(*s, s->params)) {
    *pdf = 0.0f;
    return 0;
  }
  EnvSample e = env_sample_with(*s, s->params, selection, jitter_u, jitter_v, nullptr);
  direction[0] = e.dir.x;
  direction[1] = e.dir.y;
  direction[2] = e.dir.z;
  rgb[0] = e.rad.x;
  rgb[1] = e.rad.y;
  rgb[2] = e.rad.z;
  *pdf = e.pdf;
  return e.pdf > 0.0f ? 1 : 0;
}

float or_reflectance(float cosine, float ratio) { return reflectance(cosine, ratio); }
void or_refract(const float d[3], const float n[3], float ratio, float out[3]) {
  v3 r = refract_vec(v3_of(d), v3_of(n), ratio);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}
float or_power_heuristic(float a, float b) { return power_heuristic(a, b); }

int32_t or_ray_color(const or_scene* s, const TH_GpuRenderParams* params, const float origin[3],
                     const float direction[3], uint32_t* rng_state, float rgb[3]) {
  if (!s || !params || !rng_state) return fail("null argument");
  Tracer tr{*s, *params, nullptr};
  v3 L = tr.ray_color(Ray{v3_of(origin), v3_of(direction)}, rng_state);
  rgb[0] = L.x;
  rgb[1] = L.y;
  rgb[2] = L.z;
  return TH_OK;
}

uint32_t or_pcg_hash(uint32_t x) { return th::pcg_hash(x); }
float or_rand01(uint32_t* state) { return th::rand01(state); }
// The row-then-texel CDF search the kernels use (th_math.h), for its check against
// the flat partition point.
uint32_t or_env_search_rows(const float* cdf, const float* rows, uint32_t w, uint32_t h, float sel) {
  return th::env_search_rows(cdf, rows, w, h, sel);
}
float or_sin(float x) { return th::fsin(x); }
float or_cos(float x) { return th::fcos(x); }
float or_atan2(float y, float x) { return th::fatan2(y, x); }
float or_acos(float x) { return th::facos(x); }

int32_t or_display_rgba8(const float* rgba, uint32_t width, uint32_t height, float exposure_stops,
                         uint8_t* out) {
  if (!rgba || !out) return fail("null argument");
  float scale = exp2f(exposure_stops);  // color.rs:23-25
  size_t n = (size_t)width * height;
  for (size_t i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) {
      float v = aces(rgba[4 * i + c] * scale);
      v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
      out[4 * i + c] = to_u8(lin_to_srgb(v));
    }
    out[4 * i + 3] = to_u8(rgba[4 * i + 3]);
  }
  return TH_OK;
}

}  // extern "C"
