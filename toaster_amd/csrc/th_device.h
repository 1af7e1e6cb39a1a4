// th_device.h -- device-side scene layout, traversal and the path integrator (gfx950).
//
// Scene layout in HBM (built on the device from the uploaded reference records by
// prepare kernels, so the C boundary stays the reference's gpu_types.rs layout):
//   cnodes : one 64-byte child-pair record per reference BVH node index.  For an
//            interior node i (left child i+1, right child first_or_right,
//            flatten.rs:72-97) it holds both children's boxes and their traversal
//            entries, so one 64-byte fetch tests two boxes.
//   slots  : one 48-byte record per leaf primitive slot, in leaf order
//            (bvh_primitives order): triangles as (v0, mat) (v1-v0, tag) (v2-v0, mat),
//            spheres as (centre, radius) (0, tag) (0, mat).  Leaf tests read 48 B
//            instead of the reference's 64 B triangle + 80 B attribute record;
//            v1-v0 and v2-v0 are the same IEEE subtractions the reference performs
//            per test (wgsl:609-610), done once.
//   tag    : bit31 kind (1 triangle), bits 29-30 attribute flags, bits 0-28 the
//            original primitive index.
// Traversal entry encoding (one u32): bit31 = 0 -> interior node index; bit31 = 1 ->
// leaf with bits 25-30 = primitive count (1..62) and bits 0-24 = first slot; a count
// field of 63 marks a large leaf whose first/count are read from the reference node.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/toaster_hip.h"
#include "th_math.h"

namespace thd {
using th::v3;

struct CNode {
  float4 a;  // L.min.xyz, L.max.x
  float4 b;  // L.max.yz, R.min.xy
  float4 c;  // R.min.z, R.max.xyz
  uint4 d;   // L entry, R entry, unused, unused
};

#define THD_LEAF_BIT 0x80000000u
#define THD_BIG_LEAF 63u
#define THD_NO_SLOT 0xffffffffu
#define THD_KIND_TRI 0x80000000u

struct DevScene {
  const CNode* cnodes;
  const float4* slots;  // 3 float4 per slot
  const TH_GpuBvhNode* nodes;
  const TH_GpuSphere* spheres;
  const TH_GpuTriangle* triangles;
  const TH_GpuTriangleAttributes* attrs;
  const TH_GpuMaterial* materials;
  const TH_GpuLight* lights;
  const float4* env;
  const uint32_t* tex;
  uint32_t node_count;
  uint32_t root_entry;
  float root_min[3], root_max[3];
  const float4* nodes4;  // 4-wide nodes (th_wide.h), 8 float4 per node
  uint32_t root4;        // 4-wide traversal entry of the root
  uint32_t node4_bytes;  // size of the nodes4 array
  // environment CDF, compact: the .w of every texel (4 B each) and of the last texel of
  // every row; null when the map's CDF is not non-decreasing (then the flat search)
  const float* env_cdf;
  const float* env_rows;
  uint32_t env_rows_w;   // row width the row array was built for
  // progressive graph replay: {frame_index, accumulated_samples} in device memory that
  // the graph advances between replays; null for an ordinary launch (params rule)
  const uint32_t* frame_state;
  // primitives the 4-wide traversal tests before the tree (their boxes are left out of
  // the 4-wide boxes): slot, the first slot of its reference leaf (the tie rank) and that
  // leaf's reference node (whose box the reference must enter to reach the primitive)
  uint32_t ext_n;
  uint32_t ext_slot[8];
  uint32_t ext_leaf[8];
  uint32_t ext_node[8];
};

// The frame seed and accumulated sample count of a launch: from the params, or (FS, a
// kernel instantiated for a progressive graph) from the graph's device state.
template <bool FS>
__device__ __forceinline__ uint32_t frame_of(const DevScene& S, const TH_GpuRenderParams& P) {
  if constexpr (FS) return S.frame_state[0];
  return P.frame_index;
}
template <bool FS>
__device__ __forceinline__ uint32_t accumulated_of(const DevScene& S, const TH_GpuRenderParams& P) {
  if constexpr (FS) return S.frame_state[1];
  return P.accumulated_samples;
}

struct Counters {
  uint64_t paths, closest_rays, shadow_rays, nodes, primrefs, tris, spheres, closest_hits,
      tri_hits, lights_scanned, env_lookups, env_samples;
};

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }

// Marks three loaded values as used here (an empty asm statement, no instruction): the
// loads that produce them are issued with the loads before it instead of inside a later
// branch, so a record costs one memory round trip.
__device__ __forceinline__ void th_keep3(float x, float y, float z) {
  __asm__ volatile("" ::"v"(x), "v"(y), "v"(z));
}
__device__ __forceinline__ void th_keep4u(const uint4& v) {
  __asm__ volatile("" ::"v"(v.x), "v"(v.y), "v"(v.z), "v"(v.w));
}

// A whole light record (96 B) in one round trip.
__device__ __forceinline__ TH_GpuLight load_light(const DevScene& S, uint32_t i) {
  const uint4* p = (const uint4*)(S.lights + i);
  TH_GpuLight L;
  uint4 w[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) w[k] = p[k];
#pragma unroll
  for (int k = 0; k < 6; ++k) th_keep4u(w[k]);
  __builtin_memcpy(&L, w, sizeof L);
  return L;
}

// A whole material record (48 B) in one round trip.
__device__ __forceinline__ TH_GpuMaterial load_material(const DevScene& S, uint32_t i) {
  const uint4* p = (const uint4*)(S.materials + i);
  uint4 a = p[0], b = p[1], c = p[2];
  th_keep4u(a);
  th_keep4u(b);
  th_keep4u(c);
  TH_GpuMaterial m;
  m.kind = a.x;
  m.texture_offset = a.y;
  m.texture_width = a.z;
  m.texture_height = a.w;
  m.albedo[0] = __uint_as_float(b.x);
  m.albedo[1] = __uint_as_float(b.y);
  m.albedo[2] = __uint_as_float(b.z);
  m.albedo[3] = __uint_as_float(b.w);
  m.params[0] = __uint_as_float(c.x);
  m.params[1] = __uint_as_float(c.y);
  m.params[2] = __uint_as_float(c.z);
  m.params[3] = __uint_as_float(c.w);
  return m;
}

// ------------------------------------------------------------------ box test
// wgsl:700-730 with t_max supplied; returns the entry distance in *t_enter.
__device__ __forceinline__ bool box_test(float mnx, float mny, float mnz, float mxx, float mxy,
                                         float mxz, const v3& o, const v3& d, const v3& inv,
                                         float tmax, float* t_enter) {
  float tmin = TH_MIN_DISTANCE;
#define THD_AXIS(DA, OA, MN, MX, IA)                                  \
  if (th::fabs_(DA) < TH_EPSILON) {                                   \
    if (OA < MN || OA > MX) return false;                             \
  } else {                                                            \
    float nv = (MN - OA) * IA, fv = (MX - OA) * IA;                   \
    tmin = th::fmax_(tmin, th::fmin_(nv, fv));                        \
    tmax = th::fmin_(tmax, th::fmax_(nv, fv));                        \
    if (tmax < tmin) return false;                                    \
  }
  THD_AXIS(d.x, o.x, mnx, mxx, inv.x)
  THD_AXIS(d.y, o.y, mny, mxy, inv.y)
  THD_AXIS(d.z, o.z, mnz, mxz, inv.z)
#undef THD_AXIS
  *t_enter = tmin;
  return true;
}

// ------------------------------------------------------------------ primitive tests
// Distance-only forms of hit_triangle / hit_sphere (wgsl:551-664); TH_FLT_MAX = miss.
__device__ __forceinline__ float tri_t(const v3& o, const v3& d, const v3& v0, const v3& e1,
                                       const v3& e2) {
  v3 p = th::cross(d, e2);
  float det = th::dot(p, e1);
  if (th::fabs_(det) < TH_EPSILON) return TH_FLT_MAX;
  float inv = 1.0f / det;
  v3 oo = th::sub(o, v0);
  float u = th::dot(oo, p) * inv;
  if (u < 0.0f || u > 1.0f) return TH_FLT_MAX;
  v3 q = th::cross(oo, e1);
  float v = th::dot(d, q) * inv;
  if (v < 0.0f || (u + v) > 1.0f) return TH_FLT_MAX;
  float t = th::dot(e2, q) * inv;
  if (t < TH_MIN_DISTANCE || t > TH_MAX_DISTANCE) return TH_FLT_MAX;
  return t;
}

__device__ __forceinline__ float sphere_t(const v3& o, const v3& d, const v3& c, float r) {
  v3 off = th::sub(o, c);
  float a = th::dot(d, d);
  float hb = th::dot(off, d);
  float cc = th::dot(off, off) - r * r;
  float disc = hb * hb - a * cc;
  if (disc < 0.0f) return TH_FLT_MAX;
  float sq = th::fsqrt(disc);
  float t = (-hb - sq) / a;
  if (t < TH_MIN_DISTANCE) {
    t = (-1.0f * hb + sq) / a;
    if (t < TH_MIN_DISTANCE) return TH_FLT_MAX;
  }
  return t;
}

__device__ __forceinline__ float slot_t(const DevScene& S, uint32_t slot, const v3& o, const v3& d,
                                        uint32_t* tag_out) {
  const float4* p = S.slots + 3u * (size_t)slot;
  float4 a = p[0], b = p[1], c = p[2];
  // all 48 bytes in one round trip: without this use the third load would wait for the
  // tag (the compiler sinks it into the triangle branch)
  th_keep3(c.x, c.y, c.z);
  uint32_t tag = fbits(b.w);
  *tag_out = tag;
  if (tag & THD_KIND_TRI)
    return tri_t(o, d, th::mk(a.x, a.y, a.z), th::mk(b.x, b.y, b.z), th::mk(c.x, c.y, c.z));
  return sphere_t(o, d, th::mk(a.x, a.y, a.z), a.w);
}

// ------------------------------------------------------------------ stack
struct Stack {
  uint2* base;  // LDS, entry k of this lane at base[k * stride]
  int stride;
  int sp;
  // entry k (k and stride below 2^23: a full-rate 24-bit multiply)
  __device__ __forceinline__ uint2& at(int k) { return base[__umul24((uint32_t)k, (uint32_t)stride)]; }
  __device__ __forceinline__ void push(uint32_t e, float t) {
    at(sp) = make_uint2(e, fbits(t));
    ++sp;
  }
  __device__ __forceinline__ uint2 pop() {
    --sp;
    return at(sp);
  }
};

__device__ __forceinline__ void leaf_range(const DevScene& S, uint32_t e, uint32_t* first,
                                           uint32_t* count) {
  uint32_t c = (e >> 25) & 63u;
  if (c == THD_BIG_LEAF) {
    const TH_GpuBvhNode& n = S.nodes[e & 0x01ffffffu];
    *first = n.first_or_right;
    *count = n.primitive_count;
  } else {
    *first = e & 0x01ffffffu;
    *count = c;
  }
}

template <bool COUNT>
struct RayStats {
  uint32_t nodes = 0, primrefs = 0, tris = 0, spheres = 0;
  __device__ __forceinline__ void node(uint32_t n) {
    if (COUNT) nodes += n;
  }
  __device__ __forceinline__ void prim(uint32_t tag) {
    if (COUNT) {
      primrefs++;
      if (tag & THD_KIND_TRI)
        tris++;
      else
        spheres++;
    }
  }
};

// ------------------------------------------------------------------ traversal
// MODE 0: closest hit in the reference visit order (wgsl:746-792: right child popped
//         first), a child box is tested against the current best and its entry
//         distance re-checked when popped -- the same nodes pass as in the
//         reference's pop-time test.
// MODE 1: closest hit, nearer child first; equal-distance candidates are resolved as
//         the reference resolves them (its first found wins: leaves in descending
//         first-slot order, ascending slot within a leaf).
// MODE 2: any hit with t < limit (shadow rays; the reference's closest hit compared
//         with the same limit gives the same answer).
template <int MODE, bool COUNT>
__device__ __forceinline__ bool traverse(const DevScene& S, const v3& o, const v3& d, float limit,
                                         Stack& st, RayStats<COUNT>& rs, float* best_t,
                                         uint32_t* best_slot) {
  float best = (MODE == 2) ? limit : TH_FLT_MAX;
  uint32_t bslot = THD_NO_SLOT, bleaf = 0;
  v3 inv = th::mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  rs.node(1);
  float te;
  if (!box_test(S.root_min[0], S.root_min[1], S.root_min[2], S.root_max[0], S.root_max[1],
                S.root_max[2], o, d, inv, best, &te)) {
    *best_t = TH_FLT_MAX;
    *best_slot = THD_NO_SLOT;
    return false;
  }
  st.sp = 0;
  uint32_t e = S.root_entry;
  for (;;) {
    if (e & THD_LEAF_BIT) {
      uint32_t first, count;
      leaf_range(S, e, &first, &count);
      for (uint32_t i = 0; i < count; ++i) {
        uint32_t tag;
        float t = slot_t(S, first + i, o, d, &tag);
        rs.prim(tag);
        if (MODE == 2) {
          if (t < limit) {
            *best_t = t;
            *best_slot = first + i;
            return true;
          }
        } else if (MODE == 0) {
          if (t < best) {
            best = t;
            bslot = first + i;
          }
        } else {
          if (t < best || (t == best && t != TH_FLT_MAX &&
                           (first > bleaf || (first == bleaf && first + i < bslot)))) {
            best = t;
            bslot = first + i;
            bleaf = first;
          }
        }
      }
    } else {
      CNode n = S.cnodes[e];
      rs.node(2);
      float tl, tr;
      bool hl = box_test(n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, o, d, inv, best, &tl);
      bool hr = box_test(n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, o, d, inv, best, &tr);
      if (MODE == 0) {
        if (hl) st.push(n.d.x, tl);
        if (hr) {
          e = n.d.y;
          continue;
        }
      } else {
        if (hl && hr) {
          bool left_first = tl < tr;
          st.push(left_first ? n.d.y : n.d.x, left_first ? tr : tl);
          e = left_first ? n.d.x : n.d.y;
          continue;
        }
        if (hl) {
          e = n.d.x;
          continue;
        }
        if (hr) {
          e = n.d.y;
          continue;
        }
      }
    }
    // pop the next entry that still passes against the current best
    bool found = false;
    while (st.sp > 0) {
      uint2 x = st.pop();
      if (__uint_as_float(x.y) <= best) {
        e = x.x;
        found = true;
        break;
      }
    }
    if (!found) break;
  }
  if (MODE == 2) {
    *best_t = TH_FLT_MAX;
    *best_slot = THD_NO_SLOT;
    return false;
  }
  *best_t = best;
  *best_slot = bslot;
  return bslot != THD_NO_SLOT;
}

// Linear fallback when the scene has no BVH (wgsl:667-698): spheres then triangles,
// first found wins ties.  Slot encodes kind<<31 | index.
template <bool COUNT>
__device__ float closest_linear(const DevScene& S, uint32_t nsph, uint32_t ntri, const v3& o,
                                const v3& d, float limit, RayStats<COUNT>& rs, uint32_t* slot) {
  float best = TH_FLT_MAX;
  uint32_t bs = THD_NO_SLOT;
  for (uint32_t i = 0; i < nsph; ++i) {
    const float* cr = S.spheres[i].center_radius;
    float t = sphere_t(o, d, th::mk(cr[0], cr[1], cr[2]), cr[3]);
    rs.prim(0);
    if (best > t) {
      best = t;
      bs = i;
    }
  }
  for (uint32_t i = 0; i < ntri; ++i) {
    const TH_GpuTriangle& tr = S.triangles[i];
    v3 v0 = th::mk(tr.v0[0], tr.v0[1], tr.v0[2]);
    v3 e1 = th::sub(th::mk(tr.v1[0], tr.v1[1], tr.v1[2]), v0);
    v3 e2 = th::sub(th::mk(tr.v2[0], tr.v2[1], tr.v2[2]), v0);
    float t = tri_t(o, d, v0, e1, e2);
    rs.prim(THD_KIND_TRI);
    if (best > t) {
      best = t;
      bs = THD_KIND_TRI | i;
    }
  }
  (void)limit;
  *slot = bs;
  return best;
}

// ------------------------------------------------------------------ hit record
struct Hit {
  float t;
  v3 p, n;
  bool front;
  uint32_t mat;
  float tu, tv;
};

// Full hit record of the winning primitive (wgsl:582-595, 636-663), recomputed once.
__device__ __forceinline__ void finalize_tri(const DevScene& S, const v3& o, const v3& d, v3 v0,
                                             v3 e1, v3 e2, uint32_t mat, uint32_t orig,
                                             uint32_t flags, float t, Hit* h) {
  v3 p = th::cross(d, e2);
  float det = th::dot(p, e1);
  float inv = 1.0f / det;
  v3 oo = th::sub(o, v0);
  float u = th::dot(oo, p) * inv;
  v3 q = th::cross(oo, e1);
  float v = th::dot(d, q) * inv;
  v3 gn = th::normalize(th::cross(e1, e2));
  bool front = th::dot(d, gn) < 0.0f;
  float w0 = 1.0f - u - v;
  v3 outward = gn;
  h->tu = 0.0f;
  h->tv = 0.0f;
  if (flags) {
    const TH_GpuTriangleAttributes& at = S.attrs[orig];
    if (flags & 1u) {
      outward = th::normalize(th::add(
          th::add(th::smul(w0, th::mk(at.n0[0], at.n0[1], at.n0[2])),
                  th::smul(u, th::mk(at.n1[0], at.n1[1], at.n1[2]))),
          th::smul(v, th::mk(at.n2[0], at.n2[1], at.n2[2]))));
      if (th::dot(outward, gn) < 0.0f) outward = th::neg(outward);
    }
    if (flags & 2u) {
      h->tu = (w0 * at.uv0[0] + u * at.uv1[0]) + v * at.uv2[0];
      h->tv = (w0 * at.uv0[1] + u * at.uv1[1]) + v * at.uv2[1];
    }
  }
  h->t = t;
  h->p = th::add(o, th::smul(t, d));
  h->n = front ? outward : th::smul(-1.0f, outward);
  h->front = front;
  h->mat = mat;
}

__device__ __forceinline__ void finalize_sphere(const v3& o, const v3& d, v3 c, float r,
                                                uint32_t mat, float t, Hit* h) {
  h->t = t;
  h->p = th::add(o, th::smul(t, d));
  v3 tn = th::divs(th::sub(h->p, c), r);
  h->front = th::dot(d, tn) < 0.0f;
  h->n = h->front ? tn : th::neg(tn);
  h->mat = mat;
  h->tu = 0.0f;
  h->tv = 0.0f;
}

__device__ __forceinline__ void finalize_slot(const DevScene& S, uint32_t slot, const v3& o,
                                              const v3& d, float t, Hit* h) {
  const float4* p = S.slots + 3u * (size_t)slot;
  float4 a = p[0], b = p[1], c = p[2];
  th_keep3(a.x, a.y, a.z);  // the whole record in one round trip
  th_keep3(a.w, c.x, c.y);
  th_keep3(c.z, b.x, b.y);
  uint32_t tag = fbits(b.w);
  uint32_t mat = fbits(c.w);
  if (tag & THD_KIND_TRI) {
    finalize_tri(S, o, d, th::mk(a.x, a.y, a.z), th::mk(b.x, b.y, b.z), th::mk(c.x, c.y, c.z), mat,
                 tag & 0x1fffffffu, (tag >> 29) & 3u, t, h);
  } else {
    finalize_sphere(o, d, th::mk(a.x, a.y, a.z), a.w, mat, t, h);
  }
}

__device__ __forceinline__ void finalize_linear(const DevScene& S, uint32_t slot, const v3& o,
                                                const v3& d, float t, Hit* h) {
  if (slot & THD_KIND_TRI) {
    uint32_t i = slot & 0x7fffffffu;
    const TH_GpuTriangle& tr = S.triangles[i];
    v3 v0 = th::mk(tr.v0[0], tr.v0[1], tr.v0[2]);
    v3 e1 = th::sub(th::mk(tr.v1[0], tr.v1[1], tr.v1[2]), v0);
    v3 e2 = th::sub(th::mk(tr.v2[0], tr.v2[1], tr.v2[2]), v0);
    finalize_tri(S, o, d, v0, e1, e2, tr.material_index, i, S.attrs[i].flags & 3u, t, h);
  } else {
    const float* cr = S.spheres[slot].center_radius;
    finalize_sphere(o, d, th::mk(cr[0], cr[1], cr[2]), cr[3], S.spheres[slot].material_index, t, h);
  }
}

}  // namespace thd
