// th_wide.h -- cooperative 4-lane traversal over a 4-wide BVH (gfx950).
//
// A path is owned by a quad of lanes (lanes 4k..4k+3 of a wave).  The quad walks a
// 4-wide hierarchy obtained by collapsing every interior node of the reference's
// median-split binary tree with its interior children (same leaves, same leaf primitive
// order): at an interior node lane j tests child j's box, at a leaf lane j tests
// primitive j, and the quad combines results with DPP quad permutes and, for the next
// entry, one ds_bpermute (the LDS crossbar, no LDS memory access).
// A scene may name up to 8 extracted primitives (DevScene::ext_*: large ones that would
// inflate every box above them); the quad tests those first, and the 4-wide boxes are
// the boxes of what remains below each node (the reference's boxes otherwise).  Children are visited nearest first; the farther hit children go on a
// per-quad LDS stack with their entry distance, re-checked against the current best
// when popped.  The closest hit is resolved exactly as the reference resolves it:
// smallest distance, ties to the reference's first-found primitive (leaves in
// descending first-slot order, ascending slot within a leaf).
//
// Record layout (node4 array, 128 B per node): 4 children x {float4(min.xyz, entry),
// float4(max.xyz, 0)}; an absent child has entry 0xffffffff.  Entries use the encoding
// of th_device.h, interior entries index node4 records.
#pragma once
#include "th_device.h"

namespace thd {

#define THD_NO_CHILD 0xffffffffu

struct Child4 {
  float4 a;  // min.xyz, entry bits
  float4 b;  // max.xyz, unused
};

// ------------------------------------------------------------------ quad exchange
template <int K>
__device__ __forceinline__ uint32_t qbcast_u(uint32_t v) {  // value of lane K of this quad
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, K * 0x55, 0xF, 0xF, false);
}
template <int K>
__device__ __forceinline__ float qbcast_f(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), K * 0x55, 0xF, 0xF, false));
}

// Value of the quad lane given by a DPP quad_perm pattern (0xB1: lane ^ 1, 0x4E: lane ^ 2,
// 0x1B: lane ^ 3).
template <int PERM>
__device__ __forceinline__ uint32_t qperm_u(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, PERM, 0xF, 0xF, false);
}

// Value held by the other quad of this lane's octet (lanes 8k..8k+7; DPP row_half_mirror,
// lane 8k+i reads lane 8k+7-i).  Meaningful where each quad holds one value in its four
// lanes and all eight lanes are active.
__device__ __forceinline__ uint32_t oswap_u(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, false);
}
__device__ __forceinline__ float oswap_f(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xF, 0xF, false));
}

// ------------------------------------------------------------------ node fetch
// Buffer resource over the node4 array (raw, 32-bit offsets; word 3 = 0x00020000, the
// gfx9 raw-buffer format).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t node_rsrc(const Child4* nodes4, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)nodes4, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ Child4 load_child(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  typedef unsigned v4u __attribute__((ext_vector_type(4)));
  v4u x = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
  v4u y = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16u, 0, 0);
  Child4 c;
  c.a = make_float4(__uint_as_float(x.x), __uint_as_float(x.y), __uint_as_float(x.z), __uint_as_float(x.w));
  c.b = make_float4(__uint_as_float(y.x), __uint_as_float(y.y), __uint_as_float(y.z), __uint_as_float(y.w));
  return c;
}

// ------------------------------------------------------------------ ray + box
struct RayPre {
  v3 o, d, inv;
  bool px, py, pz;  // |d_axis| < 1e-8: parallel axis (wgsl:712-717)
  bool anypar;
  float pinf, ninf;  // +inf / -inf, opaque to constant folding (mn2 / mx2)
};

__device__ __forceinline__ RayPre make_ray(const v3& o, const v3& d) {
  RayPre r;
  r.o = o;
  r.d = d;
  r.inv = th::mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  r.px = th::fabs_(d.x) < TH_EPSILON;
  r.py = th::fabs_(d.y) < TH_EPSILON;
  r.pz = th::fabs_(d.z) < TH_EPSILON;
  r.anypar = r.px || r.py || r.pz;
  // an empty asm hides the constants: the compiler would otherwise turn med3 against
  // +-inf back into fmaxf / fminf with their NaN-quieting instructions
  r.pinf = __builtin_inff();
  r.ninf = -__builtin_inff();
  __asm__("" : "+s"(r.pinf));
  __asm__("" : "+s"(r.ninf));
  return r;
}

// Branch-free form of intersects_aabb (wgsl:700-730): the per-axis early exits of the
// reference only ever shrink [t_min, t_max], so testing t_min <= t_max once at the end
// accepts exactly the same boxes; a parallel axis contributes (-inf, inf) when the
// origin lies in the slab and (inf, -inf) otherwise.
// min / max as v_med3 against -inf / +inf: exact for non-NaN operands (every slab value
// of a ray with no parallel axis is finite; a parallel axis's values are replaced below),
// without the NaN-quieting instructions fminf / fmaxf need in IEEE mode.
__device__ __forceinline__ float mn2(const RayPre& r, float a, float b) { return __builtin_amdgcn_fmed3f(a, b, r.ninf); }
__device__ __forceinline__ float mx2(const RayPre& r, float a, float b) { return __builtin_amdgcn_fmed3f(a, b, r.pinf); }

__device__ __forceinline__ bool box_bf(float mnx, float mny, float mnz, float mxx, float mxy,
                                       float mxz, const RayPre& r, float tmax, float* te) {
  const float INF = __builtin_inff();
  float nx = (mnx - r.o.x) * r.inv.x, fx = (mxx - r.o.x) * r.inv.x;
  float ny = (mny - r.o.y) * r.inv.y, fy = (mxy - r.o.y) * r.inv.y;
  float nz = (mnz - r.o.z) * r.inv.z, fz = (mxz - r.o.z) * r.inv.z;
  float ax = mn2(r, nx, fx), bx = mx2(r, nx, fx);
  float ay = mn2(r, ny, fy), by = mx2(r, ny, fy);
  float az = mn2(r, nz, fz), bz = mx2(r, nz, fz);
  if (__builtin_expect(!r.anypar, 1)) {  // (almost) every ray: no parallel axis
    float lo = mx2(r, mx2(r, TH_MIN_DISTANCE, ax), mx2(r, ay, az));
    float hi = mn2(r, mn2(r, tmax, bx), mn2(r, by, bz));
    *te = lo;
    return !(hi < lo);
  }
  if (r.px) {
    bool in = !(r.o.x < mnx || r.o.x > mxx);
    ax = in ? -INF : INF;
    bx = in ? INF : -INF;
  }
  if (r.py) {
    bool in = !(r.o.y < mny || r.o.y > mxy);
    ay = in ? -INF : INF;
    by = in ? INF : -INF;
  }
  if (r.pz) {
    bool in = !(r.o.z < mnz || r.o.z > mxz);
    az = in ? -INF : INF;
    bz = in ? INF : -INF;
  }
  float lo = __builtin_fmaxf(__builtin_fmaxf(TH_MIN_DISTANCE, ax), __builtin_fmaxf(ay, az));
  float hi = __builtin_fminf(__builtin_fminf(tmax, bx), __builtin_fminf(by, bz));
  *te = lo;
  return !(hi < lo);
}

// ------------------------------------------------------------------ traversal
// a[B + lq] with compile-time indices (a run-time index into an array passed by value
// would go through scratch memory)
template <int B>
__device__ __forceinline__ uint32_t quad_pick(const uint32_t (&a)[8], uint32_t lq) {
  uint32_t v = a[B];
  v = lq == 1u ? a[B + 1] : v;
  v = lq == 2u ? a[B + 2] : v;
  v = lq == 3u ? a[B + 3] : v;
  return v;
}

// anyhit false: closest hit (reference tie-breaking).  anyhit true: any hit with
// t < limit.  Every value that steers control flow is identical in the four lanes of the
// quad; different quads of a wave may run different modes in the same loop.  A quad
// with `active` false returns at once (no hit).  The first `cached` 4-wide nodes
// (breadth-first order: the top levels every ray visits) may be served from a copy in LDS.
__device__ __forceinline__ bool traverse_quad_rt(const DevScene& S, const Child4* __restrict__ nodes4,
                                                 uint32_t root4, bool active, bool anyhit,
                                                 const v3& o, const v3& d, float limit, Stack& st,
                                                 float* best_t, uint32_t* best_slot,
                                                 const Child4* lds_nodes, uint32_t cached) {
  const uint32_t lq = threadIdx.x & 3u;
  const float INF = __builtin_inff();
  float best = anyhit ? limit : TH_FLT_MAX;
  uint32_t bslot = THD_NO_SLOT, bleaf = 0;
  RayPre r = make_ray(o, d);
  float te;
  *best_t = TH_FLT_MAX;
  *best_slot = THD_NO_SLOT;
  const __amdgpu_buffer_rsrc_t nrs = node_rsrc(nodes4, S.node4_bytes);
  if (!active) return false;
  // the scene's extracted primitives first (the 4-wide boxes leave them out), lane j of
  // the quad testing primitives j and j + 4; candidates merged in lane order with the
  // reference tie rule
#pragma unroll
  for (int base = 0; base < 8; base += 4) {
    if ((uint32_t)base < S.ext_n) {
      const uint32_t k = (uint32_t)base + lq;
      const uint32_t slot = base == 0 ? quad_pick<0>(S.ext_slot, lq) : quad_pick<4>(S.ext_slot, lq);
      const uint32_t leaf = base == 0 ? quad_pick<0>(S.ext_leaf, lq) : quad_pick<4>(S.ext_leaf, lq);
      const uint32_t node = base == 0 ? quad_pick<0>(S.ext_node, lq) : quad_pick<4>(S.ext_node, lq);
      float t = INF;
      if (k < S.ext_n) {
        // the reference reaches the primitive only through its leaf's box (and the
        // ancestors', whose slab intervals contain the leaf's)
        const float4* nb = (const float4*)(S.nodes + node);
        const float4 mn = nb[0], mx = nb[1];
        uint32_t tag;
        const float tp = slot_t(S, slot, o, d, &tag);
        float tl;
        if (box_bf(mn.x, mn.y, mn.z, mx.x, mx.y, mx.z, r, TH_FLT_MAX, &tl)) t = tp;
      }
      if (anyhit) {
        const bool h = t < limit;
        if (qbcast_u<0>(h) | qbcast_u<1>(h) | qbcast_u<2>(h) | qbcast_u<3>(h)) {
          *best_t = 0.0f;
          *best_slot = slot;
          return true;
        }
      } else {
        const float tq[4] = {qbcast_f<0>(t), qbcast_f<1>(t), qbcast_f<2>(t), qbcast_f<3>(t)};
        const uint32_t sq[4] = {qbcast_u<0>(slot), qbcast_u<1>(slot), qbcast_u<2>(slot), qbcast_u<3>(slot)};
        const uint32_t lfq[4] = {qbcast_u<0>(leaf), qbcast_u<1>(leaf), qbcast_u<2>(leaf), qbcast_u<3>(leaf)};
#pragma unroll
        for (int w = 0; w < 4; ++w)
          if (tq[w] < TH_FLT_MAX &&
              (tq[w] < best || (tq[w] == best && (lfq[w] > bleaf || (lfq[w] == bleaf && sq[w] < bslot))))) {
            best = tq[w];
            bslot = sq[w];
            bleaf = lfq[w];
          }
      }
    }
  }
  if (!box_bf(S.root_min[0], S.root_min[1], S.root_min[2], S.root_max[0], S.root_max[1], S.root_max[2], r,
              best, &te)) {
    if (anyhit) return false;
    *best_t = best;
    *best_slot = bslot;
    return bslot != THD_NO_SLOT;
  }
  st.sp = 0;
  uint32_t e = root4;
  for (;;) {
    if (e & THD_LEAF_BIT) {
      uint32_t first, count;
      leaf_range(S, e, &first, &count);
      for (uint32_t base = 0; base < count; base += 4u) {
        uint32_t k = base + lq;
        float t = INF;
        if (k < count) {
          uint32_t tag;
          t = slot_t(S, first + k, o, d, &tag);
        }
        if (anyhit) {
          bool h = t < limit;
          uint32_t any = qbcast_u<0>(h) | qbcast_u<1>(h) | qbcast_u<2>(h) | qbcast_u<3>(h);
          if (any) {
            *best_t = 0.0f;
            *best_slot = first;
            return true;
          }
        } else {
          // quad argmin over (t, lane): the smallest slot wins among equal t in a leaf
          float t0 = qbcast_f<0>(t), t1 = qbcast_f<1>(t), t2 = qbcast_f<2>(t), t3 = qbcast_f<3>(t);
          uint32_t w = 0;
          float tw = t0;
          if (t1 < tw) { tw = t1; w = 1; }
          if (t2 < tw) { tw = t2; w = 2; }
          if (t3 < tw) { tw = t3; w = 3; }
          if (tw < TH_FLT_MAX) {
            uint32_t slot = first + base + w;
            if (tw < best || (tw == best && (first > bleaf || (first == bleaf && slot < bslot)))) {
              best = tw;
              bslot = slot;
              bleaf = first;
            }
          }
        }
      }
    } else {
      // LDS (top levels) or a buffer load: two distinct load forms, so the compiler
      // cannot merge them into one flat load that waits on both memory counters
      Child4 c;
      if (!(e < cached)) c = load_child(nrs, (4u * e + lq) * (uint32_t)sizeof(Child4));
      else c = lds_nodes[4u * e + lq];
      th_keep3(c.b.x, c.b.y, c.b.z);  // both halves of the child in one round trip
      uint32_t ce = __float_as_uint(c.a.w);
      float tE = INF;
      if (ce != THD_NO_CHILD) {
        float t;
        if (box_bf(c.a.x, c.a.y, c.a.z, c.b.x, c.b.y, c.b.z, r, best, &t)) tE = t;
      }
      // Visit order by key = entry-distance bits with the two low bits replaced by the
      // lane: unique in the quad, nearest first up to the last two mantissa bits
      // (distances are positive or +inf, whose bits order as the floats).  The order
      // only changes the work done: the closest hit and its tie-breaking do not depend
      // on it.  A missed child keys at or above MISS.
      const uint32_t MISS = 0x7f800000u;
      uint32_t key = (__float_as_uint(tE) & ~3u) | lq;
      // the other three lanes' keys (lane ^ 1, ^ 2, ^ 3 of the quad: no compare with itself)
      const uint32_t kx1 = qperm_u<0xB1>(key), kx2 = qperm_u<0x4E>(key), kx3 = qperm_u<0x1B>(key);
      // hit children of the quad: its four bits of the wave's hit mask
      const uint64_t hm = __ballot(key < MISS);
      const uint32_t nh = (uint32_t)__builtin_popcount((uint32_t)(hm >> (threadIdx.x & 60u)) & 0xfu);
      if (nh > 0u) {
        uint32_t rank = (kx1 < key) + (kx2 < key) + (kx3 < key);
        uint32_t w = __builtin_elementwise_min(__builtin_elementwise_min(key, kx1),
                                               __builtin_elementwise_min(kx2, kx3)) & 3u;
        // entry of lane w of the quad through the LDS crossbar (ds_bpermute: no memory
        // access; one instruction in place of four broadcasts and three selects, which
        // measured 2 % faster on the headline frame and the terrain passes)
        const uint32_t ew = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(((threadIdx.x & 60u) | w) << 2), (int)ce);
        // the other hit children, farthest deepest
        const int nsp = st.sp + (int)nh - 1;  // the stack top after the pushes
        if (key < MISS && rank > 0u) st.at(nsp - (int)rank) = make_uint2(ce, __float_as_uint(tE));
        st.sp = nsp;
        e = ew;
        continue;
      }
    }
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
  if (anyhit) return false;
  *best_t = best;
  *best_slot = bslot;
  return bslot != THD_NO_SLOT;
}

// Any hit with t < limit among the scene's extracted primitives, tested by one lane:
// the tests of traverse_quad_rt's first step in the same arithmetic (the primitive's
// distance, counted only through its reference leaf's box), so a ray this finds blocked
// is one the quad traversal reports occluded.  Shading kernels resolve such shadow rays
// at once instead of queueing them.
__device__ __forceinline__ bool lane_extracted_anyhit(const DevScene& S, const v3& o, const v3& d, float limit) {
  if (S.ext_n == 0u) return false;
  const RayPre r = make_ray(o, d);
#pragma unroll 1
  for (uint32_t k = 0; k < S.ext_n; ++k) {
    const float4* nb = (const float4*)(S.nodes + S.ext_node[k]);
    const float4 mn = nb[0], mx = nb[1];
    uint32_t tag;
    const float tp = slot_t(S, S.ext_slot[k], o, d, &tag);
    float tl;
    if (box_bf(mn.x, mn.y, mn.z, mx.x, mx.y, mx.z, r, TH_FLT_MAX, &tl) && tp < limit) return true;
  }
  return false;
}

// MODE 1: closest hit.  MODE 2: any hit with t < limit.
template <int MODE>
__device__ __forceinline__ bool traverse_quad(const DevScene& S, const Child4* __restrict__ nodes4,
                                              uint32_t root4, const v3& o, const v3& d, float limit,
                                              Stack& st, float* best_t, uint32_t* best_slot,
                                              const Child4* lds_nodes = nullptr,
                                              uint32_t cached = 0) {
  return traverse_quad_rt(S, nodes4, root4, true, MODE == 2, o, d, limit, st, best_t, best_slot,
                          lds_nodes, cached);
}

}  // namespace thd
