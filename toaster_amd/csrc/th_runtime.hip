// th_runtime.hip -- device runtime of libtoaster_hip: contexts, scene upload, prepare
// kernels, the path-tracing megakernel and the C ABI declared in include/toaster_hip.h.
//
// Replaces crates/toaster-gpu/src/{device.rs:49-100, buffers.rs:117-271,
// dispatch.rs:28-60, readback.rs:8-25} and the WGSL entry point
// shaders/pathtrace_spheres.wgsl:508-548.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "th_internal.h"

using namespace thd;

// ====================================================================== kernels
// Reference nodes -> child-pair records (one per reference node index).
__global__ void k_prepare_cnodes(const TH_GpuBvhNode* __restrict__ nodes, uint32_t count,
                                 CNode* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const TH_GpuBvhNode& n = nodes[i];
  CNode c;
  if (n.primitive_count > 0u) {
    c.a = make_float4(0, 0, 0, 0);
    c.b = c.a;
    c.c = c.a;
    c.d = make_uint4(0, 0, 0, 0);
  } else {
    uint32_t li = i + 1u, ri = n.first_or_right;
    const TH_GpuBvhNode& L = nodes[li];
    const TH_GpuBvhNode& R = nodes[ri];
    c.a = make_float4(L.min[0], L.min[1], L.min[2], L.max[0]);
    c.b = make_float4(L.max[1], L.max[2], R.min[0], R.min[1]);
    c.c = make_float4(R.min[2], R.max[0], R.max[1], R.max[2]);
    c.d = make_uint4(th_entry_of(L, li), th_entry_of(R, ri), 0u, 0u);
  }
  out[i] = c;
}

// Reference primitive refs -> 48-byte leaf-ordered slots.
__global__ void k_prepare_slots(const TH_GpuPrimitiveRef* __restrict__ prims, uint32_t count,
                                const TH_GpuSphere* __restrict__ spheres,
                                const TH_GpuTriangle* __restrict__ tris,
                                const TH_GpuTriangleAttributes* __restrict__ attrs,
                                float4* __restrict__ slots) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  TH_GpuPrimitiveRef r = prims[i];
  float4 a, b, c;
  if (r.kind == 1u) {
    const TH_GpuTriangle& t = tris[r.index];
    uint32_t tag = THD_KIND_TRI | ((attrs[r.index].flags & 3u) << 29) | (r.index & 0x1fffffffu);
    a = make_float4(t.v0[0], t.v0[1], t.v0[2], __uint_as_float(t.material_index));
    b = make_float4(t.v1[0] - t.v0[0], t.v1[1] - t.v0[1], t.v1[2] - t.v0[2], __uint_as_float(tag));
    c = make_float4(t.v2[0] - t.v0[0], t.v2[1] - t.v0[1], t.v2[2] - t.v0[2],
                    __uint_as_float(t.material_index));
  } else {
    const TH_GpuSphere& s = spheres[r.index];
    a = make_float4(s.center_radius[0], s.center_radius[1], s.center_radius[2], s.center_radius[3]);
    b = make_float4(0, 0, 0, __uint_as_float(r.index & 0x1fffffffu));
    c = make_float4(0, 0, 0, __uint_as_float(s.material_index));
  }
  slots[3u * i] = a;
  slots[3u * i + 1u] = b;
  slots[3u * i + 2u] = c;
}

// Refit, one level of the reference tree (deepest first): a leaf's box from its
// primitives, an interior node's from its two children, as the packer computes them
// (th_pack.cpp collect_prims / bounds_of): triangle vertex min/max -/+ 1e-5, sphere
// center -/+ radius, folded with "strictly smaller / larger replaces" in primitive
// order; left child before right keeps that order.  Topology fields are not touched.
__device__ __forceinline__ float refit_min(float acc, float x) { return x < acc ? x : acc; }
__device__ __forceinline__ float refit_max(float acc, float x) { return acc < x ? x : acc; }

__global__ void k_refit_level(TH_GpuBvhNode* __restrict__ nodes, const uint32_t* __restrict__ order,
                              uint32_t count, const TH_GpuPrimitiveRef* __restrict__ prims,
                              const TH_GpuSphere* __restrict__ spheres, const TH_GpuTriangle* __restrict__ tris) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t ni = order[i];
  TH_GpuBvhNode& n = nodes[ni];
  const float INF = __builtin_inff();
  float mn[3] = {INF, INF, INF}, mx[3] = {-INF, -INF, -INF};
  if (n.primitive_count > 0u) {
    for (uint32_t k = 0; k < n.primitive_count; ++k) {
      const TH_GpuPrimitiveRef r = prims[n.first_or_right + k];
      if (r.kind == 1u) {
        const TH_GpuTriangle& t = tris[r.index];
        for (int a = 0; a < 3; ++a) {
          const float lo = refit_min(refit_min(t.v0[a], t.v1[a]), t.v2[a]) - 1e-5f;
          const float hi = refit_max(refit_max(t.v0[a], t.v1[a]), t.v2[a]) + 1e-5f;
          mn[a] = refit_min(mn[a], lo);
          mx[a] = refit_max(mx[a], hi);
        }
      } else {
        const float* cr = spheres[r.index].center_radius;
        for (int a = 0; a < 3; ++a) {
          mn[a] = refit_min(mn[a], cr[a] - cr[3]);
          mx[a] = refit_max(mx[a], cr[a] + cr[3]);
        }
      }
    }
  } else {
    const TH_GpuBvhNode& L = nodes[ni + 1u];
    const TH_GpuBvhNode& R = nodes[n.first_or_right];
    for (int a = 0; a < 3; ++a) {
      mn[a] = refit_min(L.min[a], R.min[a]);
      mx[a] = refit_max(L.max[a], R.max[a]);
    }
  }
  for (int a = 0; a < 3; ++a) {
    n.min[a] = mn[a];
    n.max[a] = mx[a];
  }
}

// Refit of the 4-wide records: each child slot takes the box of its reference node and
// keeps its entry.
__global__ void k_refit_wide(Child4* __restrict__ nodes4, const uint32_t* __restrict__ src, uint32_t count,
                             const TH_GpuBvhNode* __restrict__ nodes) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t b = src[i];
  if (b == THD_NO_CHILD) return;
  const TH_GpuBvhNode& g = nodes[b];
  Child4 c = nodes4[i];
  c.a = make_float4(g.min[0], g.min[1], g.min[2], c.a.w);
  c.b = make_float4(g.max[0], g.max[1], g.max[2], c.b.w);
  nodes4[i] = c;
}

// Summed surface area of all node boxes (the weight of a SAH cost), one block, fixed
// order: per-thread strided sums, then a tree reduction in LDS.
__global__ void k_area_sum(const TH_GpuBvhNode* __restrict__ nodes, uint32_t count, float* __restrict__ out) {
  __shared__ float part[256];
  float acc = 0.0f;
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) {
    const TH_GpuBvhNode& n = nodes[i];
    const float dx = n.max[0] - n.min[0], dy = n.max[1] - n.min[1], dz = n.max[2] - n.min[2];
    acc = acc + 2.0f * (dx * dy + dy * dz + dz * dx);
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t w = blockDim.x / 2u; w > 0u; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] = part[threadIdx.x] + part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0];
}

// Pixel of a lane: each 64-lane wave covers one 8x8 tile (coherent primary rays).
__device__ __forceinline__ bool tile_pixel(uint32_t W, uint32_t H, uint32_t* x, uint32_t* y) {
  uint32_t lane = threadIdx.x & 63u;
  uint32_t tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  uint32_t tiles_x = (W + 7u) >> 3;
  *x = (tile % tiles_x) * 8u + (lane & 7u);
  *y = (tile / tiles_x) * 8u + (lane >> 3);
  return *x < W && *y < H;
}

// The megakernel: one lane per pixel, params.samples samples from the pixel's
// reference PCG stream, then the progressive running average (wgsl:508-548).
template <int TMODE, bool COUNT, bool FS = false>
__global__ void __launch_bounds__(64) k_render_pixel(DevScene S, TH_GpuRenderParams P,
                                                     TH_GpuCamera cam, float4* __restrict__ accum,
                                                     const float* __restrict__ srgb_lut,
                                                     unsigned long long* counters) {
  extern __shared__ uint2 lds_stack[];
  uint32_t x, y;
  if (!tile_pixel(P.width, P.height, &x, &y)) return;
  Stack st{lds_stack + threadIdx.x, (int)blockDim.x, 0};
  LaneCounters<COUNT> lc;
  Integrator<TMODE, COUNT> it{S, P, srgb_lut, st, lc};
  uint32_t idx = y * P.width + x;
  uint32_t s = th::pcg_hash(idx ^ (frame_of<FS>(S, P) * 9781u));
  float cr = 0.0f, cg = 0.0f, cb = 0.0f, ca = 0.0f;
  for (uint32_t i = 0; i < P.samples; ++i) {
    v3 o, d;
    camera_ray(cam, P, x, y, &s, &o, &d);
    v3 L = it.ray_color(o, d, &s);
    cr = cr + L.x;
    cg = cg + L.y;
    cb = cb + L.z;
    ca = ca + 1.0f;
  }
  if (COUNT) lc.paths += P.samples;
  float ns = (float)P.samples;
  float4 b = make_float4(cr / ns, cg / ns, cb / ns, ca / ns);
  if (accumulated_of<FS>(S, P) == 0u) {
    accum[idx] = b;
  } else {
    float prev = (float)accumulated_of<FS>(S, P);
    float tot = prev + ns;
    float4 o = accum[idx];
    accum[idx] = make_float4((o.x * prev + b.x * ns) / tot, (o.y * prev + b.y * ns) / tot,
                             (o.z * prev + b.z * ns) / tot, (o.w * prev + b.w * ns) / tot);
  }
  lc.flush(counters);
}

// Per-sample RNG variant: G consecutive lanes share one pixel and split its samples
// (lane j takes samples j, j+G, ...); sample s of pixel idx draws from the stream
// seeded pcg_hash((idx ^ frame*9781) + s*0x9E3779B9), which for s = 0 is the
// reference seed.  The G partial sums are combined in a fixed tree order.
template <int TMODE, bool COUNT, int G, bool FS = false>
__global__ void __launch_bounds__(64) k_render_sample(DevScene S, TH_GpuRenderParams P,
                                                      TH_GpuCamera cam, float4* __restrict__ accum,
                                                      const float* __restrict__ srgb_lut,
                                                      unsigned long long* counters) {
  extern __shared__ uint2 lds_stack[];
  uint32_t lane = threadIdx.x & 63u;
  uint32_t gid = (blockIdx.x * blockDim.x + threadIdx.x) / G;  // pixel slot
  uint32_t j = lane % G;
  // pixels per wave = 64/G, laid out as tiles of (8 x 8/G)
  uint32_t ppw = 64u / G;
  uint32_t wave = gid / ppw, pw = gid % ppw;
  uint32_t th_ = 8u / (G >= 8 ? 8u : (uint32_t)G);  // tile height
  uint32_t tw = ppw / th_;
  uint32_t tiles_x = (P.width + tw - 1u) / tw;
  uint32_t x = (wave % tiles_x) * tw + pw % tw;
  uint32_t y = (wave / tiles_x) * th_ + pw / tw;
  bool active = x < P.width && y < P.height;
  Stack st{lds_stack + threadIdx.x, (int)blockDim.x, 0};
  LaneCounters<COUNT> lc;
  Integrator<TMODE, COUNT> it{S, P, srgb_lut, st, lc};
  uint32_t idx = y * P.width + x;
  float cr = 0.0f, cg = 0.0f, cb = 0.0f, ca = 0.0f;
  if (active) {
    uint32_t base = idx ^ (frame_of<FS>(S, P) * 9781u);
    for (uint32_t smp = j; smp < P.samples; smp += G) {
      uint32_t s = th::pcg_hash(base + smp * 0x9E3779B9u);
      v3 o, d;
      camera_ray(cam, P, x, y, &s, &o, &d);
      v3 L = it.ray_color(o, d, &s);
      cr = cr + L.x;
      cg = cg + L.y;
      cb = cb + L.z;
      ca = ca + 1.0f;
      if (COUNT) lc.paths++;
    }
  }
#pragma unroll
  for (int off = 1; off < G; off <<= 1) {
    cr = cr + __shfl_xor(cr, off, 64);
    cg = cg + __shfl_xor(cg, off, 64);
    cb = cb + __shfl_xor(cb, off, 64);
    ca = ca + __shfl_xor(ca, off, 64);
  }
  if (active && j == 0u) {
    float ns = (float)P.samples;
    float4 b = make_float4(cr / ns, cg / ns, cb / ns, ca / ns);
    if (accumulated_of<FS>(S, P) == 0u) {
      accum[idx] = b;
    } else {
      float prev = (float)accumulated_of<FS>(S, P);
      float tot = prev + ns;
      float4 o = accum[idx];
      accum[idx] = make_float4((o.x * prev + b.x * ns) / tot, (o.y * prev + b.y * ns) / tot,
                               (o.z * prev + b.z * ns) / tot, (o.w * prev + b.w * ns) / tot);
    }
  }
  lc.flush(counters);
}

// Closest hit of each pixel's first camera ray (parity query).
template <int TMODE>
__global__ void __launch_bounds__(64) k_primary(DevScene S, TH_GpuRenderParams P, TH_GpuCamera cam,
                                                uint32_t* __restrict__ slot_out,
                                                float* __restrict__ t_out) {
  extern __shared__ uint2 lds_stack[];
  uint32_t x, y;
  if (!tile_pixel(P.width, P.height, &x, &y)) return;
  Stack st{lds_stack + threadIdx.x, (int)blockDim.x, 0};
  uint32_t idx = y * P.width + x;
  uint32_t s = th::pcg_hash(idx ^ (frame_of<false>(S, P) * 9781u));
  v3 o, d;
  camera_ray(cam, P, x, y, &s, &o, &d);
  RayStats<false> rs;
  float t;
  uint32_t slot;
  if (S.node_count == 0u)
    t = closest_linear<false>(S, P.sphere_count, P.triangle_count, o, d, TH_FLT_MAX, rs, &slot);
  else
    traverse<TMODE, false>(S, o, d, TH_FLT_MAX, st, rs, &t, &slot);
  slot_out[idx] = slot;
  t_out[idx] = t;
}

// Quad kernels: four lanes own one path (th_wide.h).  G quads share a pixel in the
// per-sample mode (sample j, j+G, ... per quad); in the reference RNG mode G = 1 and the
// quad runs the pixel's samples in stream order.  A wave covers 16/G pixels.
// LDS layout of the quad kernels: `cached` 4-wide nodes, then the per-quad stacks.
template <bool PER_SAMPLE, int G, bool FS = false>
__global__ void __launch_bounds__(256) k_render_quad(DevScene S, TH_GpuRenderParams P,
                                                     TH_GpuCamera cam, float4* __restrict__ accum,
                                                     const float* __restrict__ srgb_lut,
                                                     uint32_t cached) {
  extern __shared__ uint2 lds_raw[];
  Child4* lds_nodes = (Child4*)lds_raw;
  for (uint32_t i = threadIdx.x; i < 4u * cached; i += blockDim.x)
    lds_nodes[i] = ((const Child4*)S.nodes4)[i];
  __syncthreads();
  uint2* lds_stack = lds_raw + 4u * cached * (sizeof(Child4) / sizeof(uint2));
  constexpr uint32_t PPW = 16u / G;
  constexpr uint32_t TW = PPW >= 8u ? 4u : (PPW >= 2u ? 2u : 1u);
  constexpr uint32_t TH = PPW / TW;
  const uint32_t lane = threadIdx.x & 63u, qw = lane >> 2;
  const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint32_t pw = qw / G, j = qw % G;
  const uint32_t tiles_x = (P.width + TW - 1u) / TW;
  const uint32_t x = (wave % tiles_x) * TW + pw % TW;
  const uint32_t y = (wave / tiles_x) * TH + pw / TW;
  const bool active = x < P.width && y < P.height;
  Stack st{lds_stack + (threadIdx.x >> 2), (int)(blockDim.x >> 2), 0};
  LaneCounters<false> lc;
  Integrator<3, false> it{S, P, srgb_lut, st, lc, lds_nodes, cached};
  const uint32_t idx = y * P.width + x;
  float cr = 0.0f, cg = 0.0f, cb = 0.0f, ca = 0.0f;
  if (active) {
    uint32_t base = idx ^ (frame_of<FS>(S, P) * 9781u);
    uint32_t s = th::pcg_hash(base);
    for (uint32_t smp = PER_SAMPLE ? j : 0u; smp < P.samples; smp += PER_SAMPLE ? G : 1u) {
      if (PER_SAMPLE) s = th::pcg_hash(base + smp * 0x9E3779B9u);
      v3 o, d;
      camera_ray(cam, P, x, y, &s, &o, &d);
      v3 L = it.ray_color(o, d, &s);
      cr = cr + L.x;
      cg = cg + L.y;
      cb = cb + L.z;
      ca = ca + 1.0f;
    }
  }
#pragma unroll
  for (int off = 4; off < 4 * G; off <<= 1) {
    cr = cr + __shfl_xor(cr, off, 64);
    cg = cg + __shfl_xor(cg, off, 64);
    cb = cb + __shfl_xor(cb, off, 64);
    ca = ca + __shfl_xor(ca, off, 64);
  }
  if (active && (lane % (4u * G)) == 0u) {
    float ns = (float)P.samples;
    float4 b = make_float4(cr / ns, cg / ns, cb / ns, ca / ns);
    if (accumulated_of<FS>(S, P) == 0u) {
      accum[idx] = b;
    } else {
      float prev = (float)accumulated_of<FS>(S, P);
      float tot = prev + ns;
      float4 o = accum[idx];
      accum[idx] = make_float4((o.x * prev + b.x * ns) / tot, (o.y * prev + b.y * ns) / tot,
                               (o.z * prev + b.z * ns) / tot, (o.w * prev + b.w * ns) / tot);
    }
  }
}

// Eight lanes per path: lanes 8k..8k+3 trace the path, lanes 8k+4..8k+7 the shadow rays
// of its previous vertex at the same time (OctetPath).  Reference RNG (G = 1): the octet
// runs the pixel's samples in stream order.  Per-sample RNG: G octets share a pixel
// (samples j, j+G, ... per octet), combined by the same butterfly as k_render_quad.  A
// wave covers 8/G pixels.  The reference-RNG variant without an environment is held to
// 96 VGPRs (5 waves per SIMD), which it fits without spilling.
template <bool ENV, bool PER_SAMPLE, int G, bool FS = false>
__global__ void __launch_bounds__(256, (ENV || PER_SAMPLE) ? 4 : 5) k_render_octet(DevScene S, TH_GpuRenderParams P,
                                                      TH_GpuCamera cam, float4* __restrict__ accum,
                                                      const float* __restrict__ srgb_lut,
                                                      uint32_t cached) {
  extern __shared__ uint2 lds_raw[];
  Child4* lds_nodes = (Child4*)lds_raw;
  for (uint32_t i = threadIdx.x; i < 4u * cached; i += blockDim.x)
    lds_nodes[i] = ((const Child4*)S.nodes4)[i];
  __syncthreads();
  uint2* lds_stack = lds_raw + 4u * cached * (sizeof(Child4) / sizeof(uint2));
  constexpr uint32_t PPW = 8u / G;
  constexpr uint32_t TW = PPW >= 4u ? 4u : (PPW >= 2u ? 2u : 1u);
  constexpr uint32_t TH = PPW / TW;
  const uint32_t lane = threadIdx.x & 63u, ow = lane >> 3;
  const bool shadow = ((lane >> 2) & 1u) != 0u;
  const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint32_t pw = ow / G, j = ow % G;
  const uint32_t tiles_x = (P.width + TW - 1u) / TW;
  const uint32_t x = (wave % tiles_x) * TW + pw % TW;
  const uint32_t y = (wave / tiles_x) * TH + pw / TW;
  const bool active = x < P.width && y < P.height;
  Stack st{lds_stack + (threadIdx.x >> 2), (int)(blockDim.x >> 2), 0};
  LaneCounters<false> lc;
  Integrator<3, false> it{S, P, srgb_lut, st, lc, lds_nodes, cached};
  const uint32_t idx = y * P.width + x;
  OctetPath<ENV, 3, false> path{it, shadow};
  if (active) {
    uint32_t base = idx ^ (frame_of<FS>(S, P) * 9781u);
    uint32_t s = th::pcg_hash(base);
    for (uint32_t smp = PER_SAMPLE ? j : 0u; smp < P.samples; smp += PER_SAMPLE ? G : 1u) {
      if (PER_SAMPLE) s = th::pcg_hash(base + smp * 0x9E3779B9u);
      v3 o, d;
      camera_ray(cam, P, x, y, &s, &o, &d);
      path.sample(o, d, &s, smp + (PER_SAMPLE ? G : 1u) >= P.samples);
    }
  }
  float cr = path.cr, cg = path.cg, cb = path.cb, ca = path.ca;
#pragma unroll
  for (int off = 8; off < 8 * G; off <<= 1) {
    cr = cr + __shfl_xor(cr, off, 64);
    cg = cg + __shfl_xor(cg, off, 64);
    cb = cb + __shfl_xor(cb, off, 64);
    ca = ca + __shfl_xor(ca, off, 64);
  }
  if (active && (lane % (8u * G)) == 0u) {
    float ns = (float)P.samples;
    float4 b = make_float4(cr / ns, cg / ns, cb / ns, ca / ns);
    if (accumulated_of<FS>(S, P) == 0u) {
      accum[idx] = b;
    } else {
      float prev = (float)accumulated_of<FS>(S, P);
      float tot = prev + ns;
      float4 o = accum[idx];
      accum[idx] = make_float4((o.x * prev + b.x * ns) / tot, (o.y * prev + b.y * ns) / tot,
                               (o.z * prev + b.z * ns) / tot, (o.w * prev + b.w * ns) / tot);
    }
  }
}

// Reference RNG, speculative: G octets per pixel.  A pixel's samples share one random
// stream (each draw replaces the state by pcg_hash of it), so sample k starts in the state
// where sample k-1 ended, and a sample is a function of its start state alone.  One octet
// (the lead) traces the pixel's next sample from its exact start; the others trace the
// sample after it from candidate starts: the state `2 + per_vertex * b` draws after the
// lead's start, i.e. the start that follows a lead whose path ends after b diffuse
// vertices, for b = max_bounces, max_bounces - 1, ...  When the lead ends, its radiance is
// added to the pixel (samples in order, so the image equals the sequential kernels' bit for
// bit); a candidate whose start equals the lead's end state is exactly the reference's
// next sample and becomes the lead, and every other octet restarts as a candidate for the
// sample after that.  With no match the next sample restarts from the known state, no
// later than the sequential order would start it.  The dispatch runs G = 2: a single
// candidate for a full-length lead; more candidates multiply the work past what the
// frame's idle lanes absorb.
#define TH_SPEC_LEAD 0u
#define TH_SPEC_CAND 1u
#define TH_SPEC_IDLE 2u

template <bool ENV, int G, bool FS = false>
__global__ void __launch_bounds__(256, 4) k_render_spec(DevScene S, TH_GpuRenderParams P, TH_GpuCamera cam,
                                                        float4* __restrict__ accum,
                                                        const float* __restrict__ srgb_lut, uint32_t cached,
                                                        uint32_t per_vertex) {
  extern __shared__ uint2 lds_raw[];
  Child4* lds_nodes = (Child4*)lds_raw;
  for (uint32_t i = threadIdx.x; i < 4u * cached; i += blockDim.x)
    lds_nodes[i] = ((const Child4*)S.nodes4)[i];
  __syncthreads();
  uint2* lds_stack = lds_raw + 4u * cached * (sizeof(Child4) / sizeof(uint2));
  constexpr uint32_t PPW = 8u / G;
  constexpr uint32_t TW = PPW >= 4u ? 4u : (PPW >= 2u ? 2u : 1u);
  constexpr uint32_t TH = PPW / TW;
  const uint32_t lane = threadIdx.x & 63u, ow = lane >> 3;
  const bool shadow = ((lane >> 2) & 1u) != 0u;
  const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint32_t pw = ow / G, j = ow % G;
  const uint32_t gl = pw * 8u * G;  // first lane of this pixel's octets
  const uint32_t tiles_x = (P.width + TW - 1u) / TW;
  const uint32_t x = (wave % tiles_x) * TW + pw % TW;
  const uint32_t y = (wave / tiles_x) * TH + pw / TW;
  const bool active = x < P.width && y < P.height;
  Stack st{lds_stack + (threadIdx.x >> 2), (int)(blockDim.x >> 2), 0};
  LaneCounters<false> lc;
  Integrator<3, false> it{S, P, srgb_lut, st, lc, lds_nodes, cached};
  const uint32_t idx = y * P.width + x;
  OctetPath<ENV, 3, false> path{it, shadow};
  const uint32_t n = P.samples, B = P.max_bounces;
  float pr = 0.0f, pg = 0.0f, pb = 0.0f, pa = 0.0f;  // pixel sums in sample order
  if (active && n > 0u) {
    uint32_t role = TH_SPEC_IDLE, start = 0u, s = 0u;
    bool begun = false, fin = false;
    uint32_t next = 0u, lq = 0u;  // the lead's sample and octet
    // (re)assigns the octets for lead octet lq starting sample `next` at state e: the lead
    // restarts unless `keep`; the others become candidates for sample next + 1
    auto assign = [&](uint32_t e, bool keep) {
      uint32_t cand[G > 1 ? G - 1 : 1];
      {
        // one walk of the stream from e records the candidate starts: rank r takes
        // b = B - r diffuse vertices (ranks with b < 0 stay idle)
        uint32_t h = e;
        const uint32_t len = 2u + per_vertex * B;
        if (G == 2) {
          // the single candidate (b = B) starts where the walk ends
          for (uint32_t t = 1u; t <= len; ++t) h = th::pcg_hash(h);
          cand[0] = h;
        } else {
#pragma unroll
          for (int r = 0; r < G - 1; ++r) cand[r] = 0u;
          for (uint32_t t = 1u; t <= len; ++t) {
            h = th::pcg_hash(h);
#pragma unroll
            for (int r = 0; r < G - 1; ++r)
              if ((uint32_t)r <= B && t == 2u + per_vertex * (B - (uint32_t)r)) cand[r] = h;
          }
        }
      }
      if (j == lq) {
        role = TH_SPEC_LEAD;
        if (!keep) {
          start = s = e;
          begun = fin = false;
          path.abandon();
        }
      } else {
        const uint32_t r = j < lq ? j : j - 1u;  // rank among the other octets
        bool use = false;
        uint32_t c = 0u;
#pragma unroll
        for (int q = 0; q < G - 1; ++q)
          if ((uint32_t)q == r) {
            use = (uint32_t)q <= B && next + 1u < n;
            c = cand[q];
          }
        role = use ? TH_SPEC_CAND : TH_SPEC_IDLE;
        start = s = c;
        begun = fin = false;
        path.abandon();
      }
    };
    assign(th::pcg_hash(idx ^ (frame_of<FS>(S, P) * 9781u)), false);
    for (;;) {
      if (role != TH_SPEC_IDLE && !fin) {
        if (!begun) {
          v3 o, d;
          camera_ray(cam, P, x, y, &s, &o, &d);
          path.begin(o, d);
          path.cr = path.cg = path.cb = path.ca = 0.0f;
          begun = true;
        }
        if (!path.step(&s, true)) fin = true;
      }
      bool done = false;
      // the lead has ended: add its radiance, hand the lead to the matching candidate
      // (which may have ended too: then add it at once, and so on)
      while (__shfl((int)fin, (int)(gl + 8u * lq), 64) != 0) {
        const int src = (int)(gl + 8u * lq);
        const uint32_t e = (uint32_t)__shfl((int)s, src, 64);
        const float lr = __shfl(path.cr, src, 64), lg = __shfl(path.cg, src, 64), lb = __shfl(path.cb, src, 64);
        pr = pr + lr;
        pg = pg + lg;
        pb = pb + lb;
        pa = pa + 1.0f;
        if (++next >= n) {
          done = true;
          break;
        }
        const bool mine = role == TH_SPEC_CAND && start == e;
        uint32_t nl = lq;
        bool found = false;
#pragma unroll
        for (int q = 0; q < G; ++q)
          if (!found && __shfl((int)mine, (int)(gl + 8u * (uint32_t)q), 64) != 0) {
            found = true;
            nl = (uint32_t)q;
          }
        lq = nl;
        assign(e, found);
      }
      if (done) break;
    }
  }
  if (active && (lane % (8u * G)) == 0u) {
    float ns = (float)P.samples;
    float4 b = make_float4(pr / ns, pg / ns, pb / ns, pa / ns);
    if (accumulated_of<FS>(S, P) == 0u) {
      accum[idx] = b;
    } else {
      float prev = (float)accumulated_of<FS>(S, P);
      float tot = prev + ns;
      float4 o = accum[idx];
      accum[idx] = make_float4((o.x * prev + b.x * ns) / tot, (o.y * prev + b.y * ns) / tot,
                               (o.z * prev + b.z * ns) / tot, (o.w * prev + b.w * ns) / tot);
    }
  }
}

// 4-wide nodes the quad and octet kernels copy into LDS: the top four levels.
#ifndef TH_QUAD_LDS_NODES
#define TH_QUAD_LDS_NODES 0u
#endif

// Reference RNG with dynamic pixel assignment.  A pixel's samples share one random
// stream, so its samples run in sequence and a statically assigned wave lasts as long as
// its slowest pixel.  Here every quad is persistent: when its pixel is done it takes the
// next one (4x4-tile order, one atomic per wave for the quads that fetch together), so a
// wave's time is its share of the frame.  Same per-pixel work and order as k_render_quad.
template <bool FS = false>
__global__ void __launch_bounds__(256) k_render_quad_dyn(DevScene S, TH_GpuRenderParams P,
                                                         TH_GpuCamera cam, float4* __restrict__ accum,
                                                         const float* __restrict__ srgb_lut,
                                                         uint32_t cached, uint32_t* __restrict__ work) {
  extern __shared__ uint2 lds_raw[];
  Child4* lds_nodes = (Child4*)lds_raw;
  for (uint32_t i = threadIdx.x; i < 4u * cached; i += blockDim.x)
    lds_nodes[i] = ((const Child4*)S.nodes4)[i];
  __syncthreads();
  uint2* lds_stack = lds_raw + 4u * cached * (sizeof(Child4) / sizeof(uint2));
  const uint32_t lane = threadIdx.x & 63u, lq = lane & 3u;
  Stack st{lds_stack + (threadIdx.x >> 2), (int)(blockDim.x >> 2), 0};
  LaneCounters<false> lc;
  Integrator<3, false> it{S, P, srgb_lut, st, lc, lds_nodes, cached};
  const uint32_t tiles_x = (P.width + 3u) / 4u;
  const uint32_t n_tiled = tiles_x * ((P.height + 3u) / 4u) * 16u;
  for (;;) {
    // quads fetching now (all four lanes active): one atomic for the wave
    const uint64_t m = __ballot(lq == 0u);
    const uint32_t first = (uint32_t)__ffsll((unsigned long long)m) - 1u;
    uint32_t base = 0;
    if (lane == first) base = atomicAdd(work, (uint32_t)__popcll(m));
    base = __shfl(base, (int)first, 64);
    uint32_t p = base + (uint32_t)__popcll(m & ((1ull << (lane & ~3u)) - 1ull));
    p = __shfl(p, (int)(lane & ~3u), 64);
    if (p >= n_tiled) break;
    const uint32_t tile = p >> 4, w = p & 15u;
    const uint32_t x = (tile % tiles_x) * 4u + (w & 3u), y = (tile / tiles_x) * 4u + (w >> 2);
    if (x >= P.width || y >= P.height) continue;
    const uint32_t idx = y * P.width + x;
    uint32_t s = th::pcg_hash(idx ^ (frame_of<FS>(S, P) * 9781u));
    float cr = 0.0f, cg = 0.0f, cb = 0.0f, ca = 0.0f;
    for (uint32_t smp = 0; smp < P.samples; ++smp) {
      v3 o, d;
      camera_ray(cam, P, x, y, &s, &o, &d);
      v3 L = it.ray_color(o, d, &s);
      cr = cr + L.x;
      cg = cg + L.y;
      cb = cb + L.z;
      ca = ca + 1.0f;
    }
    if (lq == 0u) {
      float ns = (float)P.samples;
      float4 b = make_float4(cr / ns, cg / ns, cb / ns, ca / ns);
      if (accumulated_of<FS>(S, P) == 0u) {
        accum[idx] = b;
      } else {
        float prev = (float)accumulated_of<FS>(S, P);
        float tot = prev + ns;
        float4 o = accum[idx];
        accum[idx] = make_float4((o.x * prev + b.x * ns) / tot, (o.y * prev + b.y * ns) / tot,
                                 (o.z * prev + b.z * ns) / tot, (o.w * prev + b.w * ns) / tot);
      }
    }
  }
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

__global__ void __launch_bounds__(64) k_primary_quad(DevScene S, TH_GpuRenderParams P,
                                                     TH_GpuCamera cam, uint32_t* __restrict__ slot_out,
                                                     float* __restrict__ t_out) {
  extern __shared__ uint2 lds_stack[];
  const uint32_t lane = threadIdx.x & 63u, qw = lane >> 2;
  const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint32_t tiles_x = (P.width + 3u) / 4u;
  const uint32_t x = (wave % tiles_x) * 4u + (qw & 3u), y = (wave / tiles_x) * 4u + (qw >> 2);
  if (x >= P.width || y >= P.height) return;
  Stack st{lds_stack + (threadIdx.x >> 2), (int)(blockDim.x >> 2), 0};
  const uint32_t idx = y * P.width + x;
  uint32_t s = th::pcg_hash(idx ^ (frame_of<false>(S, P) * 9781u));
  v3 o, d;
  camera_ray(cam, P, x, y, &s, &o, &d);
  float t;
  uint32_t slot;
  traverse_quad<1>(S, (const Child4*)S.nodes4, S.root4, o, d, TH_FLT_MAX, st, &t, &slot);
  if ((lane & 3u) == 0u) {
    slot_out[idx] = slot;
    t_out[idx] = t;
  }
}

// Display conversion on the device (color.rs:42-55, image_output.rs:59-92).
// Random streams of given pixels as the render kernels seed them: the reference stream
// pcg_hash(idx ^ frame*9781) (wgsl:519), or the per-sample stream of sample s,
// pcg_hash((idx ^ frame*9781) + s*0x9E3779B9); then `draws` draws of random_f32.
__global__ void k_rng_draws(const uint32_t* __restrict__ pixels, uint32_t count, uint32_t frame_index,
                            uint32_t per_sample, uint32_t sample, uint32_t draws, uint32_t* __restrict__ states,
                            float* __restrict__ values) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t idx = pixels[i];
  uint32_t st = per_sample ? th::pcg_hash((idx ^ (frame_index * 9781u)) + sample * 0x9E3779B9u)
                           : th::pcg_hash(idx ^ (frame_index * 9781u));
  for (uint32_t k = 0; k < draws; ++k) {
    const float v = th::rand01(&st);
    states[(size_t)i * draws + k] = st;
    values[(size_t)i * draws + k] = v;
  }
}

__global__ void k_display(const float4* __restrict__ accum, uint32_t n, float scale,
                          const float* __restrict__ thr, uchar4* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 p = accum[i];
  out[i] = make_uchar4(th_display_channel(p.x, scale, thr), th_display_channel(p.y, scale, thr),
                       th_display_channel(p.z, scale, thr), th_u8(p.w));
}

// ====================================================================== host side
namespace {


template <class T>
int32_t dalloc(T** p, size_t count) {
  *p = nullptr;
  size_t bytes = std::max<size_t>(count, 1) * sizeof(T);  // 1-element dummies (buffers.rs)
  TH_CHECK(hipMalloc((void**)p, bytes));
  TH_CHECK(hipMemset(*p, 0, bytes));
  return TH_OK;
}

template <class T>
int32_t upload(th_context* ctx, T* dst, const T* src, size_t count) {
  if (count == 0) return TH_OK;
  if (!src) return th_set_error(TH_E_INVALID, "null array with nonzero count");
  TH_CHECK(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  return TH_OK;
}

void free_scene_buffers(th_scene* s) {
  void* ptrs[] = {s->d_nodes, s->d_prims, s->d_spheres, s->d_tris, s->d_attrs, s->d_mats,
                  s->d_lights, s->d_env, s->d_tex, s->d_cnodes, s->d_slots, s->d_accum,
                  s->d_nodes4, s->d_refit_order, s->d_wide_src, s->d_accum_save, s->d_area,
                  s->d_rgba8, s->d_env_cdf, s->d_env_rows};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

// Tree depth (for the traversal stack size) from the reference layout.
uint32_t bvh_depth(const TH_GpuBvhNode* nodes, uint32_t count) {
  if (count == 0) return 0;
  std::vector<uint32_t> depth(count, 0);
  uint32_t maxd = 1;
  depth[0] = 1;
  for (uint32_t i = 0; i < count; ++i) {  // pre-order: parents precede children
    if (nodes[i].primitive_count == 0) {
      uint32_t l = i + 1, r = nodes[i].first_or_right;
      if (l < count) depth[l] = depth[i] + 1;
      if (r < count) depth[r] = depth[i] + 1;
    }
    maxd = std::max(maxd, depth[i]);
  }
  return maxd;
}

// Structural validation of a reference-layout BVH before any kernel reads it.
int32_t validate_bvh(const th_scene_desc* d) {
  uint32_t n = d->bvh_node_count, np = d->bvh_primitive_count;
  if (np != d->sphere_count + d->triangle_count && n > 0)
    return th_set_error(TH_E_INVALID, "bvh primitive count must equal sphere + triangle count");
  for (uint32_t i = 0; i < n; ++i) {
    const TH_GpuBvhNode& b = d->bvh_nodes[i];
    if (b.primitive_count > 0) {
      if ((uint64_t)b.first_or_right + b.primitive_count > np)
        return th_set_error(TH_E_INVALID, "bvh leaf range out of bounds");
      if (i >= (1u << 25) && (b.first_or_right >= (1u << 25) || b.primitive_count >= THD_BIG_LEAF))
        return th_set_error(TH_E_UNSUPPORTED, "leaf beyond the 2^25 traversal-entry range");
    } else {
      if (i + 1 >= n || b.first_or_right >= n || b.first_or_right <= i + 1)
        return th_set_error(TH_E_INVALID, "bvh interior child index out of bounds");
    }
  }
  for (uint32_t i = 0; i < np; ++i) {
    const TH_GpuPrimitiveRef& r = d->bvh_primitives[i];
    if (r.kind > 1u) return th_set_error(TH_E_INVALID, "bvh primitive kind must be 0 or 1");
    if ((r.kind == 0u && r.index >= d->sphere_count) || (r.kind == 1u && r.index >= d->triangle_count))
      return th_set_error(TH_E_INVALID, "bvh primitive index out of range");
  }
  return TH_OK;
}

// Texture and environment ranges are checked against the pixel counts of the arrays the
// kernels will read: the descriptor's at create, the uploaded ones at an update (an update
// never replaces texels or environment pixels).
int32_t validate_refs(const th_scene_desc* d, uint64_t tex_count, uint64_t env_count) {
  for (uint32_t i = 0; i < d->sphere_count; ++i)
    if (d->spheres[i].material_index >= d->material_count)
      return th_set_error(TH_E_INVALID, "sphere material index out of range");
  for (uint32_t i = 0; i < d->triangle_count; ++i)
    if (d->triangles[i].material_index >= d->material_count)
      return th_set_error(TH_E_INVALID, "triangle material index out of range");
  for (uint32_t i = 0; i < d->light_count; ++i)
    if (d->lights[i].material_index >= d->material_count || d->lights[i].kind > 1u)
      return th_set_error(TH_E_INVALID, "light record out of range");
  for (uint32_t i = 0; i < d->material_count; ++i) {
    const TH_GpuMaterial& m = d->materials[i];
    if (m.texture_width && m.texture_height &&
        (uint64_t)m.texture_offset + (uint64_t)m.texture_width * m.texture_height > tex_count)
      return th_set_error(TH_E_INVALID, "material texture out of range");
  }
  const TH_GpuRenderParams& p = d->params;
  if (p.background_kind > 2u) return th_set_error(TH_E_INVALID, "unknown background kind");
  if (p.background_kind == 2u &&
      (p.environment_width == 0 || p.environment_height == 0 ||
       (uint64_t)p.environment_width * p.environment_height != env_count))
    return th_set_error(TH_E_INVALID, "environment background requires an environment map");
  if (p.light_count != d->light_count) return th_set_error(TH_E_MISMATCH, "light count mismatch");
  return TH_OK;
}

int32_t validate_desc(const th_scene_desc* d, uint64_t tex_count, uint64_t env_count) {
  if (!d) return th_set_error(TH_E_INVALID, "null scene descriptor");
  if (d->sphere_count + (uint64_t)d->triangle_count == 0)
    return th_set_error(TH_E_INVALID, "GPU renderer requires at least one object");
  if (d->material_count == 0 || !d->materials)
    return th_set_error(TH_E_INVALID, "GPU renderer requires at least one material");
  if ((d->sphere_count && !d->spheres) || (d->triangle_count && (!d->triangles || !d->triangle_attributes)) ||
      (d->light_count && !d->lights) || (d->bvh_node_count && (!d->bvh_nodes || !d->bvh_primitives)) ||
      (d->texture_pixel_count && !d->texture_pixels) ||
      (d->environment_pixel_count && !d->environment_pixels))
    return th_set_error(TH_E_INVALID, "null array with nonzero count");
  if (d->params.width == 0 || d->params.height == 0)
    return th_set_error(TH_E_INVALID, "render dimensions must be nonzero");
  if (d->params.sphere_count != d->sphere_count || d->params.triangle_count != d->triangle_count ||
      d->params.material_count != d->material_count || d->params.bvh_node_count != d->bvh_node_count ||
      d->params.bvh_primitive_count != d->bvh_primitive_count)
    return th_set_error(TH_E_MISMATCH, "params counts disagree with the arrays");
  int32_t rc = validate_refs(d, tex_count, env_count);
  if (rc) return rc;
  return validate_bvh(d);
}

int32_t validate_desc(const th_scene_desc* d) {
  if (!d) return th_set_error(TH_E_INVALID, "null scene descriptor");
  return validate_desc(d, d->texture_pixel_count, d->environment_pixel_count);
}

// Collapses the reference binary tree into the 4-wide tree of th_wide.h (breadth-first
// node order, so the upper levels are contiguous): the children of a 4-wide node are
// the children of a binary interior node, with each interior child replaced by its own
// two children.  Returns the root entry and the number of 4-wide levels.
float host_bits_to_float(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// `box` (optional): the box of each reference node for the 4-wide records, 6 floats per
// node (min, max), NaN in box[6i] for a node with nothing left to test (left out).
void build_wide(const TH_GpuBvhNode* nodes, uint32_t n, std::vector<Child4>& out, uint32_t* root4,
                uint32_t* depth4, std::vector<uint32_t>* src, const std::vector<float>* box) {
  out.clear();
  src->clear();
  *depth4 = 0;
  if (n == 0) {
    *root4 = 0;
    return;
  }
  if (nodes[0].primitive_count > 0) {
    *root4 = th_entry_of(nodes[0], 0u);
    return;
  }
  std::vector<std::pair<uint32_t, uint32_t>> q;  // (binary node, level)
  q.push_back({0u, 1u});
  for (size_t qi = 0; qi < q.size(); ++qi) {
    uint32_t b = q[qi].first, lvl = q[qi].second;
    *depth4 = std::max(*depth4, lvl);
    uint32_t kids[4], nk = 0;
    uint32_t bc[2] = {b + 1u, nodes[b].first_or_right};
    for (uint32_t c : bc) {
      if (nodes[c].primitive_count == 0) {
        kids[nk++] = c + 1u;
        kids[nk++] = nodes[c].first_or_right;
      } else {
        kids[nk++] = c;
      }
    }
    for (uint32_t k = 0; k < 4; ++k) {
      Child4 ch;
      const float* bx = (k < nk && box) ? &(*box)[6 * (size_t)kids[k]] : nullptr;
      if (k < nk && !(bx && bx[0] != bx[0])) {
        const TH_GpuBvhNode& g = nodes[kids[k]];
        uint32_t entry;
        if (g.primitive_count == 0) {
          entry = (uint32_t)q.size();
          q.push_back({kids[k], lvl + 1u});
        } else {
          entry = th_entry_of(g, kids[k]);
        }
        ch.a = bx ? make_float4(bx[0], bx[1], bx[2], host_bits_to_float(entry))
                  : make_float4(g.min[0], g.min[1], g.min[2], host_bits_to_float(entry));
        ch.b = bx ? make_float4(bx[3], bx[4], bx[5], 0.0f) : make_float4(g.max[0], g.max[1], g.max[2], 0.0f);
        src->push_back(kids[k]);
      } else {
        ch.a = make_float4(0, 0, 0, host_bits_to_float(THD_NO_CHILD));
        ch.b = make_float4(0, 0, 0, 0);
        src->push_back(THD_NO_CHILD);
      }
      out.push_back(ch);
    }
  }
  *root4 = 0u;
}

// Compact copies of the environment CDF for the row-then-texel search: the CDF (.w of
// every texel) and its value at the end of every row.  Built only for a non-decreasing
// CDF of the map's own width x height (what the packer writes); any other map keeps the
// flat search over the texels, which gives the same index for a non-decreasing CDF.
int32_t prepare_env(th_scene* s, const th_scene_desc* d) {
  const uint32_t W = d->params.environment_width, H = d->params.environment_height;
  const uint32_t n = d->environment_pixel_count;
  if (n == 0 || W == 0 || H == 0 || (uint64_t)W * H != n) return TH_OK;
  std::vector<float> cdf(n), rows(H);
  float prev = -3.4028235e38f;
  for (uint32_t i = 0; i < n; ++i) {
    const float c = d->environment_pixels[4 * (size_t)i + 3];
    if (!(c >= prev)) return TH_OK;  // decreasing or NaN: flat search
    cdf[i] = prev = c;
  }
  for (uint32_t j = 0; j < H; ++j) rows[j] = cdf[(size_t)(j + 1) * W - 1];
  int32_t rc;
  if ((rc = dalloc(&s->d_env_cdf, n)) || (rc = dalloc(&s->d_env_rows, H))) return rc;
  TH_CHECK(hipMemcpy(s->d_env_cdf, cdf.data(), n * sizeof(float), hipMemcpyHostToDevice));
  TH_CHECK(hipMemcpy(s->d_env_rows, rows.data(), H * sizeof(float), hipMemcpyHostToDevice));
  s->env_rows_w = W;
  return TH_OK;
}

// Primitives tested before the 4-wide tree.  A primitive whose box is large against the
// scene sits deep in a median-split tree (the split follows centroids) and inflates the
// box of every ancestor: a roof or ground quad over a height field makes nearly every ray
// enter those boxes, all the way down.  Candidates are the (at most 8) largest primitives
// whose box surface area exceeds 1/64 of the root's; their removal is kept when the boxes
// of the tree without them shrink by at least 2 root areas per candidate in total (the
// expected node visits a ray saves, to first order, against one more primitive test per
// candidate).  The 4-wide records then carry, per reference node, the box of what remains
// below it; a node with nothing left is left out.  The candidates stay in their leaves
// (where a second test finds the same distance and rank, so closest hits and their ties
// are unchanged).  Before a candidate's hit counts, the traversal tests the box of its
// reference leaf (as the reference's slab test computes it, without a far bound): the
// reference reaches the primitive only through that box and its ancestors, whose slab
// intervals contain the leaf's, so a ray the leaf's rounding turns away misses the
// primitive in both.  Boxes are folded from each primitive's packer box (triangle vertex
// min/max -+ 1e-5, sphere center -+ radius).
static void select_extracted(th_scene* s, const th_scene_desc* d, std::vector<float>& box) {
  s->ext_n = 0;
  box.clear();
  const uint32_t nn = d->bvh_node_count, np = d->bvh_primitive_count;
  if (nn < 3 || np == 0) return;
  auto sa = [](const float* mn, const float* mx) {
    const double a = (double)mx[0] - mn[0], b = (double)mx[1] - mn[1], c = (double)mx[2] - mn[2];
    return 2.0 * (a * b + b * c + a * c);
  };
  std::vector<float> pb(6 * (size_t)np);
  std::vector<double> psa(np);
  for (uint32_t q = 0; q < np; ++q) {
    const TH_GpuPrimitiveRef& r = d->bvh_primitives[q];
    float* b = &pb[6 * (size_t)q];
    if (r.kind == 1u && r.index < d->triangle_count) {
      const TH_GpuTriangle& t = d->triangles[r.index];
      for (int a = 0; a < 3; ++a) {
        b[a] = std::min(std::min(t.v0[a], t.v1[a]), t.v2[a]) - 1e-5f;
        b[3 + a] = std::max(std::max(t.v0[a], t.v1[a]), t.v2[a]) + 1e-5f;
      }
    } else if (r.kind == 0u && r.index < d->sphere_count) {
      const TH_GpuSphere& sp = d->spheres[r.index];
      for (int a = 0; a < 3; ++a) {
        b[a] = sp.center_radius[a] - sp.center_radius[3];
        b[3 + a] = sp.center_radius[a] + sp.center_radius[3];
      }
    } else {
      return;  // unknown reference: keep the tree as it is
    }
    psa[q] = sa(b, b + 3);
  }
  const double root = sa(d->bvh_nodes[0].min, d->bvh_nodes[0].max);
  if (!(root > 0.0)) return;
  std::vector<uint32_t> cand;
  for (uint32_t q = 0; q < np; ++q)
    if (psa[q] > root / 64.0) cand.push_back(q);
  std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) { return psa[a] > psa[b]; });
  if (cand.size() > 8) cand.resize(8);
  if (cand.empty()) return;
  std::vector<uint8_t> ext(np, 0);
  for (uint32_t q : cand) ext[q] = 1;
  // boxes without the candidates, children before parents (pre-order: reverse index order)
  const float NaN = __builtin_nanf("");
  box.assign(6 * (size_t)nn, NaN);
  double gain = 0.0;
  for (uint32_t i = nn; i-- > 0;) {
    const TH_GpuBvhNode& n = d->bvh_nodes[i];
    float* b = &box[6 * (size_t)i];
    bool any = false;
    auto fold = [&](const float* c) {
      for (int a = 0; a < 3; ++a) {
        b[a] = any ? std::min(b[a], c[a]) : c[a];
        b[3 + a] = any ? std::max(b[3 + a], c[3 + a]) : c[3 + a];
      }
      any = true;
    };
    if (n.primitive_count > 0) {
      for (uint32_t k = 0; k < n.primitive_count; ++k) {
        const uint32_t q = n.first_or_right + k;
        if (q < np && !ext[q]) fold(&pb[6 * (size_t)q]);
      }
    } else {
      for (uint32_t c : {i + 1u, n.first_or_right}) {
        const float* cb = &box[6 * (size_t)c];
        if (cb[0] == cb[0]) fold(cb);
      }
    }
    gain += sa(n.min, n.max) - (any ? sa(b, b + 3) : 0.0);
  }
  if (gain / root < 2.0 * (double)cand.size()) {
    box.clear();
    return;
  }
  for (uint32_t i = 0; i < nn; ++i) {
    const TH_GpuBvhNode& n = d->bvh_nodes[i];
    if (n.primitive_count == 0) continue;
    for (uint32_t k = 0; k < n.primitive_count; ++k) {
      const uint32_t q = n.first_or_right + k;
      if (q < np && ext[q]) {
        for (uint32_t j = 0; j < cand.size(); ++j)
          if (cand[j] == q) {
            s->ext_slot[j] = q;
            s->ext_leaf[j] = n.first_or_right;
            s->ext_node[j] = i;
          }
      }
    }
  }
  s->ext_n = (uint32_t)cand.size();
}

int32_t prepare_bvh(th_scene* s, const th_scene_desc* d) {
  th_context* ctx = s->ctx;
  s->node_count = d->bvh_node_count;
  s->prim_count = d->bvh_primitive_count;
  s->ext_n = 0;  // chosen again below for a new tree
  if (s->node_count) {
    const TH_GpuBvhNode& r = d->bvh_nodes[0];
    s->root_entry = th_entry_of(r, 0u);
    for (int k = 0; k < 3; ++k) {
      s->root_min[k] = r.min[k];
      s->root_max[k] = r.max[k];
    }
    s->max_depth = bvh_depth(d->bvh_nodes, d->bvh_node_count);
    std::vector<Child4> w;
    std::vector<uint32_t> wsrc;
    std::vector<float> rbox;
    select_extracted(s, d, rbox);
    build_wide(d->bvh_nodes, d->bvh_node_count, w, &s->root4, &s->depth4, &wsrc, rbox.empty() ? nullptr : &rbox);
    s->node4_count = (uint32_t)(w.size() / 4);
    if (!w.empty()) {
      TH_CHECK(hipMemcpy(s->d_nodes4, w.data(), w.size() * sizeof(Child4), hipMemcpyHostToDevice));
      TH_CHECK(hipMemcpy(s->d_wide_src, wsrc.data(), wsrc.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    // refit order: node indices grouped by level, deepest level first (pre-order: a
    // parent precedes its children, so one pass assigns every level)
    {
      const uint32_t nn = d->bvh_node_count;
      std::vector<uint32_t> level(nn, 0u);
      for (uint32_t i = 0; i < nn; ++i)
        if (d->bvh_nodes[i].primitive_count == 0u) {
          level[i + 1u] = level[i] + 1u;
          level[d->bvh_nodes[i].first_or_right] = level[i] + 1u;
        }
      const uint32_t levels = s->max_depth;
      std::vector<uint32_t> cnt(levels + 1u, 0u), order(nn);
      for (uint32_t i = 0; i < nn; ++i) cnt[levels - 1u - level[i]]++;
      s->refit_level_end.assign(levels, 0u);
      uint32_t acc = 0;
      for (uint32_t k = 0; k < levels; ++k) {
        uint32_t c = cnt[k];
        cnt[k] = acc;
        acc += c;
        s->refit_level_end[k] = acc;
      }
      for (uint32_t i = 0; i < nn; ++i) order[cnt[levels - 1u - level[i]]++] = i;
      TH_CHECK(hipMemcpy(s->d_refit_order, order.data(), nn * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(k_prepare_cnodes, dim3((s->node_count + 255) / 256), dim3(256), 0, ctx->stream,
                       s->d_nodes, s->node_count, s->d_cnodes);
    TH_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_area_sum, dim3(1), dim3(256), 0, ctx->stream, s->d_nodes, s->node_count, s->d_area);
    TH_CHECK(hipGetLastError());
    TH_CHECK(hipMemcpyAsync(&s->area_build, s->d_area, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    TH_CHECK(hipStreamSynchronize(ctx->stream));
    s->area_now = s->area_build;
  }
  if (s->prim_count) {
    hipLaunchKernelGGL(k_prepare_slots, dim3((s->prim_count + 255) / 256), dim3(256), 0, ctx->stream,
                       s->d_prims, s->prim_count, s->d_spheres, s->d_tris, s->d_attrs, s->d_slots);
    TH_CHECK(hipGetLastError());
  }
  return TH_OK;
}

// Device refit of the uploaded tree from the scene's current spheres and triangles:
// slots, reference nodes level by level, child-pair records, 4-wide records, and the
// root box the launches take by value (read back: one 24-byte copy).
int32_t refit_bvh(th_scene* s) {
  th_context* ctx = s->ctx;
  hipLaunchKernelGGL(k_prepare_slots, dim3((s->prim_count + 255) / 256), dim3(256), 0, ctx->stream,
                     s->d_prims, s->prim_count, s->d_spheres, s->d_tris, s->d_attrs, s->d_slots);
  TH_CHECK(hipGetLastError());
  uint32_t begin = 0;
  for (uint32_t end : s->refit_level_end) {
    if (end > begin)
      hipLaunchKernelGGL(k_refit_level, dim3((end - begin + 255) / 256), dim3(256), 0, ctx->stream,
                         s->d_nodes, s->d_refit_order + begin, end - begin, s->d_prims, s->d_spheres,
                         s->d_tris);
    begin = end;
  }
  TH_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_prepare_cnodes, dim3((s->node_count + 255) / 256), dim3(256), 0, ctx->stream,
                     s->d_nodes, s->node_count, s->d_cnodes);
  TH_CHECK(hipGetLastError());
  if (s->node4_count) {
    const uint32_t nc = 4u * s->node4_count;
    hipLaunchKernelGGL(k_refit_wide, dim3((nc + 255) / 256), dim3(256), 0, ctx->stream, s->d_nodes4,
                       s->d_wide_src, nc, s->d_nodes);
    TH_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_area_sum, dim3(1), dim3(256), 0, ctx->stream, s->d_nodes, s->node_count, s->d_area);
  TH_CHECK(hipGetLastError());
  TH_GpuBvhNode root;
  TH_CHECK(hipMemcpyAsync(&root, s->d_nodes, sizeof root, hipMemcpyDeviceToHost, ctx->stream));
  TH_CHECK(hipMemcpyAsync(&s->area_now, s->d_area, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  TH_CHECK(hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < 3; ++k) {
    s->root_min[k] = root.min[k];
    s->root_max[k] = root.max[k];
  }
  return TH_OK;
}

int32_t check_render_params_impl(const th_scene* s, const TH_GpuRenderParams* p) {
  if (!p) return th_set_error(TH_E_INVALID, "null params");
  if (p->width != s->width || p->height != s->height)
    return th_set_error(TH_E_MISMATCH, "render dimensions differ from the scene's output buffer");
  if (p->samples == 0) return th_set_error(TH_E_INVALID, "samples must be nonzero");
  if (p->sphere_count != s->params.sphere_count || p->triangle_count != s->params.triangle_count ||
      p->material_count != s->params.material_count || p->bvh_node_count != s->node_count ||
      p->bvh_primitive_count != s->prim_count || p->light_count != s->params.light_count)
    return th_set_error(TH_E_MISMATCH, "params counts differ from the uploaded scene");
  if (p->background_kind > 2u) return th_set_error(TH_E_INVALID, "unknown background kind");
  if (p->background_kind == 2u &&
      ((uint64_t)p->environment_width * p->environment_height != s->env_count || s->env_count == 0))
    return th_set_error(TH_E_MISMATCH, "environment dimensions differ from the uploaded map");
  return TH_OK;
}

}  // namespace

DevScene dev_scene(const th_scene* s) {
  DevScene S;
  S.cnodes = s->d_cnodes;
  S.slots = s->d_slots;
  S.nodes = s->d_nodes;
  S.spheres = s->d_spheres;
  S.triangles = s->d_tris;
  S.attrs = s->d_attrs;
  S.materials = s->d_mats;
  S.lights = s->d_lights;
  S.env = s->d_env;
  S.tex = s->d_tex;
  S.node_count = s->node_count;
  S.root_entry = s->root_entry;
  S.nodes4 = (const float4*)s->d_nodes4;
  S.root4 = s->root4;
  S.node4_bytes = (uint32_t)((size_t)s->node4_count * 4u * sizeof(Child4));
  S.env_cdf = s->d_env_cdf;
  S.env_rows = s->d_env_rows;
  S.env_rows_w = s->env_rows_w;
  S.frame_state = nullptr;
  for (int k = 0; k < 3; ++k) {
    S.root_min[k] = s->root_min[k];
    S.root_max[k] = s->root_max[k];
  }
  S.ext_n = s->ext_n;
  for (int k = 0; k < 8; ++k) {
    S.ext_slot[k] = s->ext_slot[k];
    S.ext_leaf[k] = s->ext_leaf[k];
    S.ext_node[k] = s->ext_node[k];
  }
  return S;
}

// Folds the `n` oldest pending launches into the statistics, in launch order, and
// returns their event pairs to the pool.  On an event error the pairs already folded
// leave the pending list, so no pair is ever both pooled and pending.
static int32_t fold_oldest(th_context* ctx, size_t n) {
  auto& pend = ctx->ev_pending;
  n = std::min(n, pend.size());
  size_t done = 0;
  int32_t rc = TH_OK;
  for (; done < n; ++done) {
    const auto pr = pend[done];
    hipError_t e = hipEventSynchronize(pr.second);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, pr.first, pr.second);
    if (e != hipSuccess) {
      rc = th_set_error(TH_E_DEVICE, std::string("launch timing: ") + hipGetErrorString(e));
      break;
    }
    ctx->stats.last_kernel_ms = ms;
    ctx->stats.total_kernel_ms += ms;
    ctx->stats.launches += 1;
    ctx->ev_pool.push_back(pr);
  }
  pend.erase(pend.begin(), pend.begin() + (std::ptrdiff_t)done);
  return rc;
}

// Per-launch event pair from a pool (every launch is timed; th_get_stats sums them).
// A caller that never reads the statistics still keeps the pending list bounded: at
// 4,096 pending launches the oldest 2,048 (long finished) are folded.
int32_t begin_timed(th_context* ctx, std::pair<hipEvent_t, hipEvent_t>* pr) {
  if (ctx->ev_pending.size() >= 4096) {
    int32_t rc = fold_oldest(ctx, 2048);
    if (rc) return rc;
  }
  if (ctx->ev_pool.empty()) {
    hipEvent_t a, b;
    TH_CHECK(hipEventCreate(&a));
    TH_CHECK(hipEventCreate(&b));
    ctx->ev_pool.push_back({a, b});
  }
  *pr = ctx->ev_pool.back();
  ctx->ev_pool.pop_back();
  TH_CHECK(hipEventRecord(pr->first, ctx->stream));
  return TH_OK;
}

int32_t end_timed(th_context* ctx, std::pair<hipEvent_t, hipEvent_t> pr) {
  hipError_t e = hipEventRecord(pr.second, ctx->stream);
  if (e != hipSuccess) {
    ctx->ev_pool.push_back(pr);
    return th_set_error(TH_E_DEVICE, std::string("hipEventRecord: ") + hipGetErrorString(e));
  }
  ctx->ev_pending.push_back(pr);
  return TH_OK;
}

int32_t fold_timing(th_context* ctx) { return fold_oldest(ctx, ctx->ev_pending.size()); }


int32_t th_check_render_params(const th_scene* s, const TH_GpuRenderParams* p) {
  return check_render_params_impl(s, p);
}
static int32_t check_render_params(const th_scene* s, const TH_GpuRenderParams* p) {
  return check_render_params_impl(s, p);
}

// ====================================================================== C ABI
extern "C" {

int32_t th_device_count(int32_t* count) {
  if (!count) return th_set_error(TH_E_INVALID, "null argument");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return TH_OK;
}

int32_t th_device_info(int32_t device, char* name, size_t name_cap, char* driver, size_t driver_cap) {
  if (!name || !driver || name_cap == 0 || driver_cap == 0) return th_set_error(TH_E_INVALID, "null argument");
  hipDeviceProp_t prop;
  TH_CHECK(hipGetDeviceProperties(&prop, device));
  int rt = 0, drv = 0;
  TH_CHECK(hipRuntimeGetVersion(&rt));
  TH_CHECK(hipDriverGetVersion(&drv));
  snprintf(name, name_cap, "%s (%s)", prop.name, prop.gcnArchName);
  snprintf(driver, driver_cap, "HIP runtime %d, driver %d", rt, drv);
  return TH_OK;
}

int32_t th_context_create(int32_t device, th_context** out) {
  if (!out) return th_set_error(TH_E_INVALID, "null argument");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return th_set_error(TH_E_NO_DEVICE, "no HIP device visible");
  if (device < 0 || device >= n) return th_set_error(TH_E_INVALID, "device index out of range");
  th_context* c = new th_context();
  c->device = device;
  memset(&c->stats, 0, sizeof c->stats);
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e == hipSuccess) e = hipMalloc((void**)&c->d_counters, 16 * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMalloc((void**)&c->d_work, 16 * sizeof(uint32_t));
  if (e == hipSuccess) e = hipDeviceGetAttribute(&c->cu_count, hipDeviceAttributeMultiprocessorCount, device);
  if (e == hipSuccess) e = hipMalloc((void**)&c->d_srgb_lut, 512 * sizeof(float));
  if (e == hipSuccess) {
    float lut[512];
    th_srgb_table(lut);
    th_srgb_thresholds(lut + 256);
    e = hipMemcpy(c->d_srgb_lut, lut, sizeof lut, hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    th_context_destroy(c);
    return th_set_error(TH_E_DEVICE, std::string("context setup: ") + hipGetErrorString(e));
  }
  int32_t rc = th_wavefront_scratch_init(c);  // with `device` current
  if (rc) {
    th_context_destroy(c);
    return rc;
  }
  *out = c;
  return TH_OK;
}

void th_context_destroy(th_context* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->d_counters) (void)hipFree(c->d_counters);
  if (c->d_work) (void)hipFree(c->d_work);
  if (c->d_srgb_lut) (void)hipFree(c->d_srgb_lut);
  th_batch_free(c->batch);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  for (auto* v : {&c->ev_pool, &c->ev_pending})
    for (auto& pr : *v) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int32_t th_scene_create(th_context* ctx, const th_scene_desc* d, uint32_t node_capacity,
                        th_scene** out) {
  if (!ctx || !out) return th_set_error(TH_E_INVALID, "null argument");
  *out = nullptr;
  int32_t rc = validate_desc(d);
  if (rc) return rc;
  uint32_t nprim = d->sphere_count + d->triangle_count;
  uint32_t cap = node_capacity ? node_capacity : (nprim ? 2 * nprim - 1 : 1);  // buffers.rs:205-211
  if (d->bvh_node_count > cap) return th_set_error(TH_E_CAPACITY, "bvh node count exceeds capacity");
  TH_CHECK(hipSetDevice(ctx->device));
  th_scene* s = new th_scene();
  s->ctx = ctx;
  s->params = d->params;
  s->node_capacity = cap;
  s->width = d->params.width;
  s->height = d->params.height;
  s->env_count = d->environment_pixel_count;
  s->tex_count = d->texture_pixel_count;
  size_t npix = (size_t)s->width * s->height;
  if ((rc = dalloc(&s->d_nodes, cap)) || (rc = dalloc(&s->d_prims, nprim)) ||
      (rc = dalloc(&s->d_spheres, d->sphere_count)) || (rc = dalloc(&s->d_tris, d->triangle_count)) ||
      (rc = dalloc(&s->d_attrs, d->triangle_count)) || (rc = dalloc(&s->d_mats, d->material_count)) ||
      (rc = dalloc(&s->d_lights, d->light_count)) || (rc = dalloc(&s->d_env, d->environment_pixel_count)) ||
      (rc = dalloc(&s->d_tex, d->texture_pixel_count)) || (rc = dalloc(&s->d_cnodes, cap)) ||
      (rc = dalloc(&s->d_slots, 3 * (size_t)nprim)) || (rc = dalloc(&s->d_accum, npix)) ||
      (rc = dalloc(&s->d_nodes4, 4 * (size_t)cap)) || (rc = dalloc(&s->d_refit_order, cap)) ||
      (rc = dalloc(&s->d_area, 1)) ||
      (rc = dalloc(&s->d_wide_src, 4 * (size_t)cap))) {
    free_scene_buffers(s);
    delete s;
    return rc;
  }
  if ((rc = upload(ctx, s->d_nodes, d->bvh_nodes, d->bvh_node_count)) ||
      (rc = upload(ctx, s->d_prims, d->bvh_primitives, d->bvh_primitive_count)) ||
      (rc = upload(ctx, s->d_spheres, d->spheres, d->sphere_count)) ||
      (rc = upload(ctx, s->d_tris, d->triangles, d->triangle_count)) ||
      (rc = upload(ctx, s->d_attrs, d->triangle_attributes, d->triangle_count)) ||
      (rc = upload(ctx, s->d_mats, d->materials, d->material_count)) ||
      (rc = upload(ctx, s->d_lights, d->lights, d->light_count)) ||
      (rc = upload(ctx, (float*)s->d_env, d->environment_pixels, 4 * (size_t)d->environment_pixel_count)) ||
      (rc = upload(ctx, s->d_tex, d->texture_pixels, d->texture_pixel_count)) ||
      (rc = prepare_env(s, d)) || (rc = prepare_bvh(s, d))) {
    (void)hipStreamSynchronize(ctx->stream);
    free_scene_buffers(s);
    delete s;
    return rc;
  }
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    free_scene_buffers(s);
    delete s;
    return th_set_error(TH_E_DEVICE, std::string("scene upload: ") + hipGetErrorString(e));
  }
  *out = s;
  return TH_OK;
}

int32_t th_scene_update(th_scene* s, uint32_t mask, const th_scene_desc* d) {
  if (!s || !d) return th_set_error(TH_E_INVALID, "null argument");
  th_context* ctx = s->ctx;
  const uint32_t known = TH_CHANGE_CAMERA | TH_CHANGE_SPHERES | TH_CHANGE_TRIANGLES | TH_CHANGE_LIGHTS |
                         TH_CHANGE_BVH | TH_CHANGE_MATERIALS | TH_CHANGE_REFIT;
  if (mask & ~known) return th_set_error(TH_E_INVALID, "unknown change bit");
  const bool refit = (mask & TH_CHANGE_REFIT) != 0u;
  if (refit && (mask & TH_CHANGE_BVH))
    return th_set_error(TH_E_INVALID, "a refit and a new bvh are exclusive");
  if (refit && !(mask & (TH_CHANGE_SPHERES | TH_CHANGE_TRIANGLES)))
    return th_set_error(TH_E_INVALID, "a refit needs new spheres or triangles");
  if (refit && s->node_count == 0)
    return th_set_error(TH_E_INVALID, "the scene has no bvh to refit");
  // path_tracer.rs:706-719: moved spheres or triangles upload the frame's rebuilt tree
  // with them whenever the frame has one; without a tree in the descriptor the uploaded
  // one would no longer match the primitives, so only a refit may keep it.
  if ((mask & (TH_CHANGE_SPHERES | TH_CHANGE_TRIANGLES)) && !refit && !(mask & TH_CHANGE_BVH)) {
    if (d->bvh_node_count > 0)
      mask |= TH_CHANGE_BVH;
    else if (s->node_count > 0)
      return th_set_error(TH_E_INVALID,
                          "moved geometry needs its rebuilt bvh (TH_CHANGE_BVH) or a refit (TH_CHANGE_REFIT)");
  }
  // path_tracer.rs:639-653: counts are fixed after setup; nodes within capacity.
  if (d->sphere_count != s->params.sphere_count || d->triangle_count != s->params.triangle_count ||
      d->material_count != s->params.material_count || d->light_count != s->params.light_count ||
      (!refit && d->bvh_primitive_count != s->prim_count) || d->params.width != s->width ||
      d->params.height != s->height)
    return th_set_error(TH_E_MISMATCH, "scene counts changed after setup");
  if (d->bvh_node_count > s->node_capacity)
    return th_set_error(TH_E_CAPACITY, "bvh node count exceeds capacity");
  // Texels and environment pixels stay as uploaded at create: a descriptor that carries
  // them must carry the same counts (a per-frame pack without them carries 0), and
  // material texture ranges are checked against the uploaded texels.
  if ((d->texture_pixel_count && d->texture_pixel_count != s->tex_count) ||
      (d->environment_pixel_count && d->environment_pixel_count != s->env_count))
    return th_set_error(TH_E_MISMATCH, "texture or environment pixel count differs from the uploaded scene");
  int32_t rc = validate_desc(d, s->tex_count, s->env_count);
  if (rc) return rc;
  TH_CHECK(hipSetDevice(ctx->device));
  bool geometry = false;
  if (mask & TH_CHANGE_SPHERES) {
    if ((rc = upload(ctx, s->d_spheres, d->spheres, d->sphere_count))) return rc;
    geometry = true;
  }
  if (mask & TH_CHANGE_TRIANGLES) {
    if ((rc = upload(ctx, s->d_tris, d->triangles, d->triangle_count)) ||
        (rc = upload(ctx, s->d_attrs, d->triangle_attributes, d->triangle_count)))
      return rc;
    geometry = true;
  }
  if (mask & TH_CHANGE_MATERIALS)
    if ((rc = upload(ctx, s->d_mats, d->materials, d->material_count))) return rc;
  if (mask & TH_CHANGE_LIGHTS)
    if ((rc = upload(ctx, s->d_lights, d->lights, d->light_count))) return rc;
  if (mask & TH_CHANGE_BVH) {
    if ((rc = upload(ctx, s->d_nodes, d->bvh_nodes, d->bvh_node_count)) ||
        (rc = upload(ctx, s->d_prims, d->bvh_primitives, d->bvh_primitive_count)))
      return rc;
    geometry = true;
  }
  s->params = d->params;
  s->generation++;
  if (geometry || refit) s->auto_valid = false;  // the fastest schedule may change with the tree
  if (refit) {
    s->params.bvh_node_count = s->node_count;
    s->params.bvh_primitive_count = s->prim_count;
    if ((rc = refit_bvh(s))) return rc;
  } else if (geometry) {
    if ((rc = prepare_bvh(s, d))) return rc;
  }
  TH_CHECK(hipStreamSynchronize(ctx->stream));
  return TH_OK;
}

int32_t th_scene_refit_quality(const th_scene* s, float* ratio) {
  if (!s || !ratio) return th_set_error(TH_E_INVALID, "null argument");
  *ratio = s->area_build > 0.0f ? s->area_now / s->area_build : 1.0f;
  return TH_OK;
}

int32_t th_scene_traversal_info(const th_scene* s, uint32_t* extracted, uint32_t* nodes4) {
  if (!s) return th_set_error(TH_E_INVALID, "null argument");
  if (extracted) *extracted = s->ext_n;
  if (nodes4) *nodes4 = s->node4_count;
  return TH_OK;
}

void th_scene_destroy(th_scene* s) {
  if (!s) return;
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  free_scene_buffers(s);
  th_wavefront_free(s->wf);
  th_guided_free(s->gw);
  delete s;
}

// The kernel launches of one render on ctx->stream (th_render's dispatch by kernel,
// traversal, schedule and random-stream mode); also what a progressive graph captures,
// with S.frame_state set.
extern "C++" {
template <bool FS>
static int32_t launch_render_t(th_context* ctx, th_scene* s, const TH_GpuRenderParams* p, const TH_GpuCamera* cam,
                               const th_render_options& o, const DevScene& S, size_t lds, uint32_t block,
                               unsigned long long* cnt) {
  if (o.kernel == TH_KERNEL_WAVEFRONT && p->max_bounces > 0) {
    int32_t rc = th_wavefront_render(ctx, s, p, cam, &o);
    if (rc) return rc;
  } else if (o.traversal == TH_TRAVERSE_WIDE && s->node_count > 0) {
    const uint32_t qblock = 256;
    uint32_t depth = 3u * s->depth4 + 2u;
    uint32_t cached = std::min<uint32_t>(s->node4_count, TH_QUAD_LDS_NODES);
    size_t lds4 = (size_t)depth * (qblock / 4) * sizeof(uint2) + (size_t)cached * 4 * sizeof(Child4);
    int G = o.rng_mode == TH_RNG_REFERENCE ? 1
            : (p->samples >= 16 ? 16 : p->samples >= 8 ? 8 : p->samples >= 4 ? 4 : p->samples >= 2 ? 2 : 1);
    uint32_t ppw = 16u / (uint32_t)G;
    uint32_t tw = ppw >= 8 ? 4 : (ppw >= 2 ? 2 : 1), tht = ppw / tw;
    uint32_t waves = ((p->width + tw - 1) / tw) * ((p->height + tht - 1) / tht);
    dim3 grid((waves + 3) / 4), blk(qblock);
#define TH_LAUNCHQ(PS, GG) \
  hipLaunchKernelGGL((k_render_quad<PS, GG, FS>), grid, blk, lds4, ctx->stream, S, *p, *cam, s->d_accum, ctx->d_srgb_lut, cached)
    if (o.rng_mode == TH_RNG_REFERENCE && o.schedule == TH_SCHEDULE_DYNAMIC) {
      if (ctx->dyn_blocks_per_cu == 0) {
        int nb = 0;
        TH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_render_quad_dyn<FS>, (int)qblock, lds4));
        ctx->dyn_blocks_per_cu = std::max(nb, 1);
      }
      // persistent: as many blocks as stay resident, but no more quads than pixels
      uint32_t pblocks = (uint32_t)(std::max(ctx->cu_count, 1) * ctx->dyn_blocks_per_cu);
      pblocks = std::min<uint32_t>(pblocks, (p->width * p->height + 63u) / 64u);
      hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->d_work, 1u);
      hipLaunchKernelGGL(k_render_quad_dyn<FS>, dim3(std::max(pblocks, 1u)), blk, lds4, ctx->stream, S, *p, *cam,
                         s->d_accum, ctx->d_srgb_lut, cached, ctx->d_work);
    } else if (o.rng_mode == TH_RNG_REFERENCE && o.schedule == TH_SCHEDULE_SPECULATIVE) {
      // two octets per pixel (one with a single sample); the draws per diffuse vertex:
      // 2 (BSDF), 3 (light pick and point) with lights, 3 (environment pick and jitter)
      // with an environment
      const int GS = p->samples >= 2 ? 2 : 1;
      const bool env = p->background_kind == 2u;
      const uint32_t per_vertex = 2u + ((p->light_count > 0u && p->total_light_weight > 0.0f) ? 3u : 0u) +
                                  (env && p->environment_width * p->environment_height > 0u ? 3u : 0u);
      uint32_t sppw = 8u / (uint32_t)GS;
      uint32_t stw = sppw >= 4 ? 4 : (sppw >= 2 ? 2 : 1), sth = sppw / stw;
      uint32_t swaves = ((p->width + stw - 1) / stw) * ((p->height + sth - 1) / sth);
      dim3 sgrid((swaves + 3u) / 4u);
#define TH_LAUNCHS(E, GG)                                                                               \
  hipLaunchKernelGGL((k_render_spec<E, GG, FS>), sgrid, blk, lds4, ctx->stream, S, *p, *cam, s->d_accum, \
                     ctx->d_srgb_lut, cached, per_vertex)
      switch (GS * (env ? -1 : 1)) {
        case 2: TH_LAUNCHS(false, 2); break;
        case 1: TH_LAUNCHS(false, 1); break;
        case -2: TH_LAUNCHS(true, 2); break;
        default: TH_LAUNCHS(true, 1); break;
      }
#undef TH_LAUNCHS
    } else if (o.schedule == TH_SCHEDULE_OCTET || o.schedule == TH_SCHEDULE_SPECULATIVE) {
      // octets per pixel: 1 with the reference RNG, else min(samples, 8) as a power of two
      int GO = o.rng_mode == TH_RNG_REFERENCE ? 1 : std::min(G, 8);
      uint32_t oppw = 8u / (uint32_t)GO;
      uint32_t otw = oppw >= 4 ? 4 : (oppw >= 2 ? 2 : 1), oth = oppw / otw;
      uint32_t owaves = ((p->width + otw - 1) / otw) * ((p->height + oth - 1) / oth);
      dim3 ogrid((owaves + 3u) / 4u);
      bool env = p->background_kind == 2u;
#define TH_LAUNCHO(E, PS, GG)                                                                       \
  hipLaunchKernelGGL((k_render_octet<E, PS, GG, FS>), ogrid, blk, lds4, ctx->stream, S, *p, *cam, s->d_accum, \
                     ctx->d_srgb_lut, cached)
      if (o.rng_mode == TH_RNG_REFERENCE) {
        if (env) TH_LAUNCHO(true, false, 1);
        else TH_LAUNCHO(false, false, 1);
      } else {
        switch (GO * (env ? -1 : 1)) {
          case 8: TH_LAUNCHO(false, true, 8); break;
          case 4: TH_LAUNCHO(false, true, 4); break;
          case 2: TH_LAUNCHO(false, true, 2); break;
          case 1: TH_LAUNCHO(false, true, 1); break;
          case -8: TH_LAUNCHO(true, true, 8); break;
          case -4: TH_LAUNCHO(true, true, 4); break;
          case -2: TH_LAUNCHO(true, true, 2); break;
          default: TH_LAUNCHO(true, true, 1); break;
        }
      }
#undef TH_LAUNCHO
    } else if (o.rng_mode == TH_RNG_REFERENCE) {
      TH_LAUNCHQ(false, 1);
    } else {
      switch (G) {
        case 16: TH_LAUNCHQ(true, 16); break;
        case 8: TH_LAUNCHQ(true, 8); break;
        case 4: TH_LAUNCHQ(true, 4); break;
        case 2: TH_LAUNCHQ(true, 2); break;
        default: TH_LAUNCHQ(true, 1); break;
      }
    }
#undef TH_LAUNCHQ
  } else if (o.rng_mode == TH_RNG_REFERENCE) {
    uint32_t tiles = ((p->width + 7) / 8) * ((p->height + 7) / 8);
    dim3 grid(tiles), blk(block);
#define TH_LAUNCH(TM, C) \
  hipLaunchKernelGGL((k_render_pixel<TM, C, FS>), grid, blk, lds, ctx->stream, S, *p, *cam, s->d_accum, ctx->d_srgb_lut, cnt)
    if (o.traversal == 0 && !o.collect_stats) TH_LAUNCH(0, false);
    else if (o.traversal == 0) TH_LAUNCH(0, true);
    else if (!o.collect_stats) TH_LAUNCH(1, false);
    else TH_LAUNCH(1, true);
#undef TH_LAUNCH
  } else {
    // G lanes per pixel, G = min(samples, 8) rounded down to a power of two
    int G = p->samples >= 8 ? 8 : (p->samples >= 4 ? 4 : (p->samples >= 2 ? 2 : 1));
    uint32_t ppw = 64u / (uint32_t)G;
    uint32_t th_ = 8u / (uint32_t)std::min(G, 8), tw = ppw / th_;
    uint32_t waves = ((p->width + tw - 1) / tw) * ((p->height + th_ - 1) / th_);
    dim3 grid(waves), blk(block);
#define TH_LAUNCH(TM, C, GG) \
  hipLaunchKernelGGL((k_render_sample<TM, C, GG, FS>), grid, blk, lds, ctx->stream, S, *p, *cam, s->d_accum, ctx->d_srgb_lut, cnt)
#define TH_LAUNCH_G(TM, C)            \
  switch (G) {                        \
    case 8: TH_LAUNCH(TM, C, 8); break; \
    case 4: TH_LAUNCH(TM, C, 4); break; \
    case 2: TH_LAUNCH(TM, C, 2); break; \
    default: TH_LAUNCH(TM, C, 1); break; \
  }
    if (o.traversal == 0 && !o.collect_stats) { TH_LAUNCH_G(0, false) }
    else if (o.traversal == 0) { TH_LAUNCH_G(0, true) }
    else if (!o.collect_stats) { TH_LAUNCH_G(1, false) }
    else { TH_LAUNCH_G(1, true) }
#undef TH_LAUNCH_G
#undef TH_LAUNCH
  }
  return TH_OK;
}

}  // extern "C++"

static int32_t launch_render(th_context* ctx, th_scene* s, const TH_GpuRenderParams* p, const TH_GpuCamera* cam,
                             const th_render_options& o, const DevScene& S, size_t lds, uint32_t block,
                             unsigned long long* cnt) {
  return S.frame_state != nullptr ? launch_render_t<true>(ctx, s, p, cam, o, S, lds, block, cnt)
                                  : launch_render_t<false>(ctx, s, p, cam, o, S, lds, block, cnt);
}

// TH_SCHEDULE_AUTO (reference RNG, 4-wide megakernel): the first render of a scene with
// these samples / bounces / background times the static, octet and speculative schedules
// (same image) on this frame -- one warm-up and two timed launches each, at seed 0 of the
// accumulation -- and keeps the fastest until the scene's geometry changes.  The trials
// overwrite the accumulation, so a render that folds into it gets it back first; their
// launches are left out of th_get_stats.
static int32_t auto_schedule(th_context* ctx, th_scene* s, const TH_GpuRenderParams* p, const TH_GpuCamera* cam,
                             th_render_options* o) {
  if (s->auto_valid && s->auto_samples == p->samples && s->auto_bounces == p->max_bounces &&
      s->auto_background == p->background_kind) {
    o->schedule = s->auto_choice;
    return TH_OK;
  }
  int32_t rc = fold_timing(ctx);  // the caller's earlier renders keep their timings
  if (rc) return rc;
  const size_t bytes = (size_t)s->width * s->height * sizeof(float4);
  const bool keep = p->accumulated_samples > 0u;
  if (keep) {
    if (!s->d_accum_save) TH_CHECK(hipMalloc((void**)&s->d_accum_save, bytes));
    TH_CHECK(hipMemcpyAsync(s->d_accum_save, s->d_accum, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  }
  TH_GpuRenderParams tp = *p;
  tp.accumulated_samples = 0u;
  const uint32_t cands[3] = {TH_SCHEDULE_STATIC, TH_SCHEDULE_OCTET, TH_SCHEDULE_SPECULATIVE};
  float best = 0.0f;
  uint32_t choice = TH_SCHEDULE_STATIC;
  for (int c = 0; c < 3; ++c) {
    th_render_options t = *o;
    t.schedule = cands[c];
    // a failed trial: the caller's accumulation comes back, the trials leave no timings
    auto fail = [&](int32_t code) {
      for (auto& pr : ctx->ev_pending) ctx->ev_pool.push_back(pr);
      ctx->ev_pending.clear();
      if (keep) (void)hipMemcpyAsync(s->d_accum, s->d_accum_save, bytes, hipMemcpyDeviceToDevice, ctx->stream);
      return code;
    };
    for (int r = 0; r < 3; ++r)
      if ((rc = th_render(ctx, s, &tp, cam, &t))) return fail(rc);
    float total = 0.0f;
    for (size_t i = 0; i < ctx->ev_pending.size(); ++i) {
      const auto& pr = ctx->ev_pending[i];
      float ms = 0.0f;
      hipError_t e = hipEventSynchronize(pr.second);
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, pr.first, pr.second);
      if (e != hipSuccess)
        return fail(th_set_error(TH_E_DEVICE, std::string("schedule trial timing: ") + hipGetErrorString(e)));
      if (i > 0) total += ms;  // the first launch warms up
    }
    for (auto& pr : ctx->ev_pending) ctx->ev_pool.push_back(pr);
    ctx->ev_pending.clear();
    if (c == 0 || total < best) {
      best = total;
      choice = cands[c];
    }
  }
  if (keep) TH_CHECK(hipMemcpyAsync(s->d_accum, s->d_accum_save, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  s->auto_valid = true;
  s->auto_samples = p->samples;
  s->auto_bounces = p->max_bounces;
  s->auto_background = p->background_kind;
  s->auto_choice = choice;
  o->schedule = choice;
  return TH_OK;
}

int32_t th_render(th_context* ctx, th_scene* s, const TH_GpuRenderParams* p, const TH_GpuCamera* cam,
                  const th_render_options* opt) {
  if (!ctx || !s || !cam) return th_set_error(TH_E_INVALID, "null argument");
  if (s->ctx != ctx) return th_set_error(TH_E_INVALID, "scene belongs to another context");
  int32_t rc = check_render_params(s, p);
  if (rc) return rc;
  th_render_options o;
  memset(&o, 0, sizeof o);
  if (opt) o = *opt;
  if (o.rng_mode > 1u || o.traversal > 2u || o.kernel > 1u || o.schedule > 4u)
    return th_set_error(TH_E_INVALID, "unknown render option");
  if (o.schedule == TH_SCHEDULE_AUTO) {
    if (o.rng_mode == TH_RNG_REFERENCE && o.traversal == TH_TRAVERSE_WIDE && o.kernel == TH_KERNEL_MEGA &&
        !o.collect_stats && s->node_count > 0) {
      if ((rc = auto_schedule(ctx, s, p, cam, &o))) return rc;
    } else {
      o.schedule = TH_SCHEDULE_STATIC;
    }
  }
  if (o.traversal == TH_TRAVERSE_WIDE && o.collect_stats)
    return th_set_error(TH_E_UNSUPPORTED, "statistics are counted with the scalar traversals");
  if (o.kernel == TH_KERNEL_WAVEFRONT && o.collect_stats)
    return th_set_error(TH_E_UNSUPPORTED, "statistics are counted by the megakernel");
  TH_CHECK(hipSetDevice(ctx->device));
  DevScene S = dev_scene(s);
  uint32_t depth = std::max<uint32_t>(s->max_depth + 1u, 2u);
  const uint32_t block = 64;
  size_t lds = (size_t)depth * block * sizeof(uint2);
  if (lds > 64 * 1024) return th_set_error(TH_E_UNSUPPORTED, "bvh too deep for the traversal stack");
  if (o.collect_stats) {  // a fill kernel, not a memset node, so a caller may capture the render
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, ctx->stream, (uint32_t*)ctx->d_counters,
                       (uint32_t)(16 * sizeof(unsigned long long) / sizeof(uint32_t)));
    TH_CHECK(hipGetLastError());
  }
  unsigned long long* cnt = ctx->d_counters;
  std::pair<hipEvent_t, hipEvent_t> evp;
  if ((rc = begin_timed(ctx, &evp))) return rc;
  if ((rc = launch_render(ctx, s, p, cam, o, S, lds, block, cnt))) {
    (void)end_timed(ctx, evp);
    return rc;
  }
  hipError_t le = hipGetLastError();
  if ((rc = end_timed(ctx, evp))) return rc;
  if (le != hipSuccess) return th_set_error(TH_E_DEVICE, std::string("render launch: ") + hipGetErrorString(le));
  ctx->counters_pending = o.collect_stats != 0;
  ctx->stats.paths = (uint64_t)p->width * p->height * p->samples;
  return TH_OK;
}

// ---------------------------------------------------------------- progressive graph
// One progressive batch captured as a hipGraph -- the render's kernels reading the frame
// seed and accumulated sample count from device memory, then a one-lane kernel that
// advances both -- so each replay is the next batch of the reference's progressive loop
// (path_tracer.rs:655-674, 900-923: frame seed = frame counter, running average over
// accumulated + samples) with no host work but the graph launch.
struct th_progressive_graph {
  th_context* ctx = nullptr;
  th_scene* scene = nullptr;
  uint32_t* d_state = nullptr;  // {frame_index, accumulated_samples}
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t samples = 0, frame_index = 0, accumulated = 0;
  uint64_t generation = 0;
};

namespace {
__global__ void k_frame_state_set(uint32_t* st, uint32_t frame_index, uint32_t accumulated) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st[0] = frame_index;
    st[1] = accumulated;
  }
}
__global__ void k_frame_state_advance(uint32_t* st, uint32_t samples) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st[0] = st[0] + 1u;
    st[1] = st[1] + samples;
  }
}
}  // namespace

void th_progressive_graph_destroy(th_progressive_graph* g) {
  if (!g) return;
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->d_state) (void)hipFree(g->d_state);
  delete g;
}

int32_t th_progressive_graph_create(th_context* ctx, th_scene* s, const TH_GpuRenderParams* p,
                                    const TH_GpuCamera* cam, const th_render_options* opt,
                                    th_progressive_graph** out) {
  if (!ctx || !s || !cam || !out) return th_set_error(TH_E_INVALID, "null argument");
  *out = nullptr;
  if (s->ctx != ctx) return th_set_error(TH_E_INVALID, "scene belongs to another context");
  int32_t rc = check_render_params(s, p);
  if (rc) return rc;
  th_render_options o;
  memset(&o, 0, sizeof o);
  if (opt) o = *opt;
  if (o.rng_mode > 1u || o.traversal > 2u || o.kernel > 1u || o.schedule > 4u)
    return th_set_error(TH_E_INVALID, "unknown render option");
  if (o.kernel != TH_KERNEL_MEGA || o.schedule == TH_SCHEDULE_AUTO || o.collect_stats)
    return th_set_error(TH_E_UNSUPPORTED, "progressive graphs capture the megakernels with a fixed schedule and "
                                          "no statistics");
  TH_CHECK(hipSetDevice(ctx->device));
  DevScene S = dev_scene(s);
  uint32_t depth = std::max<uint32_t>(s->max_depth + 1u, 2u);
  const uint32_t block = 64;
  size_t lds = (size_t)depth * block * sizeof(uint2);
  if (lds > 64 * 1024) return th_set_error(TH_E_UNSUPPORTED, "bvh too deep for the traversal stack");
  th_progressive_graph* g = new th_progressive_graph();
  g->ctx = ctx;
  g->scene = s;
  g->samples = p->samples;
  g->frame_index = p->frame_index;
  g->accumulated = p->accumulated_samples;
  g->generation = s->generation;
  hipError_t e = hipMalloc((void**)&g->d_state, 2 * sizeof(uint32_t));
  if (e != hipSuccess) {
    delete g;
    return th_set_error(TH_E_DEVICE, std::string("progressive graph state: ") + hipGetErrorString(e));
  }
  hipLaunchKernelGGL(k_frame_state_set, dim3(1), dim3(64), 0, ctx->stream, g->d_state, p->frame_index,
                     p->accumulated_samples);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    th_progressive_graph_destroy(g);
    return th_set_error(TH_E_DEVICE, std::string("progressive graph capture: ") + hipGetErrorString(e));
  }
  S.frame_state = g->d_state;
  rc = launch_render(ctx, s, p, cam, o, S, lds, block, ctx->d_counters);
  hipLaunchKernelGGL(k_frame_state_advance, dim3(1), dim3(64), 0, ctx->stream, g->d_state, p->samples);
  hipError_t le = hipGetLastError();
  e = hipStreamEndCapture(ctx->stream, &g->graph);  // always ends the capture
  if (rc) {
    th_progressive_graph_destroy(g);
    return rc;
  }
  if (e == hipSuccess && le != hipSuccess) e = le;
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    th_progressive_graph_destroy(g);
    return th_set_error(TH_E_DEVICE, std::string("progressive graph capture: ") + hipGetErrorString(e));
  }
  *out = g;
  return TH_OK;
}

int32_t th_progressive_graph_replay(th_progressive_graph* g, uint32_t frames) {
  if (!g) return th_set_error(TH_E_INVALID, "null argument");
  if (g->generation != g->scene->generation)
    return th_set_error(TH_E_MISMATCH, "the scene was updated after the graph was captured");
  TH_CHECK(hipSetDevice(g->ctx->device));
  for (uint32_t f = 0; f < frames; ++f) {
    TH_CHECK(hipGraphLaunch(g->exec, g->ctx->stream));
    g->frame_index += 1u;
    g->accumulated += g->samples;
  }
  return TH_OK;
}

int32_t th_progressive_graph_state(const th_progressive_graph* g, uint32_t* frame_index, uint32_t* accumulated) {
  if (!g || !frame_index || !accumulated) return th_set_error(TH_E_INVALID, "null argument");
  *frame_index = g->frame_index;
  *accumulated = g->accumulated;
  return TH_OK;
}

int32_t th_synchronize(th_context* ctx) {
  if (!ctx) return th_set_error(TH_E_INVALID, "null argument");
  TH_CHECK(hipSetDevice(ctx->device));
  TH_CHECK(hipStreamSynchronize(ctx->stream));
  return TH_OK;
}

int32_t th_readback(th_context* ctx, th_scene* s, float* rgba, size_t bytes) {
  if (!ctx || !s || !rgba) return th_set_error(TH_E_INVALID, "null argument");
  size_t need = (size_t)s->width * s->height * 4 * sizeof(float);
  if (bytes != need) return th_set_error(TH_E_INVALID, "readback buffer size mismatch");
  TH_CHECK(hipSetDevice(ctx->device));
  TH_CHECK(hipMemcpyAsync(rgba, s->d_accum, need, hipMemcpyDeviceToHost, ctx->stream));
  TH_CHECK(hipStreamSynchronize(ctx->stream));
  return TH_OK;
}

int32_t th_upload_accumulation(th_context* ctx, th_scene* s, const float* rgba, size_t bytes) {
  if (!ctx || !s || !rgba) return th_set_error(TH_E_INVALID, "null argument");
  size_t need = (size_t)s->width * s->height * 4 * sizeof(float);
  if (bytes != need) return th_set_error(TH_E_INVALID, "accumulation buffer size mismatch");
  TH_CHECK(hipSetDevice(ctx->device));
  TH_CHECK(hipMemcpyAsync(s->d_accum, rgba, need, hipMemcpyHostToDevice, ctx->stream));
  TH_CHECK(hipStreamSynchronize(ctx->stream));
  return TH_OK;
}

int32_t th_get_stats(const th_context* cctx, th_stats* out) {
  if (!cctx || !out) return th_set_error(TH_E_INVALID, "null argument");
  th_context* ctx = const_cast<th_context*>(cctx);
  TH_CHECK(hipSetDevice(ctx->device));
  int32_t rc = fold_timing(ctx);
  if (rc) return rc;
  if (ctx->counters_pending) {
    unsigned long long h[16];
    TH_CHECK(hipMemcpy(h, ctx->d_counters, sizeof h, hipMemcpyDeviceToHost));
    th_stats& s = ctx->stats;
    s.paths = h[0];
    s.closest_rays = h[1];
    s.shadow_rays = h[2];
    s.rays = h[1] + h[2];
    s.nodes_visited = h[3];
    s.primref_tests = h[4];
    s.triangle_tests = h[5];
    s.sphere_tests = h[6];
    s.closest_hits = h[7];
    s.triangle_hits = h[8];
    s.lights_scanned = h[9];
    s.env_lookups = h[10];
    s.env_samples = h[11];
    s.closest_nodes_visited = h[12];
    s.closest_primref_tests = h[13];
    s.closest_triangle_tests = h[14];
    s.closest_sphere_tests = h[15];
    ctx->counters_pending = false;
  }
  *out = ctx->stats;
  return TH_OK;
}

int32_t th_reset_stats(th_context* ctx) {
  if (!ctx) return th_set_error(TH_E_INVALID, "null argument");
  TH_CHECK(hipSetDevice(ctx->device));
  int32_t rc = fold_timing(ctx);
  if (rc) return rc;
  memset(&ctx->stats, 0, sizeof ctx->stats);
  return TH_OK;
}

int32_t th_rng_draws(th_context* ctx, const uint32_t* pixels, uint32_t count, uint32_t frame_index,
                     uint32_t rng_mode, uint32_t sample, uint32_t draws, uint32_t* states_out,
                     float* values_out) {
  if (!ctx || !pixels || !states_out || !values_out || count == 0 || draws == 0)
    return th_set_error(TH_E_INVALID, "null argument or empty query");
  if (rng_mode > 1u) return th_set_error(TH_E_INVALID, "unknown rng mode");
  if ((uint64_t)count * draws > (1u << 24)) return th_set_error(TH_E_INVALID, "query larger than 2^24 draws");
  TH_CHECK(hipSetDevice(ctx->device));
  const size_t n = (size_t)count * draws;
  uint32_t *d_pix = nullptr, *d_st = nullptr;
  float* d_v = nullptr;
  hipError_t e = hipMalloc((void**)&d_pix, count * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc((void**)&d_st, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc((void**)&d_v, n * sizeof(float));
  if (e == hipSuccess) e = hipMemcpyAsync(d_pix, pixels, count * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_rng_draws, dim3((count + 255) / 256), dim3(256), 0, ctx->stream, d_pix, count,
                       frame_index, rng_mode, sample, draws, d_st, d_v);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(states_out, d_st, n * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(values_out, d_v, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (d_pix) (void)hipFree(d_pix);
  if (d_st) (void)hipFree(d_st);
  if (d_v) (void)hipFree(d_v);
  if (e != hipSuccess) return th_set_error(TH_E_DEVICE, std::string("rng draws: ") + hipGetErrorString(e));
  return TH_OK;
}

int32_t th_primary_hits(th_context* ctx, th_scene* s, const TH_GpuRenderParams* p,
                        const TH_GpuCamera* cam, uint32_t traversal, uint32_t* slot_out,
                        float* distance_out, size_t count) {
  if (!ctx || !s || !cam || !slot_out || !distance_out)
    return th_set_error(TH_E_INVALID, "null argument");
  int32_t rc = check_render_params(s, p);
  if (rc) return rc;
  size_t npix = (size_t)p->width * p->height;
  if (count != npix) return th_set_error(TH_E_INVALID, "output size must equal width*height");
  if (traversal > 2u) return th_set_error(TH_E_INVALID, "unknown traversal");
  TH_CHECK(hipSetDevice(ctx->device));
  uint32_t* d_slot = nullptr;
  float* d_t = nullptr;
  TH_CHECK(hipMalloc((void**)&d_slot, npix * sizeof(uint32_t)));
  hipError_t e = hipMalloc((void**)&d_t, npix * sizeof(float));
  if (e != hipSuccess) {
    (void)hipFree(d_slot);
    return th_set_error(TH_E_DEVICE, "hipMalloc");
  }
  DevScene S = dev_scene(s);
  uint32_t depth = std::max<uint32_t>(s->max_depth + 1u, 2u);
  size_t lds = (size_t)depth * 64 * sizeof(uint2);
  uint32_t tiles = ((p->width + 7) / 8) * ((p->height + 7) / 8);
  if (traversal == 2 && s->node_count > 0) {
    size_t lds4 = (size_t)(3u * s->depth4 + 2u) * 16 * sizeof(uint2);
    uint32_t waves = ((p->width + 3) / 4) * ((p->height + 3) / 4);
    hipLaunchKernelGGL(k_primary_quad, dim3(waves), dim3(64), lds4, ctx->stream, S, *p, *cam, d_slot, d_t);
  } else if (traversal == 0)
    hipLaunchKernelGGL(k_primary<0>, dim3(tiles), dim3(64), lds, ctx->stream, S, *p, *cam, d_slot, d_t);
  else
    hipLaunchKernelGGL(k_primary<1>, dim3(tiles), dim3(64), lds, ctx->stream, S, *p, *cam, d_slot, d_t);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(slot_out, d_slot, npix * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(distance_out, d_t, npix * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d_slot);
  (void)hipFree(d_t);
  if (e != hipSuccess) return th_set_error(TH_E_DEVICE, std::string("primary hits: ") + hipGetErrorString(e));
  return TH_OK;
}

int32_t th_display_rgba8_device(th_context* ctx, th_scene* s, float exposure_stops, uint8_t* out,
                                size_t out_bytes) {
  if (!ctx || !s || !out) return th_set_error(TH_E_INVALID, "null argument");
  size_t npix = (size_t)s->width * s->height;
  if (out_bytes != npix * 4) return th_set_error(TH_E_INVALID, "output size must be width*height*4");
  TH_CHECK(hipSetDevice(ctx->device));
  if (!s->d_rgba8) TH_CHECK(hipMalloc((void**)&s->d_rgba8, npix * 4));  // kept with the scene
  float scale = exp2f(exposure_stops);
  hipLaunchKernelGGL(k_display, dim3((npix + 255) / 256), dim3(256), 0, ctx->stream, s->d_accum,
                     (uint32_t)npix, scale, ctx->d_srgb_lut + 256, s->d_rgba8);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, s->d_rgba8, npix * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return th_set_error(TH_E_DEVICE, std::string("display: ") + hipGetErrorString(e));
  return TH_OK;
}

}  // extern "C"
