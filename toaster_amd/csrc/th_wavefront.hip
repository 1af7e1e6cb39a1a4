// th_wavefront.hip -- wavefront path tracer: the reference estimator
// (shaders/pathtrace_spheres.wgsl:508-548, 805-916) split into stages that each run
// over a compacted queue of live paths.
//
// Per iteration: extend (closest hit, 4 lanes per ray) -> shade (one lane per path:
// one bounce of the reference loop body, emits up to two shadow rays) -> shadow (any
// hit, 4 lanes per ray) -> finish (adds visible direct light in the reference's order,
// completes samples, regenerates the next camera path of the same pixel from the same
// random stream, and compacts live paths into the next queue with a wave ballot plus
// one atomic per wave).  Queues are split into WF_NSEG segments with a counter each:
// block b reads and appends to segment b % WF_NSEG, so a slot stays in its segment (the
// slots k with (k / 64) % WF_NSEG equal to it) and the append atomics spread over as
// many addresses (with one counter per queue they serialise in the memory system).  A path slot owns one pixel (reference streams: all its samples
// in stream order) or one of G sample groups of a pixel (per-sample streams, combined
// in a fixed tree order by the resolve kernel).  Every value a stage hands to the next
// goes through HBM in structure-of-arrays form; the deferred direct-light terms are the
// exact products the reference adds, added in the reference's order, so the image is
// the megakernel's image bit for bit.
#include "th_internal.h"

namespace {

#define WF_INCLUDE_EMISSIVE 1u
#define WF_PREV_DIFFUSE 2u
#define WF_DONE 4u
#define WF_PEND_LIGHT 8u
#define WF_PEND_ENV 16u
#define WF_C(c) ((c) * 32u)  // queue counters 128 bytes apart: no two share a cache line
#define WF_NSEG 16u
#define WF_QC(q, sg) WF_C(4u + (q) * WF_NSEG + (sg))  // queue q: 0, 1 live, 2 shadow
#define WF_CNT_LINES (4u + 3u * WF_NSEG)

struct WfBufs {
  float4 *o, *d, *thr, *rad, *acc;  // o.w = prev bsdf pdf, d.w = flag bits
  uint4* st;                        // rng, next sample index, bounce, pixel
  uint2* hit;                       // t bits, slot
  float4 *sho, *sh0, *c0, *sh1, *c1;
  uint32_t* vis;                     // 2 per slot
  uint32_t *q0, *q1, *qs;            // live queues (ping-pong), shadow queue; in segments
  uint32_t* cnt;                     // WF_QC counters
  uint32_t seg_cap;                  // slots per segment of q0, q1; qs holds twice that
};

struct WfArgs {
  uint32_t n_slots;  // npix * G
  uint32_t G;
  uint32_t per_sample;
};

__device__ __forceinline__ uint32_t wf_seed(const TH_GpuRenderParams& P, const WfArgs& A,
                                            uint32_t pixel, uint32_t smp) {
  uint32_t base = pixel ^ (P.frame_index * 9781u);
  return A.per_sample ? th::pcg_hash(base + smp * 0x9E3779B9u) : th::pcg_hash(base);
}

// Starts sample `smp` of slot k: camera ray (two draws), fresh path state.
__device__ __forceinline__ void wf_begin_sample(const WfBufs& B, const TH_GpuRenderParams& P,
                                                const TH_GpuCamera& cam, uint32_t k, uint32_t pixel,
                                                uint32_t smp, uint32_t rng) {
  uint32_t x = pixel % P.width, y = pixel / P.width;
  v3 o, d;
  camera_ray(cam, P, x, y, &rng, &o, &d);
  B.o[k] = make_float4(o.x, o.y, o.z, 0.0f);
  B.d[k] = make_float4(d.x, d.y, d.z, __uint_as_float(WF_INCLUDE_EMISSIVE));
  B.thr[k] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
  B.rad[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  B.st[k] = make_uint4(rng, smp, 0u, pixel);
}

// Appends `live` lanes' values to a queue: one atomic per wave, order kept within a wave.
__device__ __forceinline__ void wf_push(bool live, uint32_t value, uint32_t* queue, uint32_t* count) {
  uint64_t m = __ballot(live);
  if (m == 0) return;
  uint32_t lane = threadIdx.x & 63u;
  uint32_t first = (uint32_t)__ffsll((unsigned long long)m) - 1u;
  uint32_t base = 0;
  if (lane == first) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = __shfl(base, (int)first, 64);
  if (live) {
    uint64_t below = m & ((1ull << lane) - 1ull);
    queue[base + (uint32_t)__popcll(below)] = value;
  }
}

__global__ void wf_reset(WfBufs B) {
  for (uint32_t c = threadIdx.x; c < WF_CNT_LINES; c += blockDim.x) B.cnt[WF_C(c)] = 0u;
}

// Appends: `la` lanes' values a, then `lb` lanes' values b, with one atomic.
__device__ __forceinline__ void wf_push2(bool la, uint32_t a, bool lb, uint32_t b, uint32_t* queue,
                                         uint32_t* count) {
  uint64_t ma = __ballot(la), mb = __ballot(lb);
  if ((ma | mb) == 0) return;
  uint32_t lane = threadIdx.x & 63u;
  uint32_t first = (uint32_t)__ffsll((unsigned long long)(ma | mb)) - 1u;
  uint32_t na = (uint32_t)__popcll(ma);
  uint32_t base = 0;
  if (lane == first) base = atomicAdd(count, na + (uint32_t)__popcll(mb));
  base = __shfl(base, (int)first, 64);
  uint64_t lt = (1ull << lane) - 1ull;
  if (la) queue[base + (uint32_t)__popcll(ma & lt)] = a;
  if (lb) queue[base + na + (uint32_t)__popcll(mb & lt)] = b;
}

// ---------------------------------------------------------------- generate
__global__ void __launch_bounds__(256) wf_generate(WfBufs B, WfArgs A, TH_GpuRenderParams P,
                                                   TH_GpuCamera cam) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t sg = (k >> 6) % WF_NSEG;  // this wave's segment (wf_reset cleared the counters)
  bool live = false;
  if (k < A.n_slots) {
    uint32_t pixel = k / A.G, j = k % A.G;
    B.acc[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint32_t smp = A.per_sample ? j : 0u;
    if (smp < P.samples) {
      wf_begin_sample(B, P, cam, k, pixel, smp, wf_seed(P, A, pixel, smp));
      live = true;
    } else {
      B.st[k] = make_uint4(0u, smp, 0u, pixel);
    }
  }
  wf_push(live, k, B.q0 + (size_t)sg * B.seg_cap, &B.cnt[WF_QC(0u, sg)]);
}

// ---------------------------------------------------------------- extend
template <int TMODE>
__global__ void __launch_bounds__(64) wf_extend(WfBufs B, DevScene S, TH_GpuRenderParams P,
                                                uint32_t live_idx, uint32_t per_quad) {
  extern __shared__ uint2 lds_stack[];
  const uint32_t sg = blockIdx.x % WF_NSEG;
  const uint32_t* q = (live_idx ? B.q1 : B.q0) + (size_t)sg * B.seg_cap;
  uint32_t n = B.cnt[WF_QC(live_idx, sg)];
  if (blockIdx.x == 0)  // shadow queue of this iteration
    for (uint32_t c = threadIdx.x; c < WF_NSEG; c += blockDim.x) B.cnt[WF_QC(2u, c)] = 0;
  uint32_t lanes_per = per_quad ? 4u : 1u;
  Stack st{lds_stack + (per_quad ? (threadIdx.x >> 2) : threadIdx.x),
           (int)(per_quad ? (blockDim.x >> 2) : blockDim.x), 0};
  const uint32_t rays = blockDim.x / lanes_per;
  uint32_t stride = (gridDim.x / WF_NSEG) * rays;
  for (uint32_t i = (blockIdx.x / WF_NSEG) * rays + threadIdx.x / lanes_per; ; i += stride) {
    // the loop bound is uniform per quad, so quads stay converged for traverse_quad
    bool act = i < n;
    if (!__any(act)) break;
    if (act) {
      uint32_t k = q[i];
      float4 o4 = B.o[k], d4 = B.d[k];
      v3 o = th::mk(o4.x, o4.y, o4.z), d = th::mk(d4.x, d4.y, d4.z);
      float t;
      uint32_t slot;
      RayStats<false> rs;
      if (S.node_count == 0u)
        t = closest_linear<false>(S, P.sphere_count, P.triangle_count, o, d, TH_FLT_MAX, rs, &slot);
      else if (TMODE == 3)
        traverse_quad<1>(S, (const Child4*)S.nodes4, S.root4, o, d, TH_FLT_MAX, st, &t, &slot);
      else
        traverse<TMODE, false>(S, o, d, TH_FLT_MAX, st, rs, &t, &slot);
      if (!per_quad || (threadIdx.x & 3u) == 0u) B.hit[k] = make_uint2(__float_as_uint(t), slot);
    }
  }
}

// ---------------------------------------------------------------- shade
// One iteration of the reference loop body (wgsl:817-911) for the hit found by extend.
// `pretest`: shadow rays an extracted primitive blocks are resolved here (the 4-wide
// traversal's own first step in the same arithmetic); the scalar traversals keep every
// shadow ray for their reference-order walk.
__global__ void __launch_bounds__(64) wf_shade(WfBufs B, DevScene S, TH_GpuRenderParams P,
                                               const float* __restrict__ srgb_lut, uint32_t live_idx,
                                               uint32_t pretest) {
  const uint32_t sg = blockIdx.x % WF_NSEG;
  const uint32_t* q = (live_idx ? B.q1 : B.q0) + (size_t)sg * B.seg_cap;
  uint32_t n = B.cnt[WF_QC(live_idx, sg)];
  if (blockIdx.x == 0)  // next live queue
    for (uint32_t c = threadIdx.x; c < WF_NSEG; c += blockDim.x) B.cnt[WF_QC(live_idx ^ 1u, c)] = 0;
  uint32_t* qs = B.qs + (size_t)sg * 2u * B.seg_cap;
  Stack st{nullptr, 0, 0};
  LaneCounters<false> lc;
  Integrator<1, false> it{S, P, srgb_lut, st, lc};
  uint32_t stride = (gridDim.x / WF_NSEG) * blockDim.x;
  for (uint32_t i = (blockIdx.x / WF_NSEG) * blockDim.x + threadIdx.x; ; i += stride) {
    bool act = i < n;
    if (!__any(act)) break;
    bool sh_light = false, sh_env = false;
    uint32_t k = 0;
    if (act) {
      k = q[i];
      float4 o4 = B.o[k], d4 = B.d[k], t4 = B.thr[k], r4 = B.rad[k];
      uint4 s4 = B.st[k];
      uint2 h = B.hit[k];
      v3 o = th::mk(o4.x, o4.y, o4.z), d = th::mk(d4.x, d4.y, d4.z);
      v3 throughput = th::mk(t4.x, t4.y, t4.z), radiance = th::mk(r4.x, r4.y, r4.z);
      uint32_t flags = __float_as_uint(d4.w) & (WF_INCLUDE_EMISSIVE | WF_PREV_DIFFUSE);
      float prev_pdf = o4.w;
      uint32_t rng = s4.x, bounce = s4.z + 1u;
      float t = __uint_as_float(h.x);
      uint32_t slot = h.y;
      bool done = false;
      if (slot == THD_NO_SLOT || t > TH_MAX_DISTANCE) {  // miss (wgsl:819-834)
        if (P.background_kind == 0u) {
          float lt = (th::normalize(d).y + 1.0f) * 0.5f;
          radiance = th::add(radiance, th::mul(throughput, th::mix(th::mk(1.0f, 1.0f, 1.0f),
                                                                    th::mk(0.35f, 0.65f, 1.0f), lt)));
        } else if (P.background_kind == 2u) {
          float mis = 1.0f;
          if (flags & WF_PREV_DIFFUSE) mis = it.power_heuristic(prev_pdf, it.env_pdf(d));
          radiance = th::add(radiance, th::muls(th::mul(throughput, it.env_lookup(d)), mis));
        }
        done = true;
      } else {
        Hit rec;
        if (S.node_count == 0u)
          finalize_linear(S, slot, o, d, t, &rec);
        else
          finalize_slot(S, slot, o, d, t, &rec);
        const TH_GpuMaterial m = load_material(S, rec.mat);
        v3 att = th::mk(0, 0, 0), nd = d;
        bool brk = false;
        if (m.kind == 0u) {
          v3 albedo = it.base_color(m, rec);
          // direct light (wgsl:406-472) with the visibility test deferred
          {
            v3 so, dir, term;
            float limit;
            // a shadow ray an extracted primitive blocks adds nothing: resolved here
            if (it.light_candidate(rec.p, rec.n, albedo, &rng, &so, &dir, &limit, &term) &&
                !(pretest && lane_extracted_anyhit(S, so, dir, limit))) {
              v3 c = th::mul(throughput, term);
              B.sho[k] = make_float4(so.x, so.y, so.z, 0.0f);
              B.sh0[k] = make_float4(dir.x, dir.y, dir.z, limit);
              B.c0[k] = make_float4(c.x, c.y, c.z, 0.0f);
              sh_light = true;
            }
          }
          // direct environment (wgsl:486-505) with the visibility test deferred
          if (P.background_kind == 2u) {
            EnvSample es;
            float sc;
            if (it.env_candidate(rec.n, &rng, &es, &sc)) {
              v3 c = th::mul(throughput, it.env_term(albedo, es, sc, it.power_heuristic(es.pdf, sc / TH_PI)));
              v3 so = th::add(rec.p, th::muls(rec.n, TH_MIN_DISTANCE));
              if (!(pretest && lane_extracted_anyhit(S, so, es.dir, 10000.001f))) {
                B.sho[k] = make_float4(so.x, so.y, so.z, 0.0f);
                B.sh1[k] = make_float4(es.dir.x, es.dir.y, es.dir.z, 10000.001f);
                B.c1[k] = make_float4(c.x, c.y, c.z, 0.0f);
                sh_env = true;
              }
            }
          }
          v3 dir = th::add(rec.n, it.random_unit_vector(&rng));
          if (th::dot(dir, dir) < TH_EPSILON) dir = rec.n;
          nd = th::normalize(dir);
          att = albedo;
          flags = WF_PREV_DIFFUSE;
          prev_pdf = th::fmax_(0.0f, th::dot(rec.n, nd)) / TH_PI;
        } else if (m.kind == 1u) {
          v3 refl = it.reflect_vec(d, rec.n);
          float rough = th::clampf(m.params[0], 0.0f, 1.0f);
          v3 dir = th::add(refl, th::smul(rough, it.random_unit_vector(&rng)));
          if (th::dot(dir, rec.n) <= 0.0f) brk = true;
          nd = th::normalize(dir);
          att = th::mk(m.albedo[0], m.albedo[1], m.albedo[2]);
          flags = WF_INCLUDE_EMISSIVE;
        } else if (m.kind == 2u) {
          float ratio = rec.front ? 1.0f / m.params[1] : m.params[1];
          v3 ud = th::normalize(d);
          float ct = th::fmin_(1.0f, th::dot(th::neg(ud), rec.n));
          float stt = th::fsqrt(1.0f - ct * ct);
          bool cannot = (ratio * stt) > 1.0f;
          v3 dir;
          if (cannot || it.reflectance(ct, ratio) > th::rand01(&rng))
            dir = it.reflect_vec(ud, rec.n);
          else
            dir = it.refract_vec(ud, rec.n, ratio);
          nd = th::normalize(dir);
          att = th::mk(1, 1, 1);
          flags = WF_INCLUDE_EMISSIVE;
        } else {
          if (m.kind == 3u && (flags & WF_INCLUDE_EMISSIVE))
            radiance = th::add(radiance, th::muls(th::mul(throughput, th::mk(m.albedo[0], m.albedo[1],
                                                                              m.albedo[2])),
                                                  m.params[2]));
          brk = true;
        }
        if (brk) {
          done = true;
        } else {
          throughput = th::mul(throughput, att);
          B.o[k] = make_float4(rec.p.x, rec.p.y, rec.p.z, prev_pdf);
          if (bounce >= P.max_bounces) done = true;
        }
        d = nd;
      }
      uint32_t f = flags | (done ? WF_DONE : 0u) | (sh_light ? WF_PEND_LIGHT : 0u) |
                   (sh_env ? WF_PEND_ENV : 0u);
      B.d[k] = make_float4(d.x, d.y, d.z, __uint_as_float(f));
      B.thr[k] = make_float4(throughput.x, throughput.y, throughput.z, 0.0f);
      B.rad[k] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
      B.st[k] = make_uint4(rng, s4.y, bounce, s4.w);
    }
    wf_push2(sh_light, 2u * k, sh_env, 2u * k + 1u, qs, &B.cnt[WF_QC(2u, sg)]);
  }
}

// ---------------------------------------------------------------- shadow
template <int TMODE>
__global__ void __launch_bounds__(64) wf_shadow(WfBufs B, DevScene S, TH_GpuRenderParams P,
                                                uint32_t per_quad) {
  extern __shared__ uint2 lds_stack[];
  const uint32_t sg = blockIdx.x % WF_NSEG;
  const uint32_t* qs = B.qs + (size_t)sg * 2u * B.seg_cap;
  uint32_t n = B.cnt[WF_QC(2u, sg)];
  uint32_t lanes_per = per_quad ? 4u : 1u;
  Stack st{lds_stack + (per_quad ? (threadIdx.x >> 2) : threadIdx.x),
           (int)(per_quad ? (blockDim.x >> 2) : blockDim.x), 0};
  const uint32_t rays = blockDim.x / lanes_per;
  uint32_t stride = (gridDim.x / WF_NSEG) * rays;
  for (uint32_t i = (blockIdx.x / WF_NSEG) * rays + threadIdx.x / lanes_per; ; i += stride) {
    bool act = i < n;
    if (!__any(act)) break;
    if (act) {
      uint32_t e = qs[i], k = e >> 1;
      float4 so = B.sho[k], sd = (e & 1u) ? B.sh1[k] : B.sh0[k];
      v3 o = th::mk(so.x, so.y, so.z), d = th::mk(sd.x, sd.y, sd.z);
      float limit = sd.w, t;
      uint32_t slot;
      RayStats<false> rs;
      if (S.node_count == 0u)
        t = closest_linear<false>(S, P.sphere_count, P.triangle_count, o, d, limit, rs, &slot);
      else if (TMODE == 3)
        traverse_quad<2>(S, (const Child4*)S.nodes4, S.root4, o, d, limit, st, &t, &slot);
      else if (TMODE == 0)
        traverse<0, false>(S, o, d, TH_FLT_MAX, st, rs, &t, &slot);
      else
        traverse<2, false>(S, o, d, limit, st, rs, &t, &slot);
      if (!per_quad || (threadIdx.x & 3u) == 0u) B.vis[e] = (t < limit) ? 0u : 1u;
    }
  }
}

// ---------------------------------------------------------------- finish
__global__ void __launch_bounds__(64) wf_finish(WfBufs B, WfArgs A, TH_GpuRenderParams P,
                                                TH_GpuCamera cam, uint32_t live_idx) {
  const uint32_t sg = blockIdx.x % WF_NSEG;
  const uint32_t* q = (live_idx ? B.q1 : B.q0) + (size_t)sg * B.seg_cap;
  uint32_t* qn = (live_idx ? B.q0 : B.q1) + (size_t)sg * B.seg_cap;
  uint32_t n = B.cnt[WF_QC(live_idx, sg)];
  uint32_t stride = (gridDim.x / WF_NSEG) * blockDim.x;
  for (uint32_t i = (blockIdx.x / WF_NSEG) * blockDim.x + threadIdx.x; ; i += stride) {
    bool act = i < n;
    if (!__any(act)) break;
    bool live = false;
    uint32_t k = 0;
    if (act) {
      k = q[i];
      float4 d4 = B.d[k];
      uint32_t f = __float_as_uint(d4.w);
      float4 r4 = B.rad[k];
      if (f & (WF_PEND_LIGHT | WF_PEND_ENV)) {
        // radiance += throughput * direct term, light first then environment (wgsl:846-847)
        if ((f & WF_PEND_LIGHT) && B.vis[2u * k]) {
          float4 c = B.c0[k];
          r4 = make_float4(r4.x + c.x, r4.y + c.y, r4.z + c.z, 0.0f);
        }
        if ((f & WF_PEND_ENV) && B.vis[2u * k + 1u]) {
          float4 c = B.c1[k];
          r4 = make_float4(r4.x + c.x, r4.y + c.y, r4.z + c.z, 0.0f);
        }
        B.rad[k] = r4;
      }
      if (f & WF_DONE) {
        // color += vec4(ray_color, 1) (wgsl:536), then the next sample of this slot
        float4 a = B.acc[k];
        B.acc[k] = make_float4(a.x + r4.x, a.y + r4.y, a.z + r4.z, a.w + 1.0f);
        uint4 s4 = B.st[k];
        uint32_t smp = s4.y + (A.per_sample ? A.G : 1u);
        if (smp < P.samples) {
          uint32_t rng = A.per_sample ? wf_seed(P, A, s4.w, smp) : s4.x;
          wf_begin_sample(B, P, cam, k, s4.w, smp, rng);
          live = P.max_bounces > 0u;
          if (!live) {  // zero-bounce paths finish at once (ray_color returns 0)
            // remaining samples contribute (0, 0, 0, 1) each
            float4 a2 = B.acc[k];
            uint32_t left = 0;
            for (uint32_t ss = smp; ss < P.samples; ss += (A.per_sample ? A.G : 1u)) ++left;
            for (uint32_t c = 0; c < left; ++c) a2.w = a2.w + 1.0f;
            B.acc[k] = a2;
          }
        }
      } else {
        B.d[k] = make_float4(d4.x, d4.y, d4.z,
                             __uint_as_float(f & (WF_INCLUDE_EMISSIVE | WF_PREV_DIFFUSE)));
        live = true;
      }
    }
    wf_push(live, k, qn, &B.cnt[WF_QC(live_idx ^ 1u, sg)]);
  }
}

// ---------------------------------------------------------------- resolve
// Batch average of the G slots of each pixel (pairwise tree in slot order) folded into
// the progressive running average (wgsl:538-547).
__global__ void __launch_bounds__(256) wf_resolve(WfBufs B, WfArgs A, TH_GpuRenderParams P,
                                                  float4* __restrict__ accum) {
  uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t npix = P.width * P.height;
  if (px >= npix) return;
  float4 part[16];
  for (uint32_t j = 0; j < A.G; ++j) part[j] = B.acc[px * A.G + j];
  for (uint32_t off = 1; off < A.G; off <<= 1)
    for (uint32_t j = 0; j < A.G; j += 2u * off)
      part[j] = make_float4(part[j].x + part[j + off].x, part[j].y + part[j + off].y,
                            part[j].z + part[j + off].z, part[j].w + part[j + off].w);
  float ns = (float)P.samples;
  float4 b = make_float4(part[0].x / ns, part[0].y / ns, part[0].z / ns, part[0].w / ns);
  if (P.accumulated_samples == 0u) {
    accum[px] = b;
  } else {
    float prev = (float)P.accumulated_samples;
    float tot = prev + ns;
    float4 o = accum[px];
    accum[px] = make_float4((o.x * prev + b.x * ns) / tot, (o.y * prev + b.y * ns) / tot,
                            (o.z * prev + b.z * ns) / tot, (o.w * prev + b.w * ns) / tot);
  }
}

}  // namespace

struct th_wavefront {
  uint32_t capacity = 0;
  WfBufs b;
  std::vector<void*> allocs;
};

// Workspace of th_render_wavefront_static: the same structure-of-arrays buffers as a
// scene's wavefront workspace, for at most WF_STATIC_SLOTS path slots, in one
// file-scope device array that every GPU holds from module load to process end.  Its
// address on the current device comes from hipGetSymbolAddress (th_context_create);
// host code never uses the array's name as a device address.
#define WF_STATIC_SLOTS 65536u
#define WF_SLOT_BYTES (10u * 16u + 16u + 8u + 8u + 4u + 4u + 8u)  // o..c1, st, hit, vis, q0, q1, qs
#define WF_STATIC_BYTES ((size_t)WF_STATIC_SLOTS * WF_SLOT_BYTES + WF_C(WF_CNT_LINES) * sizeof(uint32_t))
static_assert(WF_STATIC_SLOTS % (64u * WF_NSEG) == 0u, "the static queues hold whole segments");
static __device__ __align__(256) unsigned char scratch[WF_STATIC_BYTES];

int32_t th_wavefront_scratch_init(th_context* ctx) {
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr)
    return th_set_error(TH_E_DEVICE, std::string("wavefront scratch address: ") + hipGetErrorString(e));
  ctx->wf_scratch = ptr;
  return TH_OK;
}

// Carves the WfBufs arrays for WF_STATIC_SLOTS slots out of the scratch array at `base`
// (the device address of `scratch`), each array 256-byte aligned.
static WfBufs wf_static_bufs(void* base) {
  const size_t n = WF_STATIC_SLOTS;
  unsigned char* p = (unsigned char*)base;
  auto take = [&](size_t bytes) {
    unsigned char* q = p;
    p += (bytes + 255u) & ~(size_t)255u;
    return (void*)q;
  };
  WfBufs b;
  b.o = (float4*)take(n * 16);
  b.d = (float4*)take(n * 16);
  b.thr = (float4*)take(n * 16);
  b.rad = (float4*)take(n * 16);
  b.acc = (float4*)take(n * 16);
  b.st = (uint4*)take(n * 16);
  b.hit = (uint2*)take(n * 8);
  b.sho = (float4*)take(n * 16);
  b.sh0 = (float4*)take(n * 16);
  b.c0 = (float4*)take(n * 16);
  b.sh1 = (float4*)take(n * 16);
  b.c1 = (float4*)take(n * 16);
  b.vis = (uint32_t*)take(n * 8);
  b.q0 = (uint32_t*)take(n * 4);
  b.q1 = (uint32_t*)take(n * 4);
  b.qs = (uint32_t*)take(n * 8);
  b.cnt = (uint32_t*)take(WF_C(WF_CNT_LINES) * sizeof(uint32_t));
  return b;
}
static_assert(WF_STATIC_BYTES % 256u == 0u, "scratch layout assumes 256-byte multiples");

void th_wavefront_free(th_wavefront* wf) {
  if (!wf) return;
  for (void* p : wf->allocs) (void)hipFree(p);
  delete wf;
}

static int32_t wf_alloc(th_wavefront* wf, void** p, size_t bytes) {
  TH_CHECK(hipMalloc(p, bytes));
  wf->allocs.push_back(*p);
  return TH_OK;
}

// Slots per queue segment for n slots: segment s holds the slots k with
// (k / 64) % WF_NSEG == s, at most this many.
static uint32_t wf_seg_cap(uint32_t n) { return (n + 64u * WF_NSEG - 1u) / (64u * WF_NSEG) * 64u; }

static int32_t wf_reserve(th_scene* s, uint32_t n) {
  if (s->wf && s->wf->capacity >= n) return TH_OK;
  th_wavefront_free(s->wf);
  s->wf = nullptr;
  th_wavefront* wf = new th_wavefront();
  wf->capacity = n;
  WfBufs& b = wf->b;
  size_t f4 = (size_t)n * sizeof(float4);
  const size_t qn = (size_t)WF_NSEG * wf_seg_cap(n);  // queue entries (whole segments)
  int32_t rc;
  if ((rc = wf_alloc(wf, (void**)&b.o, f4)) || (rc = wf_alloc(wf, (void**)&b.d, f4)) ||
      (rc = wf_alloc(wf, (void**)&b.thr, f4)) || (rc = wf_alloc(wf, (void**)&b.rad, f4)) ||
      (rc = wf_alloc(wf, (void**)&b.acc, f4)) || (rc = wf_alloc(wf, (void**)&b.st, (size_t)n * sizeof(uint4))) ||
      (rc = wf_alloc(wf, (void**)&b.hit, (size_t)n * sizeof(uint2))) ||
      (rc = wf_alloc(wf, (void**)&b.sho, f4)) || (rc = wf_alloc(wf, (void**)&b.sh0, f4)) ||
      (rc = wf_alloc(wf, (void**)&b.c0, f4)) || (rc = wf_alloc(wf, (void**)&b.sh1, f4)) ||
      (rc = wf_alloc(wf, (void**)&b.c1, f4)) ||
      (rc = wf_alloc(wf, (void**)&b.vis, 2 * (size_t)n * sizeof(uint32_t))) ||
      (rc = wf_alloc(wf, (void**)&b.q0, qn * sizeof(uint32_t))) ||
      (rc = wf_alloc(wf, (void**)&b.q1, qn * sizeof(uint32_t))) ||
      (rc = wf_alloc(wf, (void**)&b.qs, 2 * qn * sizeof(uint32_t))) ||
      (rc = wf_alloc(wf, (void**)&b.cnt, WF_C(WF_CNT_LINES) * sizeof(uint32_t)))) {
    th_wavefront_free(wf);
    return rc;
  }
  s->wf = wf;
  return TH_OK;
}

// Sample groups per pixel of a wavefront render (per-sample streams: up to 16, fewer
// when the image would need more than 2^24 slots) and the slot count.
static int32_t wf_shape(const TH_GpuRenderParams* p, const th_render_options* o, uint32_t* G_out,
                        uint32_t* n_out) {
  uint32_t npix = p->width * p->height;
  uint32_t G = 1;
  if (o->rng_mode == TH_RNG_PER_SAMPLE) {
    G = p->samples >= 16 ? 16 : p->samples >= 8 ? 8 : p->samples >= 4 ? 4 : p->samples >= 2 ? 2 : 1;
    while (G > 1 && (uint64_t)npix * G > (16u << 20)) G >>= 1;  // bound the workspace
  }
  uint64_t n64 = (uint64_t)npix * G;
  if (n64 > 0x7fffffffull) return th_set_error(TH_E_UNSUPPORTED, "image too large for the wavefront");
  *G_out = G;
  *n_out = (uint32_t)n64;
  return TH_OK;
}

static int32_t wf_launch(th_context* ctx, th_scene* s, const TH_GpuRenderParams* p, const TH_GpuCamera* cam,
                         const th_render_options* o, const WfBufs& B, uint32_t n, uint32_t G);

int32_t th_wavefront_render(th_context* ctx, th_scene* s, const TH_GpuRenderParams* p,
                            const TH_GpuCamera* cam, const th_render_options* o) {
  uint32_t G, n;
  int32_t rc = wf_shape(p, o, &G, &n);
  if (rc) return rc;
  if ((rc = wf_reserve(s, n))) return rc;
  return wf_launch(ctx, s, p, cam, o, s->wf->b, n, G);
}

extern "C" int32_t th_render_wavefront_static(th_context* ctx, th_scene* s, const TH_GpuRenderParams* p,
                                              const TH_GpuCamera* cam, const th_render_options* opt) {
  if (!ctx || !s || !cam) return th_set_error(TH_E_INVALID, "null argument");
  if (s->ctx != ctx) return th_set_error(TH_E_INVALID, "scene belongs to another context");
  int32_t rc = th_check_render_params(s, p);
  if (rc) return rc;
  th_render_options o;
  memset(&o, 0, sizeof o);
  if (opt) o = *opt;
  if (o.rng_mode > 1u || o.traversal > 2u) return th_set_error(TH_E_INVALID, "unknown render option");
  if (o.collect_stats) return th_set_error(TH_E_UNSUPPORTED, "statistics are counted by the megakernel");
  if (ctx->wf_scratch == nullptr) return th_set_error(TH_E_DEVICE, "no wavefront scratch on this context");
  if (p->max_bounces == 0u) return th_set_error(TH_E_UNSUPPORTED, "zero-bounce renders use th_render");
  uint32_t G, n;
  if ((rc = wf_shape(p, &o, &G, &n))) return rc;
  if (n > WF_STATIC_SLOTS)
    return th_set_error(TH_E_CAPACITY, "the static wavefront workspace holds 65536 path slots "
                                       "(width * height, times the sample groups with per-sample streams)");
  TH_CHECK(hipSetDevice(ctx->device));
  return wf_launch(ctx, s, p, cam, &o, wf_static_bufs(ctx->wf_scratch), n, G);
}

static int32_t wf_launch(th_context* ctx, th_scene* s, const TH_GpuRenderParams* p, const TH_GpuCamera* cam,
                         const th_render_options* o, const WfBufs& Bq, uint32_t n, uint32_t G) {
  uint32_t npix = p->width * p->height;
  WfBufs B = Bq;
  B.seg_cap = wf_seg_cap(n);
  WfArgs A{n, G, o->rng_mode == TH_RNG_PER_SAMPLE ? 1u : 0u};
  DevScene S = dev_scene(s);
  const bool quad = o->traversal == TH_TRAVERSE_WIDE && s->node_count > 0;
  const uint32_t block = 64;
  uint32_t depth = quad ? 3u * s->depth4 + 2u : std::max<uint32_t>(s->max_depth + 1u, 2u);
  size_t lds = (size_t)depth * (quad ? block / 4 : block) * sizeof(uint2);
  // grids grid-stride over the queues, each a multiple of WF_NSEG (block b serves segment
  // b % WF_NSEG).  The closest-hit grid holds one wave per queue round of the first bounce,
  // so the dispatcher balances the waves' unequal costs (terrain 1280x720, 16 spp:
  // 45.8 -> 39.1 ms against 8192 persistent waves); the others stay persistent (shadow
  // grids of 32,768 waves or one per 16 rays measured within noise or 1 % slower).
  auto segs = [](uint32_t g) { return (g + WF_NSEG - 1u) / WF_NSEG * WF_NSEG; };
  uint32_t grid_trace = segs((n * (quad ? 4u : 1u) + block - 1) / block);
  uint32_t grid_scalar = segs(std::min<uint32_t>((n + block - 1) / block, 4096u));
  uint32_t grid_shadow = segs(std::min<uint32_t>((2u * n * (quad ? 4u : 1u) + block - 1) / block, 8192u));
  uint32_t per_sample_count = A.per_sample ? (p->samples + G - 1) / G : p->samples;
  uint32_t iters = per_sample_count * std::max<uint32_t>(p->max_bounces, 1u);
  hipLaunchKernelGGL(wf_reset, dim3(1), dim3(64), 0, ctx->stream, B);
  hipLaunchKernelGGL(wf_generate, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, B, A, *p, *cam);
  for (uint32_t itn = 0; itn < iters; ++itn) {
    uint32_t li = itn & 1u;
    if (quad)
      hipLaunchKernelGGL(wf_extend<3>, dim3(grid_trace), dim3(block), lds, ctx->stream, B, S, *p, li, 1u);
    else if (o->traversal == TH_TRAVERSE_REFERENCE)
      hipLaunchKernelGGL(wf_extend<0>, dim3(grid_trace), dim3(block), lds, ctx->stream, B, S, *p, li, 0u);
    else
      hipLaunchKernelGGL(wf_extend<1>, dim3(grid_trace), dim3(block), lds, ctx->stream, B, S, *p, li, 0u);
    hipLaunchKernelGGL(wf_shade, dim3(grid_scalar), dim3(block), 0, ctx->stream, B, S, *p,
                       (const float*)ctx->d_srgb_lut, li, quad ? 1u : 0u);
    if (quad)
      hipLaunchKernelGGL(wf_shadow<3>, dim3(grid_shadow), dim3(block), lds, ctx->stream, B, S, *p, 1u);
    else if (o->traversal == TH_TRAVERSE_REFERENCE)
      hipLaunchKernelGGL(wf_shadow<0>, dim3(grid_shadow), dim3(block), lds, ctx->stream, B, S, *p, 0u);
    else
      hipLaunchKernelGGL(wf_shadow<1>, dim3(grid_shadow), dim3(block), lds, ctx->stream, B, S, *p, 0u);
    hipLaunchKernelGGL(wf_finish, dim3(grid_scalar), dim3(block), 0, ctx->stream, B, A, *p, *cam, li);
  }
  hipLaunchKernelGGL(wf_resolve, dim3((npix + 255) / 256), dim3(256), 0, ctx->stream, B, A, *p,
                     s->d_accum);
  TH_CHECK(hipGetLastError());
  return TH_OK;
}
