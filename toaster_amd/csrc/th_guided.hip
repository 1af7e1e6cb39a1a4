// th_guided.hip -- path tracing with the learned guiding distribution.
//
// A wavefront over one sample of every pixel per pass: path k of pass s is pixel k's
// sample s, seeded as th_render's per-sample mode.  Per bounce:
//   extend   closest hit, 4 lanes per ray over the 4-wide BVH;
//   shade A  the reference loop body (shaders/pathtrace_spheres.wgsl:817-911) for
//            misses, emitters, metals and dielectrics; a diffuse vertex is parked
//            (position, normal, albedo) and appended to the guide queue;
//   infer    the fused guide MLP on MFMA over the queue (th_guide.hip);
//   shade B  diffuse vertices: direct light and environment terms with deferred
//            visibility, then the continuation by one-sample MIS between the guide
//            (probability alpha) and the reference's cosine lobe;
//   shadow   any hit, 4 lanes per ray.
// The shade kernels compact continuing paths straight into the next live queue and
// finished ones into a done queue; a path's deferred direct terms are added (in the
// reference's order, light then environment) when the next shade A picks it up, or by
// the retire kernel that closes the pass for finished paths.  After each pass the
// training records of the pass receive their weights and the guide takes its Adam
// steps on them.
//
// Estimator.  The continuation density is p(w) = alpha q(w) + (1 - alpha) cos/pi, the
// throughput takes albedo (cos/pi) / p(w), and both the environment term's MIS weight
// and an escaping ray's MIS weight use p: unbiased for any network, and with alpha = 0
// every operation and random draw is the reference's (the image equals th_render's
// per-sample image with one sample group).
//
// Training target.  A record at vertex v stores the features, the sampled direction w,
// the throughput beta_v before the bounce and (set by finish) the path radiance L_v
// after v's direct terms.  When the path completes with radiance L, the record's weight
// is lum((L - L_v) / beta_v) = f cos L_i(w) / p(w): the Monte Carlo weight of the product
// target f cos L_i under the sampling density, which the loss -sum w log q / sum w fits.
#include <algorithm>
#include <vector>

#include "th_guide.h"
#include "th_internal.h"

namespace {

#define GD_INCLUDE_EMISSIVE 1u
#define GD_PREV_DIFFUSE 2u
#define GD_DONE 4u
#define GD_PEND_LIGHT 8u
#define GD_PEND_ENV 16u
#define GD_NO_RECORD 0xffffffffu

struct GdBufs {
  float4 *o, *d, *thr, *rad, *acc;  // o.w previous continuation pdf, d.w flags, thr.w record
  uint32_t* rng;
  uint2* hit;
  float4 *vp, *vn, *vw;  // guide-queue row: (p, albedo.x), (n, albedo.y), (wo, albedo.z)
  float4 *sho, *sh0, *c0, *sh1, *c1;
  uint32_t* vis;
  uint32_t *q0, *q1, *qs, *qg;  // queues in GD_NSEG segments (see GD_QC)
  uint32_t* cnt;  // GD_C: [0] paths at the first bounce, [5] diffuse total, [6] rows the
                  // network evaluated, [8 + c] records of chunk c; GD_QC: queue segments
  uint32_t seg_cap;  // paths per segment of q0, q1, qg (and its rows); qs holds twice that
  uint16_t* O;    // guide outputs, TG_OUT binary16 values per guide-queue row
  uint16_t* rX;   // training records: BF16 features, 64 wide (zero beyond 48)
  float *rdir, *rw;
  uint16_t* tX;   // the records of each chunk made contiguous for training (rows of rX,
  float* tdir;    // rdir; rw is written at the same contiguous rows)
  float4 *rbeta, *rL;
  uint32_t* rpath;
  uint32_t rcap;  // chunks x chunk capacity
};

#define GD_MAX_CHUNKS 16u
// scalar counter lines: 0-3 queues, 5 diffuse vertices, 6 network rows, 8..23 chunk totals,
// then the record-rate state and the training statistics
#define GD_DPP_PREV 24u    // diffuse-vertex count at the last training pass's end
#define GD_DPP 26u         // diffuse vertices of the last training pass (persists across renders)
#define GD_STATS 27u       // three words: samples offered, samples trained, steps applied
#define GD_REC_ALL 30u     // this pass records every eligible vertex
#define GD_REC_THRESH 31u  // ... or those whose hash is below this
#define GD_C(c) ((c) * 32u)  // queue counters 128 bytes apart: no two share a cache line
// Queues in segments.  A wave that appends to a queue takes its slots with one atomic on
// a counter, and with one counter for the whole queue those same-address atomics
// serialise in the memory system (the first bounce's shading launch on a 1280x720 frame
// took 173 us with them, 24 us with the slots taken without atomics).  Each
// queue is therefore split into GD_NSEG segments with a counter each: block b appends to
// and consumes from segment b % GD_NSEG (grids are multiples of GD_NSEG), so a path stays
// in one segment for the whole pass, a segment never holds more than seg_cap paths, and
// its state is read back by the same XCD (blocks are dealt to the 8 XCDs round-robin).
// Queue q: 0 and 1 the live queues, 2 the shadow-ray queue, 3 the guide queue.
#ifndef GD_NSEG
#define GD_NSEG 16u
#endif
#define GD_QC(q, sg) GD_C(32u + (q) * GD_NSEG + (sg))
// Training records of chunk c from segment sg: their own counter (records of one chunk
// from every segment on one counter serialise the same way).
#define GD_RC(c, sg) GD_C(32u + 4u * GD_NSEG + (c) * GD_NSEG + (sg))
#define GD_CNT_LINES (32u + 4u * GD_NSEG + GD_MAX_CHUNKS * GD_NSEG)

struct GdArgs {
  uint32_t n;           // paths per pass (pixels)
  uint32_t smp;         // sample index of the pass
  uint32_t bounce;      // bounce of this iteration, 0-based
  float alpha;          // guide probability (0: cosine lobe only)
  uint32_t defer;       // diffuse vertices go through the guide queue
  uint32_t record;      // record training samples
  uint32_t env_nee;     // environment sample + shadow ray at diffuse vertices
  uint32_t nchunks;     // records are spread uniformly at random over this many chunks
  uint32_t chunk_cap;   // capacity of one chunk (chunk c holds rows c * chunk_cap ...), in
                        // GD_NSEG pieces of chunk_cap / GD_NSEG, one per queue segment
  float bmin[3], bscale[3];
};

// Appends `live` lanes' values to a queue: one atomic per wave, order kept within a wave.
// Returns the slot of this lane (valid when live).
__device__ __forceinline__ uint32_t gd_push(bool live, uint32_t value, uint32_t* queue, uint32_t* count,
                                            uint32_t cap = 0xffffffffu) {
  uint64_t m = __ballot(live);
  if (m == 0) return 0;
  uint32_t lane = threadIdx.x & 63u;
  uint32_t first = (uint32_t)__ffsll((unsigned long long)m) - 1u;
  uint32_t base = 0;
  if (lane == first) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = __shfl(base, (int)first, 64);
  uint32_t idx = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
  if (live && queue && idx < cap) queue[idx] = value;
  return idx;
}

// Two pushes onto one queue with one atomic: `la` lanes' values a, then `lb` lanes'
// values b.
__device__ __forceinline__ void gd_push2(bool la, uint32_t a, bool lb, uint32_t b, uint32_t* queue,
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

// Counters of a pass (and, with `all`, of the render: everything but the record-rate
// estimate GD_DPP, which persists with the workspace).  Thread 0 then sets the pass's
// record rate: a vertex is recorded with probability target / (the last training pass's
// diffuse vertices), `init_dpp` standing in before any pass has counted them.
__global__ void gd_reset(GdBufs B, uint32_t all, uint32_t target, uint32_t init_dpp) {
  const uint32_t t = threadIdx.x;
  if (t < 4u || (t >= 8u && t < 8u + GD_MAX_CHUNKS) || (all && t < 32u && t != GD_DPP)) B.cnt[GD_C(t)] = 0u;
  if (all && t < 3u) B.cnt[GD_C(GD_STATS) + t] = 0u;
  for (uint32_t c = t; c < (4u + GD_MAX_CHUNKS) * GD_NSEG; c += blockDim.x) B.cnt[GD_C(32u + c)] = 0u;
  if (t == 0u) {
    uint32_t dpp = B.cnt[GD_C(GD_DPP)];
    if (dpp == 0u) B.cnt[GD_C(GD_DPP)] = dpp = init_dpp;
    const double pr = (double)target / (double)(dpp > 0u ? dpp : 1u);
    B.cnt[GD_C(GD_REC_ALL)] = pr >= 1.0 ? 1u : 0u;
    B.cnt[GD_C(GD_REC_THRESH)] = (uint32_t)fmin(pr * 4294967296.0, 4294967295.0);
  }
}

// ---------------------------------------------------------------- generate
__global__ void __launch_bounds__(256) gd_generate(GdBufs B, GdArgs A, TH_GpuRenderParams P,
                                                   TH_GpuCamera cam) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < A.n) {
    if (P.max_bounces == 0u) {  // ray_color returns zero (wgsl:536)
      float4 a = A.smp == 0u ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : B.acc[k];
      a.w = a.w + 1.0f;
      B.acc[k] = a;
    } else {
      if (A.smp == 0u) B.acc[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      uint32_t rng = th::pcg_hash((k ^ (P.frame_index * 9781u)) + A.smp * 0x9E3779B9u);
      v3 o, d;
      camera_ray(cam, P, k % P.width, k / P.width, &rng, &o, &d);
      B.o[k] = make_float4(o.x, o.y, o.z, 0.0f);
      B.d[k] = make_float4(d.x, d.y, d.z, __uint_as_float(GD_INCLUDE_EMISSIVE));
      B.rng[k] = rng;
      // throughput (1, 1, 1) with no record and zero radiance: the first bounce's shading
      // starts from these constants instead of reading them back
    }
  }
  // every path is live at the first bounce: that bounce's queue is the identity (path i
  // at position i), so only its length is stored
  if (k == 0u) B.cnt[GD_C(0)] = P.max_bounces == 0u ? 0u : A.n;
}

// ---------------------------------------------------------------- extend / shadow
// `first`: the first bounce, whose queue is the identity (every path, in path order).
// Later bounces: block b takes rays of live-queue segment b % GD_NSEG, 16 per wave.
__global__ void __launch_bounds__(64) gd_extend(GdBufs B, DevScene S, TH_GpuRenderParams P, uint32_t li,
                                                uint32_t first) {
  extern __shared__ uint2 lds_stack[];
  const uint32_t sg = blockIdx.x % GD_NSEG;
  const uint32_t* q = (li ? B.q1 : B.q0) + (size_t)sg * B.seg_cap;
  const uint32_t n = first ? B.cnt[GD_C(0)] : B.cnt[GD_QC(li, sg)];
  if (blockIdx.x == 0)  // queues filled later in this iteration
    for (uint32_t c = threadIdx.x; c < GD_NSEG; c += blockDim.x) {
      B.cnt[GD_QC(2u, c)] = 0;
      B.cnt[GD_QC(3u, c)] = 0;
      B.cnt[GD_QC(li ^ 1u, c)] = 0;
    }
  Stack st{lds_stack + (threadIdx.x >> 2), (int)(blockDim.x >> 2), 0};
  const uint32_t rays = blockDim.x / 4u;
  const uint32_t start = first ? (blockIdx.x * blockDim.x + threadIdx.x) / 4u
                               : (blockIdx.x / GD_NSEG) * rays + threadIdx.x / 4u;
  const uint32_t stride = first ? gridDim.x * rays : (gridDim.x / GD_NSEG) * rays;
  for (uint32_t i = start;; i += stride) {
    bool act = i < n;  // uniform per quad
    if (!__any(act)) break;
    if (act) {
      uint32_t k = first ? i : q[i];
      float4 o4 = B.o[k], d4 = B.d[k];
      v3 o = th::mk(o4.x, o4.y, o4.z), d = th::mk(d4.x, d4.y, d4.z);
      float t;
      uint32_t slot;
      if (S.node_count == 0u) {
        RayStats<false> rs;
        t = closest_linear<false>(S, P.sphere_count, P.triangle_count, o, d, TH_FLT_MAX, rs, &slot);
      } else {
        traverse_quad<1>(S, (const Child4*)S.nodes4, S.root4, o, d, TH_FLT_MAX, st, &t, &slot);
      }
      if ((threadIdx.x & 3u) == 0u) B.hit[k] = make_uint2(__float_as_uint(t), slot);
    }
  }
}

// Shadow rays of shadow-queue segment b % GD_NSEG, 16 per wave.
__global__ void __launch_bounds__(64) gd_shadow(GdBufs B, DevScene S, TH_GpuRenderParams P) {
  extern __shared__ uint2 lds_stack[];
  const uint32_t sg = blockIdx.x % GD_NSEG;
  const uint32_t* qs = B.qs + (size_t)sg * 2u * B.seg_cap;
  const uint32_t n = B.cnt[GD_QC(2u, sg)];
  Stack st{lds_stack + (threadIdx.x >> 2), (int)(blockDim.x >> 2), 0};
  const uint32_t rays = blockDim.x / 4u;
  const uint32_t stride = (gridDim.x / GD_NSEG) * rays;
  for (uint32_t i = (blockIdx.x / GD_NSEG) * rays + threadIdx.x / 4u;; i += stride) {
    bool act = i < n;
    if (!__any(act)) break;
    if (act) {
      uint32_t e = qs[i], k = e >> 1;
      float4 so = B.sho[k], sd = (e & 1u) ? B.sh1[k] : B.sh0[k];
      v3 o = th::mk(so.x, so.y, so.z), d = th::mk(sd.x, sd.y, sd.z);
      float limit = sd.w, t;
      uint32_t slot;
      if (S.node_count == 0u) {
        RayStats<false> rs;
        t = closest_linear<false>(S, P.sphere_count, P.triangle_count, o, d, limit, rs, &slot);
      } else {
        traverse_quad<2>(S, (const Child4*)S.nodes4, S.root4, o, d, limit, st, &t, &slot);
      }
      if ((threadIdx.x & 3u) == 0u) B.vis[e] = (t < limit) ? 0u : 1u;
    }
  }
}

// ---------------------------------------------------------------- diffuse vertex
// Direct light and environment terms of a diffuse vertex (wgsl:846-847) with their
// visibility deferred, then the continuation.  Without a mixture (or alpha = 0) the
// continuation is the reference's cosine lobe (wgsl:849-857), bit for bit; otherwise
// one-sample MIS between the guide and the cosine lobe.  Returns false when the
// continuation has zero contribution (a guide direction below the surface).
//
// The last vertex of a path (`last`: no continuation will be traced, wgsl:812) keeps
// the reference's environment sample with the reference's weight (power heuristic
// against the cosine lobe, wgsl:496-503) whatever alpha and env_nee are: the reference
// estimator credits that share of the next bounce, so every variant here estimates the
// same quantity as the reference (without it, an estimator that drops the environment
// sample, or weighs it against the mixture, converges to a slightly different image).
// MIX: `m` holds the vertex's decoded mixture (a compile-time flag rather than a null
// test: comparing the address of a private variable with null keeps it out of registers).
template <bool MIX>
__device__ __forceinline__ bool shade_diffuse(const GdBufs& B, Integrator<1, false>& it,
                                              const TH_GpuRenderParams& P, uint32_t k, const v3& p,
                                              const v3& n, const v3& albedo, const tg::Mixture* m,
                                              float alpha, bool env_nee, bool last, v3& throughput,
                                              uint32_t& rng, v3& nd, float& pdf, bool& sh_light,
                                              bool& sh_env) {
  const bool guided = MIX && alpha > 0.0f;
  {
    v3 so, dir, term;
    float limit;
    // a shadow ray an extracted primitive blocks adds nothing: resolved here, not queued
    if (it.light_candidate(p, n, albedo, &rng, &so, &dir, &limit, &term) &&
        !lane_extracted_anyhit(it.S, so, dir, limit)) {
      v3 c = th::mul(throughput, term);
      B.sho[k] = make_float4(so.x, so.y, so.z, 0.0f);
      B.sh0[k] = make_float4(dir.x, dir.y, dir.z, limit);
      B.c0[k] = make_float4(c.x, c.y, c.z, 0.0f);
      sh_light = true;
    }
  }
  if (P.background_kind == 2u && (env_nee || last)) {
    EnvSample es;
    float sc;
    if (it.env_candidate(n, &rng, &es, &sc)) {
      float bp = sc / TH_PI;
      if (guided && !last) bp = alpha * tg::mixture_pdf(*m, es.dir) + (1.0f - alpha) * bp;
      v3 c = th::mul(throughput, it.env_term(albedo, es, sc, it.power_heuristic(es.pdf, bp)));
      v3 so = th::add(p, th::muls(n, TH_MIN_DISTANCE));
      if (!lane_extracted_anyhit(it.S, so, es.dir, 10000.001f)) {
        B.sho[k] = make_float4(so.x, so.y, so.z, 0.0f);
        B.sh1[k] = make_float4(es.dir.x, es.dir.y, es.dir.z, 10000.001f);
        B.c1[k] = make_float4(c.x, c.y, c.z, 0.0f);
        sh_env = true;
      }
    }
  }
  if (!guided) {
    v3 dir = th::add(n, it.random_unit_vector(&rng));
    if (th::dot(dir, dir) < TH_EPSILON) dir = n;
    nd = th::normalize(dir);
    throughput = th::mul(throughput, albedo);
    pdf = th::fmax_(0.0f, th::dot(n, nd)) / TH_PI;
    return true;
  }
  if (th::rand01(&rng) < alpha) {
    float u0 = th::rand01(&rng);
    float u1 = th::rand01(&rng);
    float u2 = th::rand01(&rng);
    nd = tg::mixture_sample(*m, u0, u1, u2);
  } else {
    v3 dir = th::add(n, it.random_unit_vector(&rng));
    if (th::dot(dir, dir) < TH_EPSILON) dir = n;
    nd = th::normalize(dir);
  }
  float c = th::dot(n, nd);
  if (!(c > 0.0f)) {
    pdf = 0.0f;
    return false;
  }
  float cp = c / TH_PI;
  pdf = alpha * tg::mixture_pdf(*m, nd) + (1.0f - alpha) * cp;
  throughput = th::mul(throughput, th::muls(albedo, cp / pdf));
  return true;
}

// Adds the deferred direct terms of the previous vertex (light first, then environment,
// wgsl:846-847) and stores the radiance a pending training record waits for.
__device__ __forceinline__ void apply_pending(const GdBufs& B, uint32_t k, uint32_t flags, float4& r4,
                                              float4& t4) {
  if (flags & (GD_PEND_LIGHT | GD_PEND_ENV)) {
    if ((flags & GD_PEND_LIGHT) && B.vis[2u * k]) {
      const float4 c = B.c0[k];
      r4 = make_float4(r4.x + c.x, r4.y + c.y, r4.z + c.z, 0.0f);
    }
    if ((flags & GD_PEND_ENV) && B.vis[2u * k + 1u]) {
      const float4 c = B.c1[k];
      r4 = make_float4(r4.x + c.x, r4.y + c.y, r4.z + c.z, 0.0f);
    }
  }
  const uint32_t rec = __float_as_uint(t4.w);
  if (rec != GD_NO_RECORD) {
    B.rL[rec] = r4;
    t4.w = __uint_as_float(GD_NO_RECORD);
  }
}

// ---------------------------------------------------------------- shade A
__global__ void __launch_bounds__(64) gd_shade_a(GdBufs B, GdArgs A, DevScene S, TH_GpuRenderParams P,
                                                 const float* __restrict__ srgb_lut, uint32_t li) {
  const bool first = A.bounce == 0u;  // identity queue
  // block b reads and appends to segment b % GD_NSEG; on the identity queue its paths k
  // are those with (k / 64) % GD_NSEG == b % GD_NSEG (the grid is a multiple of GD_NSEG)
  const uint32_t sg = blockIdx.x % GD_NSEG;
  const uint32_t* q = (li ? B.q1 : B.q0) + (size_t)sg * B.seg_cap;
  const uint32_t n = first ? B.cnt[GD_C(0)] : B.cnt[GD_QC(li, sg)];
  uint32_t* qlive = (li ? B.q0 : B.q1) + (size_t)sg * B.seg_cap;
  uint32_t* qsh = B.qs + (size_t)sg * 2u * B.seg_cap;
  const uint32_t gbase = sg * B.seg_cap;  // this segment's guide-queue rows
  Stack st{nullptr, 0, 0};
  LaneCounters<false> lc;
  Integrator<1, false> it{S, P, srgb_lut, st, lc};
  const uint32_t start = first ? blockIdx.x * blockDim.x + threadIdx.x
                               : (blockIdx.x / GD_NSEG) * blockDim.x + threadIdx.x;
  const uint32_t stride = first ? gridDim.x * blockDim.x : (gridDim.x / GD_NSEG) * blockDim.x;
  for (uint32_t i = start;; i += stride) {
    bool act = i < n;
    if (!__any(act)) break;
    bool sh_light = false, sh_env = false, park = false, live = false;
    uint32_t k = 0;
    float4 pk_p, pk_n, pk_w;  // the parked vertex
    if (act) {
      k = first ? i : q[i];
      float4 o4 = B.o[k], d4 = B.d[k];
      float4 t4 = first ? make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(GD_NO_RECORD)) : B.thr[k];
      float4 r4 = first ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : B.rad[k];
      apply_pending(B, k, __float_as_uint(d4.w), r4, t4);
      uint32_t rng = B.rng[k];
      uint2 h = B.hit[k];
      v3 d = th::mk(d4.x, d4.y, d4.z);
      v3 throughput = th::mk(t4.x, t4.y, t4.z), radiance = th::mk(r4.x, r4.y, r4.z);
      uint32_t flags = __float_as_uint(d4.w) & (GD_INCLUDE_EMISSIVE | GD_PREV_DIFFUSE);
      float prev_pdf = o4.w;
      float t = __uint_as_float(h.x);
      uint32_t slot = h.y;
      bool done = false;
      v3 nd = d;
      float4 o_next = o4;
      if (slot == THD_NO_SLOT || t > TH_MAX_DISTANCE) {  // miss (wgsl:819-834)
        if (P.background_kind == 0u) {
          float lt = (th::normalize(d).y + 1.0f) * 0.5f;
          radiance = th::add(radiance, th::mul(throughput, th::mix(th::mk(1.0f, 1.0f, 1.0f),
                                                                    th::mk(0.35f, 0.65f, 1.0f), lt)));
        } else if (P.background_kind == 2u) {
          float mis = 1.0f;  // escaping after a diffuse vertex: MIS with its environment sample
          if ((flags & GD_PREV_DIFFUSE) && A.env_nee) mis = it.power_heuristic(prev_pdf, it.env_pdf(d));
          radiance = th::add(radiance, th::muls(th::mul(throughput, it.env_lookup(d)), mis));
        }
        done = true;
      } else {
        v3 o = th::mk(o4.x, o4.y, o4.z);
        Hit rec;
        if (S.node_count == 0u)
          finalize_linear(S, slot, o, d, t, &rec);
        else
          finalize_slot(S, slot, o, d, t, &rec);
        const TH_GpuMaterial m = load_material(S, rec.mat);
        v3 att = th::mk(0, 0, 0);
        bool brk = false;
        if (m.kind == 0u) {
          v3 albedo = it.base_color(m, rec);
          if (A.defer) {
            pk_p = make_float4(rec.p.x, rec.p.y, rec.p.z, albedo.x);
            pk_n = make_float4(rec.n.x, rec.n.y, rec.n.z, albedo.y);
            pk_w = make_float4(-d.x, -d.y, -d.z, albedo.z);
            park = true;
          } else {
            float pdf;
            shade_diffuse<false>(B, it, P, k, rec.p, rec.n, albedo, nullptr, 0.0f, A.env_nee != 0u,
                          A.bounce + 1u >= P.max_bounces, throughput, rng, nd, pdf, sh_light, sh_env);
            flags = GD_PREV_DIFFUSE;
            o_next = make_float4(rec.p.x, rec.p.y, rec.p.z, pdf);
            if (A.bounce + 1u >= P.max_bounces) done = true;
          }
        } else {
          if (m.kind == 1u) {
            v3 refl = it.reflect_vec(d, rec.n);
            float rough = th::clampf(m.params[0], 0.0f, 1.0f);
            v3 dir = th::add(refl, th::smul(rough, it.random_unit_vector(&rng)));
            if (th::dot(dir, rec.n) <= 0.0f) brk = true;
            nd = th::normalize(dir);
            att = th::mk(m.albedo[0], m.albedo[1], m.albedo[2]);
            flags = GD_INCLUDE_EMISSIVE;
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
            flags = GD_INCLUDE_EMISSIVE;
          } else {
            if (m.kind == 3u && (flags & GD_INCLUDE_EMISSIVE))
              radiance = th::add(radiance, th::muls(th::mul(throughput, th::mk(m.albedo[0], m.albedo[1],
                                                                                m.albedo[2])),
                                                    m.params[2]));
            brk = true;
          }
          if (brk) {
            done = true;
          } else {
            throughput = th::mul(throughput, att);
            o_next = make_float4(rec.p.x, rec.p.y, rec.p.z, prev_pdf);
            if (A.bounce + 1u >= P.max_bounces) done = true;
          }
        }
      }
      if (!park) {
        uint32_t f = flags | (done ? GD_DONE : 0u) | (sh_light ? GD_PEND_LIGHT : 0u) |
                     (sh_env ? GD_PEND_ENV : 0u);
        B.o[k] = o_next;
        B.d[k] = make_float4(nd.x, nd.y, nd.z, __uint_as_float(f));
        B.thr[k] = make_float4(throughput.x, throughput.y, throughput.z, t4.w);
        B.rng[k] = rng;
        live = !done;
      } else if (first) {
        B.thr[k] = t4;  // the parked path's throughput, read by the second shading kernel
      }
      B.rad[k] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
    }
    gd_push2(sh_light, 2u * k, sh_env, 2u * k + 1u, qsh, &B.cnt[GD_QC(2u, sg)]);
    gd_push(live, k, qlive, &B.cnt[GD_QC(li ^ 1u, sg)]);
    const uint32_t gi = gbase + gd_push(park, k, B.qg + gbase, &B.cnt[GD_QC(3u, sg)]);
    if (park) {  // vertex data in queue order: contiguous rows for inference and shade B
      B.vp[gi] = pk_p;
      B.vn[gi] = pk_n;
      B.vw[gi] = pk_w;
    }
  }
}

// ---------------------------------------------------------------- shade B
__global__ void __launch_bounds__(64) gd_shade_b(GdBufs B, GdArgs A, DevScene S, TH_GpuRenderParams P,
                                                 const float* __restrict__ srgb_lut, uint32_t li) {
  // block b takes guide-queue segment b % GD_NSEG and appends to that segment's queues
  const uint32_t sg = blockIdx.x % GD_NSEG, jb = blockIdx.x / GD_NSEG;
  const uint32_t n = B.cnt[GD_QC(3u, sg)];
  const uint32_t gbase = sg * B.seg_cap;
  uint32_t* qlive = (li ? B.q0 : B.q1) + (size_t)sg * B.seg_cap;
  uint32_t* qsh = B.qs + (size_t)sg * 2u * B.seg_cap;
  if (jb == 0 && threadIdx.x == 0) {
    atomicAdd(&B.cnt[GD_C(5)], n);
    if (A.alpha > 0.0f) atomicAdd(&B.cnt[GD_C(6)], n);  // rows the network evaluated
  }
  Stack st{nullptr, 0, 0};
  LaneCounters<false> lc;
  Integrator<1, false> it{S, P, srgb_lut, st, lc};
  const bool use_guide = A.alpha > 0.0f;
  const uint32_t stride = (gridDim.x / GD_NSEG) * blockDim.x;
  for (uint32_t t = jb * blockDim.x + threadIdx.x;; t += stride) {
    bool act = t < n;
    if (!__any(act)) break;
    const uint32_t i = gbase + t;  // guide-queue row
    bool sh_light = false, sh_env = false, rec_ok = false, live = false;
    uint32_t k = 0;
    v3 p, nrm, wo, nd, beta;
    if (act) {
      k = B.qg[i];
      const float4 gp = B.vp[i], gn = B.vn[i], gw = B.vw[i], t4 = B.thr[k];
      p = th::mk(gp.x, gp.y, gp.z);
      nrm = th::mk(gn.x, gn.y, gn.z);
      wo = th::mk(gw.x, gw.y, gw.z);
      v3 albedo = th::mk(gp.w, gn.w, gw.w);
      v3 throughput = th::mk(t4.x, t4.y, t4.z);
      beta = throughput;
      uint32_t rng = B.rng[k];
      tg::Mixture mix;
      if (use_guide) {
        float o[TG_OUT];
        const uint4* o8 = (const uint4*)(B.O + (size_t)i * TG_OUT);  // 80-byte rows, 8 values per load
#pragma unroll
        for (int j = 0; j < TG_OUT / 8; ++j) {
          const uint4 v = o8[j];
          const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            o[8 * j + 2 * q] = tg::h2f((uint16_t)(w[q] & 0xffffu));
            o[8 * j + 2 * q + 1] = tg::h2f((uint16_t)(w[q] >> 16));
          }
        }
        tg::decode(o, &mix);
      }
      float pdf;
      // alpha = 0 (use_guide false) leaves `mix` unread
      bool cont = shade_diffuse<true>(B, it, P, k, p, nrm, albedo, &mix, A.alpha,
                                A.env_nee != 0u, A.bounce + 1u >= P.max_bounces, throughput, rng, nd, pdf,
                                sh_light, sh_env);
      bool done = !cont || A.bounce + 1u >= P.max_bounces;
      if (A.record && !done) {
        uint32_t h = th::pcg_hash((k * 0x9E3779B9u) ^ (A.bounce * 0x85EBCA6Bu) ^
                                  (A.smp * 0xC2B2AE35u + P.frame_index));
        rec_ok = B.cnt[GD_C(GD_REC_ALL)] != 0u || h < B.cnt[GD_C(GD_REC_THRESH)];
      }
      uint32_t f = GD_PREV_DIFFUSE | (done ? GD_DONE : 0u) | (sh_light ? GD_PEND_LIGHT : 0u) |
                   (sh_env ? GD_PEND_ENV : 0u);
      B.o[k] = make_float4(p.x, p.y, p.z, pdf);
      B.d[k] = make_float4(nd.x, nd.y, nd.z, __uint_as_float(f));
      B.thr[k] = make_float4(throughput.x, throughput.y, throughput.z, __uint_as_float(GD_NO_RECORD));
      B.rng[k] = rng;
      live = !done;
    }
    gd_push2(sh_light, 2u * k, sh_env, 2u * k + 1u, qsh, &B.cnt[GD_QC(2u, sg)]);
    gd_push(live, k, qlive, &B.cnt[GD_QC(li ^ 1u, sg)]);
    // chunk drawn at random so every training step sees an unbiased subset; slots are
    // handed out with one atomic per (wave, chunk), all of a wave's chunks at once: lane c
    // takes chunk c's slots (a chain of atomics, each waiting for the last, had cost a
    // wave one memory round trip per chunk)
    uint32_t r = 0xffffffffu;
    const uint32_t myc =
        rec_ok ? th::pcg_hash(k ^ (A.bounce << 27) ^ (A.smp * 0x2545F491u)) % A.nchunks : 0xffffffffu;
    if (__ballot(rec_ok) != 0) {
      const uint32_t lane = threadIdx.x & 63u;
      uint64_t mymask = 0, lanemask = 0;  // lanes of this lane's chunk / of chunk `lane`
      for (uint32_t c = 0; c < A.nchunks; ++c) {
        const uint64_t m = __ballot(myc == c);
        if (myc == c) mymask = m;
        if (lane == c) lanemask = m;
      }
      uint32_t base = 0;
      if (lanemask != 0) base = atomicAdd(&B.cnt[GD_RC(lane, sg)], (uint32_t)__popcll(lanemask));
      const uint32_t cb = __shfl(base, rec_ok ? (int)myc : (int)lane, 64);
      if (rec_ok) {
        const uint32_t piece = A.chunk_cap / GD_NSEG;  // this segment's piece of the chunk
        const uint32_t idx = cb + (uint32_t)__popcll(mymask & ((1ull << lane) - 1ull));
        if (idx < piece) r = myc * A.chunk_cap + sg * piece + idx;
      }
    }
    if (r != 0xffffffffu) {
      B.rpath[r] = k;
      v3 p01 = th::mk((p.x - A.bmin[0]) * A.bscale[0], (p.y - A.bmin[1]) * A.bscale[1],
                      (p.z - A.bmin[2]) * A.bscale[2]);
      // the features of tg::encode_row, two at a time at compile-time positions (no
      // per-lane array, which would live in scratch memory)
      uint4* x = (uint4*)(B.rX + (size_t)r * 64u);  // 128-byte row, eight 16-byte stores
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        uint32_t h[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int j0 = 8 * q + 2 * t;
          if (j0 < 6 * TG_POS_OCT) {
            float fs, fc;
            tg::encode_pos_pair(j0, p01, &fs, &fc);
            h[t] = tg::f2bf2(fs, fc);
          } else if (j0 < TG_IN) {
            h[t] = tg::f2bf2(tg::encode_feature(j0, p01, nrm, wo), tg::encode_feature(j0 + 1, p01, nrm, wo));
          } else {
            h[t] = 0u;
          }
        }
        x[q] = make_uint4(h[0], h[1], h[2], h[3]);
      }
      B.rdir[3u * r] = nd.x;
      B.rdir[3u * r + 1u] = nd.y;
      B.rdir[3u * r + 2u] = nd.z;
      B.rbeta[r] = make_float4(beta.x, beta.y, beta.z, 0.0f);
      B.thr[k].w = __uint_as_float(r);  // finish stores the radiance after this vertex
    }
  }
}

// ---------------------------------------------------------------- retire
// Paths finished during the pass: their last deferred direct terms, then
// color += vec4(ray_color, 1) (wgsl:536).
// Every path of a pass finishes exactly once (n = 0 when there are no bounces: the
// generate kernel has counted those samples).
__global__ void __launch_bounds__(256) gd_retire(GdBufs B, uint32_t n) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    float4 r4 = B.rad[k], t4 = B.thr[k];
    apply_pending(B, k, __float_as_uint(B.d[k].w), r4, t4);
    B.rad[k] = r4;
    const float4 a = B.acc[k];
    B.acc[k] = make_float4(a.x + r4.x, a.y + r4.y, a.z + r4.z, a.w + 1.0f);
  }
}

// ---------------------------------------------------------------- records, resolve
// Weight of each record of the pass: lum((L - L_v) / beta_v), zero where not finite.
// The chunk's records, in GD_NSEG pieces (one per queue segment), are gathered into
// contiguous rows of tX, tdir and rw (piece 0's first, then piece 1's, ...), and the
// chunk's total is left in GD_C(8 + c) for the training step.  One thread per destination
// row.
__global__ void __launch_bounds__(256) gd_weights(GdBufs B, uint32_t nchunks, uint32_t chunk_cap) {
  __shared__ uint32_t pre[GD_MAX_CHUNKS][GD_NSEG + 1];  // pieces' first rows; [GD_NSEG]: total
  const uint32_t piece = chunk_cap / GD_NSEG;
  if (blockIdx.x == 0 && threadIdx.x == 0) {  // the record-rate estimate for the next passes
    const uint32_t total = B.cnt[GD_C(5)], d = total - B.cnt[GD_C(GD_DPP_PREV)];
    B.cnt[GD_C(GD_DPP_PREV)] = total;
    B.cnt[GD_C(GD_DPP)] = d > 0u ? d : 1u;
  }
  for (uint32_t c = threadIdx.x; c < nchunks; c += blockDim.x) {
    uint32_t acc = 0;
    for (uint32_t j = 0; j < GD_NSEG; ++j) {
      pre[c][j] = acc;
      acc += min(B.cnt[GD_RC(c, j)], piece);
    }
    pre[c][GD_NSEG] = acc;
    if (blockIdx.x == 0) B.cnt[GD_C(8u + c)] = acc;
  }
  __syncthreads();
  const uint32_t n = nchunks * chunk_cap;
  for (uint32_t d = blockIdx.x * blockDim.x + threadIdx.x; d < n; d += gridDim.x * blockDim.x) {
    const uint32_t c = d / chunk_cap, o = d - c * chunk_cap;
    if (o >= pre[c][GD_NSEG]) continue;
    uint32_t sg = 0;  // the piece holding row o: the last whose first row is <= o
    for (uint32_t j = 1; j < GD_NSEG; ++j)
      if (pre[c][j] <= o) sg = j;
    const uint32_t r = c * chunk_cap + sg * piece + (o - pre[c][sg]);
    float4 L = B.rad[B.rpath[r]], Lv = B.rL[r], b = B.rbeta[r];
    float v0 = b.x > 0.0f ? (L.x - Lv.x) / b.x : 0.0f;
    float v1 = b.y > 0.0f ? (L.y - Lv.y) / b.y : 0.0f;
    float v2 = b.z > 0.0f ? (L.z - Lv.z) / b.z : 0.0f;
    float w = 0.2126f * v0 + 0.7152f * v1 + 0.0722f * v2;
    B.rw[d] = (w >= 0.0f && w < 3.0e38f) ? w : 0.0f;
    const uint4* src = (const uint4*)(B.rX + (size_t)r * 64u);
    uint4* to = (uint4*)(B.tX + (size_t)d * 64u);
#pragma unroll
    for (int q = 0; q < 8; ++q) to[q] = src[q];
    B.tdir[3u * d] = B.rdir[3u * r];
    B.tdir[3u * d + 1u] = B.rdir[3u * r + 1u];
    B.tdir[3u * d + 2u] = B.rdir[3u * r + 2u];
  }
}

// Sample average (wgsl:537) folded into the progressive running average (wgsl:538-547).
__global__ void __launch_bounds__(256) gd_resolve(GdBufs B, GdArgs A, TH_GpuRenderParams P,
                                                  float4* __restrict__ accum) {
  uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  if (px >= A.n) return;
  float4 s = B.acc[px];
  float ns = (float)P.samples;
  float4 b = make_float4(s.x / ns, s.y / ns, s.z / ns, s.w / ns);
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

struct th_guided_ws {
  uint32_t capacity = 0, rcap = 0;
  GdBufs b;
  std::vector<void*> allocs;
};

void th_guided_free(th_guided_ws* w) {
  if (!w) return;
  for (void* p : w->allocs) (void)hipFree(p);
  delete w;
}

static int32_t gd_alloc(th_guided_ws* w, void** p, size_t bytes) {
  TH_CHECK(hipMalloc(p, bytes));
  w->allocs.push_back(*p);
  TH_CHECK(hipMemset(*p, 0, bytes));
  return TH_OK;
}

// Paths per queue segment for a pass over n paths: the first bounce's identity queue hands
// segment s the paths k with (k / 64) % GD_NSEG == s, at most this many.
static uint32_t gd_seg_cap(uint32_t n) { return (n + 64u * GD_NSEG - 1u) / (64u * GD_NSEG) * 64u; }

static int32_t gd_reserve(th_scene* s, uint32_t n, uint32_t rcap) {
  if (s->gw && s->gw->capacity >= n && s->gw->rcap >= rcap) return TH_OK;
  TH_CHECK(hipStreamSynchronize(s->ctx->stream));
  th_guided_free(s->gw);
  s->gw = nullptr;
  th_guided_ws* w = new th_guided_ws();
  w->capacity = n;
  w->rcap = rcap;
  GdBufs& b = w->b;
  const size_t f4 = (size_t)n * sizeof(float4), u = (size_t)n * sizeof(uint32_t);
  const size_t rows = (size_t)GD_NSEG * gd_seg_cap(n);  // queue positions (a multiple of 64)
  const size_t qf4 = rows * sizeof(float4), qu = rows * sizeof(uint32_t);
  int32_t rc;
  if ((rc = gd_alloc(w, (void**)&b.o, f4)) || (rc = gd_alloc(w, (void**)&b.d, f4)) ||
      (rc = gd_alloc(w, (void**)&b.thr, f4)) || (rc = gd_alloc(w, (void**)&b.rad, f4)) ||
      (rc = gd_alloc(w, (void**)&b.acc, f4)) || (rc = gd_alloc(w, (void**)&b.rng, u)) ||
      (rc = gd_alloc(w, (void**)&b.hit, (size_t)n * sizeof(uint2))) ||
      (rc = gd_alloc(w, (void**)&b.vp, qf4)) || (rc = gd_alloc(w, (void**)&b.vn, qf4)) ||
      (rc = gd_alloc(w, (void**)&b.vw, qf4)) || (rc = gd_alloc(w, (void**)&b.sho, f4)) ||
      (rc = gd_alloc(w, (void**)&b.sh0, f4)) || (rc = gd_alloc(w, (void**)&b.c0, f4)) ||
      (rc = gd_alloc(w, (void**)&b.sh1, f4)) || (rc = gd_alloc(w, (void**)&b.c1, f4)) ||
      (rc = gd_alloc(w, (void**)&b.vis, 2 * u)) || (rc = gd_alloc(w, (void**)&b.q0, qu)) ||
      (rc = gd_alloc(w, (void**)&b.q1, qu)) || (rc = gd_alloc(w, (void**)&b.qs, 2 * qu)) ||
      (rc = gd_alloc(w, (void**)&b.qg, qu)) ||
      (rc = gd_alloc(w, (void**)&b.cnt, GD_C(GD_CNT_LINES) * sizeof(uint32_t))) ||
      (rc = gd_alloc(w, (void**)&b.O, rows * TG_OUT * sizeof(uint16_t))) ||
      (rc = gd_alloc(w, (void**)&b.rX, (size_t)rcap * 64 * sizeof(uint16_t))) ||
      (rc = gd_alloc(w, (void**)&b.rdir, (size_t)rcap * 3 * sizeof(float))) ||
      (rc = gd_alloc(w, (void**)&b.tX, (size_t)rcap * 64 * sizeof(uint16_t))) ||
      (rc = gd_alloc(w, (void**)&b.tdir, (size_t)rcap * 3 * sizeof(float))) ||
      (rc = gd_alloc(w, (void**)&b.rw, (size_t)rcap * sizeof(float))) ||
      (rc = gd_alloc(w, (void**)&b.rbeta, (size_t)rcap * sizeof(float4))) ||
      (rc = gd_alloc(w, (void**)&b.rL, (size_t)rcap * sizeof(float4))) ||
      (rc = gd_alloc(w, (void**)&b.rpath, (size_t)rcap * sizeof(uint32_t)))) {
    th_guided_free(w);
    return rc;
  }
  b.rcap = rcap;
  s->gw = w;
  return TH_OK;
}

extern "C" int32_t th_render_guided(th_context* ctx, th_scene* s, th_guide* g, const TH_GpuRenderParams* p,
                                    const TH_GpuCamera* cam, const th_guided_options* opt,
                                    th_guided_stats* stats) {
  if (!ctx || !s || !g || !p || !cam || !opt) return th_set_error(TH_E_INVALID, "null argument");
  if (s->ctx != ctx || guide_context(g) != ctx)
    return th_set_error(TH_E_INVALID, "scene or guide belongs to another context");
  int32_t rc = th_check_render_params(s, p);
  if (rc) return rc;
  if (!(opt->alpha >= 0.0f && opt->alpha <= 0.95f))
    return th_set_error(TH_E_INVALID, "alpha must lie in [0, 0.95]");
  if (opt->flags & ~(TH_GUIDED_NO_ENV_NEE | TH_GUIDED_ALPHA_FROM_START | TH_GUIDED_TIME_NETWORK))
    return th_set_error(TH_E_INVALID, "unknown guided flag");
  const bool time_net = (opt->flags & TH_GUIDED_TIME_NETWORK) != 0u;
  const uint32_t batch = opt->batch ? opt->batch : (1u << 18);
  const uint32_t steps = opt->steps_per_pass ? opt->steps_per_pass : 1u;
  const uint32_t wrate = opt->warmup_steps ? std::max<uint32_t>(opt->warmup_rate, 1u) : steps;
  if (batch % 32u != 0u || batch > (1u << 22) || steps > 64u || wrate > 64u)
    return th_set_error(TH_E_INVALID, "batch must be a multiple of 32 up to 2^22, steps at most 64");
  const uint64_t npix64 = (uint64_t)p->width * p->height;
  if (npix64 > (1u << 26)) return th_set_error(TH_E_UNSUPPORTED, "image too large for the guided wavefront");
  const uint32_t npix = (uint32_t)npix64;
  const bool train = opt->train != 0u;
  const uint32_t max_chunks = std::max(steps, wrate);
  if (max_chunks > GD_MAX_CHUNKS) return th_set_error(TH_E_INVALID, "at most 16 steps per pass");
  // one chunk per step: capacity 1.5 x batch absorbs the spread of the random split (and of
  // the segments' shares: GD_NSEG pieces of a multiple of 32 records each)
  const uint32_t chunk_cap = (batch + batch / 2u + 32u * GD_NSEG - 1u) / (32u * GD_NSEG) * (32u * GD_NSEG);
  if (train && (uint64_t)chunk_cap * max_chunks > (1u << 26)) return th_set_error(TH_E_INVALID, "batch x steps too large");
  const uint32_t rcap = train ? chunk_cap * max_chunks : 32u;
  TH_CHECK(hipSetDevice(ctx->device));
  if ((rc = gd_reserve(s, npix, rcap))) return rc;
  th_guided_ws* w = s->gw;
  GdBufs B = w->b;
  B.seg_cap = gd_seg_cap(npix);
  DevScene S = dev_scene(s);
  const uint32_t block = 64;
  const uint32_t depth = 3u * s->depth4 + 2u;
  const size_t lds = (size_t)depth * (block / 4) * sizeof(uint2);
  // every grid a multiple of GD_NSEG: block b serves queue segment b % GD_NSEG
  auto segs = [](uint32_t g) { return (g + GD_NSEG - 1u) / GD_NSEG * GD_NSEG; };
  // The closest-hit grid holds one wave per 16 paths, so on the first bounce each wave
  // traces one round of rays and the dispatcher balances the waves' unequal costs: on the
  // 1280x720 terrain 1.16x faster than 8192 persistent waves (2.55 -> 2.21 ms per plain
  // pass).  Later bounces' surplus waves find their segment empty and exit at once.
  // The shadow grid stays at 8192 waves (a larger one measured slower in the guided passes)
  // and the shading grid at 4096 (8192 or 14,400 within noise).
  const uint32_t grid_trace = segs((npix * 4u + block - 1) / block);
  const uint32_t grid_scalar = segs(std::min<uint32_t>((npix + block - 1) / block, 4096u));
  const uint32_t grid_shadow = segs(std::min<uint32_t>((2u * npix * 4u + block - 1) / block, 8192u));
  GdArgs A;
  memset(&A, 0, sizeof A);
  A.n = npix;
  for (int a = 0; a < 3; ++a) {
    A.bmin[a] = s->root_min[a];
    float e = s->root_max[a] - s->root_min[a];
    A.bscale[a] = e > 0.0f ? 1.0f / e : 0.0f;
  }
  th_guided_stats st;
  memset(&st, 0, sizeof st);
  // record-rate estimate before any training pass has counted diffuse vertices
  const uint32_t init_dpp = (uint32_t)std::min<uint64_t>(
      (uint64_t)npix * std::max<uint32_t>(p->max_bounces, 2u) / 2u, 0xffffffffu);
  std::pair<hipEvent_t, hipEvent_t> evp;
  if ((rc = begin_timed(ctx, &evp))) return rc;
  hipStream_t sm = ctx->stream;
  // TH_GUIDED_TIME_NETWORK: an event pair around every inference and training launch
  std::vector<hipEvent_t> ev_infer, ev_train;
  auto mark = [&](std::vector<hipEvent_t>& v) {
    if (!time_net) return;
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) == hipSuccess) {
      if (hipEventRecord(e, sm) == hipSuccess)
        v.push_back(e);
      else
        (void)hipEventDestroy(e);
    }
  };
  auto span_ms = [&](std::vector<hipEvent_t>& v) {
    float total = 0.0f;
    for (size_t i = 0; i + 1 < v.size(); i += 2) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, v[i], v[i + 1]) == hipSuccess) total += ms;
    }
    for (hipEvent_t e : v) (void)hipEventDestroy(e);
    v.clear();
    return total;
  };
  hipLaunchKernelGGL(gd_reset, dim3(1), dim3(64), 0, sm, B, 1u, 0u, init_dpp);
  uint32_t cnt_host[GD_C(32)];
  const uint32_t steps0 = guide_steps(g);  // steps applied before this call
  uint32_t applied_host = 0;
  for (uint32_t smp = 0; smp < p->samples; ++smp) {
    const uint32_t taken = guide_steps(g);
    const bool warm = taken < opt->warmup_steps;
    const uint32_t pass_steps = warm ? wrate : steps;
    // after the warm-up, train on every train_every-th pass only
    const bool train_pass = train && (warm || opt->train_every <= 1u || (smp % opt->train_every) == 0u);
    const uint64_t target = (uint64_t)batch * pass_steps;
    A.smp = smp;
    A.nchunks = pass_steps;
    A.chunk_cap = chunk_cap;
    const bool untrained = taken == 0u && !(opt->flags & TH_GUIDED_ALPHA_FROM_START);
    A.alpha = untrained ? 0.0f
              : (opt->alpha_ramp && taken < opt->alpha_ramp ? opt->alpha * (float)taken / (float)opt->alpha_ramp
                                                             : opt->alpha);
    A.record = train_pass ? 1u : 0u;
    A.defer = (A.alpha > 0.0f || train_pass) ? 1u : 0u;
    // environment sampling at diffuse vertices as the caller asks, whatever alpha is, so
    // the guide's gain can be measured against the same estimator without it
    A.env_nee = (opt->flags & TH_GUIDED_NO_ENV_NEE) ? 0u : 1u;
    hipLaunchKernelGGL(gd_reset, dim3(1), dim3(64), 0, sm, B, 0u,
                       train_pass ? (uint32_t)std::min<uint64_t>(target, 0xffffffffu) : 0u, init_dpp);
    hipLaunchKernelGGL(gd_generate, dim3((npix + 255) / 256), dim3(256), 0, sm, B, A, *p, *cam);
    for (uint32_t b = 0; b < p->max_bounces; ++b) {
      const uint32_t li = b & 1u;
      A.bounce = b;
      hipLaunchKernelGGL(gd_extend, dim3(grid_trace), dim3(block), lds, sm, B, S, *p, li, b == 0u ? 1u : 0u);
      hipLaunchKernelGGL(gd_shade_a, dim3(grid_scalar), dim3(block), 0, sm, B, A, S, *p,
                         (const float*)ctx->d_srgb_lut, li);
      if (A.defer) {
        if (A.alpha > 0.0f) {
          mark(ev_infer);
          rc = guide_infer(g, B.cnt + GD_QC(3u, 0u), GD_QC(3u, 1u) - GD_QC(3u, 0u), GD_NSEG, B.seg_cap,
                           B.vp, B.vn, B.vw, s->root_min, s->root_max, B.O, sm);
          mark(ev_infer);
          if (rc) {
            (void)end_timed(ctx, evp);
            (void)span_ms(ev_infer);
            (void)span_ms(ev_train);
            return rc;
          }
        }
        hipLaunchKernelGGL(gd_shade_b, dim3(grid_scalar), dim3(block), 0, sm, B, A, S, *p,
                           (const float*)ctx->d_srgb_lut, li);
      }
      hipLaunchKernelGGL(gd_shadow, dim3(grid_shadow), dim3(block), lds, sm, B, S, *p);
    }
    hipLaunchKernelGGL(gd_retire, dim3(std::min<uint32_t>((npix + 255) / 256, 2048u)), dim3(256), 0, sm, B,
                       p->max_bounces == 0u ? 0u : npix);
    if (train_pass) {
      hipLaunchKernelGGL(gd_weights, dim3(std::min<uint32_t>((pass_steps * chunk_cap + 255) / 256, 4096u)),
                         dim3(256), 0, sm, B, pass_steps, chunk_cap);
      // the steps read their chunk's record count on the device (a step on fewer than
      // 1024 records is skipped there), so the host queues the next pass while this one
      // runs -- except while the schedule (warm-up rate, alpha ramp, the untrained
      // network's alpha 0) still depends on how many steps have applied: then it reads
      // that count back after the pass
      for (uint32_t j = 0; j < pass_steps; ++j) {
        const size_t off = (size_t)j * chunk_cap;
        mark(ev_train);
        rc = guide_train(g, B.tX + off * 64u, B.tdir + off * 3u, B.rw + off, B.cnt + GD_C(8u + j), chunk_cap,
                         1024u, B.cnt + GD_C(GD_STATS));
        mark(ev_train);
        if (rc) {
          (void)end_timed(ctx, evp);
          (void)span_ms(ev_infer);
          (void)span_ms(ev_train);
          return rc;
        }
      }
      const bool settled = taken > 0u && taken >= opt->warmup_steps && taken >= opt->alpha_ramp;
      if (settled) {
        guide_set_steps(g, taken + pass_steps);  // the schedule no longer depends on it
      } else {
        TH_CHECK(hipMemcpyAsync(&applied_host, B.cnt + GD_C(GD_STATS) + 2u, 4, hipMemcpyDeviceToHost, sm));
        TH_CHECK(hipStreamSynchronize(sm));
        guide_set_steps(g, steps0 + applied_host);
      }
    }
  }
  hipLaunchKernelGGL(gd_resolve, dim3((npix + 255) / 256), dim3(256), 0, sm, B, A, *p, s->d_accum);
  hipError_t le = hipGetLastError();
  if ((rc = end_timed(ctx, evp))) return rc;
  if (le != hipSuccess) return th_set_error(TH_E_DEVICE, std::string("guided launch: ") + hipGetErrorString(le));
  ctx->counters_pending = false;
  ctx->stats.paths = npix64 * p->samples;
  if (train || stats) {
    TH_CHECK(hipMemcpyAsync(cnt_host, B.cnt, sizeof cnt_host, hipMemcpyDeviceToHost, sm));
    if (train && (rc = guide_last_loss(g, &st.last_loss))) return rc;
    TH_CHECK(hipStreamSynchronize(sm));
    if (cnt_host[GD_C(GD_STATS) + 2u] == 0u) st.last_loss = 0.0f;  // no step applied in this render
    guide_set_steps(g, steps0 + cnt_host[GD_C(GD_STATS) + 2u]);
  }
  if (time_net) {
    TH_CHECK(hipStreamSynchronize(sm));
    st.infer_ms = span_ms(ev_infer);
    st.train_ms = span_ms(ev_train);
  }
  if (stats) {
    st.diffuse_vertices = cnt_host[GD_C(5)];
    st.inferred_rows = cnt_host[GD_C(6)];
    st.records = cnt_host[GD_C(GD_STATS)];
    st.trained_samples = cnt_host[GD_C(GD_STATS) + 1u];
    st.train_steps = cnt_host[GD_C(GD_STATS) + 2u];
    *stats = st;
  }
  return TH_OK;
}
