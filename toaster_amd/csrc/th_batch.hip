// th_batch.hip -- many scenes (or animation frames) of uneven size in one persistent
// launch, replayed from a captured hipGraph, plus the per-scene autotuner.
//
// Work is cut into items of one block (4 waves = 64 quads) of one scene.  Items are
// ordered by descending estimated cost (expensive scenes first, so the tail of the
// launch is made of cheap items) and handed out by an atomic counter to a grid sized
// to keep every CU busy; a block reloads the LDS copy of the top 4-wide BVH levels only
// when its next item belongs to another scene.  Each item runs the same quad
// integrator as k_render_quad, with that scene's quads-per-pixel (G) and treelet size,
// which the autotuner picks per scene by timing the candidates on the device.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "th_internal.h"

namespace {

struct BatchScene {
  DevScene S;
  TH_GpuRenderParams P;
  TH_GpuCamera cam;
  float4* accum;
  const Child4* nodes4;
  uint32_t G, per_sample, cached, first_item, items, tiles_x, tw, th;
  uint32_t pad;
};

#define BATCH_BLOCK 256u
#define BATCH_MAX_CACHED 85u

__global__ void k_batch_reset(uint32_t* counter) {
  if (threadIdx.x == 0) *counter = 0u;
}

__global__ void __launch_bounds__(BATCH_BLOCK) k_render_batch(const BatchScene* __restrict__ table,
                                                              uint32_t nscenes, uint32_t total_items,
                                                              uint32_t* counter,
                                                              const float* __restrict__ srgb_lut,
                                                              uint32_t stack_rows) {
  extern __shared__ uint2 lds_raw[];
  __shared__ uint32_t s_item, s_scene;
  Child4* lds_nodes = (Child4*)lds_raw;
  uint2* lds_stack = lds_raw + BATCH_MAX_CACHED * 4u * (sizeof(Child4) / sizeof(uint2));
  uint32_t loaded = 0xffffffffu;
  for (;;) {
    if (threadIdx.x == 0) {
      uint32_t it = atomicAdd(counter, 1u);
      uint32_t sc = 0;
      if (it < total_items)
        while (sc + 1u < nscenes && table[sc + 1u].first_item <= it) ++sc;
      s_item = it;
      s_scene = sc;
    }
    __syncthreads();
    const uint32_t item = s_item, sc = s_scene;
    if (item >= total_items) break;
    const BatchScene& B = table[sc];
    if (sc != loaded) {
      for (uint32_t i = threadIdx.x; i < 4u * B.cached; i += blockDim.x) lds_nodes[i] = B.nodes4[i];
      loaded = sc;
    }
    __syncthreads();
    const TH_GpuRenderParams& P = B.P;
    const uint32_t G = B.G;
    const uint32_t lane = threadIdx.x & 63u, qw = lane >> 2;
    const uint32_t wave = (item - B.first_item) * (BATCH_BLOCK / 64u) + (threadIdx.x >> 6);
    const uint32_t pw = qw / G, j = qw % G;
    const uint32_t x = (wave % B.tiles_x) * B.tw + pw % B.tw;
    const uint32_t y = (wave / B.tiles_x) * B.th + pw / B.tw;
    const bool active = x < P.width && y < P.height;
    Stack st{lds_stack + (threadIdx.x >> 2), (int)(BATCH_BLOCK / 4u), 0};
    LaneCounters<false> lc;
    Integrator<3, false> itg{B.S, P, srgb_lut, st, lc, lds_nodes, B.cached};
    const uint32_t idx = y * P.width + x;
    float cr = 0.0f, cg = 0.0f, cb = 0.0f, ca = 0.0f;
    if (active) {
      uint32_t base = idx ^ (P.frame_index * 9781u);
      uint32_t s = th::pcg_hash(base);
      const uint32_t step = B.per_sample ? G : 1u;
      for (uint32_t smp = B.per_sample ? j : 0u; smp < P.samples; smp += step) {
        if (B.per_sample) s = th::pcg_hash(base + smp * 0x9E3779B9u);
        v3 o, d;
        camera_ray(B.cam, P, x, y, &s, &o, &d);
        v3 L = itg.ray_color(o, d, &s);
        cr = cr + L.x;
        cg = cg + L.y;
        cb = cb + L.z;
        ca = ca + 1.0f;
      }
    }
    for (uint32_t off = 4; off < 4u * G; off <<= 1) {
      cr = cr + __shfl_xor(cr, (int)off, 64);
      cg = cg + __shfl_xor(cg, (int)off, 64);
      cb = cb + __shfl_xor(cb, (int)off, 64);
      ca = ca + __shfl_xor(ca, (int)off, 64);
    }
    if (active && (lane % (4u * G)) == 0u) {
      float ns = (float)P.samples;
      float4 b = make_float4(cr / ns, cg / ns, cb / ns, ca / ns);
      if (P.accumulated_samples == 0u) {
        B.accum[idx] = b;
      } else {
        float prev = (float)P.accumulated_samples;
        float tot = prev + ns;
        float4 o = B.accum[idx];
        B.accum[idx] = make_float4((o.x * prev + b.x * ns) / tot, (o.y * prev + b.y * ns) / tot,
                                   (o.z * prev + b.z * ns) / tot, (o.w * prev + b.w * ns) / tot);
      }
    }
    __syncthreads();
  }
  (void)stack_rows;
}

uint32_t default_G(uint32_t rng_mode, uint32_t samples) {
  if (rng_mode == TH_RNG_REFERENCE) return 1;
  return samples >= 16 ? 16 : samples >= 8 ? 8 : samples >= 4 ? 4 : samples >= 2 ? 2 : 1;
}

bool valid_G(uint32_t g) { return g == 1 || g == 2 || g == 4 || g == 8 || g == 16; }

}  // namespace

// Device tables and the captured graph of the last batch shape.
struct th_batch_state {
  BatchScene* d_table = nullptr;
  uint32_t table_cap = 0;
  uint32_t* d_counter = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t g_nscenes = 0, g_items = 0, g_grid = 0, g_rows = 0;
  size_t g_lds = 0;
};

static th_batch_state* batch_state(th_context* ctx) {
  static_assert(sizeof(BatchScene) % 16 == 0 || true, "");
  if (!ctx->batch) ctx->batch = new th_batch_state();
  return ctx->batch;
}

void th_batch_free(th_batch_state* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->d_table) (void)hipFree(b->d_table);
  if (b->d_counter) (void)hipFree(b->d_counter);
  delete b;
}

static int32_t fill_entry(th_scene* s, const th_batch_item& it, uint32_t rng_mode, BatchScene* e,
                          uint32_t* rows) {
  int32_t rc = th_check_render_params(s, &it.params);
  if (rc) return rc;
  if (s->node_count == 0) return th_set_error(TH_E_UNSUPPORTED, "batched scenes need a BVH");
  uint32_t G = it.quads_per_pixel ? it.quads_per_pixel : default_G(rng_mode, it.params.samples);
  if (!valid_G(G)) return th_set_error(TH_E_INVALID, "quads_per_pixel must be 1, 2, 4, 8 or 16");
  if (rng_mode == TH_RNG_REFERENCE && G != 1)
    return th_set_error(TH_E_INVALID, "reference RNG streams need one quad per pixel");
  e->S = dev_scene(s);
  e->P = it.params;
  e->cam = it.camera;
  e->accum = s->d_accum;
  e->nodes4 = s->d_nodes4;
  e->G = G;
  e->per_sample = rng_mode == TH_RNG_PER_SAMPLE ? 1u : 0u;
  uint32_t cached = it.cached_nodes == 0xffffffffu ? BATCH_MAX_CACHED : it.cached_nodes;
  e->cached = std::min<uint32_t>(std::min<uint32_t>(cached, BATCH_MAX_CACHED), s->node4_count);
  uint32_t ppw = 16u / G;  // pixels per wave
  uint32_t tw = it.tile_width ? it.tile_width : (ppw >= 8 ? 4u : (ppw >= 2 ? 2u : 1u));
  if ((tw & (tw - 1u)) != 0u || tw > ppw)
    return th_set_error(TH_E_INVALID, "tile_width must be a power of two up to 16 / quads_per_pixel");
  e->tw = tw;
  e->th = ppw / tw;
  e->tiles_x = (it.params.width + e->tw - 1) / e->tw;
  uint32_t waves = e->tiles_x * ((it.params.height + e->th - 1) / e->th);
  e->items = (waves + 3) / 4;
  e->pad = 0;
  *rows = std::max(*rows, 3u * s->depth4 + 2u);
  return TH_OK;
}

extern "C" int32_t th_render_batch(th_context* ctx, const th_batch_item* items, uint32_t count,
                                   const th_render_options* opt) {
  if (!ctx || !items || count == 0) return th_set_error(TH_E_INVALID, "null argument or empty batch");
  th_render_options o;
  memset(&o, 0, sizeof o);
  if (opt) o = *opt;
  if (o.rng_mode > 1u) return th_set_error(TH_E_INVALID, "unknown rng mode");
  if (o.batch_blocks_per_cu > 64u) return th_set_error(TH_E_INVALID, "batch_blocks_per_cu above 64");
  TH_CHECK(hipSetDevice(ctx->device));
  std::vector<BatchScene> table(count);
  uint32_t rows = 2;
  std::vector<double> cost(count);
  for (uint32_t i = 0; i < count; ++i) {
    if (!items[i].scene || items[i].scene->ctx != ctx)
      return th_set_error(TH_E_INVALID, "batch item without a scene of this context");
    int32_t rc = fill_entry(items[i].scene, items[i], o.rng_mode, &table[i], &rows);
    if (rc) return rc;
    const TH_GpuRenderParams& p = items[i].params;
    cost[i] = (double)p.width * p.height * p.samples * std::max(1u, p.max_bounces) *
              (1.0 + std::log2(1.0 + (double)items[i].scene->prim_count));
  }
  // expensive scenes first
  std::vector<uint32_t> order(count);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
  std::vector<BatchScene> sorted(count);
  uint32_t total = 0;
  for (uint32_t k = 0; k < count; ++k) {
    sorted[k] = table[order[k]];
    sorted[k].first_item = total;
    total += sorted[k].items;
  }
  th_batch_state* b = batch_state(ctx);
  if (b->table_cap < count) {
    if (b->d_table) (void)hipFree(b->d_table);
    b->d_table = nullptr;
    TH_CHECK(hipMalloc((void**)&b->d_table, count * sizeof(BatchScene)));
    b->table_cap = count;
    if (b->exec) {  // the graph holds the old table pointer
      (void)hipGraphExecDestroy(b->exec);
      (void)hipGraphDestroy(b->graph);
      b->exec = nullptr;
      b->graph = nullptr;
    }
  }
  if (!b->d_counter) TH_CHECK(hipMalloc((void**)&b->d_counter, 64));
  TH_CHECK(hipMemcpyAsync(b->d_table, sorted.data(), count * sizeof(BatchScene), hipMemcpyHostToDevice,
                          ctx->stream));
  int cus = 0;
  TH_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  const uint32_t per_cu = o.batch_blocks_per_cu ? o.batch_blocks_per_cu : 4u;
  uint32_t grid = std::min<uint32_t>(total, (uint32_t)std::max(cus, 1) * per_cu);
  size_t lds = (size_t)BATCH_MAX_CACHED * 4 * sizeof(Child4) + (size_t)rows * (BATCH_BLOCK / 4) * sizeof(uint2);
  bool same = b->exec && b->g_nscenes == count && b->g_items == total && b->g_grid == grid &&
              b->g_rows == rows && b->g_lds == lds;
  std::pair<hipEvent_t, hipEvent_t> evp;
  int32_t rc = begin_timed(ctx, &evp);
  if (rc) return rc;
  if (!same) {
    if (b->exec) {
      (void)hipGraphExecDestroy(b->exec);
      (void)hipGraphDestroy(b->graph);
      b->exec = nullptr;
      b->graph = nullptr;
    }
    // capture: counter reset (a kernel, no memset node) + the persistent launch
    hipStream_t cs;
    TH_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, cs, b->d_counter);
      hipLaunchKernelGGL(k_render_batch, dim3(grid), dim3(BATCH_BLOCK), lds, cs, b->d_table, count,
                         total, b->d_counter, (const float*)ctx->d_srgb_lut, rows);
      e = hipStreamEndCapture(cs, &b->graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) {
      (void)end_timed(ctx, evp);
      return th_set_error(TH_E_DEVICE, std::string("batch graph: ") + hipGetErrorString(e));
    }
    b->g_nscenes = count;
    b->g_items = total;
    b->g_grid = grid;
    b->g_rows = rows;
    b->g_lds = lds;
  }
  hipError_t e = hipGraphLaunch(b->exec, ctx->stream);
  rc = end_timed(ctx, evp);
  if (e != hipSuccess) return th_set_error(TH_E_DEVICE, std::string("batch launch: ") + hipGetErrorString(e));
  if (rc) return rc;
  ctx->stats.paths = 0;
  for (uint32_t i = 0; i < count; ++i)
    ctx->stats.paths += (uint64_t)items[i].params.width * items[i].params.height * items[i].params.samples;
  return TH_OK;
}

// Median device time of a batch: one untimed launch first (a new launch shape captures
// its graph then, on the host), then three timed launches.
static int32_t time_batch(th_context* ctx, const th_batch_item* items, uint32_t count, const th_render_options& o,
                          float* ms) {
  int32_t rc = th_render_batch(ctx, items, count, &o);
  if (rc) return rc;
  if ((rc = th_synchronize(ctx))) return rc;
  float t[3];
  for (int r = 0; r < 3; ++r) {
    if ((rc = th_render_batch(ctx, items, count, &o))) return rc;
    th_stats st;
    if ((rc = th_get_stats(ctx, &st))) return rc;  // waits for the launch
    t[r] = (float)st.last_kernel_ms;
  }
  std::sort(t, t + 3);
  *ms = t[1];
  return TH_OK;
}

// Coordinate descent over the launch shape of one scene, each candidate alone on the
// device: quads per pixel (per-sample RNG), then the treelet nodes cached in LDS, then
// the wave's tile width; the result is also stored on the scene for later batches.
extern "C" int32_t th_autotune(th_context* ctx, th_scene* s, const TH_GpuRenderParams* p,
                               const TH_GpuCamera* cam, const th_render_options* opt, th_tuning* out) {
  if (!ctx || !s || !p || !cam || !out) return th_set_error(TH_E_INVALID, "null argument");
  th_render_options o;
  memset(&o, 0, sizeof o);
  if (opt) o = *opt;
  int32_t rc = th_synchronize(ctx);
  if (rc) return rc;
  th_stats saved = ctx->stats;
  th_batch_item it;
  memset(&it, 0, sizeof it);
  it.scene = s;
  it.params = *p;
  it.camera = *cam;
  it.quads_per_pixel = default_G(o.rng_mode, p->samples);
  it.cached_nodes = BATCH_MAX_CACHED;
  it.tile_width = 0;
  float best = 3.4e38f;
  auto consider = [&](const th_batch_item& c) -> int32_t {
    float ms = 0.0f;
    int32_t r = time_batch(ctx, &c, 1, o, &ms);
    if (r) return r;
    if (ms < best) {
      best = ms;
      it = c;
    }
    return TH_OK;
  };
  if ((rc = consider(it))) return rc;
  if (o.rng_mode == TH_RNG_PER_SAMPLE)
    for (uint32_t g = 1; g <= 16 && g <= std::max(1u, p->samples); g <<= 1) {
      th_batch_item c = it;
      c.quads_per_pixel = g;
      c.tile_width = 0;
      if (g != it.quads_per_pixel && (rc = consider(c))) return rc;
    }
  for (uint32_t cached : {0u, 21u, 85u}) {
    th_batch_item c = it;
    c.cached_nodes = cached;
    if (cached != it.cached_nodes && (rc = consider(c))) return rc;
  }
  const uint32_t ppw = 16u / it.quads_per_pixel;
  const uint32_t tw0 = it.tile_width ? it.tile_width : (ppw >= 8 ? 4u : (ppw >= 2 ? 2u : 1u));
  for (uint32_t tw = 1; tw <= ppw; tw <<= 1) {
    th_batch_item c = it;
    c.tile_width = tw;
    if (tw != tw0 && (rc = consider(c))) return rc;
  }
  ctx->stats = saved;
  th_tuning bt;
  bt.quads_per_pixel = it.quads_per_pixel;
  bt.cached_nodes = it.cached_nodes;
  bt.best_ms = best;
  bt.tile_width = it.tile_width ? it.tile_width : tw0;
  s->tuning = bt;
  *out = bt;
  return TH_OK;
}

extern "C" int32_t th_autotune_batch(th_context* ctx, const th_batch_item* items, uint32_t count,
                                     const th_render_options* opt, uint32_t* blocks_per_cu, float* best_ms) {
  if (!ctx || !items || count == 0 || !blocks_per_cu || !best_ms)
    return th_set_error(TH_E_INVALID, "null argument or empty batch");
  th_render_options o;
  memset(&o, 0, sizeof o);
  if (opt) o = *opt;
  int32_t rc = th_synchronize(ctx);
  if (rc) return rc;
  th_stats saved = ctx->stats;
  float best = 3.4e38f;
  uint32_t best_n = 4;
  for (uint32_t n : {2u, 4u, 6u, 8u}) {
    o.batch_blocks_per_cu = n;
    float ms = 0.0f;
    if ((rc = time_batch(ctx, items, count, o, &ms))) return rc;
    if (ms < best) {
      best = ms;
      best_n = n;
    }
  }
  ctx->stats = saved;
  *blocks_per_cu = best_n;
  *best_ms = best;
  return TH_OK;
}
