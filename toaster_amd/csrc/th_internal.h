// th_internal.h -- runtime objects shared by the translation units of libtoaster_hip.
#pragma once
#include <hip/hip_runtime.h>
#include <string.h>

#include <string>
#include <utility>
#include <vector>

#include "th_common.h"
#include "th_integrator.h"

using namespace thd;

#define TH_CHECK(expr)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return th_set_error(TH_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct th_wavefront;    // wavefront workspace (th_wavefront.hip)
struct th_guided_ws;    // guided wavefront workspace (th_guided.hip)
struct th_batch_state;  // batch tables and captured graph (th_batch.hip)

struct th_context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // unused spare pair
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool, ev_pending;
  unsigned long long* d_counters = nullptr;
  bool counters_pending = false;
  th_stats stats;
  float* d_srgb_lut = nullptr;
  th_batch_state* batch = nullptr;
  uint32_t* d_work = nullptr;  // work counters of persistent launches
  int cu_count = 0;            // compute units of the device
  int dyn_blocks_per_cu = 0;   // resident blocks of k_render_quad_dyn (0 = not yet queried)
  void* wf_scratch = nullptr;  // device address of the static wavefront workspace on `device`
};

struct th_scene {
  th_context* ctx = nullptr;
  TH_GpuRenderParams params;  // counts fixed at create
  uint32_t node_capacity = 0;
  uint32_t width = 0, height = 0;
  TH_GpuBvhNode* d_nodes = nullptr;
  TH_GpuPrimitiveRef* d_prims = nullptr;
  TH_GpuSphere* d_spheres = nullptr;
  TH_GpuTriangle* d_tris = nullptr;
  TH_GpuTriangleAttributes* d_attrs = nullptr;
  TH_GpuMaterial* d_mats = nullptr;
  TH_GpuLight* d_lights = nullptr;
  float4* d_env = nullptr;
  float* d_env_cdf = nullptr;   // compact CDF (null: not non-decreasing)
  float* d_env_rows = nullptr;  // CDF at the end of every row
  uint32_t env_rows_w = 0;
  uint32_t* d_tex = nullptr;
  CNode* d_cnodes = nullptr;
  Child4* d_nodes4 = nullptr;
  uint32_t root4 = 0, depth4 = 0, node4_count = 0;
  // primitives the 4-wide traversal tests before the tree (see prepare_bvh)
  uint32_t ext_n = 0, ext_slot[8] = {0, 0, 0, 0, 0, 0, 0, 0}, ext_leaf[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t ext_node[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  float4* d_slots = nullptr;
  float4* d_accum = nullptr;
  uint32_t node_count = 0, prim_count = 0, env_count = 0, tex_count = 0;
  uint32_t root_entry = 0;
  float root_min[3] = {0, 0, 0}, root_max[3] = {0, 0, 0};
  uint32_t max_depth = 0;
  // refit: reference node indices deepest level first (refit_level_end[k] ends level k of
  // that order) and, per 4-wide child slot, its reference node (0xffffffff: none)
  uint32_t* d_refit_order = nullptr;
  std::vector<uint32_t> refit_level_end;
  uint32_t* d_wide_src = nullptr;
  // TH_SCHEDULE_AUTO: the schedule chosen for (samples, bounces, background), and a copy
  // of the accumulation kept while the candidates are timed
  // summed surface area of all node boxes: as built, and after the last refit
  float area_build = 0.0f, area_now = 0.0f;
  float* d_area = nullptr;
  bool auto_valid = false;
  uint32_t auto_samples = 0, auto_bounces = 0, auto_background = 0, auto_choice = 0;
  float4* d_accum_save = nullptr;
  uchar4* d_rgba8 = nullptr;  // display staging (width*height), allocated on first use
  th_wavefront* wf = nullptr;  // allocated by the first wavefront render
  th_guided_ws* gw = nullptr;  // allocated by the first guided render
  th_tuning tuning = {0, 0, 0.0f, 0};
  uint64_t generation = 0;  // bumped by every th_scene_update (progressive graphs check it)
};


DevScene dev_scene(const th_scene* s);
int32_t begin_timed(th_context* ctx, std::pair<hipEvent_t, hipEvent_t>* pr);
int32_t end_timed(th_context* ctx, std::pair<hipEvent_t, hipEvent_t> pr);
int32_t fold_timing(th_context* ctx);

// Wavefront renderer (th_wavefront.hip).
int32_t th_wavefront_render(th_context* ctx, th_scene* s, const TH_GpuRenderParams* p,
                            const TH_GpuCamera* cam, const th_render_options* o);
void th_wavefront_free(th_wavefront* wf);
int32_t th_wavefront_scratch_init(th_context* ctx);
void th_batch_free(th_batch_state* b);
int32_t th_check_render_params(const th_scene* s, const TH_GpuRenderParams* p);

// Guided wavefront (th_guided.hip) and the network kernels it drives (th_guide.hip).
void th_guided_free(th_guided_ws* w);
// Inference over a queue in `nseg` segments: segment s holds d_count[s * count_stride]
// rows starting at row s * seg_cap of vp, vn, vw and O (TG_OUT binary16 outputs per row).
int32_t guide_infer(th_guide* g, const uint32_t* d_count, uint32_t count_stride, uint32_t nseg,
                    uint32_t seg_cap, const float4* vp, const float4* vn, const float4* vw,
                    const float* bmin, const float* bmax, uint16_t* O, hipStream_t st);
int32_t guide_last_loss(th_guide* g, float* out);
// One training step on min(*d_count, cap) rows (a no-op below min_count) with the count
// read on the device; d_stats (optional) accumulates samples offered, trained, steps applied.
int32_t guide_train(th_guide* g, const uint16_t* X, const float* dirs, const float* w, const uint32_t* d_count,
                    uint32_t cap, uint32_t min_count, uint32_t* d_stats);
uint32_t guide_steps(const th_guide* g);
void guide_set_steps(th_guide* g, uint32_t steps);
th_context* guide_context(const th_guide* g);
