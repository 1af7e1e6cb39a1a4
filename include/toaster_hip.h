/*
 * toaster_hip.h -- C ABI of the MI355X path-tracing library (libtoaster_hip.so).
 *
 * The library replaces the wgpu device runtime and WGSL megakernel of the
 * reference renderer (crates/toaster-gpu + shaders/pathtrace_spheres.wgsl) behind
 * a plain C boundary: plain pointers, sizes and int32 status codes, no C++ or torch
 * types.  Every function returns TH_OK (0) on success or a negative TH_E* code;
 * the message of the last failure is returned by th_last_error().
 *
 * Two groups of entry points:
 *   1. host packing  -- th_pack_scene() turns a renderer-neutral scene into the
 *      flat GPU records (reference: crates/toaster-gpu/src/scene_upload.rs:60-76,
 *      scene_to_gpu).  It needs no GPU.
 *   2. device runtime -- contexts, scene upload/update, render, readback, stats
 *      (reference: crates/toaster-gpu/src/{device,buffers,dispatch,readback}.rs
 *      and path_tracer.rs:518-849).
 *
 * Record layouts are byte-identical to crates/toaster-gpu/src/gpu_types.rs:8-195
 * (sizes asserted below exactly as gpu_types.rs:200-211 asserts them).
 */
#ifndef TOASTER_HIP_H
#define TOASTER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TH_ABI_VERSION 1

/* ---------------------------------------------------------------- status */
#define TH_OK 0
#define TH_E_INVALID (-1)     /* null pointer, bad size, bad count, bad enum   */
#define TH_E_DEVICE (-2)      /* HIP runtime failure                          */
#define TH_E_NO_DEVICE (-3)   /* no GPU visible                               */
#define TH_E_MISMATCH (-4)    /* counts changed after create, size mismatch   */
#define TH_E_CAPACITY (-5)    /* node count above the allocated capacity      */
#define TH_E_UNSUPPORTED (-6) /* option not supported by this build           */

/* ------------------------------------------------------ GPU record layouts */
/* gpu_types.rs:8-56 GpuRenderParams (96 B) */
typedef struct TH_GpuRenderParams {
    uint32_t width, height, samples, max_bounces;
    uint32_t sphere_count, triangle_count, material_count, frame_index;
    uint32_t background_kind; /* 0 sky, 1 black, 2 environment (scene_upload.rs:257-261) */
    uint32_t light_count;
    float total_light_weight;
    uint32_t _pad0;
    uint32_t environment_width, environment_height;
    float environment_intensity, environment_rotation_degrees;
    uint32_t accumulated_samples;
    uint32_t _pad1[3];
    uint32_t bvh_node_count, bvh_primitive_count;
    uint32_t _pad2[2];
} TH_GpuRenderParams;

/* gpu_types.rs:58-70 GpuCamera (64 B) */
typedef struct TH_GpuCamera {
    float origin[4];
    float lower_left_corner[4];
    float horizontal[4];
    float vertical[4];
} TH_GpuCamera;

/* gpu_types.rs:72-83 GpuSphere (32 B) */
typedef struct TH_GpuSphere {
    float center_radius[4];
    uint32_t material_index;
    uint32_t _pad0[3];
} TH_GpuSphere;

/* gpu_types.rs:85-100 GpuTriangle (64 B) */
typedef struct TH_GpuTriangle {
    float v0[4], v1[4], v2[4];
    uint32_t material_index;
    uint32_t _pad0[3];
} TH_GpuTriangle;

/* gpu_types.rs:102-123 GpuTriangleAttributes (80 B); flags bit0 normals, bit1 uvs */
typedef struct TH_GpuTriangleAttributes {
    float n0[4], n1[4], n2[4];
    float uv0[2], uv1[2], uv2[2];
    uint32_t flags;
    uint32_t _pad0;
} TH_GpuTriangleAttributes;

/* gpu_types.rs:125-143 GpuMaterial (48 B); kind 0 diffuse 1 metal 2 dielectric 3 emissive */
typedef struct TH_GpuMaterial {
    uint32_t kind, texture_offset, texture_width, texture_height;
    float albedo[4];
    float params[4]; /* roughness, ior, emission strength, unused */
} TH_GpuMaterial;

/* gpu_types.rs:145-167 GpuLight (96 B); kind 0 triangle, 1 sphere */
typedef struct TH_GpuLight {
    uint32_t kind, material_index;
    uint32_t _pad0[2];
    float v0[4], v1[4], v2[4];
    float center_radius[4];
    float area_weight_cumulative[4];
} TH_GpuLight;

/* gpu_types.rs:169-184 GpuBvhNode (48 B); primitive_count > 0 <=> leaf */
typedef struct TH_GpuBvhNode {
    float min[4];
    float max[4];
    uint32_t first_or_right;
    uint32_t primitive_count;
    uint32_t _pad[2];
} TH_GpuBvhNode;

/* gpu_types.rs:186-194 GpuPrimitiveRef (8 B); kind 0 sphere, 1 triangle */
typedef struct TH_GpuPrimitiveRef {
    uint32_t kind;
    uint32_t index;
} TH_GpuPrimitiveRef;

/* ------------------------------------------------ renderer-neutral scene */
/* Mirrors toaster-scene's Scene (crates/toaster-scene/src/scene.rs:19-90),
 * Material (material.rs:7-38), Sphere/Triangle/TriangleAttributes (object.rs:7-36)
 * and EnvironmentMap (environment.rs:9-24). */
#define TH_MAT_DIFFUSE 0
#define TH_MAT_METAL 1
#define TH_MAT_DIELECTRIC 2
#define TH_MAT_EMISSIVE 3
#define TH_MAT_TEXTURED_DIFFUSE 4

#define TH_BG_SKY 0
#define TH_BG_BLACK 1
#define TH_BG_ENVIRONMENT 2

typedef struct th_material {
    uint32_t kind;          /* TH_MAT_* */
    float color[3];         /* albedo (diffuse/metal/textured) or emission color */
    float roughness;        /* metal */
    float ior;              /* dielectric */
    float strength;         /* emissive */
    uint32_t texture_index; /* textured diffuse */
} th_material;

typedef struct th_sphere {
    float center[3];
    float radius;
    uint32_t material_index;
} th_sphere;

typedef struct th_triangle {
    float v[3][3];
    uint32_t material_index;
} th_triangle;

typedef struct th_triangle_attributes {
    uint32_t has_normals, has_tex_coords;
    float normals[3][3];
    float tex_coords[3][2];
} th_triangle_attributes;

typedef struct th_texture {
    uint32_t width, height;
    const uint8_t* rgba8; /* width*height*4 bytes, sRGB */
} th_texture;

typedef struct th_environment {
    uint32_t width, height;
    const float* rgb; /* width*height*3 linear radiance, row-major */
    float intensity;
    float rotation_degrees; /* normalised with rem_euclid(360) by the packer */
} th_environment;

typedef struct th_scene_input {
    float camera_position[3], camera_look_at[3], camera_up[3];
    float fov_degrees;
    uint32_t width, height, samples, max_bounces, background; /* TH_BG_* */
    const th_material* materials;
    uint32_t material_count;
    const th_sphere* spheres;
    uint32_t sphere_count;
    const th_triangle* triangles;
    uint32_t triangle_count;
    const th_triangle_attributes* triangle_attributes; /* NULL or triangle_count entries */
    const th_texture* textures;
    uint32_t texture_count;
    const th_environment* environment; /* NULL when absent */
} th_scene_input;

/* ------------------------------------------------------ packed GPU scene */
/* SceneGpuData (scene_upload.rs:17-40): params, camera and 9 flat arrays. */
typedef struct th_scene_desc {
    TH_GpuRenderParams params;
    TH_GpuCamera camera;
    const TH_GpuSphere* spheres;
    const TH_GpuTriangle* triangles;
    const TH_GpuTriangleAttributes* triangle_attributes;
    const TH_GpuMaterial* materials;
    const TH_GpuLight* lights;
    const TH_GpuBvhNode* bvh_nodes;
    const TH_GpuPrimitiveRef* bvh_primitives;
    const uint32_t* texture_pixels;
    const float* environment_pixels; /* float4 per texel: rgb + CDF in w */
    uint32_t sphere_count, triangle_count, material_count, light_count;
    uint32_t bvh_node_count, bvh_primitive_count, texture_pixel_count, environment_pixel_count;
} th_scene_desc;

typedef struct th_packed_scene th_packed_scene; /* opaque, owns the arrays */

/* Packs a scene: validation (scene_upload.rs:84-92), camera (352-368), spheres,
 * triangles, attributes, texture atlas, materials (371-404), power-weighted light
 * list (412-475), median-split BVH (296-340 + crates/toaster-bvh) and the
 * environment importance CDF (environment.rs:245-278 via 231-246).
 * include_static_pixels: bit 0 set packs texels and environment pixels (0 mirrors
 * scene_to_gpu_frame: texture/env omitted); TH_PACK_NO_BVH leaves the BVH arrays empty,
 * for geometry that moved in a scene refitted on the device (TH_CHANGE_REFIT), which the
 * reference does not have (it rebuilds; README.md:159-166 names refitting as its next
 * optimisation). */
#define TH_PACK_NO_BVH 2
int32_t th_pack_scene(const th_scene_input* scene, int32_t include_static_pixels,
                      th_packed_scene** out);
int32_t th_packed_view(const th_packed_scene* packed, th_scene_desc* out);
void th_packed_free(th_packed_scene* packed);

/* Camera basis exactly as make_camera (scene_upload.rs:352-368). */
int32_t th_make_camera(const float position[3], const float look_at[3], const float up[3],
                       float fov_degrees, float aspect, TH_GpuCamera* out);

/* Rebuilds only the BVH of a packed scene from replaced geometry (per-frame
 * rebuild of path_tracer.rs:611-638); counts must be unchanged. */
int32_t th_build_bvh(const TH_GpuSphere* spheres, uint32_t sphere_count,
                     const TH_GpuTriangle* triangles, uint32_t triangle_count,
                     TH_GpuBvhNode* nodes_out, uint32_t node_capacity, uint32_t* node_count,
                     TH_GpuPrimitiveRef* prims_out, uint32_t prim_capacity);

/* ------------------------------------------------------ device runtime */
typedef struct th_context th_context;
typedef struct th_scene th_scene;

/* Random-number mode of a render.
 * TH_RNG_REFERENCE: one PCG stream per pixel shared by all samples of a dispatch,
 *   seeded pcg_hash(idx ^ frame_index*9781) (pathtrace_spheres.wgsl:519); output is
 *   per-pixel comparable with the reference estimator.
 * TH_RNG_PER_SAMPLE: an independent stream per (pixel, sample); same estimator,
 *   statistically equivalent, more parallel work per dispatch. */
#define TH_RNG_REFERENCE 0
#define TH_RNG_PER_SAMPLE 1

/* Kernel organisation of a render. */
#define TH_KERNEL_MEGA 0      /* one fused kernel, one lane per pixel              */
#define TH_KERNEL_WAVEFRONT 1 /* extend / shade / shadow stages with compaction   */

/* Traversal order for closest-hit queries. */
#define TH_TRAVERSE_REFERENCE 0 /* visit order of wgsl:746-792 (right child first) */
#define TH_TRAVERSE_ORDERED 1   /* nearer child first, reference tie-breaking      */
#define TH_TRAVERSE_WIDE 2      /* 4 lanes per path over the 4-wide collapsed tree,
                                   nearer children first, reference tie-breaking   */

typedef struct th_render_options {
    uint32_t rng_mode;       /* TH_RNG_*                                         */
    uint32_t kernel;         /* TH_KERNEL_*                                      */
    uint32_t traversal;      /* TH_TRAVERSE_*                                    */
    uint32_t block_size;     /* 0 = autotuned / default                          */
    uint32_t collect_stats;  /* 1 = count rays / nodes / tests (slower launch)   */
    uint32_t schedule;       /* TH_SCHEDULE_*                                    */
    uint32_t batch_blocks_per_cu; /* th_render_batch: persistent blocks per compute  */
                                  /*   unit pulling from the work queue (0 = 4)     */
    uint32_t reserved;
} th_render_options;

/* Pixel-to-lane assignment of the 4-lane (TH_TRAVERSE_WIDE) megakernel.
 * TH_SCHEDULE_STATIC: a quad of lanes per path, one 4x4 pixel tile per wave (reference
 *   RNG), or G quads per pixel (per-sample RNG).
 * TH_SCHEDULE_DYNAMIC: reference RNG only; persistent quads take the next pixel from a
 *   counter when done (per-sample RNG: as STATIC).
 * TH_SCHEDULE_OCTET: eight lanes per path: one quad traces the path while the other
 *   traces the shadow rays of the path's previous vertex; one 4x2 pixel tile per wave
 *   (reference RNG), or G = min(samples, 8) octets per pixel (per-sample RNG).
 * TH_SCHEDULE_SPECULATIVE: reference RNG: two octets per pixel; one traces the pixel's
 *   next sample, the other the sample after it from the start of the pixel's random
 *   stream that follows a full-length path; kept when it equals the first sample's end
 *   state, restarted otherwise (per-sample RNG: as OCTET).  Faster where a frame leaves
 *   lanes idle (long per-pixel chains), slower where every pixel is busy.
 * TH_SCHEDULE_AUTO: reference RNG: the first render of a scene with given samples,
 *   bounces and background times STATIC, OCTET and SPECULATIVE on that frame (9 extra
 *   launches; an accumulation being folded into is kept) and uses the fastest until the
 *   scene's geometry changes (per-sample RNG or other traversals: as STATIC).
 * All schedules give the same image for the same RNG mode (per-sample RNG: the same
 * group count G, i.e. samples < 16). */
#define TH_SCHEDULE_STATIC 0
#define TH_SCHEDULE_DYNAMIC 1
#define TH_SCHEDULE_OCTET 2
#define TH_SCHEDULE_SPECULATIVE 3
#define TH_SCHEDULE_AUTO 4

typedef struct th_stats {
    double last_kernel_ms;    /* device time of the last th_render (hipEvents)    */
    double total_kernel_ms;   /* summed device time of renders since th_reset_stats */
    uint64_t launches;        /* renders timed into total_kernel_ms               */
    uint64_t paths;           /* camera paths traced by the last render           */
    uint64_t rays;            /* closest-hit + shadow rays cast                   */
    uint64_t closest_rays;
    uint64_t shadow_rays;
    uint64_t nodes_visited;   /* reference-layout nodes whose box was tested      */
    uint64_t primref_tests;   /* leaf primitive slots tested                      */
    uint64_t triangle_tests;
    uint64_t sphere_tests;
    uint64_t closest_hits;    /* closest-hit rays that hit a surface              */
    uint64_t triangle_hits;   /* ... of which on a triangle                       */
    uint64_t lights_scanned;  /* light records read by emitter selection          */
    uint64_t env_lookups;     /* bilinear environment evaluations                 */
    uint64_t env_samples;     /* environment CDF searches                         */
    uint64_t closest_nodes_visited;  /* the node / primitive counts above for the  */
    uint64_t closest_primref_tests;  /*   closest-hit rays alone (the rest are     */
    uint64_t closest_triangle_tests; /*   shadow rays)                             */
    uint64_t closest_sphere_tests;
} th_stats;

int32_t th_device_count(int32_t* count);
/* Identity of a device for benchmark reports: "<name> (<architecture>)" and the HIP
 * runtime and driver versions, each NUL-terminated and truncated to its capacity.
 * Stands in for the adapter information the reference reports (wgpu AdapterInfo,
 * crates/toaster-cli/src/benchmark.rs:64-76: name, backend, device type, driver). */
int32_t th_device_info(int32_t device, char* name, size_t name_cap, char* driver, size_t driver_cap);
int32_t th_context_create(int32_t device, th_context** out);
/* Scenes and guides of a context must be destroyed before the context. */
void th_context_destroy(th_context* ctx);
/* Message of the last failure on this thread (never NULL). */
const char* th_last_error(void);

/* Uploads a packed scene (buffers.rs:117-271).  The library copies every array;
 * the caller keeps ownership.  node_capacity = 0 selects 2N-1 (buffers.rs:205-211). */
int32_t th_scene_create(th_context* ctx, const th_scene_desc* desc, uint32_t node_capacity,
                        th_scene** out);
/* Per-frame update of the arrays whose bit is set (path_tracer.rs:676-720); a bit not
 * defined below is rejected (TH_E_INVALID) before anything is uploaded.  Counts are fixed
 * after create (path_tracer.rs:639-653): texels and environment pixels are never
 * replaced, so a descriptor carrying either must carry the uploaded count (TH_E_MISMATCH
 * otherwise; a per-frame pack without them carries 0) and material texture ranges are
 * checked against the uploaded texels.  Moved spheres or triangles upload the
 * descriptor's tree with them, as path_tracer.rs:706-719 does (TH_CHANGE_BVH implied); a
 * descriptor without a tree for a scene that has one is refused (TH_E_INVALID) unless
 * the update is a refit. */
#define TH_CHANGE_CAMERA 1u
#define TH_CHANGE_SPHERES 2u
#define TH_CHANGE_TRIANGLES 4u
#define TH_CHANGE_LIGHTS 8u
#define TH_CHANGE_BVH 16u
#define TH_CHANGE_MATERIALS 32u
/* With TH_CHANGE_SPHERES and/or TH_CHANGE_TRIANGLES: keep the uploaded tree's primitive
 * order and leaf ranges and recompute every node box on the device from the new
 * primitives (the packer's bounds: triangle vertex min/max -/+ 1e-5, sphere center -/+
 * radius; unions in the packer's order, so unmoved geometry refits to the same boxes bit
 * for bit).  The descriptor's BVH arrays are not read (pack with TH_PACK_NO_BVH).  No
 * counterpart in the reference, which rebuilds the tree whenever geometry moves
 * (docs/architecture.md:118-121); exclusive with TH_CHANGE_BVH.  Closest hits equal a
 * rebuilt tree's except where two primitives tie exactly in distance (the reference's
 * tie order follows the tree). */
#define TH_CHANGE_REFIT 64u
int32_t th_scene_update(th_scene* scene, uint32_t change_mask, const th_scene_desc* desc);
/* How far refits have loosened the tree: the summed surface area of all node boxes (the
 * weight of a SAH cost) after the last refit over that of the tree as last built or
 * uploaded (1 = as built; the same device reduction computes both).  No counterpart in the
 * reference, which rebuilds every frame; a caller rebuilds (TH_CHANGE_BVH) once it passes
 * a bound of its choice. */
int32_t th_scene_refit_quality(const th_scene* scene, float* ratio);
/* The 4-wide traversal structure derived from the uploaded tree: *extracted = primitives
 * tested before the tree by every 4-wide kernel (large primitives whose boxes would
 * inflate their ancestors'; 0-8), *nodes4 = 4-wide nodes.  Either pointer may be null.
 * No counterpart in the reference (its shader walks the uploaded binary tree). */
int32_t th_scene_traversal_info(const th_scene* scene, uint32_t* extracted, uint32_t* nodes4);
void th_scene_destroy(th_scene* scene);

/* One dispatch: params.samples samples per pixel folded into the progressive
 * average with params.accumulated_samples (wgsl:538-547).  Asynchronous. */
int32_t th_render(th_context* ctx, th_scene* scene, const TH_GpuRenderParams* params,
                  const TH_GpuCamera* camera, const th_render_options* options);
/* th_render with options->kernel = TH_KERNEL_WAVEFRONT, allocating nothing: no
 * counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * wavefront's queues, counters and per-path state live in one static device array of
 * 13.0 MiB (65,536 path slots) that every GPU the library runs on holds until the
 * process ends; renders needing more slots (width * height, times the sample groups
 * with per-sample streams: min(samples, 16) rounded down to a power of two) return
 * TH_E_CAPACITY before any device work.  Same image as th_render's wavefront for the
 * same options, folded into the scene's accumulation buffer; max_bounces must be
 * nonzero; no timing or counters.  All calls on one GPU share the array: calls that
 * overlap on one GPU (several contexts or threads) race. */
int32_t th_render_wavefront_static(th_context* ctx, th_scene* scene, const TH_GpuRenderParams* params,
                                   const TH_GpuCamera* camera, const th_render_options* options);
/* Progressive batches replayed from a captured hipGraph (the progressive loop of
 * path_tracer.rs:655-674, 900-923 as one graph per batch shape; no counterpart in the
 * reference, which re-records a dispatch per frame).  _create captures one batch of
 * params->samples samples with the given options -- the megakernel with a fixed schedule
 * (not TH_SCHEDULE_AUTO, not the wavefront, no statistics) reading the frame seed and
 * accumulated sample count from device memory, then a kernel advancing both -- starting
 * at params->frame_index and params->accumulated_samples.  Each _replay frame launches the
 * graph once: batch k renders with frame_index = start + k and folds into the scene's
 * accumulation over accumulated = start + k * samples, exactly as th_render would.
 * Asynchronous; _state returns the next frame index and accumulated count.  The graph
 * holds the scene as captured: after any th_scene_update, _replay returns
 * TH_E_MISMATCH (capture a new graph).  Destroy graphs before their scene. */
typedef struct th_progressive_graph th_progressive_graph;
int32_t th_progressive_graph_create(th_context* ctx, th_scene* scene, const TH_GpuRenderParams* params,
                                    const TH_GpuCamera* camera, const th_render_options* options,
                                    th_progressive_graph** out);
int32_t th_progressive_graph_replay(th_progressive_graph* graph, uint32_t frames);
int32_t th_progressive_graph_state(const th_progressive_graph* graph, uint32_t* frame_index,
                                   uint32_t* accumulated_samples);
void th_progressive_graph_destroy(th_progressive_graph* graph);
/* Blocks, then copies width*height float4 linear-HDR pixels (readback.rs:8-25). */
int32_t th_readback(th_context* ctx, th_scene* scene, float* rgba, size_t bytes);
/* Overwrites the accumulation buffer (progressive resume from a checkpoint). */
int32_t th_upload_accumulation(th_context* ctx, th_scene* scene, const float* rgba, size_t bytes);
int32_t th_synchronize(th_context* ctx);
/* Waits for pending renders, then reports timing and (if collected) counters.  Every
 * launch is timed by an event pair folded in launch order; a caller that never reads the
 * statistics still holds at most 4,096 unfolded pairs (the oldest 2,048 are folded, after
 * waiting for them, when the bound is reached). */
int32_t th_get_stats(const th_context* ctx, th_stats* out);
int32_t th_reset_stats(th_context* ctx);

/* Closest hit of every pixel's first camera ray (first two draws of the pixel's
 * reference stream): per pixel the leaf slot index (UINT32_MAX = miss) and the hit
 * distance (3.4028235e38 = miss).  Parity query for traversal + intersection. */
int32_t th_primary_hits(th_context* ctx, th_scene* scene, const TH_GpuRenderParams* params,
                        const TH_GpuCamera* camera, uint32_t traversal, uint32_t* slot_out,
                        float* distance_out, size_t count);

/* The random stream of each listed pixel as the render kernels seed it, evaluated on the
 * device: TH_RNG_REFERENCE seeds pcg_hash(idx ^ frame_index*9781) (wgsl:519),
 * TH_RNG_PER_SAMPLE seeds pcg_hash((idx ^ frame_index*9781) + sample*0x9E3779B9); then
 * `draws` draws of random_f32 (wgsl:189-199), per pixel the state after each draw and
 * the drawn value (count x draws each, row per pixel; count x draws <= 2^24).  Parity
 * query for the generator the kernels use; no counterpart in the reference. */
int32_t th_rng_draws(th_context* ctx, const uint32_t* pixels, uint32_t count, uint32_t frame_index,
                     uint32_t rng_mode, uint32_t sample, uint32_t draws, uint32_t* states_out,
                     float* values_out);

/* ---------------------------------------------------------- batching */
/* One entry of a batched launch: a scene (or one animation frame uploaded as its own
 * scene), its dispatch params/camera and its launch shape.  No counterpart in the
 * reference (one scene per dispatch there). */
typedef struct th_batch_item {
    th_scene* scene;
    TH_GpuRenderParams params;
    TH_GpuCamera camera;
    uint32_t quads_per_pixel; /* per-sample RNG: 1, 2, 4, 8, 16 quads share a pixel; 0 = default */
    uint32_t cached_nodes;    /* 4-wide nodes copied to LDS (<= 85); UINT32_MAX = default   */
    uint32_t tile_width;      /* pixels per row of a wave's tile: a power of two dividing   */
                              /*   16 / quads_per_pixel (the tile is that many wide and the */
                              /*   rest high); 0 = default (4, 2 or 1)                      */
    uint32_t reserved;
} th_batch_item;

typedef struct th_tuning {
    uint32_t quads_per_pixel;
    uint32_t cached_nodes;
    float best_ms;            /* median device time of the chosen shape                      */
    uint32_t tile_width;
} th_tuning;

/* Renders every item in one persistent launch (work items handed out by an atomic
 * counter, most expensive scenes first), replayed from a captured hipGraph when the
 * batch shape repeats.  Asynchronous; rng_mode of options applies to all items. */
int32_t th_render_batch(th_context* ctx, const th_batch_item* items, uint32_t count,
                        const th_render_options* options);
/* Times the launch-shape candidates for one scene on the device and returns the
 * fastest (also remembered by the scene): quads per pixel (per-sample RNG), then the
 * treelet nodes cached in LDS (0, 21, 85), then the wave's tile width, each chosen with
 * the others held (coordinate descent).  Every candidate is launched once untimed (its
 * graph is captured then) and timed as the median of three launches. */
int32_t th_autotune(th_context* ctx, th_scene* scene, const TH_GpuRenderParams* params,
                    const TH_GpuCamera* camera, const th_render_options* options, th_tuning* out);
/* Times a whole batch (items with their launch shapes) at 2, 4, 6 and 8 persistent blocks
 * per compute unit and returns the fastest count (for options->batch_blocks_per_cu) and
 * its median time; no counterpart in the reference. */
int32_t th_autotune_batch(th_context* ctx, const th_batch_item* items, uint32_t count,
                          const th_render_options* options, uint32_t* blocks_per_cu, float* best_ms);

/* ---------------------------------------------------------- path guiding */
/* Online-trained guiding distribution (no counterpart in the reference): an MLP
 * 48 -> 64 -> 64 -> 64 -> 40 (ReLU, BF16 weights, FP32 accumulation on MFMA) mapping the
 * encoded (position, normal, outgoing direction) of a diffuse vertex to an 8-lobe vMF
 * mixture over incident directions, combined with cosine sampling by one-sample MIS. */
#define TH_GUIDE_FEATURES 48
#define TH_GUIDE_OUTPUTS 40
#define TH_GUIDE_PARAMS 14056
typedef struct th_guide th_guide;
int32_t th_guide_create(th_context* ctx, uint32_t seed, th_guide** out);
void th_guide_destroy(th_guide* guide);
int32_t th_guide_get_params(th_guide* guide, float* out, uint32_t count);
int32_t th_guide_set_params(th_guide* guide, const float* params, uint32_t count);
int32_t th_guide_set_hyper(th_guide* guide, float lr, float beta1, float beta2, float eps);
/* Network outputs (count x 40) for host-supplied features (count x 48). */
int32_t th_guide_forward(th_guide* guide, const float* features, uint32_t count, float* out);
/* Network outputs (count x 40, rounded to binary16 as the renderer stores them) of the
 * guided renderer's fused inference kernel (the encoding computed on the device, then the
 * four layers) for host-supplied diffuse
 * vertices: count x 9 floats (position, normal, outgoing direction), with the position
 * normalised to [0,1]^3 by the box [bmin, bmax] as the renderer normalises it by the
 * scene bounds.  For checking that kernel; the renderer calls it on its own queues. */
int32_t th_guide_infer_vertices(th_guide* guide, const float* vertices, uint32_t count,
                                const float bmin[3], const float bmax[3], float* out);
/* One Adam step (update != 0) or the gradient only on host-supplied samples: loss
 * -sum_b w_b log q_b(dir_b) / sum_b w_b; loss and gradient (14056 floats) optional. */
int32_t th_guide_train_batch(th_guide* guide, const float* features, const float* dirs,
                             const float* weights, uint32_t count, int32_t update, float* loss,
                             float* grad);
/* One v_mfma_f32_16x16x32_bf16 with this library's fragment mapping: C = A(16x32) B(32x16). */

/* Guided rendering: `params->samples` passes of one sample per pixel (per-sample random
 * streams, TH_RNG_PER_SAMPLE), folded into the scene's accumulation buffer like
 * th_render.  At diffuse vertices the continuation direction comes from the guide with
 * probability alpha and from the reference's cosine lobe otherwise; the estimator
 * weights by the one-sample MIS density alpha q + (1 - alpha) cos/pi and uses the same
 * density in the environment MIS, so it stays unbiased for any network.  With train != 0
 * every pass records diffuse vertices (features, sampled direction, and the weight
 * f cos L_i / pdf once the path completes) and runs Adam steps on them; alpha is treated
 * as 0 until the guide has taken its first step (unless TH_GUIDED_ALPHA_FROM_START).  Every schedule depends only on the
 * guide's step count, so the estimator of each pass is fixed before the pass runs.  With alpha = 0 and train = 0 the image
 * equals th_render's TH_RNG_PER_SAMPLE image with one sample group. */
typedef struct th_guided_options {
    float alpha;              /* guide probability at diffuse vertices, [0, 0.95]      */
    uint32_t train;           /* 1 = record samples and train after every pass        */
    uint32_t steps_per_pass;  /* Adam steps per pass (0 = 1)                           */
    uint32_t batch;           /* samples per step (0 = 2^18)                           */
    uint32_t flags;           /* TH_GUIDED_*                                           */
    uint32_t warmup_steps;    /* while the guide has taken fewer Adam steps than this,  */
    uint32_t warmup_rate;     /*   take warmup_rate steps per pass instead             */
    uint32_t alpha_ramp;      /* alpha grows linearly over the guide's first alpha_ramp */
                              /*   steps (0 = full alpha after the first step)         */
    uint32_t train_every;     /* after the warm-up, record and train on every n-th pass */
                              /*   of the call only (0 or 1 = every pass)              */
    uint32_t reserved[3];
} th_guided_options;
/* At every diffuse vertex but a path's last (also while alpha is 0, so the estimator
 * with and without the guide can be compared), estimate environment light by the
 * continuation alone (no environment sample and shadow ray; an escaping continuation
 * counts with weight 1).  Without this flag the reference's environment sample and MIS
 * are kept.  A path's last diffuse vertex, whose continuation is never traced, always
 * takes the reference's environment sample weighted against the cosine lobe
 * (wgsl:496-503), so every alpha and flag estimates the reference's image. */
#define TH_GUIDED_NO_ENV_NEE 1u
/* Apply alpha from the first pass even while the guide has taken no Adam step, so an
 * untrained network (as created or as set by th_guide_set_params) can be measured;
 * with alpha_ramp > 0 the ramp still starts from 0.  Other flag bits are rejected. */
#define TH_GUIDED_ALPHA_FROM_START 2u
/* Time every inference and training launch of the network with a HIP event pair
 * (infer_ms, train_ms of the statistics); for measurement, it adds two event records per
 * launch. */
#define TH_GUIDED_TIME_NETWORK 4u

typedef struct th_guided_stats {
    uint64_t diffuse_vertices; /* deferred diffuse vertices over the call               */
    uint64_t records;          /* training samples recorded                            */
    uint32_t train_steps;      /* Adam steps taken by this call                        */
    float last_loss;           /* loss of the last step                                */
    uint64_t inferred_rows;    /* network evaluations (one per guided diffuse vertex)  */
    uint64_t trained_samples;  /* samples in the Adam steps taken                      */
    float infer_ms, train_ms;  /* device time of inference / training launches         */
                               /*   (TH_GUIDED_TIME_NETWORK; 0 otherwise)              */
    uint64_t reserved[2];
} th_guided_stats;

int32_t th_render_guided(th_context* ctx, th_scene* scene, th_guide* guide,
                         const TH_GpuRenderParams* params, const TH_GpuCamera* camera,
                         const th_guided_options* options, th_guided_stats* stats);
int32_t th_guide_mfma_selftest(th_context* ctx, const float* a, const float* b, float* c);
/* One v_mfma_f32_32x32x16_bf16 with the fragment mapping of the renderer's inference
 * kernel: C (32x32) = A (32x16) B (16x32), inputs rounded to BF16.  No counterpart in the
 * reference. */
int32_t th_guide_mfma32_selftest(th_context* ctx, const float* a, const float* b, float* c);
/* The device's float -> BF16 rounding (the gfx950 conversion instruction, used by every
 * network kernel) against integer round-to-nearest-even over all 2^32 float patterns:
 * *mismatches counts non-NaN inputs that differ (expected 0), *nan_mismatches (may be
 * null) NaN inputs whose bits differ.  No counterpart in the reference (its renderer has
 * no network); a check of the rounding the guide's oracle emulates. */
int32_t th_bf16_selftest(th_context* ctx, uint64_t* mismatches, uint64_t* nan_mismatches);

/* ---- measured device ceilings (diagnostics; no counterpart in the reference, whose
 * benchmarks report no roofline) ----
 * th_measure_copy: a streaming copy of `bytes` (a nonzero multiple of 16) between two
 * device buffers it allocates (read_only != 0: a streaming read of one), `reps` times
 * after one untimed round; *gbs = bytes read + written per second / 1e9 (the HBM
 * bandwidth a kernel can reach).
 * th_measure_mfma: `iters` rounds of four independent BF16 MFMAs per wave on every SIMD
 * (`waves_per_simd` waves each, 1..8), shape 0 = v_mfma_f32_16x16x32_bf16 (training
 * kernels), 1 = v_mfma_f32_32x32x16_bf16 (inference kernel), after one untimed launch;
 * *tflops = dense FLOP/s / 1e12; *exact (may be null) = 1 when every accumulator holds
 * the exact all-ones product sum. */
int32_t th_measure_copy(th_context* ctx, uint64_t bytes, uint32_t reps, uint32_t read_only, double* gbs);
int32_t th_measure_mfma(th_context* ctx, uint32_t shape, uint32_t waves_per_simd, uint32_t iters, double* tflops,
                        uint32_t* exact);

/* Linear HDR float4 -> RGBA8 with exposure, ACES fitted and sRGB OETF
 * (image_output.rs:59-92, color.rs:23-75).  Host implementation. */
int32_t th_display_rgba8(const float* rgba, uint32_t width, uint32_t height, float exposure_stops,
                         uint8_t* out, size_t out_bytes);
/* The same conversion on the device from a scene's accumulation buffer. */
int32_t th_display_rgba8_device(th_context* ctx, th_scene* scene, float exposure_stops,
                                uint8_t* out, size_t out_bytes);

#ifdef __cplusplus
} /* extern "C" */

static_assert(sizeof(TH_GpuRenderParams) == 96, "GpuRenderParams");
static_assert(sizeof(TH_GpuCamera) == 64, "GpuCamera");
static_assert(sizeof(TH_GpuSphere) == 32, "GpuSphere");
static_assert(sizeof(TH_GpuTriangle) == 64, "GpuTriangle");
static_assert(sizeof(TH_GpuTriangleAttributes) == 80, "GpuTriangleAttributes");
static_assert(sizeof(TH_GpuMaterial) == 48, "GpuMaterial");
static_assert(sizeof(TH_GpuLight) == 96, "GpuLight");
static_assert(sizeof(TH_GpuBvhNode) == 48, "GpuBvhNode");
static_assert(sizeof(TH_GpuPrimitiveRef) == 8, "GpuPrimitiveRef");
#endif

#endif /* TOASTER_HIP_H */
