/* oracle.h -- CPU restatement of the reference path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links or calls this code;
 * only tests/, the smoke check and the CPU-baseline leg of bench.py load it, as the
 * checker of the HIP path.
 *
 * What is restated (reference file:line):
 *   packing      crates/toaster-gpu/src/scene_upload.rs:79-475 (materials, lights,
 *                camera, env packing), crates/toaster-core/src/light.rs:6-13
 *   BVH          crates/toaster-bvh/src/{aabb.rs:7-111, bvh.rs:77-154, flatten.rs:54-97}
 *   environment  crates/toaster-scene/src/environment.rs:245-278 (CDF build)
 *   estimator    shaders/pathtrace_spheres.wgsl:189-937 ("pcg" mode: per-pixel PCG
 *                stream, draw order and formulas of the reference GPU kernel)
 *   CPU renderer crates/toaster-cpu/src/{integrator.rs:47-389, intersect.rs:28-338,
 *                light.rs:126-171}, toaster-scene/src/environment.rs:132-242,
 *                toaster-bvh/src/aabb.rs:74-111 ("toaster_cpu" mode, or_render_cpu: its
 *                recursion and draw order, sampling routines and traversal, on libm and
 *                a SplitMix64 stream per pixel; for converged comparison)
 *   display      crates/toaster-core/src/color.rs:23-75, toaster-gpu/src/image_output.rs:59-92
 *
 * Pinning (tests/test_oracle_known_answers.py, tests/test_reference_scenes.py): the
 * reference's own known-answer tests for intersection, AABB, BVH, lights, environment,
 * integrator, camera, packing and display, plus closed forms (BVH == brute force, env pdf
 * integrates to 1, white furnace, transcendental ulp bounds).
 * The PCG generator (wgsl:189-199, 519) is pinned by known answers computed from the
 * shader's text (tests/golden/make_pcg_vectors.py, tests/test_pcg_vectors.py).  The two
 * estimator modes are pinned against each other at convergence (tests/
 * test_oracle_cpu_mode.py), the toaster_cpu mode also by closed forms.
 * Parity unpinned:
 *   or_render_ex rng_mode 1 (per-sample streams) -- this project's extension, no
 *     reference counterpart; it shares every formula with rng_mode 0.
 *   toaster-cpu's StdRng (rand 0.9.4 / rand_chacha 0.9.0) stream -- not restated (the
 *     toaster_cpu mode agrees with the CPU renderer in distribution, not per pixel).
 */
#ifndef TOASTER_ORACLE_H
#define TOASTER_ORACLE_H
#include "../include/toaster_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_scene or_scene;

typedef struct or_counters {
    uint64_t paths, closest_rays, shadow_rays;
    uint64_t nodes_visited, primref_tests, triangle_tests, sphere_tests;
    uint64_t closest_hits, triangle_hits, lights_scanned, env_lookups, env_samples;
} or_counters;

typedef struct or_hit {
    float distance;
    float point[3];
    float normal[3];
    uint32_t front_face;
    uint32_t material_index;
    float tex_coord[2];
    uint32_t slot; /* BVH primitive slot (or kind<<31|index for linear); UINT32_MAX = miss */
} or_hit;

/* Packs a scene exactly as scene_to_gpu does; returns 0 or a negative code. */
int32_t or_pack(const th_scene_input* scene, or_scene** out);
int32_t or_view(const or_scene* s, th_scene_desc* out);
void or_free(or_scene* s);
const char* or_last_error(void);

/* Renders rows [row_begin, row_end) of one dispatch in pcg mode into rgba (float4 per
 * pixel, whole image; rows outside the range untouched).  threads <= 1: serial. */
int32_t or_render(const or_scene* s, const TH_GpuRenderParams* params, const TH_GpuCamera* camera,
                  uint32_t row_begin, uint32_t row_end, int32_t threads, float* rgba,
                  or_counters* counters);

/* As or_render; rng_mode 1 = per-sample streams pcg_hash((idx ^ frame*9781) + s*0x9E3779B9)
 * with samples split over `group` partial sums combined as a butterfly. */
int32_t or_render_ex(const or_scene* s, const TH_GpuRenderParams* params, const TH_GpuCamera* camera,
                     uint32_t row_begin, uint32_t row_end, int32_t threads, uint32_t rng_mode,
                     uint32_t group, float* rgba, or_counters* counters);

/* First camera ray of each pixel (draws 1-2 of its stream): closest hit slot + t. */
int32_t or_primary_hits(const or_scene* s, const TH_GpuRenderParams* params,
                        const TH_GpuCamera* camera, uint32_t* slot_out, float* distance_out);

/* Closest hit of one ray through the packed scene (use_bvh=0: linear scan). */
int32_t or_closest(const or_scene* s, const float origin[3], const float direction[3],
                   int32_t use_bvh, or_hit* out);

/* Single-primitive queries in the reference's CPU form (intersect.rs:28-148):
 * inclusive [tmin, tmax]; return 1 on hit. */
int32_t or_intersect_sphere(const float origin[3], const float direction[3], const float center[3],
                            float radius, float tmin, float tmax, or_hit* out);
int32_t or_intersect_triangle(const float origin[3], const float direction[3],
                              const float v[9], const TH_GpuTriangleAttributes* attributes,
                              float tmin, float tmax, or_hit* out);
/* aabb.rs:74-111 with an explicit inverse direction. */
int32_t or_aabb_intersects(const float mn[3], const float mx[3], const float origin[3],
                           const float inverse_direction[3], float tmin, float tmax);

/* Environment helpers over a packed scene (wgsl:280-381). */
int32_t or_env_lookup(const or_scene* s, const float direction[3], float rgb[3]);
float or_env_pdf(const or_scene* s, const float direction[3]);
/* Importance sample from explicit selection/jitter values. Returns 1 if pdf > 0. */
int32_t or_env_sample(const or_scene* s, float selection, float jitter_u, float jitter_v,
                      float direction[3], float rgb[3], float* pdf);

/* Material helpers (wgsl:919-937 / integrator.rs:352-369). */
float or_reflectance(float cosine, float ratio);
void or_refract(const float d[3], const float n[3], float ratio, float out[3]);
float or_power_heuristic(float a, float b);

/* Radiance of one ray with an explicit PCG state (wgsl:805-916); state is updated. */
int32_t or_ray_color(const or_scene* s, const TH_GpuRenderParams* params, const float origin[3],
                     const float direction[3], uint32_t* rng_state, float rgb[3]);

/* PCG helpers (wgsl:189-207). */
uint32_t or_pcg_hash(uint32_t x);
float or_sin(float x);
float or_cos(float x);
float or_atan2(float y, float x);
float or_acos(float x);

/* Display conversion (color.rs:42-55, image_output.rs:59-92). */
int32_t or_display_rgba8(const float* rgba, uint32_t width, uint32_t height, float exposure_stops,
                         uint8_t* out);

#ifdef __cplusplus
}
#endif
#endif
