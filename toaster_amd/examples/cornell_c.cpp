// A C/C++ caller of libtoaster_hip.so through include/toaster_hip.h alone: the Cornell
// box of toaster_amd.scenes.cornell packed, uploaded, rendered with the reference RNG,
// read back and converted to RGBA8, written as a binary PPM.  What a host program
// (or a Rust crate binding the same entry points) does in place of toaster-gpu's
// renderer: scene_to_gpu, buffer creation, dispatch, readback, pixels_to_rgba_image.
// With MOVED.ppm it then moves the spheres, packs that frame without a tree, refits the
// uploaded tree on the device (TH_CHANGE_REFIT) and renders it with the automatically
// chosen schedule.
//
//   cornell_c OUT.ppm [width height samples] [MOVED.ppm]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "toaster_hip.h"

namespace {

int fail(const char* what) {
  std::fprintf(stderr, "cornell_c: %s: %s\n", what, th_last_error());
  return 1;
}

// linear HDR of the scene's accumulation -> RGBA8 -> binary PPM
int write_ppm(th_context* ctx, th_scene* scene, uint32_t width, uint32_t height, const char* path) {
  std::vector<float> hdr((size_t)width * height * 4);
  if (th_readback(ctx, scene, hdr.data(), hdr.size() * sizeof(float))) return fail("readback");
  std::vector<uint8_t> rgba((size_t)width * height * 4);
  if (th_display_rgba8(hdr.data(), width, height, 0.0f, rgba.data(), rgba.size())) return fail("display");
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    std::fprintf(stderr, "cornell_c: cannot write %s\n", path);
    return 1;
  }
  std::fprintf(f, "P6\n%u %u\n255\n", width, height);
  for (size_t i = 0; i < (size_t)width * height; ++i) std::fwrite(&rgba[4 * i], 1, 3, f);
  std::fclose(f);
  return 0;
}

th_triangle tri(const float a[3], const float b[3], const float c[3], uint32_t m) {
  th_triangle t;
  std::memcpy(t.v[0], a, sizeof(float) * 3);
  std::memcpy(t.v[1], b, sizeof(float) * 3);
  std::memcpy(t.v[2], c, sizeof(float) * 3);
  t.material_index = m;
  return t;
}

// two triangles (p0 p1 p2), (p0 p2 p3)
void quad(std::vector<th_triangle>& out, const float p0[3], const float p1[3], const float p2[3],
          const float p3[3], uint32_t m) {
  out.push_back(tri(p0, p1, p2, m));
  out.push_back(tri(p0, p2, p3, m));
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: cornell_c OUT.ppm [width height samples]\n");
    return 2;
  }
  const uint32_t width = argc > 4 ? (uint32_t)std::atoi(argv[2]) : 320u;
  const uint32_t height = argc > 4 ? (uint32_t)std::atoi(argv[3]) : 180u;
  const uint32_t samples = argc > 4 ? (uint32_t)std::atoi(argv[4]) : 4u;
  const char* moved_path = argc == 3 ? argv[2] : (argc > 5 ? argv[5] : nullptr);

  const th_material mats[] = {
      {TH_MAT_DIFFUSE, {0.73f, 0.73f, 0.73f}, 0.0f, 0.0f, 0.0f, 0},
      {TH_MAT_DIFFUSE, {0.65f, 0.05f, 0.05f}, 0.0f, 0.0f, 0.0f, 0},
      {TH_MAT_DIFFUSE, {0.12f, 0.45f, 0.15f}, 0.0f, 0.0f, 0.0f, 0},
      {TH_MAT_EMISSIVE, {1.0f, 1.0f, 1.0f}, 0.0f, 0.0f, 15.0f, 0},
      {TH_MAT_DIELECTRIC, {1.0f, 1.0f, 1.0f}, 0.0f, 1.5f, 0.0f, 0},
  };
  const float x0 = -1.0f, x1 = 1.0f, y0 = 0.0f, y1 = 2.0f, z0 = -2.0f, z1 = 0.0f;
  std::vector<th_triangle> tris;
  {
    const float a[3] = {x0, y0, z1}, b[3] = {x0, y0, z0}, c[3] = {x1, y0, z0}, d[3] = {x1, y0, z1};
    quad(tris, a, b, c, d, 0);  // floor
  }
  {
    const float a[3] = {x0, y1, z1}, b[3] = {x1, y1, z1}, c[3] = {x1, y1, z0}, d[3] = {x0, y1, z0};
    quad(tris, a, b, c, d, 0);  // ceiling
  }
  {
    const float a[3] = {x0, y0, z0}, b[3] = {x0, y1, z0}, c[3] = {x1, y1, z0}, d[3] = {x1, y0, z0};
    quad(tris, a, b, c, d, 0);  // back
  }
  {
    const float a[3] = {x0, y0, z1}, b[3] = {x0, y1, z1}, c[3] = {x0, y1, z0}, d[3] = {x0, y0, z0};
    quad(tris, a, b, c, d, 1);  // left
  }
  {
    const float a[3] = {x1, y0, z0}, b[3] = {x1, y1, z0}, c[3] = {x1, y1, z1}, d[3] = {x1, y0, z1};
    quad(tris, a, b, c, d, 2);  // right
  }
  {
    const float s = 0.125f, h = 1.99f;
    const float a[3] = {-s, h, -1 - s}, b[3] = {s, h, -1 - s}, c[3] = {s, h, -1 + s}, d[3] = {-s, h, -1 + s};
    quad(tris, a, b, c, d, 3);  // light
  }
  const th_sphere spheres[] = {{{-0.45f, 0.3f, -1.2f}, 0.3f, 0}, {{0.45f, 0.3f, -0.8f}, 0.3f, 4}};

  th_scene_input in;
  std::memset(&in, 0, sizeof in);
  const float pos[3] = {0.0f, 1.0f, 3.3f}, look[3] = {0.0f, 1.0f, -1.0f}, up[3] = {0.0f, 1.0f, 0.0f};
  std::memcpy(in.camera_position, pos, sizeof pos);
  std::memcpy(in.camera_look_at, look, sizeof look);
  std::memcpy(in.camera_up, up, sizeof up);
  in.fov_degrees = 40.0f;
  in.width = width;
  in.height = height;
  in.samples = samples;
  in.max_bounces = 4;
  in.background = TH_BG_BLACK;
  in.materials = mats;
  in.material_count = 5;
  in.spheres = spheres;
  in.sphere_count = 2;
  in.triangles = tris.data();
  in.triangle_count = (uint32_t)tris.size();

  th_packed_scene* packed = nullptr;
  if (th_pack_scene(&in, 1, &packed)) return fail("pack");
  th_scene_desc desc;
  if (th_packed_view(packed, &desc)) return fail("packed view");
  th_context* ctx = nullptr;
  if (th_context_create(0, &ctx)) return fail("context");
  th_scene* scene = nullptr;
  if (th_scene_create(ctx, &desc, 0, &scene)) return fail("scene");

  th_render_options opt;
  std::memset(&opt, 0, sizeof opt);
  opt.rng_mode = TH_RNG_REFERENCE;
  opt.kernel = TH_KERNEL_MEGA;
  opt.traversal = TH_TRAVERSE_WIDE;
  opt.schedule = TH_SCHEDULE_OCTET;
  TH_GpuRenderParams params = desc.params;
  params.frame_index = 0;
  params.accumulated_samples = 0;
  if (th_render(ctx, scene, &params, &desc.camera, &opt)) return fail("render");
  if (write_ppm(ctx, scene, width, height, argv[1])) return 1;
  th_stats st;
  if (th_get_stats(ctx, &st)) return fail("stats");
  std::printf("cornell %ux%u, %u spp: %.4f ms, %llu paths\n", width, height, samples, st.last_kernel_ms,
              (unsigned long long)st.paths);

  if (moved_path) {
    // the next frame of an animation: both spheres moved, the tree refitted on the device
    th_sphere moved[2] = {spheres[0], spheres[1]};
    moved[0].center[0] = spheres[0].center[0] + 0.1f;
    moved[1].center[1] = spheres[1].center[1] + 0.1f;
    th_scene_input in2 = in;
    in2.spheres = moved;
    th_packed_scene* packed2 = nullptr;
    if (th_pack_scene(&in2, 1 | TH_PACK_NO_BVH, &packed2)) return fail("pack moved frame");
    th_scene_desc desc2;
    if (th_packed_view(packed2, &desc2)) return fail("packed view");
    if (th_scene_update(scene, TH_CHANGE_SPHERES | TH_CHANGE_LIGHTS | TH_CHANGE_REFIT, &desc2))
      return fail("refit");
    th_render_options auto_opt = opt;
    auto_opt.schedule = TH_SCHEDULE_AUTO;
    // the render parameters keep the uploaded tree's counts (the moved frame has no tree)
    if (th_render(ctx, scene, &params, &desc2.camera, &auto_opt)) return fail("render moved frame");
    if (write_ppm(ctx, scene, width, height, moved_path)) return 1;
    std::printf("moved frame refitted and rendered\n");
    th_packed_free(packed2);
  }

  th_scene_destroy(scene);
  th_context_destroy(ctx);
  th_packed_free(packed);
  return 0;
}
