"""ctypes mirror of include/toaster_hip.h and the loader of libtoaster_hip.so.

The library is loaded from this package directory only; a missing library is an
error (there is no fallback path).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtoaster_hip.so")

TH_OK = 0
TH_E_INVALID = -1
TH_E_DEVICE = -2
TH_E_NO_DEVICE = -3
TH_E_MISMATCH = -4
TH_E_CAPACITY = -5
TH_E_UNSUPPORTED = -6

TH_MAT_DIFFUSE, TH_MAT_METAL, TH_MAT_DIELECTRIC, TH_MAT_EMISSIVE, TH_MAT_TEXTURED_DIFFUSE = 0, 1, 2, 3, 4
TH_BG_SKY, TH_BG_BLACK, TH_BG_ENVIRONMENT = 0, 1, 2
TH_RNG_REFERENCE, TH_RNG_PER_SAMPLE = 0, 1
TH_KERNEL_MEGA, TH_KERNEL_WAVEFRONT = 0, 1
TH_TRAVERSE_REFERENCE, TH_TRAVERSE_ORDERED, TH_TRAVERSE_WIDE = 0, 1, 2
TH_SCHEDULE_STATIC, TH_SCHEDULE_DYNAMIC, TH_SCHEDULE_OCTET, TH_SCHEDULE_SPECULATIVE, TH_SCHEDULE_AUTO = 0, 1, 2, 3, 4

TH_GUIDE_FEATURES, TH_GUIDE_OUTPUTS, TH_GUIDE_PARAMS = 48, 40, 14056
TH_GUIDED_NO_ENV_NEE = 1
TH_GUIDED_ALPHA_FROM_START = 2
TH_GUIDED_TIME_NETWORK = 4

TH_CHANGE_CAMERA = 1
TH_CHANGE_SPHERES = 2
TH_CHANGE_TRIANGLES = 4
TH_CHANGE_LIGHTS = 8
TH_CHANGE_BVH = 16
TH_CHANGE_MATERIALS = 32
TH_CHANGE_REFIT = 64
TH_PACK_NO_BVH = 2

u32, f32 = C.c_uint32, C.c_float


class GpuRenderParams(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("samples", u32), ("max_bounces", u32),
                ("sphere_count", u32), ("triangle_count", u32), ("material_count", u32),
                ("frame_index", u32), ("background_kind", u32), ("light_count", u32),
                ("total_light_weight", f32), ("_pad0", u32), ("environment_width", u32),
                ("environment_height", u32), ("environment_intensity", f32),
                ("environment_rotation_degrees", f32), ("accumulated_samples", u32),
                ("_pad1", u32 * 3), ("bvh_node_count", u32), ("bvh_primitive_count", u32),
                ("_pad2", u32 * 2)]


class GpuCamera(C.Structure):
    _fields_ = [("origin", f32 * 4), ("lower_left_corner", f32 * 4), ("horizontal", f32 * 4),
                ("vertical", f32 * 4)]


# numpy views of the record arrays (gpu_types.rs:8-195)
SPHERE_DTYPE = np.dtype([("center_radius", "<f4", 4), ("material_index", "<u4"), ("_pad0", "<u4", 3)])
TRIANGLE_DTYPE = np.dtype([("v0", "<f4", 4), ("v1", "<f4", 4), ("v2", "<f4", 4),
                           ("material_index", "<u4"), ("_pad0", "<u4", 3)])
ATTR_DTYPE = np.dtype([("n0", "<f4", 4), ("n1", "<f4", 4), ("n2", "<f4", 4), ("uv0", "<f4", 2),
                       ("uv1", "<f4", 2), ("uv2", "<f4", 2), ("flags", "<u4"), ("_pad0", "<u4")])
MATERIAL_DTYPE = np.dtype([("kind", "<u4"), ("texture_offset", "<u4"), ("texture_width", "<u4"),
                           ("texture_height", "<u4"), ("albedo", "<f4", 4), ("params", "<f4", 4)])
LIGHT_DTYPE = np.dtype([("kind", "<u4"), ("material_index", "<u4"), ("_pad0", "<u4", 2),
                        ("v0", "<f4", 4), ("v1", "<f4", 4), ("v2", "<f4", 4),
                        ("center_radius", "<f4", 4), ("area_weight_cumulative", "<f4", 4)])
NODE_DTYPE = np.dtype([("min", "<f4", 4), ("max", "<f4", 4), ("first_or_right", "<u4"),
                       ("primitive_count", "<u4"), ("_pad", "<u4", 2)])
PRIMREF_DTYPE = np.dtype([("kind", "<u4"), ("index", "<u4")])
for _dt, _size in ((SPHERE_DTYPE, 32), (TRIANGLE_DTYPE, 64), (ATTR_DTYPE, 80), (MATERIAL_DTYPE, 48),
                   (LIGHT_DTYPE, 96), (NODE_DTYPE, 48), (PRIMREF_DTYPE, 8)):
    assert _dt.itemsize == _size


class Material(C.Structure):
    _fields_ = [("kind", u32), ("color", f32 * 3), ("roughness", f32), ("ior", f32),
                ("strength", f32), ("texture_index", u32)]


class Sphere(C.Structure):
    _fields_ = [("center", f32 * 3), ("radius", f32), ("material_index", u32)]


class Triangle(C.Structure):
    _fields_ = [("v", (f32 * 3) * 3), ("material_index", u32)]


class TriangleAttributes(C.Structure):
    _fields_ = [("has_normals", u32), ("has_tex_coords", u32), ("normals", (f32 * 3) * 3),
                ("tex_coords", (f32 * 2) * 3)]


class Texture(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("rgba8", C.POINTER(C.c_uint8))]


class Environment(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("rgb", C.POINTER(f32)), ("intensity", f32),
                ("rotation_degrees", f32)]


class SceneInput(C.Structure):
    _fields_ = [("camera_position", f32 * 3), ("camera_look_at", f32 * 3), ("camera_up", f32 * 3),
                ("fov_degrees", f32), ("width", u32), ("height", u32), ("samples", u32),
                ("max_bounces", u32), ("background", u32),
                ("materials", C.c_void_p), ("material_count", u32),
                ("spheres", C.c_void_p), ("sphere_count", u32),
                ("triangles", C.c_void_p), ("triangle_count", u32),
                ("triangle_attributes", C.c_void_p),
                ("textures", C.c_void_p), ("texture_count", u32),
                ("environment", C.c_void_p)]


class SceneDesc(C.Structure):
    _fields_ = [("params", GpuRenderParams), ("camera", GpuCamera),
                ("spheres", C.c_void_p), ("triangles", C.c_void_p),
                ("triangle_attributes", C.c_void_p), ("materials", C.c_void_p),
                ("lights", C.c_void_p), ("bvh_nodes", C.c_void_p),
                ("bvh_primitives", C.c_void_p), ("texture_pixels", C.c_void_p),
                ("environment_pixels", C.c_void_p),
                ("sphere_count", u32), ("triangle_count", u32), ("material_count", u32),
                ("light_count", u32), ("bvh_node_count", u32), ("bvh_primitive_count", u32),
                ("texture_pixel_count", u32), ("environment_pixel_count", u32)]


class Tuning(C.Structure):
    _fields_ = [("quads_per_pixel", u32), ("cached_nodes", u32), ("best_ms", f32), ("tile_width", u32)]


class BatchItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("params", GpuRenderParams), ("camera", GpuCamera),
                ("quads_per_pixel", u32), ("cached_nodes", u32), ("tile_width", u32), ("reserved", u32)]


class RenderOptions(C.Structure):
    _fields_ = [("rng_mode", u32), ("kernel", u32), ("traversal", u32), ("block_size", u32),
                ("collect_stats", u32), ("schedule", u32), ("batch_blocks_per_cu", u32), ("reserved", u32)]


class Stats(C.Structure):
    _fields_ = [("last_kernel_ms", C.c_double), ("total_kernel_ms", C.c_double),
                ("launches", C.c_uint64)] + [
        (n, C.c_uint64) for n in ("paths", "rays", "closest_rays", "shadow_rays", "nodes_visited",
                                  "primref_tests", "triangle_tests", "sphere_tests", "closest_hits",
                                  "triangle_hits", "lights_scanned", "env_lookups", "env_samples",
                                  "closest_nodes_visited", "closest_primref_tests", "closest_triangle_tests",
                                  "closest_sphere_tests")]


# name -> (restype, argtypes) for every entry point declared in include/toaster_hip.h
P = C.c_void_p
class GuidedOptions(C.Structure):
    _fields_ = [("alpha", f32), ("train", u32), ("steps_per_pass", u32), ("batch", u32),
                ("flags", u32), ("warmup_steps", u32), ("warmup_rate", u32), ("alpha_ramp", u32),
                ("train_every", u32), ("reserved", u32 * 3)]


class GuidedStats(C.Structure):
    _fields_ = [("diffuse_vertices", C.c_uint64), ("records", C.c_uint64), ("train_steps", u32),
                ("last_loss", f32), ("inferred_rows", C.c_uint64), ("trained_samples", C.c_uint64),
                ("infer_ms", f32), ("train_ms", f32), ("reserved", C.c_uint64 * 2)]


SIGNATURES = {
    "th_pack_scene": (C.c_int32, [P, C.c_int32, C.POINTER(P)]),
    "th_packed_view": (C.c_int32, [P, C.POINTER(SceneDesc)]),
    "th_packed_free": (None, [P]),
    "th_make_camera": (C.c_int32, [P, P, P, f32, f32, C.POINTER(GpuCamera)]),
    "th_build_bvh": (C.c_int32, [P, u32, P, u32, P, u32, C.POINTER(u32), P, u32]),
    "th_device_count": (C.c_int32, [C.POINTER(C.c_int32)]),
    "th_device_info": (C.c_int32, [C.c_int32, C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t]),
    "th_scene_refit_quality": (C.c_int32, [P, C.POINTER(f32)]),
    "th_scene_traversal_info": (C.c_int32, [P, C.POINTER(u32), C.POINTER(u32)]),
    "th_context_create": (C.c_int32, [C.c_int32, C.POINTER(P)]),
    "th_context_destroy": (None, [P]),
    "th_last_error": (C.c_char_p, []),
    "th_scene_create": (C.c_int32, [P, C.POINTER(SceneDesc), u32, C.POINTER(P)]),
    "th_scene_update": (C.c_int32, [P, u32, C.POINTER(SceneDesc)]),
    "th_scene_destroy": (None, [P]),
    "th_render": (C.c_int32, [P, P, C.POINTER(GpuRenderParams), C.POINTER(GpuCamera),
                              C.POINTER(RenderOptions)]),
    "th_render_wavefront_static": (C.c_int32, [P, P, C.POINTER(GpuRenderParams), C.POINTER(GpuCamera),
                                               C.POINTER(RenderOptions)]),
    "th_progressive_graph_create": (C.c_int32, [P, P, C.POINTER(GpuRenderParams), C.POINTER(GpuCamera),
                                                C.POINTER(RenderOptions), C.POINTER(P)]),
    "th_progressive_graph_replay": (C.c_int32, [P, u32]),
    "th_progressive_graph_state": (C.c_int32, [P, C.POINTER(u32), C.POINTER(u32)]),
    "th_progressive_graph_destroy": (None, [P]),
    "th_readback": (C.c_int32, [P, P, P, C.c_size_t]),
    "th_upload_accumulation": (C.c_int32, [P, P, P, C.c_size_t]),
    "th_synchronize": (C.c_int32, [P]),
    "th_get_stats": (C.c_int32, [P, C.POINTER(Stats)]),
    "th_reset_stats": (C.c_int32, [P]),
    "th_rng_draws": (C.c_int32, [P, P, u32, u32, u32, u32, u32, P, P]),
    "th_primary_hits": (C.c_int32, [P, P, C.POINTER(GpuRenderParams), C.POINTER(GpuCamera), u32,
                                    P, P, C.c_size_t]),
    "th_render_batch": (C.c_int32, [P, C.POINTER(BatchItem), u32, C.POINTER(RenderOptions)]),
    "th_autotune": (C.c_int32, [P, P, C.POINTER(GpuRenderParams), C.POINTER(GpuCamera),
                                C.POINTER(RenderOptions), C.POINTER(Tuning)]),
    "th_autotune_batch": (C.c_int32, [P, C.POINTER(BatchItem), u32, C.POINTER(RenderOptions), C.POINTER(u32),
                                      C.POINTER(f32)]),
    "th_display_rgba8": (C.c_int32, [P, u32, u32, f32, P, C.c_size_t]),
    "th_display_rgba8_device": (C.c_int32, [P, P, f32, P, C.c_size_t]),
    "th_guide_create": (C.c_int32, [P, u32, C.POINTER(P)]),
    "th_guide_destroy": (None, [P]),
    "th_guide_get_params": (C.c_int32, [P, P, u32]),
    "th_guide_set_params": (C.c_int32, [P, P, u32]),
    "th_guide_set_hyper": (C.c_int32, [P, f32, f32, f32, f32]),
    "th_guide_forward": (C.c_int32, [P, P, u32, P]),
    "th_guide_infer_vertices": (C.c_int32, [P, P, u32, P, P, P]),
    "th_guide_train_batch": (C.c_int32, [P, P, P, P, u32, C.c_int32, P, P]),
    "th_guide_mfma_selftest": (C.c_int32, [P, P, P, P]),
    "th_guide_mfma32_selftest": (C.c_int32, [P, P, P, P]),
    "th_bf16_selftest": (C.c_int32, [P, P, P]),
    "th_measure_copy": (C.c_int32, [P, C.c_uint64, u32, u32, C.POINTER(C.c_double)]),
    "th_measure_mfma": (C.c_int32, [P, u32, u32, u32, C.POINTER(C.c_double), C.POINTER(u32)]),
    "th_render_guided": (C.c_int32, [P, P, P, C.POINTER(GpuRenderParams), C.POINTER(GpuCamera),
                                     C.POINTER(GuidedOptions), C.POINTER(GuidedStats)]),
}

_lib = None


class ToasterError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"toaster_hip error {code}: {message}")
        self.code = code


def lib():
    """Loads libtoaster_hip.so from the package directory (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != TH_OK:
        msg = lib().th_last_error()
        raise ToasterError(code, msg.decode() if msg else "")
    return code
