"""Renderer-neutral scene model and packing into the reference GPU records.

Scene mirrors toaster-scene's Scene (crates/toaster-scene/src/scene.rs:19-90) with
numpy-backed primitive arrays; pack() is scene_to_gpu
(crates/toaster-gpu/src/scene_upload.rs:60-76), run natively by th_pack_scene.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _abi as A

MATERIAL_IN_DTYPE = np.dtype([("kind", "<u4"), ("color", "<f4", 3), ("roughness", "<f4"),
                              ("ior", "<f4"), ("strength", "<f4"), ("texture_index", "<u4")])
SPHERE_IN_DTYPE = np.dtype([("center", "<f4", 3), ("radius", "<f4"), ("material_index", "<u4")])
TRIANGLE_IN_DTYPE = np.dtype([("v", "<f4", (3, 3)), ("material_index", "<u4")])
ATTR_IN_DTYPE = np.dtype([("has_normals", "<u4"), ("has_tex_coords", "<u4"),
                          ("normals", "<f4", (3, 3)), ("tex_coords", "<f4", (3, 2))])
assert MATERIAL_IN_DTYPE.itemsize == C.sizeof(A.Material)
assert SPHERE_IN_DTYPE.itemsize == C.sizeof(A.Sphere)
assert TRIANGLE_IN_DTYPE.itemsize == C.sizeof(A.Triangle)
assert ATTR_IN_DTYPE.itemsize == C.sizeof(A.TriangleAttributes)


def diffuse(albedo):
    return (A.TH_MAT_DIFFUSE, tuple(albedo), 0.0, 0.0, 0.0, 0)


def metal(albedo, roughness):
    return (A.TH_MAT_METAL, tuple(albedo), roughness, 0.0, 0.0, 0)


def dielectric(ior):
    return (A.TH_MAT_DIELECTRIC, (1.0, 1.0, 1.0), 0.0, ior, 0.0, 0)


def emissive(color, strength):
    return (A.TH_MAT_EMISSIVE, tuple(color), 0.0, 0.0, strength, 0)


def textured_diffuse(albedo, texture_index):
    return (A.TH_MAT_TEXTURED_DIFFUSE, tuple(albedo), 0.0, 0.0, 0.0, texture_index)


@dataclass
class EnvironmentMap:
    """environment.rs:9-24: linear RGB texels (H, W, 3), intensity, yaw in degrees."""
    rgb: np.ndarray
    intensity: float = 1.0
    rotation_degrees: float = 0.0


@dataclass
class Scene:
    camera_position: tuple = (0.0, 0.0, 1.0)
    camera_look_at: tuple = (0.0, 0.0, 0.0)
    camera_up: tuple = (0.0, 1.0, 0.0)
    fov_degrees: float = 45.0
    width: int = 320
    height: int = 180
    samples: int = 4
    max_bounces: int = 4
    background: int = A.TH_BG_SKY
    materials: np.ndarray = field(default_factory=lambda: np.zeros(0, MATERIAL_IN_DTYPE))
    spheres: np.ndarray = field(default_factory=lambda: np.zeros(0, SPHERE_IN_DTYPE))
    triangles: np.ndarray = field(default_factory=lambda: np.zeros(0, TRIANGLE_IN_DTYPE))
    attributes: Optional[np.ndarray] = None
    textures: List[np.ndarray] = field(default_factory=list)  # (H, W, 4) uint8 sRGB
    environment: Optional[EnvironmentMap] = None
    name: str = "scene"

    def input(self):
        """Builds the th_scene_input struct; the returned keep-alive list owns buffers."""
        keep = []

        def arr(a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            return a

        s = A.SceneInput()
        s.camera_position[:] = [float(v) for v in self.camera_position]
        s.camera_look_at[:] = [float(v) for v in self.camera_look_at]
        s.camera_up[:] = [float(v) for v in self.camera_up]
        s.fov_degrees = float(self.fov_degrees)
        s.width, s.height = int(self.width), int(self.height)
        s.samples, s.max_bounces = int(self.samples), int(self.max_bounces)
        s.background = int(self.background)
        m = arr(self.materials, MATERIAL_IN_DTYPE)
        s.materials, s.material_count = (m.ctypes.data if len(m) else None), len(m)
        sp = arr(self.spheres, SPHERE_IN_DTYPE)
        s.spheres, s.sphere_count = (sp.ctypes.data if len(sp) else None), len(sp)
        tr = arr(self.triangles, TRIANGLE_IN_DTYPE)
        s.triangles, s.triangle_count = (tr.ctypes.data if len(tr) else None), len(tr)
        if self.attributes is not None and len(tr):
            at = arr(self.attributes, ATTR_IN_DTYPE)
            assert len(at) == len(tr), "triangle attribute count must match triangle count"
            s.triangle_attributes = at.ctypes.data
        if self.textures:
            texs = (A.Texture * len(self.textures))()
            for i, t in enumerate(self.textures):
                t = arr(t, np.uint8)
                texs[i].height, texs[i].width = t.shape[0], t.shape[1]
                texs[i].rgba8 = t.ctypes.data_as(C.POINTER(C.c_uint8))
            keep.append(texs)
            s.textures, s.texture_count = C.addressof(texs), len(self.textures)
        if self.environment is not None:
            e = self.environment
            rgb = arr(e.rgb, np.float32)
            env = A.Environment()
            env.height, env.width = rgb.shape[0], rgb.shape[1]
            env.rgb = rgb.ctypes.data_as(C.POINTER(C.c_float))
            env.intensity = float(e.intensity)
            env.rotation_degrees = float(e.rotation_degrees)
            keep.append(env)
            s.environment = C.addressof(env)
        return s, keep

    @property
    def paths_per_frame(self):
        return self.width * self.height * self.samples


def _copy(ptr, count, dtype):
    if count == 0 or not ptr:
        return np.zeros(0, dtype)
    buf = (C.c_char * (count * dtype.itemsize)).from_address(ptr)
    return np.frombuffer(buf, dtype=dtype).copy()  # one copy out of the packer's memory


class PackedScene:
    """SceneGpuData (scene_upload.rs:17-40): params, camera and the flat record arrays."""

    def __init__(self, desc: A.SceneDesc):
        self.params = A.GpuRenderParams.from_buffer_copy(desc.params)
        self.camera = A.GpuCamera.from_buffer_copy(desc.camera)
        self.spheres = _copy(desc.spheres, desc.sphere_count, A.SPHERE_DTYPE)
        self.triangles = _copy(desc.triangles, desc.triangle_count, A.TRIANGLE_DTYPE)
        self.attributes = _copy(desc.triangle_attributes, desc.triangle_count, A.ATTR_DTYPE)
        self.materials = _copy(desc.materials, desc.material_count, A.MATERIAL_DTYPE)
        self.lights = _copy(desc.lights, desc.light_count, A.LIGHT_DTYPE)
        self.nodes = _copy(desc.bvh_nodes, desc.bvh_node_count, A.NODE_DTYPE)
        self.prims = _copy(desc.bvh_primitives, desc.bvh_primitive_count, A.PRIMREF_DTYPE)
        self.texels = _copy(desc.texture_pixels, desc.texture_pixel_count, np.dtype("<u4"))
        self.env = _copy(desc.environment_pixels, desc.environment_pixel_count * 4,
                         np.dtype("<f4")).reshape(-1, 4)

    def desc(self):
        """A SceneDesc pointing into this object's arrays (valid while it lives)."""
        d = A.SceneDesc()
        d.params = self.params
        d.camera = self.camera
        for name, a in (("spheres", self.spheres), ("triangles", self.triangles),
                        ("triangle_attributes", self.attributes), ("materials", self.materials),
                        ("lights", self.lights), ("bvh_nodes", self.nodes),
                        ("bvh_primitives", self.prims), ("texture_pixels", self.texels),
                        ("environment_pixels", self.env)):
            setattr(d, name, a.ctypes.data if a.size else None)
        d.sphere_count, d.triangle_count = len(self.spheres), len(self.triangles)
        d.material_count, d.light_count = len(self.materials), len(self.lights)
        d.bvh_node_count, d.bvh_primitive_count = len(self.nodes), len(self.prims)
        d.texture_pixel_count, d.environment_pixel_count = len(self.texels), len(self.env)
        return d


def pack(scene: Scene, include_static_pixels=True, bvh=True) -> PackedScene:
    """scene_to_gpu: validation, camera, records, lights, BVH and environment CDF
    (bvh=False: no BVH arrays, for a device refit of moved geometry)."""
    L = A.lib()
    inp, keep = scene.input()
    h = C.c_void_p()
    flags = (1 if include_static_pixels else 0) | (0 if bvh else A.TH_PACK_NO_BVH)
    A.check(L.th_pack_scene(C.byref(inp), flags, C.byref(h)))
    try:
        d = A.SceneDesc()
        A.check(L.th_packed_view(h, C.byref(d)))
        return PackedScene(d)
    finally:
        L.th_packed_free(h)
        del keep
