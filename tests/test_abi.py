"""The C ABI without a GPU: exported symbols, host packing parity with the oracle, and
argument validation that must fail before any device work (path_tracer.rs:1027-1116)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from toaster_amd import _abi as A, scenes
from toaster_amd.scene import pack

from .helpers import scene
from toaster_amd.scene import diffuse

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            names += re.findall(r"^\s*(?:int32_t|void|const char\*)\s+(th_\w+)\s*\(", src, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = A.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) <= set(A.SIGNATURES), set(names) - set(A.SIGNATURES)


def test_record_sizes_match_reference():  # gpu_types.rs:200-211
    sizes = [C.sizeof(A.GpuRenderParams), C.sizeof(A.GpuCamera), A.SPHERE_DTYPE.itemsize,
             A.TRIANGLE_DTYPE.itemsize, A.ATTR_DTYPE.itemsize, A.MATERIAL_DTYPE.itemsize,
             A.LIGHT_DTYPE.itemsize, A.NODE_DTYPE.itemsize, A.PRIMREF_DTYPE.itemsize]
    assert sizes == [96, 64, 32, 64, 80, 48, 96, 48, 8]


@pytest.mark.parametrize("make", [lambda: scenes.cornell(), lambda: scenes.bvh_stress(2048),
                                  lambda: scenes.bvh_stress(8192),
                                  lambda: scenes.mixed_environment(width=32, height=18, samples=1,
                                                                   env_size=(64, 32))])
def test_host_packing_is_byte_identical_to_oracle(make):
    sc = make()
    a, b = pack(sc), oracle.OracleScene(sc).packed
    for name in ("spheres", "triangles", "attributes", "materials", "lights", "nodes", "prims",
                 "texels", "env"):
        x, y = getattr(a, name), getattr(b, name)
        assert x.shape == y.shape and x.tobytes() == y.tobytes(), name
    assert bytes(a.params) == bytes(b.params)
    assert bytes(a.camera) == bytes(b.camera)


def test_cornell_packing_counts():  # scene_upload.rs:553-575 (Cornell layout)
    p = pack(scenes.cornell())
    assert len(p.triangles) == 12 and len(p.spheres) == 2
    assert len(p.lights) == 2 and (p.lights["kind"] == 0).all()
    assert p.params.light_count == 2 and p.params.total_light_weight > 0
    assert p.params.bvh_node_count == len(p.nodes) > 0
    assert p.params.bvh_primitive_count == len(p.prims) == 14


def test_bvh_build_entry_point_matches_packer():
    p = pack(scenes.bvh_stress(2048))
    n = len(p.spheres) + len(p.triangles)
    nodes = np.zeros(2 * n, A.NODE_DTYPE)
    prims = np.zeros(n, A.PRIMREF_DTYPE)
    cnt = C.c_uint32()
    A.check(A.lib().th_build_bvh(p.spheres.ctypes.data, len(p.spheres), p.triangles.ctypes.data,
                                 len(p.triangles), nodes.ctypes.data, len(nodes), C.byref(cnt),
                                 prims.ctypes.data, n))
    assert cnt.value == len(p.nodes)
    assert nodes[:cnt.value].tobytes() == p.nodes.tobytes()
    assert prims.tobytes() == p.prims.tobytes()


def test_packing_rejects_invalid_scenes():  # scene_upload.rs:84-92, 226-228
    with pytest.raises(A.ToasterError):
        pack(scene([diffuse((1, 1, 1))]))  # no objects
    with pytest.raises(A.ToasterError):
        pack(scene([], sph=[((0, 0, -2), 1.0, 0)]))  # no materials
    with pytest.raises(A.ToasterError):
        pack(scene([diffuse((1, 1, 1))], sph=[((0, 0, -2), 1.0, 3)]))  # bad material index
    with pytest.raises(A.ToasterError):
        pack(scene([diffuse((1, 1, 1))], sph=[((0, 0, -2), 1.0, 0)], background=A.TH_BG_ENVIRONMENT))


def test_packing_without_bvh_for_a_refit():
    """TH_PACK_NO_BVH: the same records and light list, no tree (for a device refit)."""
    sc = scenes.bvh_stress(512, width=32, height=16)
    full, bare = pack(sc), pack(sc, bvh=False)
    assert len(full.nodes) > 0 and len(bare.nodes) == 0 and len(bare.prims) == 0
    assert bare.params.bvh_node_count == 0 and bare.params.bvh_primitive_count == 0
    for a, b in ((full.triangles, bare.triangles), (full.spheres, bare.spheres), (full.lights, bare.lights),
                 (full.materials, bare.materials)):
        assert a.tobytes() == b.tobytes()
    inp, keep = sc.input()
    h = C.c_void_p()
    assert A.lib().th_pack_scene(C.byref(inp), 4, C.byref(h)) == A.TH_E_INVALID  # unknown flag


def test_null_arguments_fail_before_device_work():
    L = A.lib()
    assert L.th_pack_scene(None, 1, None) == A.TH_E_INVALID
    assert L.th_context_create(0, None) == A.TH_E_INVALID
    assert L.th_render(None, None, None, None, None) == A.TH_E_INVALID
    assert L.th_readback(None, None, None, 0) == A.TH_E_INVALID
    assert L.th_scene_create(None, None, 0, None) == A.TH_E_INVALID
    assert b"null" in L.th_last_error()


def test_host_display_matches_reference_bytes():  # image_output.rs:212-219
    from toaster_amd import display_rgba8
    px = np.array([[[-1.0, 0.5, 2.0, 1.0], [0.0, 0.25, 0.75, 0.0]]], np.float32)
    assert display_rgba8(px).ravel().tolist() == [0, 206, 245, 255, 0, 165, 223, 0]


def test_host_display_matches_oracle_random():
    from toaster_amd import display_rgba8
    rng = np.random.default_rng(2)
    px = rng.lognormal(-1, 2, (64, 64, 4)).astype(np.float32)
    px[..., 3] = rng.uniform(0, 1.2, (64, 64))
    for stops in (-1.0, 0.0, 1.5):
        assert np.array_equal(display_rgba8(px, stops), oracle.display_rgba8(px, stops))
