"""The packer against the reference's own packing tests, on the reference's scene files
(copied from scenes/ into tests/golden/reference_scenes, loaded by toaster_amd.loader):
scene_upload.rs:481-509 (power-weighted light CDF), 512-551 (materials scene), 553-575
(Cornell box triangles).  Each assertion restates the reference test's."""
import math
import os

import numpy as np

from toaster_amd import pack
from toaster_amd.loader import load_scene

D = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_scenes")


def _packed(name):
    sc = load_scene(os.path.join(D, name))
    return sc, pack(sc)


def test_packs_power_weighted_light_cdf_and_selection_pdfs():  # scene_upload.rs:481-509
    _, p = _packed("017_power_weighted_lights.json")
    assert len(p.lights) == 2
    dim = p.lights[0]["area_weight_cumulative"]
    bright = p.lights[1]["area_weight_cumulative"]
    total = p.params.total_light_weight
    assert abs(dim[0] - 8.0) < 1e-6
    assert abs(dim[1] - 0.8) < 1e-6
    assert abs(dim[2] - dim[1]) < 1e-6
    assert abs(bright[0] - 4.0 * math.pi * 0.2 * 0.2) < 1e-6
    assert abs(bright[1] - bright[0] * 20.0) < 1e-5
    assert abs(bright[2] - total) < 1e-5
    dsel, bsel = dim[1] / total, bright[1] / total
    assert abs(dsel + bsel - 1.0) < 1e-6
    assert bsel > 0.9 and bsel > dsel * 10.0
    assert abs(dsel / dim[0] - 0.1 / total) < 1e-6
    assert abs(bsel / bright[0] - 20.0 / total) < 1e-6


def test_loads_materials_scene_into_gpu_layout():  # scene_upload.rs:512-551
    sc, p = _packed("002_materials.json")
    assert (p.params.width, p.params.height) == (sc.width, sc.height)
    assert (p.params.samples, p.params.max_bounces) == (sc.samples, sc.max_bounces)
    assert (p.params.sphere_count, p.params.triangle_count, p.params.material_count) == (
        len(sc.spheres), len(sc.triangles), len(sc.materials))
    f32 = lambda v: np.asarray(v, np.float32)  # the reference compares f32 values
    assert np.array_equal(f32(list(p.camera.origin)), f32([0.0, 1.4, 6.0, 0.0]))
    assert np.array_equal(f32(p.spheres[3]["center_radius"]), f32([1.2, 0.5, 0.0, 0.5]))
    m = p.materials
    assert m[2]["kind"] == 2 and m[2]["params"][1] == np.float32(1.5)
    assert m[3]["kind"] == 1 and m[3]["params"][0] == np.float32(0.08)
    assert m[4]["kind"] == 3 and m[4]["params"][2] == np.float32(6.0)
    assert len(p.triangles) == len(sc.triangles)
    assert len(p.lights) == 1 and p.params.light_count == 1
    assert p.params.total_light_weight > 0.0
    assert p.lights[0]["kind"] == 1 and p.lights[0]["material_index"] == 4


def test_loads_cornell_box_triangles_into_gpu_layout():  # scene_upload.rs:553-575
    _, p = _packed("003_cornell_box.json")
    t = p.triangles
    assert len(t) == 12
    assert np.array_equal(t[0]["v0"], np.float32([-2.0, 0.0, 0.0, 0.0]))
    assert np.array_equal(t[0]["v1"], np.float32([2.0, 0.0, -4.0, 0.0]))
    assert t[0]["material_index"] == 0 and t[6]["material_index"] == 1
    assert t[8]["material_index"] == 2 and t[10]["material_index"] == 3
    assert len(p.lights) == 2 and p.params.light_count == 2
    assert p.params.total_light_weight > 0.0
    assert all(l["kind"] == 0 for l in p.lights)
    assert len(p.nodes) > 0
    assert p.params.bvh_node_count == len(p.nodes)
    assert p.params.bvh_primitive_count == len(p.prims)
