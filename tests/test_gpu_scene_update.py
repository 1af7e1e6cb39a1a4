"""Validation of th_scene_update through the C ABI (path_tracer.rs:639-653, 676-720).

- Texels and environment pixels are uploaded once, at create; an update whose descriptor
  carries a different count of either is refused with TH_E_MISMATCH before anything is
  uploaded, and material texture ranges are checked against the uploaded texels (a
  per-frame pack without texels still validates against them).
- Moved spheres or triangles upload the frame's rebuilt tree with them when the
  descriptor has one, as path_tracer.rs:706-719 does, so the image equals a freshly
  uploaded scene's bit for bit; a descriptor without a tree is refused (TH_E_INVALID)
  unless the update is a device refit.
"""
import ctypes as C
import dataclasses

import numpy as np
import pytest

from toaster_amd import GpuScene, RenderOptions, _abi as A, scenes
from toaster_amd.scene import pack

from . import test_gpu_edge_cases as edge

pytestmark = pytest.mark.gpu


def _update_rc(gs, packed, mask):
    d = packed.desc()
    return A.lib().th_scene_update(gs._h, int(mask), C.byref(d))


def test_update_rejects_a_larger_texture_than_uploaded(gpu_ctx):
    sc = edge.textured_floor()
    print("case update texture count mismatch", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    gs.render()
    before = gs.readback().copy()
    big = pack(sc)
    n0 = len(big.texels)
    assert n0 > 0
    big.texels = np.concatenate([big.texels, np.full(4096, 0xFFFFFFFF, np.uint32)])
    big.materials = big.materials.copy()
    big.materials["texture_width"][0] = 64
    big.materials["texture_height"][0] = 64
    assert _update_rc(gs, big, A.TH_CHANGE_MATERIALS) == A.TH_E_MISMATCH
    # same pixel count, material range past the uploaded texels: invalid
    wide = pack(sc)
    wide.materials = wide.materials.copy()
    wide.materials["texture_width"][0] = 64
    wide.materials["texture_height"][0] = 64
    assert _update_rc(gs, wide, A.TH_CHANGE_MATERIALS) == A.TH_E_INVALID
    # a per-frame pack without texels validates against the uploaded ones
    frame = pack(sc, include_static_pixels=False)
    assert len(frame.texels) == 0
    assert _update_rc(gs, frame, A.TH_CHANGE_MATERIALS | A.TH_CHANGE_CAMERA) == A.TH_OK
    frame.materials = frame.materials.copy()
    frame.materials["texture_width"][0] = 64
    frame.materials["texture_height"][0] = 64
    assert _update_rc(gs, frame, A.TH_CHANGE_MATERIALS) == A.TH_E_INVALID
    # nothing refused was uploaded: the image is unchanged
    gs.render()
    after = gs.readback()
    gs.close()
    assert np.array_equal(before.view(np.uint32), after.view(np.uint32))


def test_update_rejects_another_environment_size(gpu_ctx):
    sc = scenes.mixed_environment(32, 18, samples=1, max_bounces=2, n_triangles=64, n_spheres=8,
                                  env_size=(64, 32))
    print("case update environment count mismatch", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    other = pack(dataclasses.replace(sc, environment=scenes.EnvironmentMap(
        scenes.sky_environment(128, 64), 1.0, 0.0)))
    assert _update_rc(gs, other, A.TH_CHANGE_CAMERA) == A.TH_E_MISMATCH
    assert _update_rc(gs, pack(sc), A.TH_CHANGE_CAMERA) == A.TH_OK
    gs.close()


def _moved(sc):
    tris = sc.triangles.copy()
    tris["v"] = tris["v"] + np.float32(0.25) * np.array([1.0, 0.5, -0.5], np.float32)
    return dataclasses.replace(sc, triangles=tris)


@pytest.mark.parametrize("opts", [RenderOptions(), RenderOptions(traversal=A.TH_TRAVERSE_WIDE,
                                                                 schedule=A.TH_SCHEDULE_SPECULATIVE)],
                         ids=["reference", "speculative"])
def test_moved_geometry_uploads_its_tree(gpu_ctx, opts):
    sc = scenes.bvh_stress(2048, width=96, height=64, samples=2)
    moved = _moved(sc)
    print("case moved triangles without the bvh bit", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    assert _update_rc(gs, pack(moved), A.TH_CHANGE_TRIANGLES | A.TH_CHANGE_LIGHTS) == A.TH_OK
    gs.packed = pack(moved)
    gs.render(options=opts, frame_index=0, accumulated_samples=0)
    img = gs.readback().copy()
    gs.close()
    fresh = GpuScene.from_scene(gpu_ctx, moved)
    fresh.render(options=opts, frame_index=0, accumulated_samples=0)
    ref = fresh.readback().copy()
    fresh.close()
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


def test_moved_geometry_without_a_tree_is_refused(gpu_ctx):
    sc = scenes.bvh_stress(2048, width=32, height=18, samples=1)
    print("case moved triangles without any tree", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    gs.render()
    before = gs.readback().copy()
    no_tree = pack(_moved(sc), bvh=False)
    assert _update_rc(gs, no_tree, A.TH_CHANGE_TRIANGLES) == A.TH_E_INVALID
    assert _update_rc(gs, no_tree, A.TH_CHANGE_SPHERES | A.TH_CHANGE_CAMERA) == A.TH_E_INVALID
    gs.render()
    after = gs.readback()
    gs.close()
    assert np.array_equal(before.view(np.uint32), after.view(np.uint32))
