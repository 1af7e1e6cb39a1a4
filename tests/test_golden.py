"""The oracle and the packer against the committed regression fixtures of tests/golden/
(written by tests/golden/make_golden.py; see its docstring).  Bit-exact."""
import os

import numpy as np

import oracle
from toaster_amd import pack, scenes

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(G, name))


def test_pcg_sequence_pixel0_frame0():
    f = _load("pcg_pixel0_frame0.npz")
    state = oracle.pcg_hash(0)
    for k in range(16):
        state = oracle.pcg_hash(state)
        assert state == int(f["states"][k])
        assert np.float32(np.float32(state) / np.float32(4294967295.0)) == f["draws"][k]
    assert np.all((f["draws"] >= 0) & (f["draws"] <= 1))


def test_bvh_arrays_frozen():
    for name, sc in (("bvh_cornell", scenes.cornell(64, 36)), ("bvh_stress_2048", scenes.bvh_stress(2048))):
        f = _load(name + ".npz")
        p = pack(sc)
        assert np.array_equal(np.ascontiguousarray(p.nodes).view(np.uint8), f["nodes"])
        assert np.array_equal(np.ascontiguousarray(p.prims).view(np.uint8), f["prims"])


def test_primary_hits_and_image_frozen():
    sc = scenes.cornell(64, 36, samples=4, max_bounces=4)
    oc = oracle.OracleScene(sc)
    slots, dist = oc.primary_hits()
    f = _load("primary_cornell_64x36.npz")
    assert np.array_equal(slots, f["slots"]) and np.array_equal(dist, f["distance"])
    img, _ = oc.render(threads=4)
    assert np.array_equal(img, _load("image_cornell_64x36_4spp.npz")["rgba"])
    oc.close()


@__import__("pytest").mark.gpu
def test_gpu_render_matches_golden_image(gpu_ctx):
    from toaster_amd import GpuScene, RenderOptions, _abi as A
    sc = scenes.cornell(64, 36, samples=4, max_bounces=4)
    ref = _load("image_cornell_64x36_4spp.npz")["rgba"]
    gs = GpuScene.from_scene(gpu_ctx, sc)
    for trav in (A.TH_TRAVERSE_REFERENCE, A.TH_TRAVERSE_WIDE):
        print(f"case golden cornell traversal={trav}", flush=True)
        gs.render(options=RenderOptions(traversal=trav), frame_index=0, accumulated_samples=0)
        img = gs.readback()
        close = np.abs(img - ref) <= 1e-4 * (1.0 + np.abs(ref))
        assert float(np.mean(np.all(close, axis=-1))) >= 0.995
    slots, dist = gs.primary_hits()
    f = _load("primary_cornell_64x36.npz")
    assert np.array_equal(slots, f["slots"])
    gs.close()
