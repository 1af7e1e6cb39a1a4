"""A scene uploaded without a BVH (bvh_node_count = 0) renders through the linear scan
of every primitive (wgsl:667-698, 797-802) in every kernel that accepts it -- the scalar
megakernels, the 4-lane request (which falls back to the scalar kernel), the wavefront
and the guided wavefront at alpha = 0 -- and matches the oracle's linear path
(tolerance of tests/test_gpu_parity.py)."""
import numpy as np
import pytest

import oracle
from toaster_amd import GpuScene, RenderOptions, _abi as A, scenes
from toaster_amd.guide import Guide, render_guided
from toaster_amd.scene import pack

from .test_gpu_edge_cases import textured_floor

pytestmark = pytest.mark.gpu

TOL_FRAC = 0.995
SCENES = {
    "cornell": lambda: scenes.cornell(64, 36, samples=4),
    "mixed_env": lambda: scenes.mixed_environment(48, 27, samples=4, max_bounces=6, n_triangles=96,
                                                  n_spheres=16, env_size=(256, 128)),
    "textured_floor": textured_floor,
}
KERNELS = {
    "reference": RenderOptions(),
    "ordered": RenderOptions(traversal=A.TH_TRAVERSE_ORDERED),
    "wide_request": RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=A.TH_SCHEDULE_SPECULATIVE),
    "wavefront": RenderOptions(traversal=A.TH_TRAVERSE_WIDE, kernel=A.TH_KERNEL_WAVEFRONT),
    "per_sample": RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE, traversal=A.TH_TRAVERSE_ORDERED),
}


def _without_tree(sc):
    p = pack(sc)
    p.nodes = p.nodes[:0]
    p.prims = p.prims[:0]
    p.params.bvh_node_count = 0
    p.params.bvh_primitive_count = 0
    return p


def _within(img, ref):
    close = np.abs(img - ref) <= 1e-4 * (1.0 + np.abs(ref))
    return float(np.mean(np.all(close, axis=-1)))


@pytest.mark.parametrize("kernel", list(KERNELS))
@pytest.mark.parametrize("name", list(SCENES))
def test_linear_scene_matches_oracle(gpu_ctx, name, kernel):
    sc = SCENES[name]()
    print(f"case linear scan {name} kernel={kernel}", flush=True)
    gs = GpuScene(gpu_ctx, _without_tree(sc))
    gs.render(options=KERNELS[kernel], frame_index=0, accumulated_samples=0)
    img = gs.readback().copy()
    gs.close()
    o = KERNELS[kernel]
    group = 1
    if o.rng_mode == A.TH_RNG_PER_SAMPLE:
        while group * 2 <= min(sc.samples, 8):
            group *= 2
    ref, _ = oracle.OracleScene(sc, bvh=False).render(threads=8, rng_mode=o.rng_mode, group=group)
    frac = _within(img, ref)
    print(f"  within tol {frac:.5f}")
    assert np.isfinite(img).all()
    assert frac >= TOL_FRAC


def test_linear_scene_guided_alpha_zero(gpu_ctx):
    sc = SCENES["mixed_env"]()
    print("case linear scan guided alpha=0", flush=True)
    gs = GpuScene(gpu_ctx, _without_tree(sc))
    g = Guide(gpu_ctx, seed=2)
    render_guided(gs, g, sc.samples, alpha=0.0, train=True, batch=1024, progressive=False, warmup_steps=0)
    img = gs.readback().copy()
    gs.close()
    g.close()
    ref, _ = oracle.OracleScene(sc, bvh=False).render(threads=8, rng_mode=A.TH_RNG_PER_SAMPLE, group=1)
    assert _within(img, ref) >= TOL_FRAC


def test_linear_scene_primary_hits(gpu_ctx):
    sc = SCENES["mixed_env"]()
    gs = GpuScene(gpu_ctx, _without_tree(sc))
    oc = oracle.OracleScene(sc, bvh=False)
    ref_slot, ref_t = oc.primary_hits()
    for trav in (A.TH_TRAVERSE_REFERENCE, A.TH_TRAVERSE_ORDERED, A.TH_TRAVERSE_WIDE):
        slot, t = gs.primary_hits(traversal=trav)
        assert np.array_equal(slot, ref_slot)
        assert np.array_equal(t, ref_t)
    gs.close()
