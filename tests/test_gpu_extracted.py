"""Primitives the 4-wide kernels test before the tree (large primitives left out of the
4-wide boxes, th_runtime.hip select_extracted) against the oracle's reference traversal.

- The roof and wall of the terrain workload and workload 3's ground quad are extracted;
  the headline frame (uniform triangle cloud) keeps every primitive in the tree.
- Exact ties across the split: the terrain with its roof quad copied five times (ten
  identical roof triangles, each copy its own material) extracts eight of them (the cap)
  and leaves two in the tree, so a ray reaching the roof meets equal distances both among
  the extracted primitives and between them and the tree.  The reference takes the first
  primitive its traversal finds (right-first walk: the later leaf, then the lower slot);
  every 4-wide kernel must pick the same one: images against the oracle on every wide
  schedule, and bit-identical primary hits for a camera looking up at the roof.
- After a device refit of moved geometry the extracted primitives are still tested first
  (the refit gives the 4-wide records the reference boxes again): images equal the
  oracle's of the moved terrain.

Tolerance as tests/test_gpu_parity.py: |gpu - oracle| <= 1e-4 (1 + |oracle|) per channel on
at least 99.5 % of pixels; primary hits bit-identical."""
import numpy as np
import pytest

import oracle
from toaster_amd import GpuScene, RenderOptions, _abi as A, scenes
from toaster_amd.guide import Guide, render_guided

pytestmark = pytest.mark.gpu

WIDE = A.TH_TRAVERSE_WIDE
WIDE_KERNELS = {
    "quad": RenderOptions(traversal=WIDE),
    "octet": RenderOptions(traversal=WIDE, schedule=A.TH_SCHEDULE_OCTET),
    "speculative": RenderOptions(traversal=WIDE, schedule=A.TH_SCHEDULE_SPECULATIVE),
    "dynamic": RenderOptions(traversal=WIDE, schedule=A.TH_SCHEDULE_DYNAMIC),
    "wavefront": RenderOptions(traversal=WIDE, kernel=A.TH_KERNEL_WAVEFRONT),
}


def tie_terrain(copies=5, look_up=False):
    """The terrain workload (128 x 64 cells) with its roof quad copied `copies` times."""
    sc = scenes.terrain(64, 36, samples=4, max_bounces=4, cells=(128, 64), env_size=(256, 128))
    t = sc.triangles
    roof = t[-4:-2]  # the roof quad's two triangles (then the wall's two)
    extra = np.concatenate([roof] * (copies - 1))
    extra["material_index"] = np.repeat(np.arange(copies - 1) % 3, 2)
    sc.triangles = np.concatenate([t, extra])
    if look_up:
        sc.camera_position = (0.3, 0.9, 0.4)
        sc.camera_look_at = (0.0, 1.6, -0.5)
    return sc


def _within(img, ref):
    close = np.abs(img - ref) <= 1e-4 * (1.0 + np.abs(ref))
    return float(np.mean(np.all(close, axis=-1)))


def test_which_scenes_extract(gpu_ctx):
    cases = {"terrain": (scenes.terrain(64, 36, samples=1, cells=(128, 64), env_size=(256, 128)), 4),
             "mixed": (scenes.mixed_environment(64, 36, samples=1, env_size=(256, 128)), 2),
             "tie_terrain": (tie_terrain(), 8),
             "stress65536": (scenes.bvh_stress(65536), 0)}
    for name, (sc, want) in cases.items():
        gs = GpuScene.from_scene(gpu_ctx, sc)
        n, nodes4 = gs.traversal_info()
        print(f"case extraction {name}: {n} extracted, {nodes4} 4-wide nodes", flush=True)
        assert n == want and nodes4 > 0
        gs.close()


@pytest.mark.parametrize("kernel", list(WIDE_KERNELS))
def test_tie_terrain_images(gpu_ctx, kernel):
    sc = tie_terrain()
    print(f"case tie terrain kernel={kernel}", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    assert gs.traversal_info()[0] == 8
    gs.render(options=WIDE_KERNELS[kernel], frame_index=0, accumulated_samples=0)
    img = gs.readback().copy()
    gs.close()
    ref, _ = oracle.OracleScene(sc).render(threads=8)
    frac = _within(img, ref)
    print(f"  within tol {frac:.5f} bit-exact {float(np.mean(np.all(img == ref, axis=-1))):.5f}")
    assert frac >= 0.995


def test_tie_terrain_per_sample_and_guided(gpu_ctx):
    sc = tie_terrain()
    gs = GpuScene.from_scene(gpu_ctx, sc)
    ref, _ = oracle.OracleScene(sc).render(threads=8, rng_mode=A.TH_RNG_PER_SAMPLE, group=4)
    print("case tie terrain per-sample quad", flush=True)
    gs.render(options=RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE, traversal=WIDE), frame_index=0,
              accumulated_samples=0)
    assert _within(gs.readback(), ref) >= 0.995
    print("case tie terrain guided alpha=0", flush=True)
    ref1, _ = oracle.OracleScene(sc).render(threads=8, rng_mode=A.TH_RNG_PER_SAMPLE, group=1)
    g = Guide(gpu_ctx, seed=2)
    render_guided(gs, g, sc.samples, alpha=0.0, train=False, progressive=False, warmup_steps=0)
    assert _within(gs.readback(), ref1) >= 0.995
    g.close()
    gs.close()


def test_tie_terrain_primary_hits_looking_up(gpu_ctx):
    sc = tie_terrain(look_up=True)
    print("case tie terrain primary hits looking up", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    assert gs.traversal_info()[0] == 8
    ref_slot, ref_t = oracle.OracleScene(sc).primary_hits()
    slot, t = gs.primary_hits(traversal=WIDE)
    print(f"  identical {np.mean(slot == ref_slot):.6f}")
    assert np.array_equal(slot, ref_slot)
    assert np.array_equal(t, ref_t)
    # most of these rays see the roof: triangles n_field, n_field + 1 and the copies
    n_field = len(sc.triangles) - 12
    prims = gs.packed.prims
    hit = ref_slot < len(prims)
    tri = np.where(hit, prims["index"][np.minimum(ref_slot, len(prims) - 1)], -1)
    roof = hit & ((tri == n_field) | (tri == n_field + 1) | (tri >= n_field + 4))
    print(f"  roof hits {roof.mean():.3f}")
    assert roof.mean() > 0.3
    gs.close()


def test_terrain_after_refit(gpu_ctx):
    """No exact ties here: a refit keeps the uploaded tree's leaf order, which decides ties
    where the reference's rebuilt tree could order them otherwise."""
    def terrain():
        return scenes.terrain(64, 36, samples=4, max_bounces=4, cells=(128, 64), env_size=(256, 128))
    sc = terrain()
    gs = GpuScene.from_scene(gpu_ctx, sc)
    assert gs.traversal_info()[0] == 4
    moved = terrain()
    v = moved.triangles["v"].copy()
    v[: len(v) // 2] += np.float32([0.0, 0.05, 0.0])  # half the field raised
    v[-4:] += np.float32([0.0, -0.1, 0.0])  # roof and wall lowered
    moved.triangles["v"] = v
    print("case terrain after a device refit", flush=True)
    gs.refit(moved)
    assert gs.traversal_info()[0] == 4
    ref, _ = oracle.OracleScene(moved).render(threads=8)
    for name in ("quad", "speculative", "wavefront"):
        gs.render(options=WIDE_KERNELS[name], frame_index=0, accumulated_samples=0)
        frac = _within(gs.readback(), ref)
        print(f"  {name}: within tol {frac:.5f}")
        assert frac >= 0.995
    gs.close()
