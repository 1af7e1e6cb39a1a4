"""Batched persistent launch: every scene of a batch renders exactly as it renders alone
(same launch shape), and the autotuner returns a valid launch shape."""
import numpy as np
import pytest

import oracle
from toaster_amd import GpuScene, RenderOptions, _abi as A, scenes
from toaster_amd.batch import BatchRenderer

pytestmark = pytest.mark.gpu


def _small_batch():
    w = scenes.batched_workload(n_scenes=3, n_frames=2, samples=8, max_bounces=4, seed=40,
                                tri_range=(128, 4096), resolutions=((64, 36), (96, 54)))
    return [f for frames in w for f in frames]


@pytest.mark.parametrize("rng", [A.TH_RNG_REFERENCE, A.TH_RNG_PER_SAMPLE])
def test_batch_matches_single_scene_renders(gpu_ctx, rng):
    items = _small_batch()
    br = BatchRenderer(gpu_ctx)
    for sc in items:
        br.add(sc)
    print(f"case batch rng={rng} items={len(items)}", flush=True)
    br.render(RenderOptions(rng_mode=rng), use_tuning=False)
    imgs = [gs.readback() for gs in br.scenes]
    for sc, img in zip(items, imgs):
        group = 1 if rng == A.TH_RNG_REFERENCE else min(16, sc.samples)
        ref, _ = oracle.OracleScene(sc).render(threads=8, rng_mode=rng, group=group)
        close = np.abs(img - ref) <= 1e-4 * (1.0 + np.abs(ref))
        frac = float(np.mean(np.all(close, axis=-1)))
        print(f"{sc.name}: within tol {frac:.5f}")
        assert frac >= 0.995
    br.close()


def test_batch_graph_replay_is_stable(gpu_ctx):
    items = _small_batch()
    br = BatchRenderer(gpu_ctx)
    for sc in items:
        br.add(sc)
    o = RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE)
    first = None
    for rep in range(3):  # replays of the captured graph, eager readbacks in between
        br.render(o, use_tuning=False)
        imgs = [gs.readback().copy() for gs in br.scenes]
        if first is None:
            first = imgs
        else:
            for a, b in zip(first, imgs):
                assert np.array_equal(a, b)
    br.close()


def test_autotune_returns_valid_shape(gpu_ctx):
    items = _small_batch()[:2]
    br = BatchRenderer(gpu_ctx)
    for sc in items:
        br.add(sc)
    tun = br.autotune(RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE))
    for g, c, ms in tun:
        assert g in (1, 2, 4, 8, 16) and c in (0, 21, 85) and ms > 0
    br.render(RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE), use_tuning=True)
    gpu_ctx.synchronize()
    br.close()
