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


def test_autotuned_batch_matches_oracle(gpu_ctx):
    """The autotuner's choices (quads per pixel, LDS treelet, tile width per scene; blocks
    per compute unit for the batch) are launch shapes: the tuned batch renders every
    scene as the oracle does with the tuned number of sample groups."""
    items = _small_batch()
    br = BatchRenderer(gpu_ctx)
    for k, sc in enumerate(items):
        br.add(sc, group=k // 2)  # the two frames of each animation share a tuning
    o = RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE)
    print("case autotuned batch", flush=True)
    tun = br.autotune(o)
    n, ms = br.autotune_batch(o)
    print(f"  tuning {tun} blocks/CU {n} ({ms:.3f} ms)")
    assert n in (2, 4, 6, 8) and ms > 0
    for (g, c, t_ms, tw), (g2, c2, _, tw2) in zip(tun[::2], tun[1::2]):
        assert (g, c, tw) == (g2, c2, tw2)
        assert g in (1, 2, 4, 8, 16) and c in (0, 21, 85) and t_ms > 0
        assert tw in (1, 2, 4, 8, 16) and tw <= 16 // g
    br.render(o, use_tuning=True)
    for sc, gs, (g, _, _, _) in zip(items, br.scenes, tun):
        ref, _ = oracle.OracleScene(sc).render(threads=8, rng_mode=A.TH_RNG_PER_SAMPLE, group=g)
        img = gs.readback()
        close = np.abs(img - ref) <= 1e-4 * (1.0 + np.abs(ref))
        frac = float(np.mean(np.all(close, axis=-1)))
        print(f"  {sc.name} G={g}: within tol {frac:.5f}")
        assert frac >= 0.995
    br.close()


@pytest.mark.parametrize("g,tw", [(1, 1), (1, 16), (2, 8), (4, 1), (16, 1)])
def test_tile_shapes_give_the_same_image(gpu_ctx, g, tw):
    """A wave's tile shape only moves pixels between lanes: the image equals the default
    shape's bit for bit."""
    sc = _small_batch()[0]
    br = BatchRenderer(gpu_ctx)
    br.add(sc)
    o = RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE)
    imgs = []
    for width in (0, tw):
        arr = br.items(use_tuning=False)
        arr[0].quads_per_pixel, arr[0].tile_width = g, width
        br.render(o, items=arr)
        imgs.append(br.scenes[0].readback().copy())
    assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
    arr = br.items(use_tuning=False)
    arr[0].quads_per_pixel, arr[0].tile_width = g, 32 // g  # wider than 16 / g pixels
    import ctypes as C
    rc = A.lib().th_render_batch(gpu_ctx.handle, arr, 1, C.byref(o.struct()))
    assert rc == A.TH_E_INVALID
    br.close()
