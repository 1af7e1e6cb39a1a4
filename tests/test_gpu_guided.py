"""Guided rendering (th_render_guided) on the GPU.

- alpha = 0: every draw and operation is the reference's, so the image equals the
  oracle's per-sample image with one sample group (same tolerance as the plain paths),
  with and without training records (the deferred diffuse path).
- alpha > 0: the estimator must stay unbiased for any network: the image mean of a
  trained-online guided render agrees with a high-spp plain render within 4 standard
  errors (measured from independent chunks).
- Training gains: at equal samples per pixel the online-trained guide lowers relRMSE
  against a high-spp plain reference on the indirect-lit terrain, compared with the
  plain estimator (the method without the network).
"""
import numpy as np
import pytest

import oracle
from toaster_amd import GpuScene, RenderOptions, _abi as A, scenes
from toaster_amd.guide import Guide, rel_rmse, render_guided

pytestmark = pytest.mark.gpu


def _within(img, ref):
    close = np.abs(img - ref) <= 1e-4 * (1.0 + np.abs(ref))
    return float(np.mean(np.all(close, axis=-1)))


CASES = {
    "cornell": lambda: scenes.cornell(64, 36, samples=6, max_bounces=4),
    "mixed_env": lambda: scenes.mixed_environment(64, 36, samples=4, max_bounces=6, n_triangles=512,
                                                  n_spheres=32, env_size=(256, 128)),
    "terrain": lambda: scenes.terrain(48, 27, samples=4, max_bounces=6, cells=(64, 32), env_size=(256, 128)),
}


@pytest.mark.parametrize("train", [False, True])
@pytest.mark.parametrize("case", sorted(CASES))
def test_alpha_zero_matches_oracle(gpu_ctx, case, train):
    sc = CASES[case]()
    gs = GpuScene.from_scene(gpu_ctx, sc)
    g = Guide(gpu_ctx, seed=2)
    print(f"case guided alpha=0 {case} train={train}", flush=True)
    st = render_guided(gs, g, sc.samples, alpha=0.0, train=train, batch=1024, progressive=False,
                       warmup_steps=0)
    img = gs.readback()
    ref, _ = oracle.OracleScene(sc).render(threads=8, rng_mode=A.TH_RNG_PER_SAMPLE, group=1)
    frac = _within(img, ref)
    print(f"  within tol {frac:.5f} stats {st}")
    assert frac >= 0.995
    if train:
        assert st["records"] > 0 and st["diffuse_vertices"] > 0
    gs.close()
    g.close()


def _terrain_small():
    return scenes.terrain(64, 36, samples=16, max_bounces=6, cells=(128, 64), env_size=(512, 256))


@pytest.mark.parametrize("env_nee", [True, False])
def test_guided_estimator_is_unbiased(gpu_ctx, env_nee):
    sc = _terrain_small()
    gs = GpuScene.from_scene(gpu_ctx, sc)
    print("case guided unbiased: plain reference", flush=True)
    gs.render(samples=4096, options=RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE,
                                                  traversal=A.TH_TRAVERSE_WIDE), frame_index=77)
    ref = gs.readback()[..., :3].mean(-1)
    g = Guide(gpu_ctx, seed=4)
    kw = dict(alpha=0.7, env_nee=env_nee, batch=4096, warmup_steps=16, warmup_rate=4, alpha_ramp=16)
    render_guided(gs, g, 32, train=True, progressive=False, **kw)
    means = []
    for c in range(16):
        gs.frame_index = 1000 + c
        render_guided(gs, g, 32, train=True, progressive=False, **kw)
        means.append(float(gs.readback()[..., :3].mean()))
    m, se = float(np.mean(means)), float(np.std(means, ddof=1) / np.sqrt(len(means)))
    r = float(ref.mean())
    print(f"  env_nee={env_nee}: guided mean {m:.5f} +- {se:.5f}, plain 4096 spp {r:.5f}")
    assert abs(m - r) <= 4 * se + 2e-3 * r
    gs.close()
    g.close()


def test_trained_guide_lowers_error_at_equal_samples(gpu_ctx):
    sc = _terrain_small()
    gs = GpuScene.from_scene(gpu_ctx, sc)
    o = RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE, traversal=A.TH_TRAVERSE_WIDE)
    print("case guided gain: plain reference 8192 spp", flush=True)
    gs.render(samples=8192, options=o, frame_index=99)
    ref = gs.readback()
    gs.render(samples=64, options=o, frame_index=5)
    plain = rel_rmse(gs.readback(), ref)
    g = Guide(gpu_ctx, seed=6)
    kw = dict(batch=1 << 14, warmup_steps=64, warmup_rate=4, alpha_ramp=64, train_every=1)
    gs.frame_index = 11
    render_guided(gs, g, 64, train=True, progressive=False, **kw)
    gs.frame_index = 12
    st = render_guided(gs, g, 64, train=True, progressive=False, **kw)
    guided = rel_rmse(gs.readback(), ref)
    print(f"  relRMSE at 64 spp: plain {plain:.4f} guided {guided:.4f} stats {st}")
    assert guided < plain
    gs.close()
    g.close()
