"""Environment importance sampling on the device (the row-then-texel CDF search over the
compact CDF) against the oracle's flat search (wgsl:345-381) on maps with
zero-probability texels and whole zero rows at the top, middle and bottom, one bright
texel, and a map whose rows are a single texel wide; every kernel, reference and
per-sample streams (tolerance of tests/test_gpu_parity.py).  The 1024x512 sun map of
the mixed-geometry workload is covered by tests/test_gpu_configs.py."""
import numpy as np
import pytest

import oracle
from toaster_amd import GpuScene, RenderOptions, _abi as A
from toaster_amd.guide import Guide, render_guided
from toaster_amd.scene import EnvironmentMap

from .helpers import scene

pytestmark = pytest.mark.gpu

TOL_FRAC = 0.995


def _map(w, h, kind):
    rng = np.random.default_rng(w * 1000 + h)
    rgb = (0.2 + rng.random((h, w, 3))).astype(np.float32)
    if kind == "zero_rows":
        rgb[: h // 8] = 0.0
        rgb[h // 2 - 2: h // 2 + 1] = 0.0
        rgb[-h // 6:] = 0.0
        rgb[:, w // 4: w // 3] = 0.0
        rgb[h // 3, w // 2] = 400.0
    return EnvironmentMap(rgb, 1.5, 30.0)


def _scene(env):
    g = 4.0
    mats = [(A.TH_MAT_DIFFUSE, (0.6, 0.55, 0.5), 0, 0, 0, 0), (A.TH_MAT_METAL, (0.9, 0.8, 0.7), 0.2, 0, 0, 0),
            (A.TH_MAT_DIELECTRIC, (1, 1, 1), 0, 1.5, 0, 0), (A.TH_MAT_DIFFUSE, (0.3, 0.6, 0.3), 0, 0, 0, 0)]
    tris = [([(-g, 0, g), (g, 0, g), (g, 0, -g)], 0), ([(-g, 0, g), (g, 0, -g), (-g, 0, -g)], 0)]
    sph = [((-0.8, 0.5, -0.5), 0.5, 1), ((0.4, 0.4, 0.2), 0.4, 2), ((1.2, 0.6, -1.0), 0.6, 3)]
    return scene(mats, sph=sph, tris=tris, background=A.TH_BG_ENVIRONMENT, env=env, width=64, height=36,
                 samples=4, max_bounces=6, cam=((0.0, 1.5, 4.0), (0.0, 0.4, 0.0)), fov=50.0)


MAPS = {"zero_rows_96x48": (96, 48, "zero_rows"), "one_column_1x64": (1, 64, "plain"),
        "one_row_128x1": (128, 1, "plain")}
KERNELS = {
    "reference": RenderOptions(),
    "quad": RenderOptions(traversal=A.TH_TRAVERSE_WIDE),
    "speculative": RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=A.TH_SCHEDULE_SPECULATIVE),
    "wavefront": RenderOptions(traversal=A.TH_TRAVERSE_WIDE, kernel=A.TH_KERNEL_WAVEFRONT),
    "per_sample_quad": RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE, traversal=A.TH_TRAVERSE_WIDE),
}


def _within(img, ref):
    close = np.abs(img - ref) <= 1e-4 * (1.0 + np.abs(ref))
    return float(np.mean(np.all(close, axis=-1)))


@pytest.mark.parametrize("kernel", list(KERNELS))
@pytest.mark.parametrize("name", list(MAPS))
def test_env_sampling_matches_oracle(gpu_ctx, name, kernel):
    sc = _scene(_map(*MAPS[name]))
    print(f"case env search {name} kernel={kernel}", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    o = KERNELS[kernel]
    gs.render(options=o, frame_index=0, accumulated_samples=0)
    img = gs.readback().copy()
    gs.close()
    group = 4 if o.rng_mode == A.TH_RNG_PER_SAMPLE else 1
    ref, _ = oracle.OracleScene(sc).render(threads=8, rng_mode=o.rng_mode, group=group)
    frac = _within(img, ref)
    print(f"  within tol {frac:.5f}")
    assert np.isfinite(img).all()
    assert frac >= TOL_FRAC


def test_env_sampling_guided_alpha_zero(gpu_ctx):
    sc = _scene(_map(*MAPS["zero_rows_96x48"]))
    gs = GpuScene.from_scene(gpu_ctx, sc)
    g = Guide(gpu_ctx, seed=2)
    render_guided(gs, g, sc.samples, alpha=0.0, train=True, batch=1024, progressive=False, warmup_steps=0)
    img = gs.readback().copy()
    gs.close()
    g.close()
    ref, _ = oracle.OracleScene(sc).render(threads=8, rng_mode=A.TH_RNG_PER_SAMPLE, group=1)
    assert _within(img, ref) >= TOL_FRAC
