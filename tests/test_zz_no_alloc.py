"""th_render_wavefront_static (the wavefront render with its workspace in a static device
array, allocating nothing) against th_render's wavefront, which allocates its workspace
with the scene: bit-identical images on two calls in a row, the second on a smaller
input, for reference and per-sample streams and both traversal kinds, also folded into
an accumulated frame; renders above the array's 65,536 path slots are refused before
any device work."""
import ctypes as C

import numpy as np
import pytest

from toaster_amd import GpuContext, GpuScene, RenderOptions, _abi as A, scenes

pytestmark = pytest.mark.gpu

MAX_SLOTS = 65536


@pytest.fixture(scope="module")
def ctx():
    c = GpuContext(0)
    yield c
    c.close()


def _static(gs, opts, frame_index=0, accumulated_samples=0):
    p = gs.params(frame_index=frame_index, accumulated_samples=accumulated_samples)
    o = opts.struct()
    return A.lib().th_render_wavefront_static(gs.ctx.handle, gs._h, C.byref(p), C.byref(gs.packed.camera),
                                              C.byref(o))


def _pair(ctx, sc, opts):
    gs = GpuScene.from_scene(ctx, sc)
    gs.render(options=RenderOptions(rng_mode=opts.rng_mode, traversal=opts.traversal,
                                    kernel=A.TH_KERNEL_WAVEFRONT), frame_index=3, accumulated_samples=0)
    gs.render(options=RenderOptions(rng_mode=opts.rng_mode, traversal=opts.traversal,
                                    kernel=A.TH_KERNEL_WAVEFRONT), frame_index=4, accumulated_samples=sc.samples)
    ref = gs.readback().copy()
    assert _static(gs, opts, frame_index=3, accumulated_samples=0) == A.TH_OK, A.lib().th_last_error()
    assert _static(gs, opts, frame_index=4, accumulated_samples=sc.samples) == A.TH_OK, A.lib().th_last_error()
    img = gs.readback().copy()
    gs.close()
    return img, ref


CASES = [
    ("reference_rng_wide", RenderOptions(traversal=A.TH_TRAVERSE_WIDE)),
    ("reference_rng_scalar", RenderOptions(traversal=A.TH_TRAVERSE_REFERENCE)),
    ("per_sample_wide", RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE, traversal=A.TH_TRAVERSE_WIDE)),
]


@pytest.mark.parametrize("name,opts", CASES, ids=[c[0] for c in CASES])
def test_static_workspace_matches_the_allocating_wavefront(ctx, name, opts):
    # first call: 320x180 (57,600 slots with reference streams; per-sample: 64x48 x 4 groups)
    if opts.rng_mode == A.TH_RNG_PER_SAMPLE:
        big = scenes.mixed_environment(128, 96, samples=4, max_bounces=6, n_triangles=512, n_spheres=32,
                                       env_size=(256, 128))
        small = scenes.cornell(48, 27, samples=4)
    else:
        big = scenes.bvh_stress(8192, width=320, height=180, samples=2)
        small = scenes.mixed_environment(80, 45, samples=3, max_bounces=6, n_triangles=256, n_spheres=16,
                                         env_size=(256, 128))
    for sc in (big, small):  # two calls in a row, the second with a smaller input
        print(f"case static wavefront {name} {sc.name} {sc.width}x{sc.height}", flush=True)
        img, ref = _pair(ctx, sc, opts)
        assert np.isfinite(img).all()
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


def test_static_workspace_refuses_too_many_slots(ctx):
    sc = scenes.cornell(320, 240, samples=1)  # 76,800 slots
    gs = GpuScene.from_scene(ctx, sc)
    gs.render(frame_index=0, accumulated_samples=0)
    before = gs.readback().copy()
    assert _static(gs, RenderOptions(traversal=A.TH_TRAVERSE_WIDE)) == A.TH_E_CAPACITY
    ps = GpuScene.from_scene(ctx, scenes.cornell(128, 128, samples=8))  # 16,384 px x 8 groups
    assert _static(ps, RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE)) == A.TH_E_CAPACITY
    ps.close()
    after = gs.readback()
    gs.close()
    assert np.array_equal(before.view(np.uint32), after.view(np.uint32))
