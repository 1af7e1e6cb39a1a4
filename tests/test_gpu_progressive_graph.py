"""Progressive batches replayed from a captured hipGraph (th_progressive_graph_*): every
replayed frame equals the same batch rendered by th_render bit for bit (frame seed =
frame counter, running average over the accumulated samples; path_tracer.rs:655-674,
900-923), checked after each replay with the program's own eager work and
synchronizations in between; the accumulated image after the replays matches the
oracle's progressive frames (tolerance of tests/test_gpu_parity.py); a graph refuses to
replay once its scene was updated; unsupported options are refused."""
import ctypes as C

import numpy as np
import pytest

import oracle
from toaster_amd import GpuScene, RenderOptions, _abi as A, scenes
from toaster_amd.scene import pack

pytestmark = pytest.mark.gpu

WIDE = A.TH_TRAVERSE_WIDE
SCHEDULES = {
    "speculative": RenderOptions(traversal=WIDE, schedule=A.TH_SCHEDULE_SPECULATIVE),
    "octet": RenderOptions(traversal=WIDE, schedule=A.TH_SCHEDULE_OCTET),
    "static": RenderOptions(traversal=WIDE),
    "dynamic": RenderOptions(traversal=WIDE, schedule=A.TH_SCHEDULE_DYNAMIC),
    "scalar_reference": RenderOptions(),
    "per_sample_quad": RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE, traversal=WIDE),
}


@pytest.mark.parametrize("name", list(SCHEDULES))
def test_graph_replays_equal_eager_frames_and_the_oracle(gpu_ctx, name):
    opts = SCHEDULES[name]
    sc = scenes.mixed_environment(64, 36, samples=2, max_bounces=6, n_triangles=256, n_spheres=16,
                                  env_size=(256, 128))
    print(f"case progressive graph {name}", flush=True)
    eager = GpuScene.from_scene(gpu_ctx, sc)
    graphed = GpuScene.from_scene(gpu_ctx, sc)
    other = GpuScene.from_scene(gpu_ctx, scenes.cornell(32, 18, samples=1))
    g = graphed.progressive_graph(2, opts)
    frames = 5
    for f in range(frames):
        eager.render_progressive(2, options=opts)
        want = eager.readback().copy()
        g.replay(1)
        other.render(frame_index=f)  # eager work between replays
        gpu_ctx.synchronize()
        got = graphed.readback().copy()
        assert graphed.frame_index == f + 1 and graphed.accumulated_samples == 2 * (f + 1)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f
    g.replay(3)  # several frames in one call
    for _ in range(3):
        eager.render_progressive(2, options=opts)
    assert np.array_equal(graphed.readback().view(np.uint32), eager.readback().view(np.uint32))
    img = graphed.readback().copy()
    g.close()
    for gs in (eager, graphed, other):
        gs.close()
    oc = oracle.OracleScene(sc)
    acc = None
    group = 2 if opts.rng_mode == A.TH_RNG_PER_SAMPLE else 1
    for f in range(frames + 3):
        p = oc.params
        p.samples, p.frame_index, p.accumulated_samples = 2, f, 2 * f
        acc, _ = oc.render(params=p, threads=8, accum=acc, rng_mode=opts.rng_mode, group=group)
    close = np.abs(img - acc) <= 1e-4 * (1.0 + np.abs(acc))
    frac = float(np.mean(np.all(close, axis=-1)))
    print(f"  {frames + 3} replayed frames vs oracle: within tol {frac:.5f}")
    assert frac >= 0.995


def test_graph_refuses_replay_after_an_update(gpu_ctx):
    sc = scenes.cornell(32, 18, samples=1)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    g = gs.progressive_graph(1, SCHEDULES["static"])
    g.replay(2)
    gs.update(pack(sc), A.TH_CHANGE_CAMERA)
    assert A.lib().th_progressive_graph_replay(g._h, 1) == A.TH_E_MISMATCH
    g.close()
    gs.close()


def test_graph_options_are_validated(gpu_ctx):
    sc = scenes.cornell(32, 18, samples=1)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    for o in (RenderOptions(kernel=A.TH_KERNEL_WAVEFRONT), RenderOptions(traversal=WIDE, schedule=A.TH_SCHEDULE_AUTO),
              RenderOptions(collect_stats=True)):
        p = gs.params(samples=1)
        h = C.c_void_p()
        rc = A.lib().th_progressive_graph_create(gpu_ctx.handle, gs._h, C.byref(p), C.byref(gs.packed.camera),
                                                 C.byref(o.struct()), C.byref(h))
        assert rc == A.TH_E_UNSUPPORTED
    gs.close()
