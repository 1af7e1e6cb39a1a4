"""A progressive session through a frame sink equals the same batches rendered by hand
(th_render with the running average), bit for bit; the sink can stop the session or set
the next batch."""
import numpy as np
import pytest

from toaster_amd import GpuScene, RenderOptions, _abi as A, scenes
from toaster_amd.progressive import FrameSink, render_progressive_with_sink

pytestmark = pytest.mark.gpu


class Recorder(FrameSink):
    def __init__(self, stop_after=None, batches=None):
        self.frames, self.stop_after, self.batches = [], stop_after, list(batches or [])

    def deliver(self, frame):
        self.frames.append((frame.index, frame.samples, frame.accumulated_samples, frame.image.copy(),
                            frame.hdr().copy()))

    def should_continue(self):
        return self.stop_after is None or len(self.frames) < self.stop_after

    def samples_for_frame(self):
        return self.batches.pop(0) if self.batches else None


@pytest.mark.parametrize("use_graph", [False, True])
def test_session_equals_manual_batches(gpu_ctx, use_graph):
    sc = scenes.cornell(samples=4)
    opts = RenderOptions(traversal=A.TH_TRAVERSE_WIDE)
    print(f"case progressive session cornell target 10 graph={use_graph}", flush=True)
    sink = Recorder()
    out = render_progressive_with_sink(gpu_ctx, sc, 10, sink, options=opts, use_graph=use_graph)
    assert out["accumulated_samples"] == 10 and out["frames"] == 3
    assert [(i, n, acc) for i, n, acc, _, _ in sink.frames] == [(0, 4, 4), (1, 4, 8), (2, 2, 10)]
    gs = GpuScene.from_scene(gpu_ctx, sc)
    acc = 0
    for i, n in enumerate((4, 4, 2)):
        gs.render(samples=n, frame_index=i, accumulated_samples=acc, options=opts)
        acc += n
        assert np.array_equal(gs.readback().view(np.uint32), sink.frames[i][4].view(np.uint32))
    assert np.array_equal(gs.display_rgba8(), sink.frames[-1][3])
    gs.close()


def test_sink_stops_and_sets_batches(gpu_ctx):
    sc = scenes.cornell(samples=4)
    print("case progressive sink control", flush=True)
    sink = Recorder(stop_after=2)
    out = render_progressive_with_sink(gpu_ctx, sc, 100, sink)
    assert out["frames"] == 2 and out["accumulated_samples"] == 8
    sink = Recorder(batches=[1, 3])
    out = render_progressive_with_sink(gpu_ctx, sc, 6, sink)
    assert [(n, acc) for _, n, acc, _, _ in sink.frames] == [(1, 1), (3, 4), (2, 6)]


def test_graph_session_follows_changing_batches(gpu_ctx):
    """Replayed batches (one captured graph per run of equal batch sizes) deliver the
    eager session's frames bit for bit while the sink changes the batch size."""
    sc = scenes.cornell(64, 36, samples=2)
    opts = RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=A.TH_SCHEDULE_SPECULATIVE)
    print("case progressive graph session with changing batches", flush=True)
    runs = []
    for use_graph in (False, True):
        sink = Recorder(batches=[2, 2, 3, 3, 1, 2])
        render_progressive_with_sink(gpu_ctx, sc, 14, sink, options=opts, use_graph=use_graph)
        runs.append(sink.frames)
    assert [(n, acc) for _, n, acc, _, _ in runs[1]] == [(2, 2), (2, 4), (3, 7), (3, 10), (1, 11), (2, 13), (1, 14)]
    for a, b in zip(*runs):
        assert a[:3] == b[:3]
        assert np.array_equal(a[4].view(np.uint32), b[4].view(np.uint32))
        assert np.array_equal(a[3], b[3])


def _moving_cornell(frame):
    from toaster_amd.progressive import EvaluatedFrame
    sc = scenes.cornell(width=96, height=64, samples=2)
    sc.spheres["center"][0, 0] = sc.spheres["center"][0, 0] + np.float32(0.05 * frame)
    return EvaluatedFrame(scene=sc, camera=False, spheres=True, triangles=False, emissive_geometry=False,
                          time_seconds=frame / 30.0)


@pytest.mark.parametrize("update", ["rebuild", "refit"])
def test_animation_frames_equal_single_renders(gpu_ctx, update):
    """An evaluated animation through a frame sink: each delivered frame is that frame's
    scene rendered on its own with frame seed = frame index, byte for byte, whether the
    moved geometry was rebuilt or refitted."""
    from toaster_amd.progressive import render_animation_with_sink
    opts = RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=A.TH_SCHEDULE_SPECULATIVE)
    print(f"case animation sink update={update}", flush=True)
    sink = Recorder()
    out = render_animation_with_sink(gpu_ctx, _moving_cornell, 3, sink, options=opts, update=update)
    assert out["frames"] == 3 and [f[0] for f in sink.frames] == [0, 1, 2]
    for index, _, _, image, _ in sink.frames:
        gs = GpuScene.from_scene(gpu_ctx, _moving_cornell(index).scene)
        gs.render(frame_index=index, accumulated_samples=0, options=opts)
        assert np.array_equal(gs.display_rgba8(), image)
        gs.close()
    assert not np.array_equal(sink.frames[0][3], sink.frames[2][3])


def test_static_independent_frames_repeat(gpu_ctx):
    """static_independent (the benchmark mode): time zero and seed zero every frame, so
    every frame is the same image; the sink can stop the loop early."""
    from toaster_amd.progressive import render_animation_with_sink
    sink = Recorder(stop_after=2)
    print("case animation static independent", flush=True)
    out = render_animation_with_sink(gpu_ctx, _moving_cornell, 5, sink, frame_mode="static_independent")
    assert out["frames"] == 2
    assert np.array_equal(sink.frames[0][3], sink.frames[1][3])
