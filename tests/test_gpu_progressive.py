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


def test_session_equals_manual_batches(gpu_ctx):
    sc = scenes.cornell(samples=4)
    opts = RenderOptions(traversal=A.TH_TRAVERSE_WIDE)
    print("case progressive session cornell target 10", flush=True)
    sink = Recorder()
    out = render_progressive_with_sink(gpu_ctx, sc, 10, sink, options=opts)
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
