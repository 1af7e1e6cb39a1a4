"""Progressive rendering into a frame sink, as the reference's renderer loop runs it with a
progressive configuration (crates/toaster-gpu/src/path_tracer.rs:516-797, 900-925).

Each frame renders the next batch of samples into the running average: the batch is the
sink's requested count (or the scene's), clamped to what remains of the target; the
frame seed is the frame counter; the accumulated image is converted to RGBA8 and handed
to the sink, which may stop the session or change the next batch.  The scene is
uploaded once (a progressive session re-uploads only the render parameters).
"""
import time
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from .renderer import GpuContext, GpuScene, RenderOptions
from .scene import Scene


def progressive_batch(requested_samples: int, accumulated_samples: int, target_samples: int) -> Optional[int]:
    """Next batch clamped to the remaining target; None once the target is reached
    (path_tracer.rs:900-925)."""
    if requested_samples <= 0:
        raise ValueError("sample batch must be greater than zero")
    if target_samples <= 0:
        raise ValueError("sample target must be greater than zero")
    if accumulated_samples > target_samples:
        raise ValueError("accumulated samples exceed the progressive target")
    if accumulated_samples == target_samples:
        return None
    return min(requested_samples, target_samples - accumulated_samples)


@dataclass
class CompletedFrame:
    """One delivered frame (path_tracer.rs:131-160).  `image` (RGBA8, H x W x 4) and
    `hdr()` (the accumulated linear float RGBA) are valid during delivery only."""
    index: int
    time_seconds: float
    width: int
    height: int
    samples: int
    accumulated_samples: int
    max_bounces: int
    timings: dict
    render_time: float
    image: np.ndarray
    hdr: Callable[[], np.ndarray] = field(repr=False, default=None)


class FrameSink:
    """Consumer of completed frames (path_tracer.rs:163-190)."""

    def deliver(self, frame: CompletedFrame) -> None:
        raise NotImplementedError

    def should_continue(self) -> bool:
        return True

    def samples_for_frame(self) -> Optional[int]:
        return None

    def record_timings(self, index: int, timings: dict) -> None:
        pass


def render_progressive_with_sink(ctx: GpuContext, scene: Scene, target_samples: int, sink: FrameSink,
                                 options: RenderOptions = None, exposure_stops: float = 0.0) -> dict:
    """Renders `scene` progressively until `target_samples` are accumulated or the sink
    stops; returns {frames, accumulated_samples, total_time}."""
    if target_samples <= 0:
        raise ValueError("sample target must be greater than zero")
    t_total = time.perf_counter()
    gs = GpuScene.from_scene(ctx, scene)
    frame, accumulated = 0, 0
    try:
        while accumulated < target_samples:
            if not sink.should_continue():
                break
            t0 = time.perf_counter()
            requested = sink.samples_for_frame()
            requested = scene.samples if requested is None else int(requested)
            if requested <= 0:
                raise ValueError("frame sink sample count must be greater than zero")
            n = progressive_batch(requested, accumulated, target_samples)
            gs.render(samples=n, frame_index=frame, accumulated_samples=accumulated, options=options)
            ctx.synchronize()
            t1 = time.perf_counter()
            image = gs.display_rgba8(exposure_stops)
            t2 = time.perf_counter()
            timings = {"dispatch_wait": t1 - t0, "conversion": t2 - t1, "output": 0.0, "total": t2 - t0}
            accumulated += n
            sink.deliver(CompletedFrame(index=frame, time_seconds=0.0, width=gs.width, height=gs.height,
                                        samples=n, accumulated_samples=accumulated,
                                        max_bounces=int(scene.max_bounces), timings=timings,
                                        render_time=t2 - t0, image=image, hdr=gs.readback))
            timings["output"] = time.perf_counter() - t2
            timings["total"] = time.perf_counter() - t0
            sink.record_timings(frame, timings)
            frame += 1
    finally:
        gs.close()
    return {"frames": frame, "accumulated_samples": accumulated, "total_time": time.perf_counter() - t_total}
