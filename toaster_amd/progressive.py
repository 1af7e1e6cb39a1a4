"""Sink-driven rendering as the reference's renderer loop runs it: progressive sessions
(a progressive configuration) and evaluated animations (crates/toaster-gpu/src/
path_tracer.rs:516-797, 891-925).

Progressive: each frame renders the next batch of samples into the running average (the
sink's requested count or the scene's, clamped to what remains of the target), the frame
seed is the frame counter, and the scene is uploaded once.  Animation: each frame is the
evaluator's scene at that frame, its changes uploaded (geometry rebuilt or refitted),
rendered on its own with its frame seed.  Either way the image is converted to RGBA8 and
handed to the sink, which may stop the loop or set the next frame's samples.
"""
import time
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import _abi as A
from .renderer import GpuContext, GpuScene, RenderOptions
from .scene import Scene, pack


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
                                 options: RenderOptions = None, exposure_stops: float = 0.0,
                                 use_graph: bool = False) -> dict:
    """Renders `scene` progressively until `target_samples` are accumulated or the sink
    stops; returns {frames, accumulated_samples, total_time}.  use_graph=True replays each
    batch from a captured hipGraph (one graph per batch size, GpuScene.progressive_graph)
    instead of launching it: same images, frame for frame."""
    if target_samples <= 0:
        raise ValueError("sample target must be greater than zero")
    t_total = time.perf_counter()
    gs = GpuScene.from_scene(ctx, scene)
    graph = {}
    frame, accumulated = 0, 0

    def render_batch(n):
        if not use_graph:
            gs.render(samples=n, frame_index=frame, accumulated_samples=accumulated, options=options)
            return
        g = graph.get("g")
        if g is None or graph["samples"] != n:  # (re)captured where the loop stands
            if g is not None:
                g.close()
            gs.frame_index, gs.accumulated_samples = frame, accumulated
            graph["g"], graph["samples"] = gs.progressive_graph(n, options), n
        graph["g"].replay(1)

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
            render_batch(n)
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


@dataclass
class EvaluatedFrame:
    """What a scene evaluator returns for one frame (toaster-scene's evaluation result as
    path_tracer.rs:609-640 consumes it): the scene at that time, which parts changed since
    the previous frame, and the animation time."""
    scene: Scene
    camera: bool = True
    spheres: bool = True
    triangles: bool = True
    emissive_geometry: bool = True
    time_seconds: float = 0.0


def render_frame_seed(frame: int, frame_mode: str) -> int:
    """path_tracer.rs:891-897: the frame counter when animated, 0 for independent
    benchmark frames."""
    if frame_mode == "animated":
        return frame
    if frame_mode == "static_independent":
        return 0
    raise ValueError("frame_mode must be 'animated' or 'static_independent'")


def render_animation_with_sink(ctx: GpuContext, evaluate: Callable[[int], EvaluatedFrame], frame_count: int,
                               sink: FrameSink, options: RenderOptions = None, frame_mode: str = "animated",
                               update: str = "rebuild", rebuild_above: Optional[float] = None,
                               exposure_stops: float = 0.0) -> dict:
    """The reference's sink-driven animation loop (path_tracer.rs:516-797 without a
    progressive configuration): the first evaluation sets the scene up; every frame is
    evaluated (time zero in static_independent mode), must keep its counts, has what
    changed uploaded -- geometry with a rebuilt tree ("rebuild", the reference's way) or
    the uploaded tree refitted on the device ("refit", optionally rebuilt once its summed
    box area passes `rebuild_above`) -- and is rendered with its frame seed, converted to
    RGBA8 and delivered.  Returns {frames, total_time}."""
    if frame_count <= 0:
        raise ValueError("animation needs at least one frame")
    if update not in ("rebuild", "refit"):
        raise ValueError("update must be 'rebuild' or 'refit'")
    render_frame_seed(0, frame_mode)  # validates the mode
    t_total = time.perf_counter()
    first = evaluate(0)
    gs = GpuScene.from_scene(ctx, first.scene)
    counts = (len(first.scene.spheres), len(first.scene.triangles), len(first.scene.materials))
    frame = 0
    try:
        while frame < frame_count:
            if not sink.should_continue():
                break
            t0 = time.perf_counter()
            ev = first if frame == 0 else evaluate(0 if frame_mode == "static_independent" else frame)
            sc = ev.scene
            if (len(sc.spheres), len(sc.triangles), len(sc.materials)) != counts:
                raise ValueError("evaluated scenes must preserve geometry and material counts")
            t1 = time.perf_counter()
            geometry = ev.spheres or ev.triangles
            if frame > 0:
                if geometry and update == "refit":
                    gs.refit(sc, rebuild_above=rebuild_above)
                else:
                    mask = ((A.TH_CHANGE_CAMERA if ev.camera else 0) | (A.TH_CHANGE_SPHERES if ev.spheres else 0)
                            | (A.TH_CHANGE_TRIANGLES if ev.triangles else 0)
                            | (A.TH_CHANGE_LIGHTS if ev.emissive_geometry else 0)
                            | (A.TH_CHANGE_BVH if geometry else 0))
                    gs.update(pack(sc), mask)
            ctx.synchronize()
            t2 = time.perf_counter()
            requested = sink.samples_for_frame()
            requested = sc.samples if requested is None else int(requested)
            if requested <= 0:
                raise ValueError("frame sink sample count must be greater than zero")
            gs.render(samples=requested, frame_index=render_frame_seed(frame, frame_mode), accumulated_samples=0,
                      options=options)
            ctx.synchronize()
            t3 = time.perf_counter()
            image = gs.display_rgba8(exposure_stops)
            t4 = time.perf_counter()
            timings = {"evaluation": t1 - t0, "scene_update_upload": t2 - t1, "dispatch_wait": t3 - t2,
                       "conversion": t4 - t3, "output": 0.0, "total": t4 - t0}
            sink.deliver(CompletedFrame(index=frame, time_seconds=ev.time_seconds, width=gs.width,
                                        height=gs.height, samples=requested, accumulated_samples=requested,
                                        max_bounces=int(sc.max_bounces), timings=timings, render_time=t4 - t0,
                                        image=image, hdr=gs.readback))
            timings["output"] = time.perf_counter() - t4
            timings["total"] = time.perf_counter() - t0
            sink.record_timings(frame, timings)
            frame += 1
    finally:
        gs.close()
    return {"frames": frame, "total_time": time.perf_counter() - t_total}
