"""GPU renderer over the C ABI, shaped like toaster-gpu's renderer interface.

Mirrors crates/toaster-gpu/src/path_tracer.rs: a context (device.rs:49-100), a scene
uploaded once with fixed counts (buffers.rs:117-271, path_tracer.rs:639-653),
per-frame dispatches that fold `samples` new samples into the progressive
running average (wgsl:538-547, path_tracer.rs:655-674), linear-HDR readback
(readback.rs:8-25) and the exposure/ACES/sRGB display stage (image_output.rs:59-92).
"""
import atexit
import ctypes as C
import weakref
from dataclasses import dataclass

import numpy as np

from . import _abi as A
from .scene import PackedScene, Scene, pack


@dataclass
class RenderOptions:
    rng_mode: int = A.TH_RNG_REFERENCE
    kernel: int = A.TH_KERNEL_MEGA
    traversal: int = A.TH_TRAVERSE_REFERENCE
    collect_stats: bool = False
    schedule: int = A.TH_SCHEDULE_STATIC
    batch_blocks_per_cu: int = 0  # th_render_batch workers per compute unit (0 = 4)

    def struct(self):
        o = A.RenderOptions()
        o.rng_mode, o.kernel, o.traversal = self.rng_mode, self.kernel, self.traversal
        o.collect_stats = 1 if self.collect_stats else 0
        o.schedule = self.schedule
        o.batch_blocks_per_cu = self.batch_blocks_per_cu
        return o


def device_count():
    n = C.c_int32()
    A.check(A.lib().th_device_count(C.byref(n)))
    return n.value


# Live handles, released in dependency order (scenes and guides before their contexts)
# while the interpreter is still whole at exit.
_dependents = weakref.WeakSet()
_contexts = weakref.WeakSet()


@atexit.register
def _close_all():
    for group in (_dependents, _contexts):
        for obj in list(group):
            try:
                obj.close()
            except Exception:
                pass


class GpuContext:
    def __init__(self, device=0):
        self._lib = A.lib()
        self._h = C.c_void_p()
        A.check(self._lib.th_context_create(int(device), C.byref(self._h)))
        self._children = weakref.WeakSet()  # scenes and guides, released before the context
        _contexts.add(self)

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        A.check(self._lib.th_synchronize(self._h))

    def stats(self):
        s = A.Stats()
        A.check(self._lib.th_get_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in A.Stats._fields_}

    def reset_stats(self):
        A.check(self._lib.th_reset_stats(self._h))

    def measure_copy(self, nbytes=1 << 30, reps=8, read_only=False):
        """Streaming-copy (or read) bandwidth of the device (th_measure_copy): GB/s read +
        written."""
        gbs = C.c_double()
        A.check(self._lib.th_measure_copy(self._h, int(nbytes), int(reps), 1 if read_only else 0, C.byref(gbs)))
        return gbs.value

    def measure_mfma(self, shape=1, waves_per_simd=2, iters=16384):
        """Dense BF16 MFMA throughput (th_measure_mfma): (TFLOP/s, every product exact).
        shape 0: v_mfma_f32_16x16x32_bf16; 1: v_mfma_f32_32x32x16_bf16."""
        tf, ok = C.c_double(), A.u32()
        A.check(self._lib.th_measure_mfma(self._h, int(shape), int(waves_per_simd), int(iters), C.byref(tf),
                                          C.byref(ok)))
        return tf.value, bool(ok.value)

    def adopt(self, child):
        """Registers an object (scene, guide) that must be released before this context."""
        self._children.add(child)

    def close(self):
        if self._h:
            for child in list(self._children):
                child.close()
            self._lib.th_context_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuScene:
    """An uploaded scene plus its accumulation buffer."""

    def __init__(self, ctx: GpuContext, packed: PackedScene, node_capacity=0):
        self.ctx = ctx
        self.packed = packed
        self._lib = ctx._lib
        self._h = C.c_void_p()
        d = packed.desc()
        A.check(self._lib.th_scene_create(ctx.handle, C.byref(d), int(node_capacity), C.byref(self._h)))
        _dependents.add(self)
        ctx.adopt(self)
        self.width, self.height = packed.params.width, packed.params.height
        self.accumulated_samples = 0
        self.frame_index = 0
        self._graphs = weakref.WeakSet()  # progressive graphs, destroyed before the scene

    @classmethod
    def from_scene(cls, ctx, scene: Scene, node_capacity=0):
        return cls(ctx, pack(scene), node_capacity)

    def update(self, packed: PackedScene, change_mask):
        d = packed.desc()
        A.check(self._lib.th_scene_update(self._h, int(change_mask), C.byref(d)))
        if change_mask & A.TH_CHANGE_REFIT:  # the uploaded tree stays: keep its counts
            packed.params.bvh_node_count = self.packed.params.bvh_node_count
            packed.params.bvh_primitive_count = self.packed.params.bvh_primitive_count
        self.packed = packed

    def refit(self, scene: Scene, rebuild_above=None):
        """Moved geometry (same primitives, materials and lights count) without a host
        BVH rebuild: new records and light list, node boxes recomputed on the device over
        the uploaded tree.  The reference rebuilds instead (docs/architecture.md:118-121).
        With `rebuild_above`, a refit that loosens the tree past that ratio of summed box
        area (refit_quality) is followed by a rebuild; returns True when it rebuilt."""
        self.update(pack(scene, bvh=False),
                    A.TH_CHANGE_SPHERES | A.TH_CHANGE_TRIANGLES | A.TH_CHANGE_LIGHTS | A.TH_CHANGE_CAMERA
                    | A.TH_CHANGE_REFIT)
        if rebuild_above is not None and self.refit_quality() > rebuild_above:
            self.update(pack(scene), A.TH_CHANGE_SPHERES | A.TH_CHANGE_TRIANGLES | A.TH_CHANGE_BVH)
            return True
        return False

    def refit_quality(self):
        """Summed node-box area after the last refit over that of the tree as built."""
        r = C.c_float()
        A.check(self._lib.th_scene_refit_quality(self._h, C.byref(r)))
        return float(r.value)

    def traversal_info(self):
        """(primitives the 4-wide kernels test before the tree, 4-wide nodes)."""
        e, n = C.c_uint32(), C.c_uint32()
        A.check(self._lib.th_scene_traversal_info(self._h, C.byref(e), C.byref(n)))
        return int(e.value), int(n.value)

    def params(self, samples=None, frame_index=None, accumulated_samples=None, max_bounces=None):
        p = A.GpuRenderParams.from_buffer_copy(self.packed.params)
        if samples is not None:
            p.samples = samples
        if max_bounces is not None:
            p.max_bounces = max_bounces
        p.frame_index = self.frame_index if frame_index is None else frame_index
        p.accumulated_samples = (self.accumulated_samples if accumulated_samples is None
                                 else accumulated_samples)
        return p

    def render(self, samples=None, frame_index=None, accumulated_samples=None,
               options: RenderOptions = None, camera=None, max_bounces=None):
        """One asynchronous dispatch (dispatch_compute_2d without the wait)."""
        p = self.params(samples, frame_index, accumulated_samples, max_bounces)
        cam = self.packed.camera if camera is None else camera
        o = (options or RenderOptions()).struct()
        A.check(self._lib.th_render(self.ctx.handle, self._h, C.byref(p), C.byref(cam), C.byref(o)))
        return p

    def render_progressive(self, samples, options=None):
        """Progressive batch: frame seed = batch counter, running average continues."""
        p = self.render(samples=samples, frame_index=self.frame_index,
                        accumulated_samples=self.accumulated_samples, options=options)
        self.accumulated_samples += p.samples
        self.frame_index += 1
        return p

    def progressive_graph(self, samples, options: RenderOptions = None):
        """A captured progressive batch of `samples` samples (th_progressive_graph_*)
        continuing this scene's frame seed and running average; replay(n) renders the
        next n batches."""
        return ProgressiveGraph(self, samples, options)

    def readback(self):
        out = np.empty((self.height, self.width, 4), np.float32)
        A.check(self._lib.th_readback(self.ctx.handle, self._h, out.ctypes.data, out.nbytes))
        return out

    def upload_accumulation(self, rgba, accumulated_samples, frame_index):
        """Restores a progressive checkpoint (accum buffer, sample count, next seed)."""
        rgba = np.ascontiguousarray(rgba, np.float32)
        A.check(self._lib.th_upload_accumulation(self.ctx.handle, self._h, rgba.ctypes.data, rgba.nbytes))
        self.accumulated_samples = int(accumulated_samples)
        self.frame_index = int(frame_index)

    def display_rgba8(self, exposure_stops=0.0):
        out = np.empty((self.height, self.width, 4), np.uint8)
        A.check(self._lib.th_display_rgba8_device(self.ctx.handle, self._h, float(exposure_stops),
                                                   out.ctypes.data, out.nbytes))
        return out

    def primary_hits(self, traversal=A.TH_TRAVERSE_REFERENCE, frame_index=0):
        n = self.width * self.height
        slots = np.empty(n, np.uint32)
        dist = np.empty(n, np.float32)
        p = self.params(frame_index=frame_index)
        A.check(self._lib.th_primary_hits(self.ctx.handle, self._h, C.byref(p),
                                          C.byref(self.packed.camera), int(traversal),
                                          slots.ctypes.data, dist.ctypes.data, n))
        return slots, dist

    def close(self):
        if self._h:
            for g in list(self._graphs):
                g.close()
            self._lib.th_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ProgressiveGraph:
    """One progressive batch captured as a hipGraph (th_progressive_graph_create); each
    replayed frame is the next batch of the scene's progressive loop, and the scene's
    frame seed and sample count follow the replays."""

    def __init__(self, scene: GpuScene, samples, options: RenderOptions = None):
        self.scene = scene
        p = scene.params(samples=samples)
        o = (options or RenderOptions()).struct()
        h = C.c_void_p()
        A.check(scene._lib.th_progressive_graph_create(scene.ctx.handle, scene._h, C.byref(p),
                                                       C.byref(scene.packed.camera), C.byref(o), C.byref(h)))
        self._h = h
        scene._graphs.add(self)

    def replay(self, frames=1):
        A.check(self.scene._lib.th_progressive_graph_replay(self._h, int(frames)))
        f, a = C.c_uint32(), C.c_uint32()
        A.check(self.scene._lib.th_progressive_graph_state(self._h, C.byref(f), C.byref(a)))
        self.scene.frame_index, self.scene.accumulated_samples = f.value, a.value

    def close(self):
        if self._h:
            self.scene._lib.th_progressive_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def display_rgba8(pixels, exposure_stops=0.0):
    """Host display conversion of linear-HDR float RGBA (pixels_to_rgba_image)."""
    pixels = np.ascontiguousarray(pixels, np.float32)
    h, w = pixels.shape[0], pixels.shape[1]
    out = np.empty((h, w, 4), np.uint8)
    A.check(A.lib().th_display_rgba8(pixels.ctypes.data, w, h, float(exposure_stops),
                                     out.ctypes.data, out.nbytes))
    return out


def render(scene: Scene, options: RenderOptions = None, device=0):
    """Renders one frame of a scene (frame seed 0) and returns linear-HDR (H, W, 4)."""
    with GpuContext(device) as ctx:
        gs = GpuScene.from_scene(ctx, scene)
        gs.render(options=options)
        img = gs.readback()
        gs.close()
        return img
