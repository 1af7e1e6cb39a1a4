"""Batched rendering of many scenes / animation frames in one persistent launch.

Each (scene, frame) is packed on the host (BVH rebuilt per frame, as the reference
does, docs/architecture.md:118-121) and uploaded as its own device scene; one
th_render_batch call renders all of them, replayed from a captured hipGraph when the
batch shape repeats; th_autotune picks each scene's launch shape.
"""
import ctypes as C

from . import _abi as A
from .renderer import GpuContext, GpuScene, RenderOptions
from .scene import pack


class BatchRenderer:
    def __init__(self, ctx: GpuContext):
        self.ctx = ctx
        self.scenes = []   # GpuScene per item
        self.tuning = []   # th_tuning per item (or None)

    def add(self, scene):
        gs = GpuScene(self.ctx, pack(scene))
        self.scenes.append(gs)
        self.tuning.append(None)
        return gs

    def autotune(self, options: RenderOptions = None):
        o = (options or RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE)).struct()
        L = A.lib()
        for i, gs in enumerate(self.scenes):
            t = A.Tuning()
            p = gs.params(frame_index=0, accumulated_samples=0)
            A.check(L.th_autotune(self.ctx.handle, gs._h, C.byref(p), C.byref(gs.packed.camera),
                                  C.byref(o), C.byref(t)))
            self.tuning[i] = (t.quads_per_pixel, t.cached_nodes, t.best_ms)
        return self.tuning

    def items(self, frame_index=0, use_tuning=True):
        arr = (A.BatchItem * len(self.scenes))()
        for i, gs in enumerate(self.scenes):
            arr[i].scene = gs._h
            arr[i].params = gs.params(frame_index=frame_index, accumulated_samples=0)
            arr[i].camera = gs.packed.camera
            tu = self.tuning[i] if use_tuning else None
            arr[i].quads_per_pixel = tu[0] if tu else 0
            arr[i].cached_nodes = tu[1] if tu else 0xFFFFFFFF
        return arr

    def render(self, options: RenderOptions = None, frame_index=0, use_tuning=True, items=None):
        o = (options or RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE)).struct()
        arr = items if items is not None else self.items(frame_index, use_tuning)
        A.check(A.lib().th_render_batch(self.ctx.handle, arr, len(arr), C.byref(o)))

    @property
    def paths(self):
        return sum(gs.width * gs.height * gs.packed.params.samples for gs in self.scenes)

    def close(self):
        for gs in self.scenes:
            gs.close()
        self.scenes = []
