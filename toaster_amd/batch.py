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
        self.groups = []   # tuning group per item (frames of one animation share one)
        self.tuning = []   # (quads_per_pixel, cached_nodes, best_ms, tile_width) per item, or None
        self.blocks_per_cu = 0  # batch-level workers per compute unit (0 = library default)

    def add(self, scene, group=None):
        """Uploads a scene (or animation frame) as a batch item; items with the same
        `group` share one tuning (default: each item its own)."""
        gs = GpuScene(self.ctx, pack(scene))
        self.scenes.append(gs)
        self.groups.append(len(self.groups) if group is None else group)
        self.tuning.append(None)
        return gs

    def autotune(self, options: RenderOptions = None):
        """th_autotune on the first item of every group (quads per pixel, LDS treelet, tile
        width by coordinate descent), applied to the group's items."""
        o = (options or RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE)).struct()
        L = A.lib()
        done = {}
        for i, gs in enumerate(self.scenes):
            g = self.groups[i]
            if g not in done:
                t = A.Tuning()
                p = gs.params(frame_index=0, accumulated_samples=0)
                A.check(L.th_autotune(self.ctx.handle, gs._h, C.byref(p), C.byref(gs.packed.camera),
                                      C.byref(o), C.byref(t)))
                done[g] = (t.quads_per_pixel, t.cached_nodes, t.best_ms, t.tile_width)
            self.tuning[i] = done[g]
        return self.tuning

    def autotune_batch(self, options: RenderOptions = None, use_tuning=True):
        """th_autotune_batch: the persistent blocks per compute unit for the whole batch."""
        o = (options or RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE)).struct()
        arr = self.items(use_tuning=use_tuning)
        n, ms = C.c_uint32(), C.c_float()
        A.check(A.lib().th_autotune_batch(self.ctx.handle, arr, len(arr), C.byref(o), C.byref(n), C.byref(ms)))
        self.blocks_per_cu = n.value
        return n.value, ms.value

    def items(self, frame_index=0, use_tuning=True):
        arr = (A.BatchItem * len(self.scenes))()
        for i, gs in enumerate(self.scenes):
            arr[i].scene = gs._h
            arr[i].params = gs.params(frame_index=frame_index, accumulated_samples=0)
            arr[i].camera = gs.packed.camera
            tu = self.tuning[i] if use_tuning else None
            arr[i].quads_per_pixel = tu[0] if tu else 0
            arr[i].cached_nodes = tu[1] if tu else 0xFFFFFFFF
            arr[i].tile_width = tu[3] if tu else 0
        return arr

    def render(self, options: RenderOptions = None, frame_index=0, use_tuning=True, items=None):
        o = (options or RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE)).struct()
        if use_tuning and self.blocks_per_cu and not o.batch_blocks_per_cu:
            o.batch_blocks_per_cu = self.blocks_per_cu
        arr = items if items is not None else self.items(frame_index, use_tuning)
        A.check(A.lib().th_render_batch(self.ctx.handle, arr, len(arr), C.byref(o)))

    @property
    def paths(self):
        return sum(gs.width * gs.height * gs.packed.params.samples for gs in self.scenes)

    def close(self):
        for gs in self.scenes:
            gs.close()
        self.scenes = []
