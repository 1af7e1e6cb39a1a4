"""Python binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: a CPU restatement of the reference hot path (see
oracle.h for the file:line list).  Only tests/, the smoke check and the CPU-baseline
leg of bench.py import this module; the shipped library never does.

Parity pinning of each oracle function (known-answer tests of the reference, closed
forms, brute force) lives in tests/test_oracle_*.py.  Functions pinned only by
restated formulas are marked there as "parity unpinned".
"""
import ctypes as C
import os
import subprocess

import numpy as np

from toaster_amd import _abi as A
from toaster_amd.scene import PackedScene, Scene

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
FLAGS = ["-O3", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-Wall",
         "-Wextra"]


def build():
    """Compiles liboracle.so with g++ (contraction off so binary32 ops match the GPU's)."""
    tmp = LIB_PATH + ".tmp"
    subprocess.run(["g++"] + FLAGS + ["-o", tmp, os.path.join(HERE, "oracle.cpp")], check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


class Counters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "paths", "closest_rays", "shadow_rays", "nodes_visited", "primref_tests",
        "triangle_tests", "sphere_tests", "closest_hits", "triangle_hits", "lights_scanned",
        "env_lookups", "env_samples")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Hit(C.Structure):
    _fields_ = [("distance", C.c_float), ("point", C.c_float * 3), ("normal", C.c_float * 3),
                ("front_face", C.c_uint32), ("material_index", C.c_uint32),
                ("tex_coord", C.c_float * 2), ("slot", C.c_uint32)]


P, F, U = C.c_void_p, C.c_float, C.c_uint32
F3 = C.POINTER(C.c_float)
_SIGS = {
    "or_pack": (C.c_int32, [P, C.POINTER(P)]),
    "or_view": (C.c_int32, [P, C.POINTER(A.SceneDesc)]),
    "or_drop_bvh": (C.c_int32, [P]),
    "or_free": (None, [P]),
    "or_last_error": (C.c_char_p, []),
    "or_render": (C.c_int32, [P, C.POINTER(A.GpuRenderParams), C.POINTER(A.GpuCamera), U, U,
                              C.c_int32, P, C.POINTER(Counters)]),
    "or_render_ex": (C.c_int32, [P, C.POINTER(A.GpuRenderParams), C.POINTER(A.GpuCamera), U, U,
                                 C.c_int32, U, U, P, C.POINTER(Counters)]),
    "or_render_cpu": (C.c_int32, [P, C.POINTER(A.GpuRenderParams), C.POINTER(A.GpuCamera), U, U, C.c_int32,
                                  P]),
    "or_primary_hits": (C.c_int32, [P, C.POINTER(A.GpuRenderParams), C.POINTER(A.GpuCamera), P, P]),
    "or_closest": (C.c_int32, [P, F3, F3, C.c_int32, C.POINTER(Hit)]),
    "or_intersect_sphere": (C.c_int32, [F3, F3, F3, F, F, F, C.POINTER(Hit)]),
    "or_intersect_triangle": (C.c_int32, [F3, F3, F3, P, F, F, C.POINTER(Hit)]),
    "or_aabb_intersects": (C.c_int32, [F3, F3, F3, F3, F, F]),
    "or_env_lookup": (C.c_int32, [P, F3, F3]),
    "or_env_pdf": (F, [P, F3]),
    "or_env_sample": (C.c_int32, [P, F, F, F, F3, F3, C.POINTER(C.c_float)]),
    "or_reflectance": (F, [F, F]),
    "or_refract": (None, [F3, F3, F, F3]),
    "or_power_heuristic": (F, [F, F]),
    "or_ray_color": (C.c_int32, [P, C.POINTER(A.GpuRenderParams), F3, F3, C.POINTER(C.c_uint32), F3]),
    "or_pcg_hash": (U, [U]),
    "or_rand01": (F, [C.POINTER(C.c_uint32)]),
    "or_env_search_rows": (U, [P, P, U, U, F]),
    "or_sin": (F, [F]),
    "or_cos": (F, [F]),
    "or_atan2": (F, [F, F]),
    "or_acos": (F, [F]),
    "or_display_rgba8": (C.c_int32, [P, U, U, F, P]),
}
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run oracle.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def _check(code):
    if code < 0:
        raise RuntimeError("oracle: " + lib().or_last_error().decode())
    return code


class OracleScene:
    """The oracle's own packing of a Scene (restating scene_to_gpu)."""

    def __init__(self, scene: Scene, bvh=True):
        """bvh=False: no tree, every ray takes the linear scan (wgsl:667-698, 797-802)."""
        self.scene = scene
        inp, keep = scene.input()
        self._h = C.c_void_p()
        _check(lib().or_pack(C.byref(inp), C.byref(self._h)))
        del keep
        if not bvh:
            _check(lib().or_drop_bvh(self._h))
        d = A.SceneDesc()
        _check(lib().or_view(self._h, C.byref(d)))
        self.packed = PackedScene(d)

    @property
    def params(self):
        return A.GpuRenderParams.from_buffer_copy(self.packed.params)

    @property
    def camera(self):
        return self.packed.camera

    def render(self, params=None, rows=None, threads=1, accum=None, counters=False, rng_mode=0,
               group=1):
        """pcg-mode render of rows [r0, r1); returns (rgba (H, W, 4), counters dict|None).
        rng_mode 1: per-sample streams with `group` partial sums (the GPU's per-sample mode)."""
        p = params if params is not None else self.params
        h, w = p.height, p.width
        out = np.zeros((h, w, 4), np.float32) if accum is None else np.ascontiguousarray(accum, np.float32).copy()
        r0, r1 = (0, h) if rows is None else rows
        cnt = Counters()
        _check(lib().or_render_ex(self._h, C.byref(p), C.byref(self.camera), r0, r1, int(threads),
                                  int(rng_mode), int(group), out.ctypes.data,
                                  C.byref(cnt) if counters else None))
        return out, (cnt.as_dict() if counters else None)

    def render_cpu(self, params=None, rows=None, threads=1, accum=None):
        """toaster-cpu mode (the CPU renderer's estimator on its own generator, for
        converged comparison only) of rows [r0, r1); returns rgba (H, W, 4)."""
        p = params if params is not None else self.params
        h, w = p.height, p.width
        out = np.zeros((h, w, 4), np.float32) if accum is None else np.ascontiguousarray(accum, np.float32).copy()
        r0, r1 = (0, h) if rows is None else rows
        _check(lib().or_render_cpu(self._h, C.byref(p), C.byref(self.camera), r0, r1, int(threads),
                                   out.ctypes.data))
        return out

    def primary_hits(self, params=None):
        p = params if params is not None else self.params
        n = p.width * p.height
        slots = np.empty(n, np.uint32)
        dist = np.empty(n, np.float32)
        _check(lib().or_primary_hits(self._h, C.byref(p), C.byref(self.camera), slots.ctypes.data,
                                     dist.ctypes.data))
        return slots, dist

    def closest(self, origin, direction, use_bvh=True):
        h = Hit()
        hit = lib().or_closest(self._h, f3(origin), f3(direction), 1 if use_bvh else 0, C.byref(h))
        return (h if hit == 1 else None), h

    def env_lookup(self, direction):
        out = (C.c_float * 3)()
        _check(lib().or_env_lookup(self._h, f3(direction), out))
        return np.array(out[:], np.float32)

    def env_pdf(self, direction):
        return lib().or_env_pdf(self._h, f3(direction))

    def env_sample(self, selection, ju, jv):
        d, rgb, pdf = (C.c_float * 3)(), (C.c_float * 3)(), C.c_float()
        ok = lib().or_env_sample(self._h, selection, ju, jv, d, rgb, C.byref(pdf))
        if ok < 0:
            _check(ok)
        return ok == 1, np.array(d[:], np.float32), np.array(rgb[:], np.float32), pdf.value

    def ray_color(self, origin, direction, state, params=None):
        p = params if params is not None else self.params
        st = C.c_uint32(state)
        out = (C.c_float * 3)()
        _check(lib().or_ray_color(self._h, C.byref(p), f3(origin), f3(direction), C.byref(st), out))
        return np.array(out[:], np.float32), st.value

    def close(self):
        if self._h:
            lib().or_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def intersect_sphere(origin, direction, center, radius, tmin=0.001, tmax=float("inf")):
    h = Hit()
    ok = lib().or_intersect_sphere(f3(origin), f3(direction), f3(center), radius, tmin, tmax, C.byref(h))
    return h if ok else None


def intersect_triangle(origin, direction, verts, attributes=None, tmin=0.001, tmax=float("inf")):
    v = (C.c_float * 9)(*np.asarray(verts, np.float32).ravel().tolist())
    at = None
    if attributes is not None:
        at = np.ascontiguousarray(attributes, A.ATTR_DTYPE)
    h = Hit()
    ok = lib().or_intersect_triangle(f3(origin), f3(direction), v,
                                     at.ctypes.data if at is not None else None, tmin, tmax, C.byref(h))
    return h if ok else None


def aabb_intersects(mn, mx, origin, inverse_direction, tmin, tmax):
    return bool(lib().or_aabb_intersects(f3(mn), f3(mx), f3(origin), f3(inverse_direction), tmin, tmax))


def display_rgba8(pixels, exposure_stops=0.0):
    pixels = np.ascontiguousarray(pixels, np.float32)
    h, w = pixels.shape[0], pixels.shape[1]
    out = np.empty((h, w, 4), np.uint8)
    _check(lib().or_display_rgba8(pixels.ctypes.data, w, h, float(exposure_stops), out.ctypes.data))
    return out


def pcg_hash(x):
    return lib().or_pcg_hash(int(x) & 0xFFFFFFFF)


def stream(seed, n):
    """n draws of random_f32 from state `seed`: (states after each draw, values)."""
    st = C.c_uint32(int(seed) & 0xFFFFFFFF)
    states, values = np.empty(n, np.uint32), np.empty(n, np.float32)
    for k in range(n):
        values[k] = lib().or_rand01(C.byref(st))
        states[k] = st.value
    return states, values


def env_search_rows(cdf, width, height, sel):
    """Row-then-texel CDF search of the kernels (th::env_search_rows) on a (H*W,) CDF."""
    cdf = np.ascontiguousarray(cdf, np.float32)
    rows = np.ascontiguousarray(cdf.reshape(height, width)[:, -1])
    return lib().or_env_search_rows(cdf.ctypes.data, rows.ctypes.data, width, height, float(sel))


def reflectance(c, r):
    return lib().or_reflectance(c, r)


def refract(d, n, ratio):
    out = (C.c_float * 3)()
    lib().or_refract(f3(d), f3(n), ratio, out)
    return np.array(out[:], np.float32)


def power_heuristic(a, b):
    return lib().or_power_heuristic(a, b)


def sin(x):
    return lib().or_sin(x)


def cos(x):
    return lib().or_cos(x)


def atan2(y, x):
    return lib().or_atan2(y, x)


def acos(x):
    return lib().or_acos(x)
