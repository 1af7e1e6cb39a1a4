"""The reference's benchmark report, schema version 1 (crates/toaster-cli/src/benchmark.rs:
13-470), for a scene file rendered on an MI355X.

`benchmark_scene` measures what the reference's `benchmark_gpu_evaluator` measures per
frame (crates/toaster-gpu/src/path_tracer.rs:463-515), frame seed 0, no accumulation:

- scene_update_upload_ms: the scene packed again (emitter list, median-split BVH rebuilt
  and flattened, records) and every array written to the device (th_scene_update with
  all change flags) -- the reference rebuilds and uploads these every frame, static
  scenes included; with update="refit" the scene is packed without a tree and the
  uploaded tree is refitted on the device instead (TH_CHANGE_REFIT; for a static scene
  the refitted boxes equal the built ones, so the image is the same);
- dispatch_wait_ms: th_render and the wait for it;
- readback_ms: th_readback of the linear-HDR float4 image;
- conversion_ms: exposure / ACES / sRGB to RGBA8, here on the GPU
  (th_display_rgba8_device, including the copy of the RGBA bytes; the reference converts
  on the CPU);
- total_ms: their sum.

The report's image digest is the SHA-256 of the last frame's RGBA bytes.  An image
rendered here matches the reference estimator restated in fp32 with correctly rounded
arithmetic; a digest from the reference's own GPU run (Vulkan, its shader compiler's
division and transcendental precision) is not expected to match byte for byte.
`compare_reports` applies the reference's comparison (benchmark.rs:314-378).
"""
import ctypes as C
import hashlib
import json
import math
import os
import time

import numpy as np

from . import _abi as A
from .loader import load_scene
from .renderer import GpuContext, GpuScene, RenderOptions
from .scene import pack

SCHEMA_VERSION = 1
TOASTER_VERSION = "0.1.0"  # the reference version whose report format this follows
ALL_CHANGES = (A.TH_CHANGE_CAMERA | A.TH_CHANGE_SPHERES | A.TH_CHANGE_TRIANGLES | A.TH_CHANGE_LIGHTS
               | A.TH_CHANGE_BVH | A.TH_CHANGE_MATERIALS)
STAGES = ("scene_update_upload_ms", "dispatch_wait_ms", "readback_ms", "conversion_ms", "total_ms")


def statistics(values):
    """min / mean / midpoint median / nearest-rank p95 / max (benchmark.rs:424-452)."""
    v = sorted(float(x) for x in values)
    if not v:
        raise ValueError("cannot summarize empty benchmark data")
    if not all(math.isfinite(x) and x >= 0.0 for x in v):
        raise ValueError("benchmark timings must be finite and nonnegative")
    n = len(v)
    median = (v[n // 2 - 1] + v[n // 2]) / 2.0 if n % 2 == 0 else v[n // 2]
    p95 = v[min(max(math.ceil(n * 0.95) - 1, 0), n - 1)]
    return {"min": v[0], "mean": sum(v) / n, "median": median, "p95": p95, "max": v[-1]}


def rgba_sha256(data: bytes) -> str:
    """Lowercase SHA-256 digest of packed RGBA bytes (benchmark.rs:459-462)."""
    return hashlib.sha256(bytes(data)).hexdigest()


def device_info(device=0):
    name = C.create_string_buffer(256)
    driver = C.create_string_buffer(256)
    A.check(A.lib().th_device_info(int(device), name, 256, driver, 256))
    return {"name": name.value.decode(), "backend": "HIP", "device_type": "DiscreteGpu",
            "driver": "AMD ROCm", "driver_info": driver.value.decode()}


def _emissive_objects(scene):
    """Emissive spheres and triangles with positive strength (benchmark.rs:465-478)."""
    mats = scene.materials
    emissive = {i for i in range(len(mats))
                if int(mats["kind"][i]) == A.TH_MAT_EMISSIVE and float(mats["strength"][i]) > 0.0}
    n = sum(1 for m in scene.spheres["material_index"] if int(m) in emissive)
    return n + sum(1 for m in scene.triangles["material_index"] if int(m) in emissive)


REFIT_CHANGES = (A.TH_CHANGE_CAMERA | A.TH_CHANGE_SPHERES | A.TH_CHANGE_TRIANGLES | A.TH_CHANGE_LIGHTS
                 | A.TH_CHANGE_MATERIALS | A.TH_CHANGE_REFIT)


def benchmark_scene(path, warmup_frames=2, measured_frames=5, options: RenderOptions = None, device=0,
                    exposure_stops=0.0, update="rebuild"):
    """Schema-v1 report of `path` (a JSON scene file) rendered `warmup_frames` +
    `measured_frames` times; `update` is "rebuild" (the reference's per-frame work) or
    "refit"."""
    if measured_frames < 1:
        raise ValueError("benchmark needs at least one measured frame")
    if update not in ("rebuild", "refit"):
        raise ValueError("update must be 'rebuild' or 'refit'")
    opts = options or RenderOptions(traversal=A.TH_TRAVERSE_WIDE)
    t_setup = time.perf_counter()
    scene = load_scene(path)
    ctx = GpuContext(device)
    gs = GpuScene.from_scene(ctx, scene)
    ctx.synchronize()
    setup_ms = (time.perf_counter() - t_setup) * 1e3
    frames, rgba = [], None
    try:
        for i in range(warmup_frames + measured_frames):
            t0 = time.perf_counter()
            if update == "refit":
                gs.update(pack(scene, bvh=False), REFIT_CHANGES)
            else:
                gs.update(pack(scene), ALL_CHANGES)
            ctx.synchronize()
            t1 = time.perf_counter()
            gs.render(frame_index=0, accumulated_samples=0, options=opts)
            ctx.synchronize()
            t2 = time.perf_counter()
            gs.readback()
            t3 = time.perf_counter()
            rgba = gs.display_rgba8(exposure_stops)
            t4 = time.perf_counter()
            if i >= warmup_frames:
                f = {"scene_update_upload_ms": (t1 - t0) * 1e3, "dispatch_wait_ms": (t2 - t1) * 1e3,
                     "readback_ms": (t3 - t2) * 1e3, "conversion_ms": (t4 - t3) * 1e3}
                f["total_ms"] = (t4 - t0) * 1e3
                frames.append(f)
        gpu = device_info(device)
    finally:
        gs.close()
        ctx.close()
    return {
        "schema_version": SCHEMA_VERSION,
        "created_at_unix_seconds": int(time.time()),
        "toaster_version": TOASTER_VERSION,
        "scene": {"path": str(path), "width": int(scene.width), "height": int(scene.height),
                  "samples": int(scene.samples), "max_bounces": int(scene.max_bounces),
                  "spheres": int(len(scene.spheres)), "triangles": int(len(scene.triangles)),
                  "materials": int(len(scene.materials)), "lights": _emissive_objects(scene)},
        "gpu": gpu,
        "configuration": {"warmup_frames": int(warmup_frames), "measured_frames": int(measured_frames),
                          "fixed_animation_time_seconds": 0.0, "fixed_frame_seed": 0},
        "setup_ms": setup_ms,
        "frames": frames,
        "summary": {k: statistics(f[k] for f in frames) for k in STAGES},
        "image": {"width": int(rgba.shape[1]), "height": int(rgba.shape[0]),
                  "rgba_sha256": rgba_sha256(np.ascontiguousarray(rgba).tobytes()), "reference_path": None},
    }


def write_report(report, path):
    """Pretty-printed JSON with a trailing newline (benchmark.rs:296-303)."""
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    with open(path, "w") as f:
        f.write(json.dumps(report, indent=2) + "\n")


def read_report(path):
    with open(path) as f:
        return json.load(f)


def _percent_change(baseline, candidate):
    return (candidate / baseline - 1.0) * 100.0 if baseline > 0.0 else None


def compare_reports(baseline, candidate, max_regression_percent=None):
    """Median changes from `baseline` to `candidate` and the compatibility flags
    (benchmark.rs:314-378, 484-508)."""
    if max_regression_percent is not None and not (math.isfinite(max_regression_percent)
                                                   and max_regression_percent >= 0.0):
        raise ValueError("maximum regression percentage must be finite and nonnegative")
    if baseline["schema_version"] != SCHEMA_VERSION or candidate["schema_version"] != SCHEMA_VERSION:
        raise ValueError(f"benchmark report schema mismatch; expected version {SCHEMA_VERSION}")
    keys = ("width", "height", "samples", "max_bounces", "spheres", "triangles", "materials", "lights")
    if any(baseline["scene"][k] != candidate["scene"][k] for k in keys):
        raise ValueError("benchmark reports use incompatible resolved scene settings or geometry counts")
    ident = lambda g: (g["name"], g["backend"], g["device_type"])
    matching_gpu = ident(baseline["gpu"]) == ident(candidate["gpu"])
    if max_regression_percent is not None and not matching_gpu:
        raise ValueError("a regression threshold requires matching GPU name, backend, and device type")
    if max_regression_percent is not None and not baseline["summary"]["total_ms"]["median"] > 0.0:
        raise ValueError("a regression threshold requires a positive baseline median total time")
    med = lambda r, k: r["summary"][k]["median"]
    deltas = {"setup_percent": _percent_change(baseline["setup_ms"], candidate["setup_ms"])}
    for k in STAGES:
        deltas[k.replace("_ms", "_percent")] = _percent_change(med(baseline, k), med(candidate, k))
    return {
        "deltas": deltas,
        "matching_gpu": matching_gpu,
        "matching_driver": (baseline["gpu"]["driver"] == candidate["gpu"]["driver"]
                            and baseline["gpu"]["driver_info"] == candidate["gpu"]["driver_info"]),
        "matching_image_checksum": baseline["image"]["rgba_sha256"] == candidate["image"]["rgba_sha256"],
        "exceeds_regression_limit": (max_regression_percent is not None and deltas["total_percent"] is not None
                                     and deltas["total_percent"] > max_regression_percent),
    }
