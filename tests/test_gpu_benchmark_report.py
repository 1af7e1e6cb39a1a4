"""A schema-v1 report of a suite scene rendered on the GPU, compared with the published
RTX 2080 Ti report of the same scene."""
import hashlib
import json
import os

import pytest

from toaster_amd import GpuContext, GpuScene, RenderOptions, _abi as A, benchmark as B
from toaster_amd.loader import load_scene, suite_paths

pytestmark = pytest.mark.gpu


def test_suite_report_and_comparison(gpu_ctx):
    path = suite_paths()[2]  # 003_triangles_8192
    print("case benchmark report 003", flush=True)
    rep = B.benchmark_scene(path, warmup_frames=1, measured_frames=3)
    assert rep["schema_version"] == 1 and len(rep["frames"]) == 3
    for k in B.STAGES:
        s = rep["summary"][k]
        assert 0.0 <= s["min"] <= s["median"] <= s["max"]
    f = rep["frames"][0]
    parts = f["scene_update_upload_ms"] + f["dispatch_wait_ms"] + f["readback_ms"] + f["conversion_ms"]
    assert f["total_ms"] == pytest.approx(parts, rel=1e-6)
    assert "gfx950" in rep["gpu"]["name"]
    # the digest is that of the frame's RGBA8 image: render the same frame again
    sc = load_scene(path)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    gs.render(frame_index=0, accumulated_samples=0, options=RenderOptions(traversal=A.TH_TRAVERSE_WIDE))
    assert hashlib.sha256(gs.display_rgba8().tobytes()).hexdigest() == rep["image"]["rgba_sha256"]
    gs.close()
    ref_path = os.path.join(os.path.dirname(__file__), "golden", "reference_benchmarks", "reports",
                            "003_triangles_8192.json")
    with open(ref_path) as fh:
        published = json.load(fh)
    cmp = B.compare_reports(published, rep)
    assert not cmp["matching_gpu"]
    print("dispatch_wait change vs RTX 2080 Ti: %.1f %%" % cmp["deltas"]["dispatch_wait_percent"])
    assert cmp["deltas"]["dispatch_wait_percent"] < 0.0  # faster than the published run


def test_refit_update_gives_the_same_report_image():
    """update="refit": the per-frame scene update refits the uploaded tree instead of
    rebuilding it; for the static suite scene the boxes are the built ones, so the image
    digest equals the rebuild report's, and the update stage is cheaper."""
    path = suite_paths()[5]  # 006_mixed_2048t_128s: spheres and triangles
    print("case benchmark report refit 006", flush=True)
    rebuild = B.benchmark_scene(path, warmup_frames=1, measured_frames=3)
    refit = B.benchmark_scene(path, warmup_frames=1, measured_frames=3, update="refit")
    assert refit["image"]["rgba_sha256"] == rebuild["image"]["rgba_sha256"]
    med = lambda r: r["summary"]["scene_update_upload_ms"]["median"]
    print("scene update ms: rebuild %.3f, refit %.3f" % (med(rebuild), med(refit)))
    with pytest.raises(ValueError):
        B.benchmark_scene(path, update="rotate")
