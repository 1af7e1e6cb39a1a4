"""The schema-v1 benchmark report (toaster_amd/benchmark.py) against the reference's own
tests of its report (crates/toaster-cli/src/benchmark.rs:596-640) and the published
RTX 2080 Ti reports of the benchmark suite (docs/benchmarks/post-bvh/rtx-2080-ti, copied
as data into tests/golden/reference_benchmarks/reports)."""
import copy
import json
import os

import pytest

from toaster_amd import benchmark as B
from toaster_amd.loader import load_scene, suite_paths

REPORTS = os.path.join(os.path.dirname(__file__), "golden", "reference_benchmarks", "reports")


def _published(name):
    with open(os.path.join(REPORTS, name + ".json")) as f:
        return json.load(f)


def test_statistics_use_conventional_median_and_nearest_rank_p95():  # benchmark.rs:596-606
    s = B.statistics([1.0, 2.0, 3.0, 4.0])
    assert s == {"min": 1.0, "mean": 2.5, "median": 2.5, "p95": 4.0, "max": 4.0}
    with pytest.raises(ValueError):
        B.statistics([])
    # odd count: the middle value; nearest rank: ceil(0.95 n)-th smallest
    assert B.statistics([5.0, 1.0, 3.0])["median"] == 3.0
    assert B.statistics(range(1, 21))["p95"] == 19.0
    with pytest.raises(ValueError):
        B.statistics([1.0, float("nan")])


def test_sha256_is_lowercase_and_stable():  # benchmark.rs:608-613
    assert B.rgba_sha256(b"toaster") == "b84b0faf393e0c0ca2f76a6095fad0d7e7a0a127b5e9cd5c43b0727d7fef6c19"


def test_published_summaries_restate_their_frames():
    """The reference's summaries are statistics() of its own frames."""
    for path in suite_paths():
        rep = _published(os.path.splitext(os.path.basename(path))[0])
        for k in B.STAGES:
            got = B.statistics(f[k] for f in rep["frames"])
            for stat, v in got.items():
                assert v == pytest.approx(rep["summary"][k][stat], rel=1e-9, abs=1e-12), (path, k, stat)


@pytest.mark.parametrize("index", range(7))
def test_scene_report_matches_published(index):
    """Resolution, samples, bounces and geometry counts as the reference reports them,
    so compare_reports accepts a published report as the baseline."""
    path = suite_paths()[index]
    sc = load_scene(path)
    rep = _published(os.path.splitext(os.path.basename(path))[0])["scene"]
    mine = {"width": sc.width, "height": sc.height, "samples": sc.samples, "max_bounces": sc.max_bounces,
            "spheres": len(sc.spheres), "triangles": len(sc.triangles), "materials": len(sc.materials),
            "lights": B._emissive_objects(sc)}
    assert mine == {k: rep[k] for k in mine}


def test_comparison_detects_regression_and_checksum_change():  # benchmark.rs:626-640
    base = _published("003_triangles_8192")
    cand = copy.deepcopy(base)
    cand["summary"]["total_ms"]["median"] = base["summary"]["total_ms"]["median"] * 1.2
    cand["image"]["rgba_sha256"] = "0" * 64
    c = B.compare_reports(base, cand, max_regression_percent=10.0)
    assert abs(c["deltas"]["total_percent"] - 20.0) < 1e-9
    assert c["exceeds_regression_limit"] and not c["matching_image_checksum"] and c["matching_gpu"]
    # another GPU: comparable, but no regression threshold
    other = copy.deepcopy(base)
    other["gpu"]["name"] = "AMD Instinct MI355X (gfx950)"
    assert not B.compare_reports(base, other)["matching_gpu"]
    with pytest.raises(ValueError):
        B.compare_reports(base, other, max_regression_percent=10.0)
    # another scene is refused
    bad = copy.deepcopy(base)
    bad["scene"]["triangles"] += 1
    with pytest.raises(ValueError):
        B.compare_reports(base, bad)


def test_report_round_trips_through_json(tmp_path):  # benchmark.rs:615-622
    rep = _published("001_triangles_128")
    p = tmp_path / "r" / "report.json"
    B.write_report(rep, str(p))
    assert p.read_text().endswith("\n")
    assert B.read_report(str(p)) == rep
