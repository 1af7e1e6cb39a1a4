"""The scene loader subset on the reference's own benchmark suite (data copied from
scenes/benchmarks/ and assets/ of the reference into tests/golden/reference_benchmarks/):
object counts as scenes/benchmarks/README.md:7-15 lists them, packing, and an oracle
render of each at reduced size."""
import numpy as np
import pytest

import oracle
from toaster_amd import _abi as A, pack
from toaster_amd.loader import PUBLISHED_DISPATCH_WAIT_MS, load_gltf, load_hdr, load_scene, suite_paths

COUNTS = {  # scenes/benchmarks/README.md:9-15 (triangles, spheres incl. the light)
    "001_triangles_128": (128, 1), "002_triangles_2048": (2048, 1), "003_triangles_8192": (8192, 1),
    "004_spheres_64": (2, 65), "005_spheres_512": (2, 513), "006_mixed_2048t_128s": (2048, 129),
    "007_environment_control": (0, 4),
}


def test_suite_counts_and_packing():
    paths = suite_paths()
    assert len(paths) == 7
    for p in paths:
        sc = load_scene(p)
        assert (len(sc.triangles), len(sc.spheres)) == COUNTS[sc.name], sc.name
        assert (sc.width, sc.height, sc.samples, sc.max_bounces) == (320, 180, 4, 4)
        pk = pack(sc)
        assert pk.params.triangle_count == len(sc.triangles)
        if sc.name != "007_environment_control":
            assert sc.background == A.TH_BG_BLACK
            assert pk.params.light_count == 1  # the emissive sphere
        assert sc.name in PUBLISHED_DISPATCH_WAIT_MS


def test_gltf_terrain_grid():
    sc = load_scene(suite_paths()[0])
    v = sc.triangles["v"]
    assert v[..., 0].min() == -4.0 and v[..., 0].max() == 4.0  # accessor min/max of terrain_grid_8
    assert sc.attributes is not None and np.all(sc.attributes["has_normals"] == 1)
    n = sc.attributes["normals"]
    assert np.allclose(np.linalg.norm(n, axis=-1), 1.0, atol=1e-6)
    assert np.all(n[..., 1] > 0.5)  # a gentle height field: normals point up


def test_hdr_decoding():
    sc = load_scene(suite_paths()[-1])
    rgb = sc.environment.rgb
    assert rgb.shape == (2, 4, 3) and np.all(rgb >= 0) and rgb.max() > 0
    assert sc.environment.intensity == 8.0 and sc.environment.rotation_degrees == 20.0


@pytest.mark.parametrize("index", range(7))
def test_oracle_renders_suite_scene(index):
    sc = load_scene(suite_paths()[index])
    sc.width, sc.height = 64, 36
    img, _ = oracle.OracleScene(sc).render(threads=8)
    assert np.isfinite(img).all() and img[..., :3].mean() > 0
