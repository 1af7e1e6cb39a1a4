"""The oracle's brute-force linear scan (wgsl:667-698 = intersect.rs:266-289) against its
BVH traversal on whole images: the reference's own test renders BVH and linear closest
hits equal (intersect.rs:483-552); here the images of a tree-less scene and of the
same scene with its median-split tree agree on >= 99.5 % of pixels (exact ties between
primitives may resolve differently: the scan keeps the first in scene order)."""
import numpy as np
import pytest

import oracle
from toaster_amd import scenes

from .test_gpu_edge_cases import textured_floor


@pytest.mark.parametrize("make", [lambda: scenes.cornell(48, 27, samples=2),
                                  lambda: scenes.mixed_environment(40, 24, samples=2, max_bounces=4, n_triangles=96,
                                                                   n_spheres=12, env_size=(64, 32)),
                                  textured_floor], ids=["cornell", "mixed_env", "textured_floor"])
def test_linear_scan_image_equals_bvh_image(make):
    sc = make()
    with_tree, _ = oracle.OracleScene(sc).render(threads=8)
    no_tree, _ = oracle.OracleScene(sc, bvh=False).render(threads=8)
    close = np.abs(no_tree - with_tree) <= 1e-4 * (1.0 + np.abs(with_tree))
    assert float(np.mean(np.all(close, axis=-1))) >= 0.995
