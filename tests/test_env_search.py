"""The row-then-texel environment CDF search of the kernels (th::env_search_rows, shared
host/device code) returns the index of the reference's flat search (wgsl:352-362: the
first texel whose CDF value exceeds the selection, i.e. the partition point of <=, else
the last texel), checked against numpy.searchsorted(side="right") on CDFs built as the
packer builds them (environment.rs:245-278: f64 running sum, cast to f32, last = 1):
zero-probability texels and rows (at the start, middle and end), selections equal to
CDF values and their neighbours, 0 and the 0.99999994 clamp, and the sun map of the
mixed-geometry workload."""
import numpy as np
import pytest

import oracle
from toaster_amd import scenes
from toaster_amd.scene import pack


def _cdf(weights):
    w = np.asarray(weights, np.float64).ravel()
    c = np.cumsum(w) / w.sum()
    c = c.astype(np.float32)
    c[-1] = 1.0
    return c


def _flat(cdf, sel):
    return min(int(np.searchsorted(cdf, np.float32(sel), side="right")), len(cdf) - 1)


def _selections(cdf, rng, n=400):
    vals = np.unique(cdf)
    pick = vals[rng.integers(0, len(vals), min(n, len(vals)))]
    sels = np.concatenate([pick, np.nextafter(pick, np.float32(0)), np.nextafter(pick, np.float32(2)),
                           rng.random(n).astype(np.float32), [0.0, 0.99999994, 1.0]]).astype(np.float32)
    return sels[(sels >= 0) & (sels <= 0.99999994)]


def _weights(case, rng):
    h, w = {"zero_rows": (16, 24), "sparse": (9, 7), "single_row": (1, 33), "single_col": (40, 1),
            "one_texel": (1, 1), "dense": (32, 64)}[case]
    x = rng.random((h, w))
    if case == "zero_rows":
        x[:3] = 0.0
        x[7:9] = 0.0
        x[-2:] = 0.0
        x[:, 5:9] = 0.0
    if case == "sparse":
        x[rng.random((h, w)) < 0.8] = 0.0
        x[4, 3] = 1.0
    return x, w, h


@pytest.mark.parametrize("case", ["zero_rows", "sparse", "single_row", "single_col", "one_texel", "dense"])
def test_row_search_equals_flat_search(case):
    rng = np.random.default_rng(sum(map(ord, case)))
    x, w, h = _weights(case, rng)
    cdf = _cdf(x)
    for sel in _selections(cdf, rng):
        assert oracle.env_search_rows(cdf, w, h, sel) == _flat(cdf, sel), (case, float(sel))


def test_row_search_on_the_sun_map():
    sc = scenes.mixed_environment(16, 9, samples=1, max_bounces=1, n_triangles=8, n_spheres=2)
    cdf = np.ascontiguousarray(pack(sc).env[:, 3])
    w, h = 1024, 512
    rng = np.random.default_rng(5)
    sun = np.flatnonzero(sc.environment.rgb[..., 0].ravel() >= 5.0e4)
    around = np.concatenate([cdf[max(i - 2, 0):i + 2] for i in sun])
    sels = np.concatenate([_selections(cdf, rng, 3000), around, np.nextafter(around, np.float32(0))])
    sels = sels[(sels >= 0) & (sels <= 0.99999994)]
    for sel in sels:
        assert oracle.env_search_rows(cdf, w, h, sel) == _flat(cdf, sel)


def test_the_check_catches_a_planted_boundary_error():
    """A restatement of the search with < in place of <= (a planted mistake) differs from
    the flat search exactly on selections equal to a CDF value, which the selections
    above include."""
    rng = np.random.default_rng(1)
    x, w, h = _weights("zero_rows", rng)
    cdf = _cdf(x)

    def planted(sel):
        return min(int(np.searchsorted(cdf, np.float32(sel), side="left")), len(cdf) - 1)

    assert any(planted(s) != _flat(cdf, s) for s in _selections(cdf, rng))
