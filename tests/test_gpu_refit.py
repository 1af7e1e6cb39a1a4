"""Device refit of moved geometry (th_scene_update with TH_CHANGE_REFIT): the uploaded
tree keeps its primitive order and leaf ranges and every box is recomputed on the GPU
from the new primitives.  The reference rebuilds the tree whenever geometry moves and
names refitting as its next optimisation (README.md:159-166, docs/architecture.md:
118-121), so the checks are: a refit of unmoved geometry changes nothing (the boxes are
the packer's, bit for bit, so every kernel's image is unchanged); after motion the
refitted scene renders what a rebuilt scene and the oracle render for the moved
geometry (image tolerance of tests/test_gpu_parity.py); the flags are validated."""
import dataclasses

import numpy as np
import pytest

import oracle
from toaster_amd import GpuScene, RenderOptions, _abi as A, scenes
from toaster_amd.scene import pack

from . import test_gpu_edge_cases as edge

pytestmark = pytest.mark.gpu

TOL_FRAC = 0.995
KERNELS = {
    "reference": RenderOptions(),
    "ordered": RenderOptions(traversal=A.TH_TRAVERSE_ORDERED),
    "quad": RenderOptions(traversal=A.TH_TRAVERSE_WIDE),
    "octet": RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=A.TH_SCHEDULE_OCTET),
    "speculative": RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=A.TH_SCHEDULE_SPECULATIVE),
    "wavefront": RenderOptions(traversal=A.TH_TRAVERSE_WIDE, kernel=A.TH_KERNEL_WAVEFRONT),
}
SCENES = {
    "stress2048": lambda: scenes.bvh_stress(2048, width=96, height=64, samples=2),
    "cornell": lambda: scenes.cornell(width=96, height=64, samples=2),
    "root_leaf": lambda: edge.root_leaf(),  # the whole tree one leaf of 70 primitives
    "big_leaf": lambda: edge.big_leaf(),    # a leaf of 70 under an interior root
}


def _moved(sc, step):
    """Rigid motion of every other triangle (rotation about the y axis through the
    scene's centre plus a shift) and a shift of every sphere, growing with `step`."""
    tris = sc.triangles.copy()
    v = tris["v"].astype(np.float64)
    c = v.reshape(-1, 3).mean(axis=0)
    a = 0.15 * step
    R = np.array([[np.cos(a), 0.0, np.sin(a)], [0.0, 1.0, 0.0], [-np.sin(a), 0.0, np.cos(a)]])
    sel = np.arange(len(tris)) % 2 == 0
    v[sel] = (v[sel] - c) @ R.T + c + np.array([0.05, 0.02, -0.04]) * step
    tris["v"] = v.astype(np.float32)
    sph = sc.spheres.copy()
    sph["center"] = sph["center"] + np.float32(0.03 * step) * np.array([1.0, -0.5, 0.5], np.float32)
    return dataclasses.replace(sc, triangles=tris, spheres=sph)


def _render(gs, opts):
    gs.render(options=opts, frame_index=0, accumulated_samples=0)
    return gs.readback().copy()


def _close(img, ref):
    close = np.abs(img - ref) <= 1e-4 * (1.0 + np.abs(ref))
    return float(np.mean(np.all(close, axis=-1)))


@pytest.mark.parametrize("name", list(SCENES))
def test_refit_of_unmoved_geometry_changes_nothing(gpu_ctx, name):
    sc = SCENES[name]()
    print(f"case refit identity {name}", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    before = {k: _render(gs, o) for k, o in KERNELS.items()}
    assert gs.refit_quality() == 1.0
    gs.refit(sc)
    assert gs.refit_quality() == 1.0  # the same boxes, summed in the same order
    for k, o in KERNELS.items():
        after = _render(gs, o)
        assert np.array_equal(before[k].view(np.uint32), after.view(np.uint32)), k
    gs.close()


@pytest.mark.parametrize("name", list(SCENES))
def test_refit_after_motion_matches_rebuild_and_oracle(gpu_ctx, name):
    sc = SCENES[name]()
    gs = GpuScene.from_scene(gpu_ctx, sc)
    still = _render(gs, KERNELS["quad"])
    for step in (1, 2, 3):  # successive frames refit the same tree further from its build
        moved = _moved(sc, step)
        print(f"case refit motion {name} step {step}", flush=True)
        gs.refit(moved)
        ref, _ = oracle.OracleScene(moved).render(threads=8)
        rebuilt = GpuScene.from_scene(gpu_ctx, moved)
        for k in ("reference", "quad", "speculative"):
            img = _render(gs, KERNELS[k])
            other = _render(rebuilt, KERNELS[k])
            f_ref, f_reb = _close(img, ref), _close(img, other)
            exact = float(np.mean(np.all(img.view(np.uint32) == other.view(np.uint32), axis=-1)))
            print(f"  {name} step {step} {k}: vs oracle {f_ref:.5f} vs rebuild {f_reb:.5f} bit-exact {exact:.5f}")
            assert f_ref >= TOL_FRAC and f_reb >= TOL_FRAC
            if k == "quad":
                assert not np.array_equal(img, still)  # the motion shows in the image
        rebuilt.close()
    gs.close()


def test_refit_flags_are_validated(gpu_ctx):
    sc = SCENES["cornell"]()
    gs = GpuScene.from_scene(gpu_ctx, sc)
    with pytest.raises(A.ToasterError):  # refit and a new tree together
        gs.update(pack(sc), A.TH_CHANGE_TRIANGLES | A.TH_CHANGE_BVH | A.TH_CHANGE_REFIT)
    with pytest.raises(A.ToasterError):  # refit without new geometry
        gs.update(pack(sc, bvh=False), A.TH_CHANGE_MATERIALS | A.TH_CHANGE_REFIT)
    with pytest.raises(A.ToasterError):  # a change bit the library does not define
        gs.update(pack(sc), A.TH_CHANGE_TRIANGLES | 128)
    img = _render(gs, KERNELS["quad"])  # the scene is unchanged by the refused updates
    ref, _ = oracle.OracleScene(sc).render(threads=8)
    assert _close(img, ref) >= TOL_FRAC
    gs.close()


def test_refit_quality_and_rebuild_policy(gpu_ctx):
    """Motion loosens a refitted tree (summed box area grows with it); refit(...,
    rebuild_above=r) rebuilds once the ratio passes r, after which the tree is as built
    and the image is a freshly uploaded scene's, bit for bit."""
    sc = SCENES["stress2048"]()
    print("case refit quality", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    ratios = []
    for step in (1, 2, 3):
        assert gs.refit(_moved(sc, step)) is False
        ratios.append(gs.refit_quality())
    print("  summed box area after 1, 2, 3 steps:", [round(r, 4) for r in ratios])
    assert 1.0 < ratios[0] < ratios[1] < ratios[2]
    moved = _moved(sc, 3)
    assert gs.refit(moved, rebuild_above=1.0) is True
    assert gs.refit_quality() == 1.0
    img = _render(gs, KERNELS["speculative"])
    fresh = GpuScene.from_scene(gpu_ctx, moved)
    ref = _render(fresh, KERNELS["speculative"])
    fresh.close()
    gs.close()
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
