"""Edge cases of the path on the GPU against the oracle: one sphere, one triangle, 1x1
and one-row images, leaves beyond the 62 primitives a packed entry counts (coincident
centroids; the root itself a leaf), exact ties between primitives, a textured material
(bilinear sRGB texels, texture.rs:53-92 / wgsl:233-270).  Each runs through the scalar
megakernel (reference and ordered traversal), the 4-lane, octet and speculative
megakernels and the wavefront.  Tolerance as tests/test_gpu_parity.py."""
import numpy as np
import pytest

import oracle
from toaster_amd import GpuScene, RenderOptions, _abi as A
from toaster_amd.scene import (ATTR_IN_DTYPE, MATERIAL_IN_DTYPE, SPHERE_IN_DTYPE, TRIANGLE_IN_DTYPE, Scene,
                               diffuse, emissive, metal, textured_diffuse)

pytestmark = pytest.mark.gpu

TOL_FRAC = 0.995
KERNELS = {
    "reference": RenderOptions(),
    "scalar": RenderOptions(traversal=A.TH_TRAVERSE_ORDERED),
    "quad": RenderOptions(traversal=A.TH_TRAVERSE_WIDE),
    "octet": RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=A.TH_SCHEDULE_OCTET),
    "speculative": RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=A.TH_SCHEDULE_SPECULATIVE),
    "wavefront": RenderOptions(traversal=A.TH_TRAVERSE_WIDE, kernel=A.TH_KERNEL_WAVEFRONT),
}


def _mats(rows):
    m = np.zeros(len(rows), MATERIAL_IN_DTYPE)
    for i, r in enumerate(rows):
        m[i] = r
    return m


def _spheres(rows):
    s = np.zeros(len(rows), SPHERE_IN_DTYPE)
    for i, r in enumerate(rows):
        s[i] = r
    return s


def _tris(verts, mats):
    t = np.zeros(len(mats), TRIANGLE_IN_DTYPE)
    t["v"] = np.asarray(verts, np.float32).reshape(-1, 3, 3)
    t["material_index"] = mats
    return t


def one_sphere():
    return Scene(camera_position=(0, 0, 3), width=64, height=36, samples=4, background=A.TH_BG_SKY,
                 materials=_mats([diffuse((0.7, 0.5, 0.3))]), spheres=_spheres([((0, 0, 0), 1.0, 0)]),
                 name="one_sphere")


def one_triangle():
    return Scene(camera_position=(0, 0, 3), width=64, height=36, samples=4, background=A.TH_BG_SKY,
                 materials=_mats([metal((0.9, 0.8, 0.7), 0.2)]),
                 triangles=_tris([[(-1, -1, 0), (1, -1, 0), (0, 1, 0)]], [0]), name="one_triangle")


def tiny(width, height):
    def make():
        sc = one_sphere()
        sc.width, sc.height, sc.samples, sc.name = width, height, 8, f"tiny{width}x{height}"
        return sc
    return make


def _coincident_tris(n, seed):
    """n triangles whose bounding boxes share one centre (the origin): the builder's
    centroids coincide, so they end in one leaf (bvh.rs:119-121).  The first three are
    exact copies with different materials: equal hit distances, the lowest index wins."""
    rng = np.random.default_rng(seed)
    verts = []
    for i in range(n):
        sx, sy, sz = rng.uniform(0.2, 0.8, 3)
        if i < 3:
            sx, sy, sz = 0.7, 0.6, 0.5
            v = [(-sx, -sy, 0.0), (sx, 0.0, -sz), (0.0, sy, sz)]
        else:
            v = [(-sx, -sy, rng.uniform(-sz, sz)), (sx, rng.uniform(-sy, sy), -sz),
                 (rng.uniform(-sx, sx), sy, sz)]
        verts.append(v)
    return _tris(verts, [i % 3 for i in range(n)])


def root_leaf():
    """70 coincident triangles and nothing else: the root is a leaf of 70 (beyond the
    62 a packed entry counts)."""
    return Scene(camera_position=(0, 0, 3), width=96, height=64, samples=4, background=A.TH_BG_SKY,
                 materials=_mats([diffuse((0.8, 0.2, 0.2)), diffuse((0.2, 0.8, 0.2)), metal((0.5, 0.5, 0.9), 0.3)]),
                 triangles=_coincident_tris(70, 5), name="root_leaf")


def big_leaf():
    """140 coincident triangles and an emissive sphere: an interior root whose left
    subtree is one leaf of 70."""
    return Scene(camera_position=(0, 0, 3), width=96, height=64, samples=4, background=A.TH_BG_BLACK,
                 materials=_mats([diffuse((0.8, 0.2, 0.2)), diffuse((0.2, 0.8, 0.2)), metal((0.5, 0.5, 0.9), 0.3),
                                  emissive((1, 1, 1), 8.0)]),
                 spheres=_spheres([((0, 1.6, 0.5), 0.4, 3)]), triangles=_coincident_tris(140, 6), name="big_leaf")


def textured_floor():
    """A floor quad with texture coordinates over a 4x4 checker (repeated twice, sRGB)
    lit by an emissive sphere and the sky."""
    tex = np.zeros((4, 4, 4), np.uint8)
    for y in range(4):
        for x in range(4):
            tex[y, x] = (230, 40, 40, 255) if (x + y) % 2 else (30, 60, 220, 255)
    verts = [[(-2, 0, -2), (2, 0, -2), (2, 0, 2)], [(-2, 0, -2), (2, 0, 2), (-2, 0, 2)]]
    att = np.zeros(2, ATTR_IN_DTYPE)
    att["has_tex_coords"] = 1
    att["tex_coords"][0] = [(0, 0), (2, 0), (2, 2)]
    att["tex_coords"][1] = [(0, 0), (2, 2), (0, 2)]
    return Scene(camera_position=(0, 2, 3), camera_look_at=(0, 0, 0), width=96, height=64, samples=4,
                 background=A.TH_BG_SKY,
                 materials=_mats([textured_diffuse((0.9, 0.9, 0.9), 0), emissive((1, 1, 1), 5.0)]),
                 spheres=_spheres([((0.5, 1.0, 0.0), 0.3, 1)]), triangles=_tris(verts, [0, 0]),
                 attributes=att, textures=[tex], name="textured_floor")


CASES = [("one_sphere", one_sphere), ("one_triangle", one_triangle),
         ("tiny1x1", tiny(1, 1)), ("tiny5x1", tiny(5, 1)), ("root_leaf", root_leaf), ("big_leaf", big_leaf),
         ("textured_floor", textured_floor)]


@pytest.mark.parametrize("kernel", list(KERNELS))
@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_edge_case_matches_oracle(gpu_ctx, name, make, kernel):
    sc = make()
    print(f"case edge {name} kernel={kernel}", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    gs.render(options=KERNELS[kernel], frame_index=0, accumulated_samples=0)
    img = gs.readback()
    gs.close()
    ref, _ = oracle.OracleScene(sc).render(threads=8)
    assert np.isfinite(img).all()
    close = np.abs(img - ref) <= 1e-4 * (1.0 + np.abs(ref))
    frac = float(np.mean(np.all(close, axis=-1)))
    print(f"  {name} {kernel}: within tol {frac:.5f}")
    assert frac >= TOL_FRAC
