"""The oracle pinned against the reference's own known-answer tests (restated) and
closed forms.  Each test names the reference test it restates.
"""
import math

import numpy as np
import pytest

import oracle
from toaster_amd import _abi as A
from toaster_amd.scene import dielectric, diffuse, emissive, metal

from .helpers import env_map, scene

SQ = math.sqrt


def vec_close(a, b, tol):
    return np.allclose(np.asarray(a, np.float64), np.asarray(b, np.float64), atol=tol, rtol=0)


# ------------------------------------------------------------------ intersection
def test_misses_sphere():  # intersect.rs:355-359
    assert oracle.intersect_sphere((0, 0, 0), (0, 1, 0), (0, 0, -1), 0.5) is None


def test_hits_nearest_surface():  # intersect.rs:361-375
    h = oracle.intersect_sphere((0, 0, 0), (0, 0, -1), (0, 0, -2), 0.5)
    assert abs(h.distance - 1.5) < 1e-6
    assert list(h.normal) == [0.0, 0.0, 1.0]
    assert h.front_face == 1


def test_inside_ray_uses_far_root():  # intersect.rs:377-385
    h = oracle.intersect_sphere((0, 0, 0), (1, 0, 0), (0, 0, 0), 1.0)
    assert abs(h.distance - 1.0) < 1e-6
    assert list(h.normal) == [-1.0, 0.0, 0.0]
    assert h.front_face == 0


def test_scene_returns_nearest_hit():  # intersect.rs:387-431
    sc = scene([diffuse((1, 1, 1)), diffuse((1, 1, 1))],
               sph=[((0, 0, -4), 1.0, 0), ((0, 0, -2), 0.5, 1)])
    oc = oracle.OracleScene(sc)
    h, _ = oc.closest((0, 0, 0), (0, 0, -1))
    assert h.material_index == 1


TRI = [(-1, -1, -2), (1, -1, -2), (0, 1, -2)]


def test_triangle_double_sided():  # intersect.rs:433-464
    f = oracle.intersect_triangle((0, 0, 0), (0, 0, -1), TRI)
    assert abs(f.distance - 2.0) < 1e-6 and f.front_face == 1
    b = oracle.intersect_triangle((0, 0, -3), (0, 0, 1), TRI)
    assert b.front_face == 0
    assert [-x for x in f.normal] == list(b.normal)


def test_triangle_misses_outside_edges():  # intersect.rs:466-480
    assert oracle.intersect_triangle((2, 2, 1), (0, 0, -1), [(0, 0, 0), (1, 0, 0), (0, 1, 0)]) is None


def _norm(v):
    v = np.asarray(v, np.float64)
    return tuple((v / np.linalg.norm(v)).astype(np.float32))


def test_bvh_matches_linear_mixed_scene():  # intersect.rs:482-552
    sc = scene([diffuse((1, 1, 1))] * 3, sph=[((-1.5, 0, -4), 0.75, 0), ((1.25, 0, -3), 0.5, 1)],
               tris=[(TRI, 2)])
    oc = oracle.OracleScene(sc)
    rays = [((0, 0, 0), (0, 0, -1)), ((0, 0, 0), _norm((-0.4, 0, -1))), ((0, 0, 0), _norm((0.45, 0, -1))),
            ((0, 0, -3), (0, 1, 0)), ((5, 5, 0), (0, 0, -1))]
    hits = 0
    for o, d in rays:
        lin, lh = oc.closest(o, d, use_bvh=False)
        acc, ah = oc.closest(o, d, use_bvh=True)
        assert (lin is None) == (acc is None)
        if lin is not None:
            hits += 1
            assert abs(lh.distance - ah.distance) < 1e-5
            assert lh.material_index == ah.material_index and lh.front_face == ah.front_face
            assert vec_close(lh.normal, ah.normal, 1e-5)
    assert hits >= 3


def test_bvh_matches_linear_random_rays():  # closed form: BVH == brute force (A9)
    from toaster_amd import scenes
    sc = scenes.bvh_stress(512)
    oc = oracle.OracleScene(sc)
    rng = np.random.default_rng(5)
    for _ in range(300):
        o = rng.uniform(-4, 4, 3)
        d = rng.normal(size=3)
        d = _norm(d)
        lin, lh = oc.closest(o, d, use_bvh=False)
        acc, ah = oc.closest(o, d, use_bvh=True)
        assert (lin is None) == (acc is None)
        if lin is not None:
            assert lh.distance == ah.distance


def test_smooth_normal_and_uv():  # intersect.rs:608-639
    at = np.zeros(1, A.ATTR_DTYPE)
    at["n0"][0, :3] = at["n1"][0, :3] = at["n2"][0, :3] = (0, 0, 1)
    at["uv0"][0] = (0, 0)
    at["uv1"][0] = (1, 0)
    at["uv2"][0] = (0.5, 1)
    at["flags"] = 3
    h = oracle.intersect_triangle((0, 0, 0), (0, 0, -1), TRI, attributes=at)
    assert vec_close(h.normal, (0, 0, 1), 1e-6)
    assert vec_close(h.tex_coord, (0.5, 0.5), 1e-6)


def test_analytic_sphere_and_plane_hits():  # closed forms
    rng = np.random.default_rng(1)
    for _ in range(200):
        c = rng.uniform(-3, 3, 3)
        r = rng.uniform(0.1, 2)
        o = c + rng.normal(size=3) * 4
        d = np.asarray(_norm(c - o + rng.normal(size=3) * 0.1 * r), np.float64)
        h = oracle.intersect_sphere(o, d, c, r)
        oc_ = o - c
        b = oc_ @ d
        disc = b * b - (oc_ @ oc_ - r * r)
        if disc < 0 or (-b - SQ(disc)) < 0.001:
            continue
        assert h is not None
        assert abs(h.distance - (-b - SQ(disc))) < 2e-5 * max(1, abs(b))
    for _ in range(200):  # plane z = k through a big triangle
        k = rng.uniform(-3, -1)
        tri = [(-50, -50, k), (50, -50, k), (0, 50, k)]
        d = np.asarray(_norm((rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), -1)), np.float64)
        h = oracle.intersect_triangle((0, 0, 0), d, tri)
        assert abs(h.distance - k / d[2]) < 1e-5 * abs(k / d[2])


# ------------------------------------------------------------------ AABB
def test_aabb_from_outside_and_inside():  # aabb.rs:127-143
    inf = float("inf")
    assert oracle.aabb_intersects((-1,) * 3, (1,) * 3, (0, 0, 3), (inf, inf, -1), 0, inf)
    assert oracle.aabb_intersects((-1,) * 3, (1,) * 3, (0, 0, 0), (1, inf, inf), 0, inf)


def test_aabb_rejects_parallel_outside():  # aabb.rs:145-155
    inf = float("inf")
    assert not oracle.aabb_intersects((-1,) * 3, (1,) * 3, (2, 0, 0), (inf, inf, -1), 0, inf)


# ------------------------------------------------------------------ BVH structure
def _leaves(nodes):
    return [n for n in nodes if n["primitive_count"] > 0]


def test_bvh_every_primitive_once_and_leaf_size():  # bvh.rs:335-350 + invariants
    from toaster_amd import scenes
    sc = scenes.mixed_environment(width=8, height=8, samples=1, env_size=(8, 4))
    p = oracle.OracleScene(sc).packed
    n = len(sc.spheres) + len(sc.triangles)
    assert len(p.prims) == n
    keys = sorted((int(k), int(i)) for k, i in zip(p.prims["kind"], p.prims["index"]))
    assert keys == sorted([(0, i) for i in range(len(sc.spheres))] + [(1, i) for i in range(len(sc.triangles))])
    assert len(p.nodes) <= 2 * n - 1
    covered = np.zeros(n, int)
    for leaf in _leaves(p.nodes):
        assert leaf["primitive_count"] <= 4
        covered[leaf["first_or_right"]:leaf["first_or_right"] + leaf["primitive_count"]] += 1
    assert (covered == 1).all()
    for i, nd in enumerate(p.nodes):  # left child follows its parent, right child after it
        if nd["primitive_count"] == 0:
            assert nd["first_or_right"] > i + 1


def test_bvh_coincident_centroids_one_leaf():  # bvh.rs:352-364
    sc = scene([diffuse((1, 1, 1))], sph=[((0, 0, -2), 0.5, 0)] * 6)
    p = oracle.OracleScene(sc).packed
    assert len(p.nodes) == 1 and p.nodes[0]["primitive_count"] == 6


def test_flatten_right_child_index():  # flatten.rs:470-497
    sc = scene([diffuse((1, 1, 1))], sph=[((x, 0, -2), 0.1, 0) for x in (0, 1, 2, 3, 4)])
    p = oracle.OracleScene(sc).packed
    assert len(p.nodes) == 3
    assert p.nodes[0]["primitive_count"] == 0 and p.nodes[0]["first_or_right"] == 2
    assert p.nodes[1]["primitive_count"] == 2 and p.nodes[2]["primitive_count"] == 3


# ------------------------------------------------------------------ lights
LIGHT_TRIS = [([(0, 0, 0), (2, 0, 0), (0, 1, 0)], 0), ([(0, 0, 1), (1, 0, 1), (0, 1, 1)], 1)]


def test_only_emissive_triangles_collected():  # light.rs:227-247
    sc = scene([emissive((1, 0.5, 0.25), 4.0), diffuse((1, 1, 1))], tris=LIGHT_TRIS, background=A.TH_BG_BLACK)
    p = oracle.OracleScene(sc).packed
    assert len(p.lights) == 1 and p.params.light_count == 1
    awc = p.lights[0]["area_weight_cumulative"]
    assert abs(awc[0] - 1.0) < 1e-6
    assert abs(awc[1] / p.params.total_light_weight - 1.0) < 1e-6


def test_power_weighted_light_cdf():  # light.rs:249-292, scene_upload.rs:481-509
    tris = [([(0, 0, 0), (4, 0, 0), (0, 4, 0)], 0), LIGHT_TRIS[1]]
    sc = scene([emissive((1, 1, 1), 0.1), diffuse((1, 1, 1)), emissive((1, 1, 1), 20.0)], tris=tris,
               sph=[((0, 2, 0), 0.2, 2)], background=A.TH_BG_BLACK)
    p = oracle.OracleScene(sc).packed
    assert len(p.lights) == 2
    dim, bright = p.lights[0]["area_weight_cumulative"], p.lights[1]["area_weight_cumulative"]
    W = p.params.total_light_weight
    assert abs(dim[0] - 8.0) < 1e-6 and abs(dim[1] - 0.8) < 1e-6 and abs(dim[2] - dim[1]) < 1e-6
    assert abs(bright[0] - 4 * math.pi * 0.2 * 0.2) < 1e-6
    assert abs(bright[1] - bright[0] * 20.0) < 1e-5
    assert abs(bright[2] - W) < 1e-5
    assert abs(dim[1] / W + bright[1] / W - 1.0) < 1e-6
    assert bright[1] / W > 0.9


def test_rec709_weights():  # toaster-core light.rs:19-30
    for color, y in (((1, 0, 0), 0.2126), ((0, 1, 0), 0.7152), ((0, 0, 1), 0.0722)):
        sc = scene([emissive(color, 3.0)], tris=[LIGHT_TRIS[0]], background=A.TH_BG_BLACK)
        p = oracle.OracleScene(sc).packed
        assert abs(p.lights[0]["area_weight_cumulative"][1] - 3.0 * y) < 1e-6


# ------------------------------------------------------------------ environment
FAR = [((0, -100, 0), 0.01, 0)]


def _env_scene(w, h, rgb, intensity=1.0, rotation=0.0):
    return oracle.OracleScene(scene([diffuse((1, 1, 1))], sph=FAR, background=A.TH_BG_ENVIRONMENT,
                                    env=env_map(w, h, rgb, intensity, rotation)))


def _dir_at_u(u):
    a = (u - 0.5) * 2 * math.pi
    return (math.cos(a), 0.0, math.sin(a))


def test_env_directions_rotation_intensity():  # environment.rs:295-316
    rgb = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1)]
    e = _env_scene(4, 1, rgb, 2.0)
    assert vec_close(e.env_lookup(_dir_at_u(0.625)), (0, 0, 2), 1e-6)
    assert vec_close(e.env_lookup(_dir_at_u(0.375)), (0, 2, 0), 1e-6)
    r = _env_scene(4, 1, rgb, 1.0, 90.0)
    assert vec_close(r.env_lookup(_dir_at_u(0.375)), (0, 0, 1), 1e-6)


def test_env_cdf_favours_bright_and_latitude():  # environment.rs:354-377
    e = _env_scene(2, 3, [(1, 1, 1), (10, 10, 10), (1, 1, 1), (1, 1, 1), (1, 1, 1), (1, 1, 1)])
    cdf = e.packed.env[:, 3].astype(np.float64)
    prob = np.diff(np.concatenate([[0.0], cdf]))
    assert prob[1] > prob[0] * 9.9
    assert prob[2] > prob[0]
    assert abs(cdf[-1] - 1.0) < 1e-6
    assert abs(prob.sum() - 1.0) < 1e-5


def test_env_sample_pdf_agree():  # environment.rs:379-408
    rgb = [(0.1,) * 3, (0.1,) * 3, (8,) * 3, (0.1,) * 3] + [(0.1,) * 3] * 4
    e = _env_scene(4, 2, rgb, 2.0, 35.0)
    ok, d, rad, pdf = e.env_sample(0.5, 0.37, 0.61)
    assert ok and pdf > 0
    assert abs(np.linalg.norm(d.astype(np.float64)) - 1) < 1e-6
    assert abs(pdf - e.env_pdf(d)) < 1e-5
    assert vec_close(rad, e.env_lookup(d), 1e-6)


def test_env_sampling_frequencies():  # environment.rs:410-447
    rgb = [(1,) * 3, (4,) * 3, (2,) * 3, (3,) * 3, (1,) * 3, (5,) * 3]
    e = _env_scene(3, 2, rgb, 1.0, 27.0)
    cdf = e.packed.env[:, 3].astype(np.float64)
    expected = np.diff(np.concatenate([[0.0], cdf]))
    n = 10000
    counts = np.zeros(6)
    rot = e.packed.params.environment_rotation_degrees
    for s in range(n):
        ok, d, _, _ = e.env_sample((s + 0.5) / n, 0.5, 0.5)
        d = d.astype(np.float64) / np.linalg.norm(d)
        u = (math.atan2(d[2], d[0]) / (2 * math.pi) + 0.5 + rot / 360.0) % 1.0
        v = math.acos(max(-1, min(1, d[1]))) / math.pi
        counts[int(v * 2) * 3 + int(u * 3)] += 1
    assert np.all(np.abs(counts / n - expected) <= 2.0 / n)


def test_black_or_disabled_env_has_no_samples():  # environment.rs:449-457
    black = _env_scene(2, 1, [(0, 0, 0)] * 2)
    ok, _, _, pdf = black.env_sample(0.5, 0.5, 0.5)
    assert not ok and pdf == 0.0
    assert black.env_pdf((1, 0, 0)) == 0.0
    off = _env_scene(1, 1, [(1, 1, 1)], 0.0)
    assert not off.env_sample(0.5, 0.5, 0.5)[0]


def test_env_upload_metadata():  # scene_upload.rs:699-721
    e = _env_scene(4, 2, [(1, 2, 3)] * 8, 8.0, 20.0)
    p = e.packed.params
    assert (p.background_kind, p.environment_width, p.environment_height) == (2, 4, 2)
    assert p.environment_intensity == 8.0 and p.environment_rotation_degrees == 20.0
    assert len(e.packed.env) == 8 and abs(e.packed.env[-1, 3] - 1.0) < 1e-6


def test_env_pdf_integrates_to_one():  # closed form: integral of pdf over the sphere
    rng = np.random.default_rng(3)
    rgb = rng.uniform(0, 4, (8, 16, 3))
    e = _env_scene(16, 8, rgb, 1.0, 33.0)
    n = 40000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    vals = np.array([e.env_pdf(x) for x in d])
    est = vals.mean() * 4 * math.pi
    assert abs(est - 1.0) < 0.03


# ------------------------------------------------------------------ integrator
def _ray(oc, o, d, state=7, **kw):
    p = oc.params
    for k, v in kw.items():
        setattr(p, k, v)
    return oc.ray_color(o, d, state, params=p)[0]


SPH_SCENE = dict(sph=[((0, 0, -2), 0.7, 0)])


def test_zero_depth_is_black():  # integrator.rs:527-537
    oc = oracle.OracleScene(scene([diffuse((1, 1, 1))], **SPH_SCENE))
    assert list(_ray(oc, (0, 0, 0), (0, 1, 0), max_bounces=0)) == [0, 0, 0]


def test_miss_returns_sky():  # integrator.rs:539-549
    oc = oracle.OracleScene(scene([diffuse((1, 1, 1))], **SPH_SCENE, max_bounces=4))
    assert vec_close(_ray(oc, (0, 0, 0), (0, 1, 0)), (0.35, 0.65, 1.0), 1e-6)


def test_miss_samples_environment():  # integrator.rs:551-562
    oc = oracle.OracleScene(scene([diffuse((1, 1, 1))], **SPH_SCENE, background=A.TH_BG_ENVIRONMENT,
                                  env=env_map(1, 1, [(3, 2, 1)], 0.5)))
    assert list(_ray(oc, (0, 0, 0), (1, 0, 0))) == [1.5, 1.0, 0.5]


def test_diffuse_albedo_linearity():  # integrator.rs:564-573
    white = oracle.OracleScene(scene([diffuse((1, 1, 1))], **SPH_SCENE, max_bounces=2))
    col = oracle.OracleScene(scene([diffuse((0.8, 0.4, 0.2))], **SPH_SCENE, max_bounces=2))
    for st in range(20):
        a = _ray(white, (0, 0, 0), (0, 0, -1), state=st)
        b = _ray(col, (0, 0, 0), (0, 0, -1), state=st)
        assert vec_close(b, a * np.array([0.8, 0.4, 0.2], np.float32), 1e-6)


def test_perfect_metal_reflects_to_sky():  # integrator.rs:575-601 (reflection direction)
    floor = [([(-50, 0, 50), (50, 0, 50), (0, 0, -50)], 0)]
    oc = oracle.OracleScene(scene([metal((0.8, 0.8, 0.8), 0.0)], tris=floor, max_bounces=2))
    d = _norm((1, -1, 0))
    c = _ray(oc, (0, 1, 0), d)
    t = (SQ(0.5) + 1) * 0.5  # reflected direction (1,1,0)/sqrt2 -> sky blend
    sky = np.array([1 - t + 0.35 * t, 1 - t + 0.65 * t, 1.0])
    assert vec_close(c, 0.8 * sky, 1e-5)


def test_metal_roughness_clamped():  # integrator.rs:603-633
    floor = [([(-50, 0, 50), (50, 0, 50), (0, 0, -50)], 0)]
    a = oracle.OracleScene(scene([metal((1, 1, 1), 1.0)], tris=floor, max_bounces=3))
    b = oracle.OracleScene(scene([metal((1, 1, 1), 5.0)], tris=floor, max_bounces=3))
    for st in range(20):
        assert list(_ray(a, (0, 1, 0), (0, -1, 0), state=st)) == list(_ray(b, (0, 1, 0), (0, -1, 0), state=st))


def test_snell():  # integrator.rs:635-641
    r = oracle.refract((0.5, -SQ(3) * 0.5, 0.0), (0, 1, 0), 1 / 1.5)
    assert abs(r[0] - 1 / 3) < 1e-6 and abs(np.linalg.norm(r) - 1) < 1e-6


def test_total_internal_reflection_path():  # integrator.rs:643-653
    # a ray inside a dielectric sphere at a grazing angle reflects (no draw consumed)
    oc = oracle.OracleScene(scene([dielectric(1.5)], sph=[((0, 0, 0), 1.0, 0)], max_bounces=1))
    c, st = oc.ray_color((0, 0, 0.9), (0.9, 0.435889894, 0.0), 11)
    assert st == 11  # cannot_refract short-circuits the Fresnel draw (wgsl:885)


def test_schlick_endpoints():  # integrator.rs:655-659
    assert abs(oracle.reflectance(0.0, 1 / 1.5) - 1.0) < 1e-6
    assert abs(oracle.reflectance(1.0, 1 / 1.5) - 0.04) < 1e-6


def test_emissive_hit():  # integrator.rs:681-691
    oc = oracle.OracleScene(scene([emissive((1, 0.5, 0.25), 4.0)], **SPH_SCENE, max_bounces=4))
    assert vec_close(_ray(oc, (0, 0, 0), (0, 0, -1)), (4, 2, 1), 1e-6)


def test_black_background():  # integrator.rs:693-700
    oc = oracle.OracleScene(scene([diffuse((1, 1, 1))], **SPH_SCENE, max_bounces=4, background=A.TH_BG_BLACK))
    assert list(_ray(oc, (0, 0, 0), (0, 1, 0))) == [0, 0, 0]


def _nee_scene(blocked):  # integrator.rs:446-515 (+ a floor to hit)
    tris = [([(-1, 2, -1), (1, 2, -1), (0, 2, 1)], 1),
            ([(-9, 0, 9), (9, 0, 9), (0, 0, -9)], 0)]
    if blocked:
        tris += [([(-3, 1, -3), (3, 1, 3), (3, 1, -3)], 0), ([(-3, 1, -3), (-3, 1, 3), (3, 1, 3)], 0)]
    return oracle.OracleScene(scene([diffuse((1, 1, 1)), emissive((1, 1, 1), 5.0)], tris=tris,
                                    background=A.TH_BG_BLACK, max_bounces=1))


def test_nee_visible_positive_and_blocked_zero():  # integrator.rs:702-734
    vis, blk = _nee_scene(False), _nee_scene(True)
    tot = np.zeros(3)
    for st in range(32):
        tot += _ray(vis, (0, 0.5, 0), (0, -1, 0), state=st)
        assert list(_ray(blk, (0, 0.5, 0), (0, -1, 0), state=st)) == [0, 0, 0]
    assert (tot > 0).all()


def test_env_nee_positive():  # integrator.rs:736-752
    floor = [([(-9, 0, 9), (9, 0, 9), (0, 0, -9)], 0)]
    oc = oracle.OracleScene(scene([diffuse((0.5, 0.5, 0.5))], tris=floor, background=A.TH_BG_ENVIRONMENT,
                                  env=env_map(1, 1, [(2, 2, 2)]), max_bounces=1))
    tot = sum(_ray(oc, (0, 1, 0), (0, -1, 0), state=s) for s in range(64))
    assert (tot > 0).all()


def test_power_heuristic():  # integrator.rs:782-788
    assert abs(oracle.power_heuristic(0.25, 0.75) + oracle.power_heuristic(0.75, 0.25) - 1) < 1e-6
    assert oracle.power_heuristic(0.0, 0.0) == 0.0


def test_white_furnace_env_mis():  # closed form: convex diffuse sphere under uniform env -> a*L
    a, L = 0.6, 1.5
    sc = scene([diffuse((a, a, a))], sph=[((0, 0, -3), 1.0, 0)], background=A.TH_BG_ENVIRONMENT,
               env=env_map(4, 2, [(L, L, L)] * 8), width=8, height=8, samples=512, max_bounces=3,
               fov=20.0)
    img, _ = oracle.OracleScene(sc).render(threads=4)
    assert abs(img[..., :3].mean() - a * L) < 0.01 * a * L


def test_camera_center_ray():  # camera.rs:113-117
    oc = oracle.OracleScene(scene([diffuse((1, 1, 1))], **SPH_SCENE))
    c = oc.camera
    ll, h, v, o = (np.array(x[:3], np.float64) for x in (c.lower_left_corner, c.horizontal, c.vertical, c.origin))
    d = ll + 0.5 * h + 0.5 * v - o
    assert vec_close(d / np.linalg.norm(d), (0, 0, -1), 1e-6)


# ------------------------------------------------------------------ display
def test_display_known_bytes():  # image_output.rs:212-219
    px = np.array([[[-1.0, 0.5, 2.0, 1.0], [0.0, 0.25, 0.75, 0.0]]], np.float32)
    assert oracle.display_rgba8(px).ravel().tolist() == [0, 206, 245, 255, 0, 165, 223, 0]


def test_display_nonfinite():  # color.rs:104-113
    px = np.array([[[np.nan, -np.inf, np.inf, 1.0]]], np.float32)
    assert oracle.display_rgba8(px).ravel().tolist()[:3] == [0, 0, 255]


def test_display_exposure_and_monotone():  # color.rs:80-102
    x = np.linspace(0, 16, 200, dtype=np.float32)
    px = np.zeros((1, len(x), 4), np.float32)
    px[0, :, 0] = x
    b = oracle.display_rgba8(px)[0, :, 0].astype(int)
    assert (np.diff(b) >= 0).all() and b[-1] < 255 + 1
    one = np.array([[[0.1, 0.1, 0.1, 1]]], np.float32)
    assert (oracle.display_rgba8(one, 1.0) == oracle.display_rgba8(one * 2, 0.0)).all()


# ------------------------------------------------------------------ scalar math
def test_transcendentals_within_few_ulp():
    x = np.linspace(-3 * math.pi, 3 * math.pi, 20001).astype(np.float32)
    ulp = lambda v: np.spacing(np.float32(max(abs(v), 1e-30)))
    for fn, ref in ((oracle.sin, math.sin), (oracle.cos, math.cos)):
        err = max(abs(fn(float(v)) - ref(float(v))) for v in x[::7])
        assert err < 3e-7
    for v in np.linspace(-1, 1, 4001).astype(np.float32):
        assert abs(oracle.acos(float(v)) - math.acos(float(v))) <= 4 * ulp(math.acos(float(v))) + 1e-7
    rng = np.random.default_rng(0)
    for y, xx in rng.normal(size=(4000, 2)).astype(np.float32):
        assert abs(oracle.atan2(float(y), float(xx)) - math.atan2(float(y), float(xx))) <= 2.5e-7
    assert oracle.atan2(0.0, -1.0) == np.float32(math.pi)


def test_pcg_hash_is_a_permutation_sample():
    # PCG32 LCG step (747796405, 2891336453) + RXS-M-XS output: a bijection on u32.
    xs = np.arange(1 << 16, dtype=np.uint64) * 65521 + 12345
    outs = {oracle.pcg_hash(int(x)) for x in xs}
    assert len(outs) == len(xs)
