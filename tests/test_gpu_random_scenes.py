"""Seeded random scenes through every kernel against the oracle: random mixes of spheres
and triangles (with and without shading normals and texture coordinates), all material
kinds (diffuse, metal, dielectric, emissive, textured diffuse), sky, black and
environment-map backgrounds, 1-6 samples and 0-6 bounces.  Image tolerance of
tests/test_gpu_parity.py.  TOASTER_EXTRA_RANDOM_SEEDS="lo:hi" adds seeds lo..hi-1 to the
single-frame, two-frame, guided and refit comparisons (a wider sweep than the default seeds)."""
import os

import numpy as np
import pytest

import oracle
from toaster_amd import GpuScene, RenderOptions, _abi as A
from toaster_amd.scene import (ATTR_IN_DTYPE, MATERIAL_IN_DTYPE, SPHERE_IN_DTYPE, TRIANGLE_IN_DTYPE,
                               EnvironmentMap, Scene)

pytestmark = pytest.mark.gpu

TOL_FRAC = 0.995
KERNELS = {
    "reference": RenderOptions(),
    "ordered": RenderOptions(traversal=A.TH_TRAVERSE_ORDERED),
    "quad": RenderOptions(traversal=A.TH_TRAVERSE_WIDE),
    "octet": RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=A.TH_SCHEDULE_OCTET),
    "speculative": RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=A.TH_SCHEDULE_SPECULATIVE),
    "auto": RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=A.TH_SCHEDULE_AUTO),
    "wavefront": RenderOptions(traversal=A.TH_TRAVERSE_WIDE, kernel=A.TH_KERNEL_WAVEFRONT),
}
SEEDS = list(range(10))
_extra = os.environ.get("TOASTER_EXTRA_RANDOM_SEEDS", "")
EXTRA_SEEDS = list(range(*map(int, _extra.split(":")))) if _extra else []


def random_scene(seed):
    rng = np.random.default_rng(1000 + seed)
    textures = [rng.integers(0, 256, (int(rng.integers(1, 9)), int(rng.integers(1, 9)), 4), dtype=np.uint8)
                for _ in range(int(rng.integers(0, 3)))]
    background = int(rng.choice([A.TH_BG_SKY, A.TH_BG_BLACK, A.TH_BG_ENVIRONMENT]))
    rows = []
    for _ in range(int(rng.integers(1, 8))):
        kind = rng.integers(0, 5 if textures else 4)
        col = tuple(float(c) for c in rng.uniform(0.05, 0.95, 3))
        if kind == 0:
            rows.append((A.TH_MAT_DIFFUSE, col, 0.0, 0.0, 0.0, 0))
        elif kind == 1:
            rows.append((A.TH_MAT_METAL, col, float(rng.uniform(0.0, 1.0)), 0.0, 0.0, 0))
        elif kind == 2:
            rows.append((A.TH_MAT_DIELECTRIC, (1.0, 1.0, 1.0), 0.0, float(rng.uniform(1.1, 2.0)), 0.0, 0))
        elif kind == 3:
            rows.append((A.TH_MAT_EMISSIVE, col, 0.0, 0.0, float(rng.uniform(1.0, 20.0)), 0))
        else:
            rows.append((A.TH_MAT_TEXTURED_DIFFUSE, col, 0.0, 0.0, 0.0, int(rng.integers(0, len(textures)))))
    if background == A.TH_BG_BLACK and not any(r[0] == A.TH_MAT_EMISSIVE for r in rows):
        rows.append((A.TH_MAT_EMISSIVE, (1.0, 0.9, 0.8), 0.0, 0.0, 10.0, 0))
    mats = np.zeros(len(rows), MATERIAL_IN_DTYPE)
    for i, r in enumerate(rows):
        mats[i] = r
    nm = len(rows)
    ns = int(rng.integers(0, 12))
    nt = int(rng.integers(0 if ns else 1, 300))
    sph = np.zeros(ns, SPHERE_IN_DTYPE)
    sph["center"] = rng.uniform(-2.0, 2.0, (ns, 3))
    sph["radius"] = rng.uniform(0.1, 0.8, ns)
    sph["material_index"] = rng.integers(0, nm, ns)
    tris = np.zeros(nt, TRIANGLE_IN_DTYPE)
    v0 = rng.uniform(-2.5, 2.5, (nt, 1, 3))
    tris["v"] = v0 + rng.uniform(-0.6, 0.6, (nt, 3, 3))
    tris["material_index"] = rng.integers(0, nm, nt)
    if background == A.TH_BG_BLACK:  # the emitter is on some object
        em = [i for i, r in enumerate(rows) if r[0] == A.TH_MAT_EMISSIVE][0]
        if ns:
            sph["material_index"][0] = em
        else:
            tris["material_index"][0] = em
    att = np.zeros(nt, ATTR_IN_DTYPE)
    if nt:
        with_n = rng.random(nt) < 0.3
        att["has_normals"] = with_n
        n = rng.normal(size=(nt, 3, 3))
        att["normals"] = n / np.linalg.norm(n, axis=-1, keepdims=True)
        att["has_tex_coords"] = rng.random(nt) < 0.5
        att["tex_coords"] = rng.uniform(-0.5, 2.0, (nt, 3, 2))
    env = None
    if background == A.TH_BG_ENVIRONMENT:
        env = EnvironmentMap(rgb=rng.uniform(0.0, 3.0, (16, 32, 3)).astype(np.float32),
                             intensity=float(rng.uniform(0.5, 2.0)), rotation_degrees=float(rng.uniform(0, 360)))
    th, ph = rng.uniform(0.2, 1.2), rng.uniform(0, 2 * np.pi)
    cam = (6.0 * np.sin(th) * np.cos(ph), 6.0 * np.cos(th), 6.0 * np.sin(th) * np.sin(ph))
    return Scene(camera_position=tuple(float(c) for c in cam), camera_look_at=(0.0, 0.0, 0.0),
                 fov_degrees=float(rng.uniform(30, 60)), width=48, height=32,
                 samples=int(rng.integers(1, 7)), max_bounces=int(rng.integers(0, 7)), background=background,
                 materials=mats, spheres=sph, triangles=tris, attributes=att if nt else None,
                 textures=textures, environment=env, name=f"random{seed}")


@pytest.mark.parametrize("seed", SEEDS + EXTRA_SEEDS)
def test_random_scene_matches_oracle(gpu_ctx, seed):
    sc = random_scene(seed)
    print(f"case random scene {seed}: {len(sc.spheres)} spheres, {len(sc.triangles)} triangles, "
          f"{len(sc.materials)} materials, background {sc.background}, {sc.samples} spp, "
          f"{sc.max_bounces} bounces", flush=True)
    ref, _ = oracle.OracleScene(sc).render(threads=8)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    for name, opts in KERNELS.items():
        print(f"  kernel {name}", flush=True)
        gs.render(options=opts, frame_index=0, accumulated_samples=0)
        img = gs.readback()
        assert np.isfinite(img).all()
        close = np.abs(img - ref) <= 1e-4 * (1.0 + np.abs(ref))
        frac = float(np.mean(np.all(close, axis=-1)))
        exact = float(np.mean(np.all(img == ref, axis=-1)))
        print(f"  random {seed} {name}: within tol {frac:.5f} bit-exact {exact:.5f}")
        assert frac >= TOL_FRAC, (name, frac)
    gs.close()


PROGRESSIVE = {
    "speculative": (A.TH_RNG_REFERENCE, RenderOptions(traversal=A.TH_TRAVERSE_WIDE,
                                                      schedule=A.TH_SCHEDULE_SPECULATIVE)),
    "octet": (A.TH_RNG_REFERENCE, RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=A.TH_SCHEDULE_OCTET)),
    "per_sample_quad": (A.TH_RNG_PER_SAMPLE, RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE,
                                                           traversal=A.TH_TRAVERSE_WIDE)),
}


@pytest.mark.parametrize("seed", SEEDS[:5] + EXTRA_SEEDS)
def test_random_scene_two_frames_accumulated(gpu_ctx, seed):
    """Two progressive frames (seed 0, then seed 1 folded into the average) against the
    oracle run the same way; per-sample RNG with the 4-lane kernel's group count."""
    sc = random_scene(seed)
    print(f"case random progressive {seed}", flush=True)
    o = oracle.OracleScene(sc)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    for name, (rng, opts) in PROGRESSIVE.items():
        group = 1
        if rng == A.TH_RNG_PER_SAMPLE:
            while group * 2 <= min(sc.samples, 16):
                group *= 2
        p0 = type(o.params).from_buffer_copy(o.params)
        p0.frame_index, p0.accumulated_samples = 0, 0
        ref, _ = o.render(params=p0, threads=8, rng_mode=rng, group=group)
        p1 = type(o.params).from_buffer_copy(o.params)
        p1.frame_index, p1.accumulated_samples = 1, sc.samples
        ref, _ = o.render(params=p1, accum=ref, threads=8, rng_mode=rng, group=group)
        gs.render(options=opts, frame_index=0, accumulated_samples=0)
        gs.render(options=opts, frame_index=1, accumulated_samples=sc.samples)
        img = gs.readback()
        close = np.abs(img - ref) <= 1e-4 * (1.0 + np.abs(ref))
        frac = float(np.mean(np.all(close, axis=-1)))
        print(f"  random {seed} {name} two frames: within tol {frac:.5f}")
        assert frac >= TOL_FRAC, (name, frac)
    gs.close()


@pytest.mark.parametrize("rng", [A.TH_RNG_REFERENCE, A.TH_RNG_PER_SAMPLE])
def test_random_scenes_in_one_batch(gpu_ctx, rng):
    """All random scenes in one persistent batched launch: each equals its oracle image."""
    from toaster_amd.batch import BatchRenderer
    items = [random_scene(s) for s in SEEDS]
    print(f"case random batch rng={rng}", flush=True)
    br = BatchRenderer(gpu_ctx)
    for sc in items:
        br.add(sc)
    br.render(RenderOptions(rng_mode=rng), use_tuning=False)
    imgs = [gs.readback() for gs in br.scenes]
    for sc, img in zip(items, imgs):
        group = 1
        if rng == A.TH_RNG_PER_SAMPLE:
            while group * 2 <= min(sc.samples, 16):
                group *= 2
        ref, _ = oracle.OracleScene(sc).render(threads=8, rng_mode=rng, group=group)
        close = np.abs(img - ref) <= 1e-4 * (1.0 + np.abs(ref))
        frac = float(np.mean(np.all(close, axis=-1)))
        print(f"  {sc.name} batched: within tol {frac:.5f}")
        assert frac >= TOL_FRAC, (sc.name, frac)
    br.close()


@pytest.mark.parametrize("seed", SEEDS + EXTRA_SEEDS)
def test_random_scene_guided_alpha_zero(gpu_ctx, seed):
    """The guided wavefront at alpha = 0 (its own shading kernels, training records on)
    is the reference estimator with per-sample streams: equal to the oracle's image."""
    from toaster_amd.guide import Guide, render_guided
    sc = random_scene(seed)
    print(f"case random guided alpha=0 {seed}", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    g = Guide(gpu_ctx, seed=3)
    render_guided(gs, g, sc.samples, alpha=0.0, train=True, batch=1024, progressive=False, warmup_steps=0)
    img = gs.readback()
    ref, _ = oracle.OracleScene(sc).render(threads=8, rng_mode=A.TH_RNG_PER_SAMPLE, group=1)
    close = np.abs(img - ref) <= 1e-4 * (1.0 + np.abs(ref))
    frac = float(np.mean(np.all(close, axis=-1)))
    print(f"  random {seed} guided alpha=0: within tol {frac:.5f}")
    assert frac >= TOL_FRAC
    g.close()
    gs.close()


@pytest.mark.parametrize("seed", SEEDS[:6] + EXTRA_SEEDS)
def test_random_scene_refit_after_motion(gpu_ctx, seed):
    """The random scene uploaded once, then moved (every other triangle rotated about the
    y axis, spheres shifted) and refitted on the device: equal to the oracle's image of the
    moved scene."""
    import dataclasses
    sc = random_scene(seed)
    print(f"case random refit {seed}", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    tris = sc.triangles.copy()
    if len(tris):
        v = tris["v"].astype(np.float64)
        c = v.reshape(-1, 3).mean(axis=0)
        R = np.array([[np.cos(0.3), 0.0, np.sin(0.3)], [0.0, 1.0, 0.0], [-np.sin(0.3), 0.0, np.cos(0.3)]])
        sel = np.arange(len(tris)) % 2 == 0
        v[sel] = (v[sel] - c) @ R.T + c
        tris["v"] = v.astype(np.float32)
    sph = sc.spheres.copy()
    sph["center"] = sph["center"] + np.float32(0.2)
    moved = dataclasses.replace(sc, triangles=tris, spheres=sph)
    gs.refit(moved)
    ref, _ = oracle.OracleScene(moved).render(threads=8)
    for name in ("reference", "speculative", "wavefront"):
        gs.render(options=KERNELS[name], frame_index=0, accumulated_samples=0)
        img = gs.readback()
        close = np.abs(img - ref) <= 1e-4 * (1.0 + np.abs(ref))
        frac = float(np.mean(np.all(close, axis=-1)))
        print(f"  random {seed} refit {name}: within tol {frac:.5f}")
        assert frac >= TOL_FRAC, (name, frac)
    gs.close()
