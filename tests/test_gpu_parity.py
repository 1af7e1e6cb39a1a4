"""GPU parity: the HIP path against the oracle on the same seeds (per-pixel PCG stream).

Tolerance (binary32 image, per channel): |gpu - oracle| <= 1e-4 * (1 + |oracle|) on at
least 99.5 % of pixels; the remainder is documented FP path divergence (a one-ulp
difference flipping a dielectric or NEE decision).  Primary-hit slots are bit-exact
wherever the second-nearest candidate is more than 1e-4 * max(1, t) further away.
"""
import numpy as np
import pytest

import oracle
from toaster_amd import GpuScene, RenderOptions, _abi as A, scenes

pytestmark = pytest.mark.gpu

TOL_FRAC = 0.995


def _compare(img, ref):
    close = np.abs(img - ref) <= 1e-4 * (1.0 + np.abs(ref))
    frac = float(np.mean(np.all(close, axis=-1)))
    exact = float(np.mean(np.all(img == ref, axis=-1)))
    return frac, exact


def _render(ctx, sc, opts=None, **kw):
    gs = GpuScene.from_scene(ctx, sc)
    gs.render(options=opts, **kw)
    img = gs.readback()
    gs.close()
    return img


TRAVERSALS = [A.TH_TRAVERSE_REFERENCE, A.TH_TRAVERSE_ORDERED, A.TH_TRAVERSE_WIDE]


@pytest.mark.parametrize("traversal", TRAVERSALS)
@pytest.mark.parametrize("name,make", [
    ("cornell", lambda: scenes.cornell()),
    ("stress2048", lambda: scenes.bvh_stress(2048)),
    ("mixed_env", lambda: scenes.mixed_environment(width=160, height=90, samples=4, max_bounces=8,
                                                   env_size=(256, 128))),
])
def test_image_matches_oracle(gpu_ctx, name, make, traversal):
    sc = make()
    print(f"case {name} traversal={traversal}", flush=True)
    img = _render(gpu_ctx, sc, RenderOptions(traversal=traversal))
    ref, _ = oracle.OracleScene(sc).render(threads=8)
    frac, exact = _compare(img, ref)
    print(f"{name} traversal={traversal}: within tol {frac:.5f} bit-exact {exact:.5f}")
    assert np.isfinite(img).all()
    assert frac >= TOL_FRAC


@pytest.mark.parametrize("schedule", [A.TH_SCHEDULE_DYNAMIC, A.TH_SCHEDULE_OCTET, A.TH_SCHEDULE_SPECULATIVE,
                                      A.TH_SCHEDULE_AUTO])
@pytest.mark.parametrize("name,make", [
    ("cornell", lambda: scenes.cornell()),
    ("stress2048", lambda: scenes.bvh_stress(2048)),
    ("mixed_env", lambda: scenes.mixed_environment(width=160, height=90, samples=4, max_bounces=8,
                                                   env_size=(256, 128))),
])
def test_schedule_matches_oracle(gpu_ctx, name, make, schedule):
    """The 4-lane kernel with other pixel-to-lane assignments (the default is one static
    tile per wave): persistent quads taking pixels from a counter, eight lanes per pixel
    with the shadow rays traced beside the path, or a pixel's samples traced at once from
    guessed starts of its random stream.  Same per-pixel results, same image."""
    sc = make()
    print(f"case {name} wide schedule={schedule}", flush=True)
    img = _render(gpu_ctx, sc, RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=schedule))
    ref, _ = oracle.OracleScene(sc).render(threads=8)
    frac, exact = _compare(img, ref)
    print(f"{name} schedule {schedule}: within tol {frac:.5f} bit-exact {exact:.5f}")
    assert frac >= TOL_FRAC


@pytest.mark.parametrize("name,make", [
    ("cornell_1spp", lambda: scenes.cornell(samples=1)),
    ("cornell_0_bounces", lambda: scenes.cornell(max_bounces=0)),
    ("cornell_1_bounce", lambda: scenes.cornell(max_bounces=1)),
    ("stress8192_odd_size", lambda: scenes.bvh_stress(8192, width=123, height=77, samples=3)),
    ("mixed_env", lambda: scenes.mixed_environment(width=160, height=90, samples=5, max_bounces=8,
                                                   env_size=(256, 128))),
    ("mixed_env_16spp", lambda: scenes.mixed_environment(width=64, height=48, samples=16, max_bounces=8,
                                                         env_size=(256, 128))),
    ("stress2048_9spp", lambda: scenes.bvh_stress(2048, width=96, height=64, samples=9)),
    ("stress2048_2spp_0_bounces", lambda: scenes.bvh_stress(2048, width=64, height=48, samples=2,
                                                            max_bounces=0)),
])
def test_octet_is_bit_identical_to_quad(gpu_ctx, name, make):
    """The octet kernel traces a vertex's shadow rays beside the next bounce (and a
    sample's last shadow rays beside the next sample's primary ray), the speculative
    kernel traces a pixel's samples at once from guessed starts of its random stream and
    restarts the mispredicted ones; both add every term in the reference's order, so their
    images equal the 4-lane kernel's bit for bit, also when folded into an accumulated
    image.  The cases cover guesses that hold (diffuse paths) and guesses that fail
    (metal, glass, escaped and zero-bounce paths), and 1 to 16 samples."""
    sc = make()
    print(f"case octet/speculative vs quad {name}", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    out = []
    for sched in (A.TH_SCHEDULE_STATIC, A.TH_SCHEDULE_OCTET, A.TH_SCHEDULE_SPECULATIVE):
        o = RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=sched)
        gs.render(options=o, frame_index=0, accumulated_samples=0)
        gs.render(options=o, frame_index=1, accumulated_samples=sc.samples)
        out.append(gs.readback().copy())
    gs.close()
    assert np.array_equal(out[0].view(np.uint32), out[1].view(np.uint32))
    assert np.array_equal(out[0].view(np.uint32), out[2].view(np.uint32))


@pytest.mark.parametrize("name,make", [
    ("cornell", lambda: scenes.cornell(width=96, height=64)),
    ("stress2048", lambda: scenes.bvh_stress(2048, width=96, height=64, samples=3)),
])
def test_auto_schedule_keeps_the_accumulation(gpu_ctx, name, make):
    """TH_SCHEDULE_AUTO times the schedules on its first render of the scene (overwriting
    the accumulation while it does) and must leave exactly what a fixed schedule leaves,
    also when that first render folds into an accumulated frame; later renders reuse the
    choice."""
    sc = make()
    print(f"case auto schedule {name}", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    static = RenderOptions(traversal=A.TH_TRAVERSE_WIDE)
    auto = RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=A.TH_SCHEDULE_AUTO)
    out = []
    for second in (static, auto):
        gs.render(options=static, frame_index=0, accumulated_samples=0)
        gs.render(options=second, frame_index=1, accumulated_samples=sc.samples)  # auto: tunes here
        gs.render(options=second, frame_index=2, accumulated_samples=2 * sc.samples)
        out.append(gs.readback().copy())
    gs.close()
    assert np.array_equal(out[0].view(np.uint32), out[1].view(np.uint32))
    # the timed trials stay out of the statistics: one launch per render
    gs = GpuScene.from_scene(gpu_ctx, sc)
    gpu_ctx.synchronize()
    gpu_ctx.reset_stats()
    gs.render(options=auto, frame_index=0, accumulated_samples=0)
    gs.render(options=auto, frame_index=1, accumulated_samples=sc.samples)
    assert gpu_ctx.stats()["launches"] == 2
    gs.close()


def test_launch_timing_stays_bounded(gpu_ctx):
    """Every launch is timed by an event pair that th_get_stats folds into the totals; a
    caller that renders 5,000 frames without reading the statistics keeps the pending
    list bounded (its oldest 2,048 are folded at 4,096) and still has every launch
    counted, with the last one's time in last_kernel_ms."""
    sc = scenes.cornell(width=16, height=16, samples=1, max_bounces=1)
    print("case 5000 small renders without stats reads", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    gpu_ctx.synchronize()
    gpu_ctx.reset_stats()
    n = 5000
    for i in range(n):
        gs.render(frame_index=i, accumulated_samples=0)
    st = gpu_ctx.stats()
    gs.close()
    assert st["launches"] == n
    assert 0.0 < st["last_kernel_ms"] <= st["total_kernel_ms"]


@pytest.mark.parametrize("traversal,schedule", [(A.TH_TRAVERSE_ORDERED, A.TH_SCHEDULE_STATIC),
                                                (A.TH_TRAVERSE_WIDE, A.TH_SCHEDULE_STATIC),
                                                (A.TH_TRAVERSE_WIDE, A.TH_SCHEDULE_OCTET),
                                                (A.TH_TRAVERSE_WIDE, A.TH_SCHEDULE_SPECULATIVE),
                                                (A.TH_TRAVERSE_WIDE, A.TH_SCHEDULE_AUTO)])
@pytest.mark.parametrize("samples", [1, 8, 16])
def test_per_sample_rng_matches_oracle(gpu_ctx, traversal, schedule, samples):
    """Per-sample RNG; the speculative schedule runs the octet kernel there (no stream to
    foresee) and the automatic one the 4-lane kernel."""
    sc = scenes.bvh_stress(2048, samples=samples)
    print(f"case per-sample traversal={traversal} schedule={schedule} spp={samples}", flush=True)
    img = _render(gpu_ctx, sc, RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE, traversal=traversal,
                                             schedule=schedule))
    # lanes/quads/octets per pixel used by the kernel: megakernel min(spp, 8), quad
    # min(spp, 16), octet min(spp, 8)
    octets = schedule in (A.TH_SCHEDULE_OCTET, A.TH_SCHEDULE_SPECULATIVE)
    cap = 16 if (traversal == A.TH_TRAVERSE_WIDE and not octets) else 8
    group = 1
    while group * 2 <= min(samples, cap):
        group *= 2
    ref, _ = oracle.OracleScene(sc).render(threads=8, rng_mode=1, group=group)
    frac, exact = _compare(img, ref)
    print(f"per-sample spp={samples} traversal={traversal}: within tol {frac:.5f} bit-exact {exact:.5f}")
    assert frac >= TOL_FRAC
    if samples == 1:  # one sample: the per-sample stream is the reference stream
        ref1, _ = oracle.OracleScene(sc).render(threads=8)
        assert _compare(img, ref1)[0] >= TOL_FRAC


@pytest.mark.parametrize("traversal", TRAVERSALS)
@pytest.mark.parametrize("rng,samples", [(A.TH_RNG_REFERENCE, 4), (A.TH_RNG_PER_SAMPLE, 8)])
@pytest.mark.parametrize("name,make", [
    ("cornell", lambda spp: scenes.cornell(samples=spp)),
    ("mixed_env", lambda spp: scenes.mixed_environment(width=160, height=90, samples=spp, max_bounces=8,
                                                       env_size=(256, 128))),
])
def test_wavefront_matches_oracle(gpu_ctx, name, make, rng, samples, traversal):
    sc = make(samples)
    print(f"case wavefront {name} rng={rng} spp={samples} traversal={traversal}", flush=True)
    img = _render(gpu_ctx, sc, RenderOptions(rng_mode=rng, traversal=traversal,
                                             kernel=A.TH_KERNEL_WAVEFRONT))
    group = 1
    if rng == A.TH_RNG_PER_SAMPLE:
        while group * 2 <= min(samples, 16):
            group *= 2
    ref, _ = oracle.OracleScene(sc).render(threads=8, rng_mode=rng, group=group)
    frac, exact = _compare(img, ref)
    print(f"wavefront {name} rng={rng} traversal={traversal}: within tol {frac:.5f} bit-exact {exact:.5f}")
    assert frac >= TOL_FRAC


def test_primary_hits_match(gpu_ctx):
    sc = scenes.bvh_stress(8192)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    oc = oracle.OracleScene(sc)
    ref_slot, ref_t = oc.primary_hits()
    for trav in TRAVERSALS:
        slot, t = gs.primary_hits(traversal=trav)
        same = slot == ref_slot
        print(f"primary traversal={trav}: identical {same.mean():.6f}")
        assert same.mean() >= 0.9999
        assert np.array_equal(t[same], ref_t[same])
    gs.close()


def test_progressive_accumulation(gpu_ctx):
    sc = scenes.cornell(width=64, height=36, samples=2)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    oc = oracle.OracleScene(sc)
    acc = None
    for frame in range(3):
        p = gs.render_progressive(2)
        ref, _ = oc.render(params=p, accum=acc)
        acc = ref
    img = gs.readback()
    gs.close()
    frac, _ = _compare(img, acc)
    assert frac >= TOL_FRAC


def test_display_device_matches_host(gpu_ctx):
    from toaster_amd import display_rgba8
    sc = scenes.cornell(width=64, height=36)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    gs.render()
    img = gs.readback()
    dev = gs.display_rgba8(0.5)
    # the staging buffer is kept with the scene: a second conversion after another
    # render and at another exposure must not see the first one's bytes
    gs.render(frame_index=1, accumulated_samples=sc.samples)
    img2 = gs.readback()
    dev2 = gs.display_rgba8(-1.0)
    gs.close()
    assert np.array_equal(dev, display_rgba8(img, 0.5))
    assert np.array_equal(dev2, display_rgba8(img2, -1.0))
    # and against the oracle's restatement of color.rs / image_output.rs directly
    assert np.array_equal(dev, oracle.display_rgba8(img, 0.5))
    assert np.array_equal(dev2, oracle.display_rgba8(img2, -1.0))


def test_display_device_matches_oracle_on_extreme_values(gpu_ctx):
    """k_display against the oracle's display transform on an accumulation holding
    exact zeros, negatives, the sRGB and ACES knees, large values, infinities and NaN
    (color.rs:61-74 rules), uploaded as a progressive checkpoint."""
    sc = scenes.cornell(width=32, height=16)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    rng = np.random.default_rng(7)
    vals = np.concatenate([np.array([0.0, -0.0, -1.0, 1e-9, 0.0031308, 0.04045, 0.18, 0.5, 1.0, 2.0, 16.0, 1e6,
                                     3.4e38, np.inf, -np.inf, np.nan], np.float32),
                           rng.exponential(0.5, 32 * 16 * 4 - 16).astype(np.float32)])
    acc = vals.reshape(16, 32, 4)
    gs.upload_accumulation(acc, sc.samples, 1)
    for stops in (0.0, 1.5, -3.0):
        assert np.array_equal(gs.display_rgba8(stops), oracle.display_rgba8(acc, stops))
    gs.close()


def test_converged_per_sample_matches_oracle_reference_stream(gpu_ctx):
    """Converged images of the two random-stream modes agree: the per-sample-stream GPU
    render and the oracle's reference-stream render, 65,536 spp each on a 32x18 Cornell
    box, differ by relRMSE <= 1e-2 (relRMSE = sqrt(mean (I - R)^2 / (R^2 + 0.01)))."""
    from toaster_amd.guide import rel_rmse
    sc = scenes.cornell(32, 18, samples=65536, max_bounces=4)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    print("case converged cornell 32x18 65536 spp", flush=True)
    for c in range(16):
        gs.render(samples=4096, frame_index=500 + c, accumulated_samples=4096 * c,
                  options=RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE, traversal=A.TH_TRAVERSE_WIDE))
    img = gs.readback()
    gs.close()
    ref, _ = oracle.OracleScene(sc).render(threads=16)
    err = rel_rmse(img, ref)
    print(f"  converged relRMSE {err:.5f}")
    assert err <= 1e-2
