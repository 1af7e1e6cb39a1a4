"""The oracle's random number generator against known answers computed from the text of
the reference shader (tests/golden/make_pcg_vectors.py reads the constants of pcg_hash,
random_f32 and the pixel seed from shaders/pathtrace_spheres.wgsl:189-199, 519 and
evaluates them with numpy uint32 arithmetic), so a wrong constant in the generator that
the oracle and the kernels share fails here; tests/test_gpu_rng.py checks the device."""
import os

import numpy as np

import oracle

from .golden import make_pcg_vectors as gen

V = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pcg_vectors.npz"))


def test_constants_are_the_shader_text():
    assert V["constants"].tolist() == [747796405, 2891336453, 28, 4, 277803737, 22, 9781]
    assert float(V["divisor"]) == 4294967295.0


def test_oracle_hash_matches_the_shader():
    got = np.array([oracle.pcg_hash(int(x)) for x in V["inputs"]], np.uint32)
    assert np.array_equal(got, V["hashes"])


def test_oracle_pixel_streams_match_the_shader():
    for fi, f in enumerate(V["frames"]):
        for pi, px in enumerate(V["pixels"]):
            seed = oracle.pcg_hash(int(px) ^ ((int(f) * 9781) & 0xFFFFFFFF))
            assert seed == int(V["seeds"][fi, pi])
            states, values = oracle.stream(seed, V["states"].shape[2])
            assert np.array_equal(states, V["states"][fi, pi])
            assert np.array_equal(values.view(np.uint32), V["values"][fi, pi].view(np.uint32))


def test_draws_lie_in_the_inclusive_unit_interval():
    v = V["values"]
    assert (v >= 0.0).all() and (v <= 1.0).all()
    # f32(u32) rounds to nearest: states within 128 of 2^32 - 1 map to exactly 1.0
    st = np.array([0xFFFFFFFF, 0xFFFFFF80, 0xFFFFFF7F, 0], np.uint32)
    assert (st.astype(np.float32) / np.float32(4294967295.0)).tolist()[:2] == [1.0, 1.0]


def test_vectors_catch_a_planted_constant():
    """The vectors discriminate: a generator with any one constant changed (a planted
    mistake in a CPU restatement) disagrees with them."""
    base = dict(zip(["mult", "inc", "s1", "s2", "word_mult", "s3", "frame_mult"],
                    [int(x) for x in V["constants"]]), divisor=float(V["divisor"]))
    assert np.array_equal(gen.make_hash(base)(V["inputs"]), V["hashes"])
    for key, delta in (("mult", 2), ("inc", 1), ("s1", 1), ("s2", 1), ("word_mult", 2), ("s3", 1)):
        bad = dict(base, **{key: base[key] + delta})
        assert not np.array_equal(gen.make_hash(bad)(V["inputs"]), V["hashes"]), key
    bad = dict(base, frame_mult=base["frame_mult"] + 2)
    seeds, _, _ = gen.streams(bad, gen.make_hash(bad), V["pixels"], V["frames"], 1)
    assert not np.array_equal(seeds, V["seeds"])
