"""Writes tests/golden/pcg_vectors.npz: known answers of the reference kernel's random
number generator, computed from the text of the reference shader, not from this
repository's code (run from the repository root with the reference checked out:
python tests/golden/make_pcg_vectors.py [path/to/pathtrace_spheres.wgsl]).

The constants are read from shaders/pathtrace_spheres.wgsl:
- pcg_hash (lines 189-193): state = input * M + I; word = ((state >> ((state >> S1) + S2))
  ^ state) * W; return (word >> S3) ^ word;
- random_f32 (lines 196-199): rng_state = pcg_hash(rng_state); f32(rng_state) / D;
- the pixel seed (line 519): rng_state = pcg_hash(idx ^ (params.frame_index * F)).

They are evaluated with numpy uint32 wrap-around arithmetic and binary32 conversion and
division (IEEE, round to nearest even).  The file holds the parsed constants, hashes of
fixed inputs and, for a set of (pixel, frame) pairs, the first 16 states and draws of
the pixel's stream.
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WGSL = "/root/reference/shaders/pathtrace_spheres.wgsl"
DRAWS = 16
PIXELS = np.array([0, 1, 2, 319, 320, 12345, 28960, 57599, 65535, 921599], np.uint32)
FRAMES = np.array([0, 1, 2, 7, 1000, 4294967295], np.uint32)


def parse(text):
    body = re.search(r"fn pcg_hash\(input: u32\) -> u32 \{(.*?)\n\}", text, re.S).group(1)
    m = re.search(r"var state = input \* (\d+)u \+ (\d+)u;", body)
    w = re.search(r"let word = \(\(state >> \(\(state >> (\d+)u\) \+ (\d+)u\)\) \^ state\) \* (\d+)u;", body)
    r = re.search(r"return \(word >> (\d+)u\) \^ word;", body)
    f = re.search(r"rng_state = pcg_hash\(idx \^ \(params\.frame_index \* (\d+)u\)\);", text)
    d = re.search(r"rng_state = pcg_hash\(rng_state\);\s*return f32\(rng_state\) / ([0-9.]+);", text)
    return {"mult": int(m.group(1)), "inc": int(m.group(2)), "s1": int(w.group(1)), "s2": int(w.group(2)),
            "word_mult": int(w.group(3)), "s3": int(r.group(1)), "frame_mult": int(f.group(1)),
            "divisor": float(d.group(1))}


def make_hash(c):
    M, I, W = np.uint32(c["mult"]), np.uint32(c["inc"]), np.uint32(c["word_mult"])
    S1, S2, S3 = np.uint32(c["s1"]), np.uint32(c["s2"]), np.uint32(c["s3"])

    def h(x):
        with np.errstate(over="ignore"):
            x = np.asarray(x, np.uint32)
            state = x * M + I
            word = ((state >> ((state >> S1) + S2)) ^ state) * W
            return (word >> S3) ^ word
    return h


def streams(c, h, pixels, frames, draws=DRAWS):
    div = np.float32(c["divisor"])
    seeds, states, values = [], [], []
    for f in frames:
        with np.errstate(over="ignore"):
            s = h(pixels ^ (np.uint32(f) * np.uint32(c["frame_mult"])))
        seeds.append(s)
        st, va = [], []
        for _ in range(draws):
            s = h(s)
            st.append(s)
            va.append(s.astype(np.float32) / div)
        states.append(np.stack(st, 1))
        values.append(np.stack(va, 1))
    return np.stack(seeds), np.stack(states), np.stack(values)


def main():
    path = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_WGSL
    with open(path) as fh:
        c = parse(fh.read())
    h = make_hash(c)
    lcg = (np.arange(512, dtype=np.uint64) * np.uint64(2654435761) + np.uint64(12345)) & np.uint64(0xFFFFFFFF)
    inputs = np.concatenate([np.arange(256, dtype=np.uint32),
                             np.array([0x7FFFFFFF, 0x80000000, 0xFFFFFFFE, 0xFFFFFFFF, 0xDEADBEEF], np.uint32),
                             lcg.astype(np.uint32)])
    seeds, states, values = streams(c, h, PIXELS, FRAMES)
    out = os.path.join(HERE, "pcg_vectors.npz")
    np.savez(out, constants=np.array([c["mult"], c["inc"], c["s1"], c["s2"], c["word_mult"], c["s3"],
                                      c["frame_mult"]], np.uint64),
             divisor=np.float64(c["divisor"]), inputs=inputs, hashes=h(inputs), pixels=PIXELS, frames=FRAMES,
             seeds=seeds, states=states, values=values)
    print("constants", c, "->", out)


if __name__ == "__main__":
    main()
