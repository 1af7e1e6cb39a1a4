"""Writes the regression fixtures of tests/golden/ from the oracle (run from the repository
root: python tests/golden/make_golden.py).  They freeze oracle outputs that the
known-answer tests have pinned, so a later change to the oracle or the packer that
alters any of them fails tests/test_golden.py:

- pcg_pixel0_frame0.npz: the first 16 states and random_f32 draws of pixel 0 at frame 0
  (seed pcg_hash(0 ^ 0*9781), shaders/pathtrace_spheres.wgsl:186-199, 519);
- bvh_cornell.npz, bvh_stress_2048.npz: packed BVH node and primitive arrays;
- primary_cornell_64x36.npz: first camera ray's closest slot and distance per pixel;
- image_cornell_64x36_4spp.npz: the 4-spp linear-HDR image (reference RNG).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from toaster_amd import pack, scenes  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def pcg_draws(n=16):
    state = oracle.pcg_hash(0)
    states, draws = [], []
    for _ in range(n):
        state = oracle.pcg_hash(state)
        states.append(state)
        draws.append(np.float32(np.float32(state) / np.float32(4294967295.0)))
    return np.array(states, np.uint32), np.array(draws, np.float32)


def bvh_arrays(sc):
    p = pack(sc)
    return np.ascontiguousarray(p.nodes).view(np.uint8), np.ascontiguousarray(p.prims).view(np.uint8)


def main():
    s, d = pcg_draws()
    np.savez(os.path.join(HERE, "pcg_pixel0_frame0.npz"), states=s, draws=d)
    for name, sc in (("bvh_cornell", scenes.cornell(64, 36)), ("bvh_stress_2048", scenes.bvh_stress(2048))):
        n, pr = bvh_arrays(sc)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), nodes=n, prims=pr)
    sc = scenes.cornell(64, 36, samples=4, max_bounces=4)
    oc = oracle.OracleScene(sc)
    slots, dist = oc.primary_hits()
    np.savez_compressed(os.path.join(HERE, "primary_cornell_64x36.npz"), slots=slots, distance=dist)
    img, _ = oc.render(threads=4)
    np.savez_compressed(os.path.join(HERE, "image_cornell_64x36_4spp.npz"), rgba=img)
    oc.close()


if __name__ == "__main__":
    main()
