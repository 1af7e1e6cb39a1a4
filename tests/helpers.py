"""Small scene builders shared by the tests."""
import numpy as np

from toaster_amd import _abi as A
from toaster_amd.scene import (MATERIAL_IN_DTYPE, SPHERE_IN_DTYPE, TRIANGLE_IN_DTYPE, EnvironmentMap,
                               Scene)


def materials(rows):
    m = np.zeros(len(rows), MATERIAL_IN_DTYPE)
    for i, r in enumerate(rows):
        m[i] = r
    return m


def spheres(rows):
    s = np.zeros(len(rows), SPHERE_IN_DTYPE)
    for i, (c, r, mi) in enumerate(rows):
        s[i] = (c, r, mi)
    return s


def triangles(rows):
    t = np.zeros(len(rows), TRIANGLE_IN_DTYPE)
    for i, (v, mi) in enumerate(rows):
        t[i]["v"] = np.asarray(v, np.float32)
        t[i]["material_index"] = mi
    return t


def scene(mats, sph=(), tris=(), background=A.TH_BG_SKY, env=None, width=1, height=1, samples=1,
          max_bounces=1, cam=((0, 0, 0), (0, 0, -1)), fov=60.0):
    return Scene(camera_position=cam[0], camera_look_at=cam[1], fov_degrees=fov, width=width,
                 height=height, samples=samples, max_bounces=max_bounces, background=background,
                 materials=materials(mats), spheres=spheres(sph) if len(sph) else spheres([]),
                 triangles=triangles(tris) if len(tris) else triangles([]), environment=env)


def env_map(w, h, rgb, intensity=1.0, rotation=0.0):
    a = np.asarray(rgb, np.float32).reshape(h, w, 3)
    return EnvironmentMap(a, intensity, rotation)
