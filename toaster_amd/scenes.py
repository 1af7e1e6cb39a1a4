"""Seeded synthetic scenes for the benchmark workloads.

Random numbers come from PCG32 (O'Neill 2014, XSH-RR output, multiplier
6364136223846793005) seeded through SplitMix64 (Steele, Lea & Flood 2014); streams
are generated vectorised with the LCG jump-ahead A_i = a^i, B_i = c * sum_{k<i} a^k.
Workload seed for workload k is 0x70A57E400000000k.
"""
import numpy as np

from . import _abi as A
from .scene import (ATTR_IN_DTYPE, MATERIAL_IN_DTYPE, SPHERE_IN_DTYPE, TRIANGLE_IN_DTYPE,
                    EnvironmentMap, Scene, dielectric, diffuse, emissive, metal)

M64 = (1 << 64) - 1
PCG_MULT = 6364136223846793005


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & M64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return x, z ^ (z >> 31)


class Pcg32:
    """PCG32 XSH-RR stream producing blocks of uniform floats in [0, 1)."""

    def __init__(self, seed):
        x, init_state = splitmix64(seed)
        _, init_seq = splitmix64(x)
        self.inc = ((init_seq << 1) | 1) & M64
        self.state = 0
        self._step()
        self.state = (self.state + init_state) & M64
        self._step()

    def _step(self):
        self.state = (self.state * PCG_MULT + self.inc) & M64

    def u32(self, n):
        n = int(n)
        if n == 0:
            return np.zeros(0, np.uint32)
        with np.errstate(over="ignore"):
            a = np.full(n, PCG_MULT, dtype=np.uint64)
            a[0] = 1
            apow = np.multiply.accumulate(a, dtype=np.uint64)  # a^i mod 2^64
            csum = np.concatenate([[np.uint64(0)], np.add.accumulate(apow[:-1], dtype=np.uint64)])
            states = apow * np.uint64(self.state) + np.uint64(self.inc) * csum
        # advance past the n states just used
        a_n = pow(PCG_MULT, n, 1 << 64)
        s_sum = (int(csum[-1]) + int(apow[-1])) & M64
        self.state = (a_n * self.state + self.inc * s_sum) & M64
        xorshifted = (((states >> np.uint64(18)) ^ states) >> np.uint64(27)) & np.uint64(0xFFFFFFFF)
        rot = (states >> np.uint64(59)).astype(np.uint64)
        xs = xorshifted.astype(np.uint64)
        out = ((xs >> rot) | (xs << ((np.uint64(32) - rot) & np.uint64(31)))) & np.uint64(0xFFFFFFFF)
        return out.astype(np.uint32)

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self.u32(n) >> np.uint32(8)).astype(np.float64) * (1.0 / (1 << 24))
        return (lo + (hi - lo) * u).astype(np.float32)


def workload_seed(k):
    return 0x70A57E4000000000 | k


def _materials(rows):
    m = np.zeros(len(rows), MATERIAL_IN_DTYPE)
    for i, r in enumerate(rows):
        m[i] = r
    return m


def _tris(verts, mats):
    t = np.zeros(len(verts), TRIANGLE_IN_DTYPE)
    t["v"] = np.asarray(verts, np.float32).reshape(-1, 3, 3)
    t["material_index"] = mats
    return t


def _quad(p0, p1, p2, p3):
    """Two triangles (p0 p1 p2), (p0 p2 p3); normal = (p1-p0) x (p2-p0)."""
    return [[p0, p1, p2], [p0, p2, p3]]


def cornell(width=320, height=180, samples=4, max_bounces=4):
    """Workload 1: Cornell-style box, 10 wall triangles + 0.25x0.25 light quad (2 triangles,
    radiance 15), one diffuse and one dielectric sphere (r = 0.3), black background."""
    grey, red, green = (0.73, 0.73, 0.73), (0.65, 0.05, 0.05), (0.12, 0.45, 0.15)
    mats = _materials([diffuse(grey), diffuse(red), diffuse(green), emissive((1, 1, 1), 15.0),
                       dielectric(1.5)])
    x0, x1, y0, y1, z0, z1 = -1.0, 1.0, 0.0, 2.0, -2.0, 0.0
    verts, mi = [], []
    for q, m in ((_quad((x0, y0, z1), (x0, y0, z0), (x1, y0, z0), (x1, y0, z1)), 0),  # floor
                 (_quad((x0, y1, z1), (x1, y1, z1), (x1, y1, z0), (x0, y1, z0)), 0),  # ceiling
                 (_quad((x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (x1, y0, z0)), 0),  # back
                 (_quad((x0, y0, z1), (x0, y1, z1), (x0, y1, z0), (x0, y0, z0)), 1),  # left
                 (_quad((x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1)), 2)):  # right
        verts += q
        mi += [m, m]
    a, h = 0.125, 1.99
    verts += _quad((-a, h, -1 - a), (a, h, -1 - a), (a, h, -1 + a), (-a, h, -1 + a))  # normal -y
    mi += [3, 3]
    sph = np.zeros(2, SPHERE_IN_DTYPE)
    sph[0] = ((-0.45, 0.3, -1.2), 0.3, 0)
    sph[1] = ((0.45, 0.3, -0.8), 0.3, 4)
    return Scene(camera_position=(0.0, 1.0, 3.3), camera_look_at=(0.0, 1.0, -1.0), fov_degrees=40.0,
                 width=width, height=height, samples=samples, max_bounces=max_bounces,
                 background=A.TH_BG_BLACK, materials=mats, spheres=sph, triangles=_tris(verts, mi),
                 name="cornell")


def random_triangles(rng, n, lo=(-2.0, -2.0, -2.0), hi=(2.0, 2.0, 2.0), offset=0.15):
    """Centroids uniform in [lo, hi], vertices = centroid + U[-offset, offset]^3; a triangle
    with |e1 x e2|^2 <= f32::EPSILON is redrawn (loader.rs:1584-1593)."""
    out = np.zeros((0, 3, 3), np.float32)
    lo, hi = np.asarray(lo, np.float32), np.asarray(hi, np.float32)
    while len(out) < n:
        k = n - len(out)
        c = rng.uniform(3 * k).reshape(k, 3) * (hi - lo) + lo
        v = c[:, None, :] + rng.uniform(9 * k, -offset, offset).reshape(k, 3, 3)
        e1, e2 = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
        cr = np.cross(e1, e2)
        ok = (cr * cr).sum(1) > np.float32(1.1920929e-7)
        out = np.concatenate([out, v[ok].astype(np.float32)])
    return out[:n]


def bvh_stress(n_triangles=2048, width=320, height=180, samples=4, max_bounces=4):
    """Workload 2: n random triangles (Lambertian albedo U[0.2, 0.8]^3) in a 4x4x4 cube plus
    one emissive sphere (r = 0.5, radiance 10) above the cloud, black background."""
    rng = Pcg32(workload_seed(2) ^ (n_triangles << 20))
    v = random_triangles(rng, n_triangles)
    alb = rng.uniform(3 * n_triangles, 0.2, 0.8).reshape(-1, 3)
    mats = np.zeros(n_triangles + 1, MATERIAL_IN_DTYPE)
    mats["kind"][:n_triangles] = A.TH_MAT_DIFFUSE
    mats["color"][:n_triangles] = alb
    mats[n_triangles] = emissive((1, 1, 1), 10.0)
    tris = np.zeros(n_triangles, TRIANGLE_IN_DTYPE)
    tris["v"] = v
    tris["material_index"] = np.arange(n_triangles)
    sph = np.zeros(1, SPHERE_IN_DTYPE)
    sph[0] = ((0.0, 3.0, 0.0), 0.5, n_triangles)
    return Scene(camera_position=(0.0, 5.5, 8.0), camera_look_at=(0.0, 0.0, 0.0), fov_degrees=45.0,
                 width=width, height=height, samples=samples, max_bounces=max_bounces,
                 background=A.TH_BG_BLACK, materials=mats, spheres=sph, triangles=tris,
                 name=f"bvh_stress_{n_triangles}")


def sky_environment(width=1024, height=512, sun_elevation=45.0, sun_azimuth=30.0,
                    sun_radius_deg=0.5, sun_radiance=5.0e4):
    """Synthetic equirect HDR: grey sky 0.5 + 1.0 * (1 - (y + 1/2) / H) plus a sun disk
    (texel centres within sun_radius_deg of the sun direction) of the given radiance.
    Texel (x, y) centre maps to theta = pi (y + 1/2) / H, phi = 2 pi ((x + 1/2) / W - 1/2)
    (environment.rs:176-181 with rotation 0)."""
    y = (np.arange(height, dtype=np.float64) + 0.5) / height
    sky = 0.5 + 1.0 * (1.0 - y)
    rgb = np.repeat(np.repeat(sky[:, None, None], width, 1), 3, 2)
    theta = np.pi * y[:, None]
    phi = 2.0 * np.pi * ((np.arange(width) + 0.5) / width - 0.5)[None, :]
    d = np.stack([np.sin(theta) * np.cos(phi), np.cos(theta) * np.ones_like(phi),
                  np.sin(theta) * np.sin(phi)], -1)
    el, az = np.radians(sun_elevation), np.radians(sun_azimuth)
    s = np.array([np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)])
    ang = np.degrees(np.arccos(np.clip(d @ s, -1, 1)))
    rgb[ang <= sun_radius_deg] = sun_radiance
    return rgb.astype(np.float32)


def mixed_environment(width=1280, height=720, samples=64, max_bounces=8, n_triangles=2048,
                      n_spheres=128, env_size=(1024, 512)):
    """Workload 3: random triangles + log-uniform spheres on a 10x10 ground quad, materials
    diffuse/metal/dielectric 50/30/20, lit by the synthetic sky + sun HDR."""
    rng = Pcg32(workload_seed(3))
    v = random_triangles(rng, n_triangles, lo=(-2, 0, -2), hi=(2, 4, 2))
    r = np.exp(rng.uniform(n_spheres, np.log(0.05), np.log(0.4)))
    cx = rng.uniform(n_spheres, -2, 2)
    cy = r + rng.uniform(n_spheres) * (4.0 - r)
    cz = rng.uniform(n_spheres, -2, 2)
    n_obj = n_triangles + n_spheres + 2
    pick = rng.uniform(n_obj)
    alb = rng.uniform(3 * n_obj, 0.2, 0.8).reshape(-1, 3)
    malb = rng.uniform(3 * n_obj, 0.5, 0.95).reshape(-1, 3)
    rough = rng.uniform(n_obj, 0.0, 0.5)
    rows = []
    for i in range(n_obj):
        if pick[i] < 0.5:
            rows.append(diffuse(alb[i]))
        elif pick[i] < 0.8:
            rows.append(metal(malb[i], float(rough[i])))
        else:
            rows.append(dielectric(1.5))
    mats = _materials(rows)
    g = 5.0
    ground = _quad((-g, 0, g), (g, 0, g), (g, 0, -g), (-g, 0, -g))
    verts = np.concatenate([v, np.asarray(ground, np.float32)])
    tris = np.zeros(len(verts), TRIANGLE_IN_DTYPE)
    tris["v"] = verts
    tris["material_index"] = np.concatenate([np.arange(n_triangles), [n_obj - 2, n_obj - 1]])
    mats[n_obj - 1] = mats[n_obj - 2]
    sph = np.zeros(n_spheres, SPHERE_IN_DTYPE)
    sph["center"] = np.stack([cx, cy, cz], 1)
    sph["radius"] = r
    sph["material_index"] = n_triangles + np.arange(n_spheres)
    env = EnvironmentMap(sky_environment(*env_size), 1.0, 0.0)
    return Scene(camera_position=(0.0, 3.0, 9.0), camera_look_at=(0.0, 1.2, 0.0), fov_degrees=40.0,
                 width=width, height=height, samples=samples, max_bounces=max_bounces,
                 background=A.TH_BG_ENVIRONMENT, materials=mats, spheres=sph, triangles=tris,
                 environment=env, name="mixed_environment")


def _hash01(ix, iz, octave, seed):
    """Lattice value in [-1, 1) from a 64-bit mix of (ix, iz, octave, seed)."""
    with np.errstate(over="ignore"):
        h = (ix.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
             ^ iz.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)
             ^ np.uint64((octave * 0x165667B19E3779F9 + seed) & 0xFFFFFFFFFFFFFFFF))
        h ^= h >> np.uint64(33)
        h *= np.uint64(0xFF51AFD7ED558CCD)
        h ^= h >> np.uint64(33)
        h *= np.uint64(0xC4CEB9FE1A85EC53)
        h ^= h >> np.uint64(33)
    return (h >> np.uint64(40)).astype(np.float64) * (2.0 / (1 << 24)) - 1.0


def fbm(x, z, seed, octaves=6, lacunarity=2.0, gain=0.5, frequency=0.75):
    """Fractional Brownian motion of value noise (smoothstep-interpolated lattice values),
    normalised by the summed octave amplitudes to [-1, 1]."""
    x, z = np.asarray(x, np.float64), np.asarray(z, np.float64)
    total, amp, norm, f = np.zeros(np.broadcast(x, z).shape), 1.0, 0.0, frequency
    for o in range(octaves):
        X, Z = x * f, z * f
        ix, iz = np.floor(X), np.floor(Z)
        tx, tz = X - ix, Z - iz
        sx, sz = tx * tx * (3 - 2 * tx), tz * tz * (3 - 2 * tz)
        ix, iz = ix.astype(np.int64), iz.astype(np.int64)
        v00, v10 = _hash01(ix, iz, o, seed), _hash01(ix + 1, iz, o, seed)
        v01, v11 = _hash01(ix, iz + 1, o, seed), _hash01(ix + 1, iz + 1, o, seed)
        v = (v00 * (1 - sx) + v10 * sx) * (1 - sz) + (v01 * (1 - sx) + v11 * sx) * sz
        total += amp * v
        norm += amp
        amp *= gain
        f *= lacunarity
    return total / norm


def terrain(width=1280, height=720, samples=256, max_bounces=8, cells=(512, 256), env_size=(1024, 512),
            amplitude=0.6):
    """Workload 5 (indirect-dominated light for path guiding): a height field of
    cells[0] x cells[1] cells (two triangles each; 262,144 at 512 x 256) over x in [-4, 4],
    z in [-2, 2], height = amplitude * 6-octave value-noise fBm (lacunarity 2, gain 0.5).
    A roof quad at y = 1.6 covers the field with a 0.5 margin and a wall closes the back
    (z = -2.5); the other three sides are open.  The synthetic sky + sun HDR has its sun
    at elevation 60 degrees on the closed side, so the roof blocks direct sun and the
    field is lit by sky light through the openings and inter-reflection."""
    nx, nz = cells
    xs = np.linspace(-4.0, 4.0, nx + 1)
    zs = np.linspace(-2.0, 2.0, nz + 1)
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    Y = amplitude * fbm(X, Z, seed=workload_seed(5))
    P = np.stack([X, Y, Z], -1).astype(np.float32)  # [nx+1, nz+1, 3]
    v00, v10, v01, v11 = P[:-1, :-1], P[1:, :-1], P[:-1, 1:], P[1:, 1:]
    t0 = np.stack([v00, v01, v11], -2).reshape(-1, 3, 3)
    t1 = np.stack([v00, v11, v10], -2).reshape(-1, 3, 3)
    field = np.stack([t0, t1], 1).reshape(-1, 3, 3)
    m, yr, zb = 4.5, 1.6, -2.5
    roof = _quad((-m, yr, zb), (m, yr, zb), (m, yr, -zb), (-m, yr, -zb))
    wall = _quad((-m, -1.0, zb), (m, -1.0, zb), (m, yr, zb), (-m, yr, zb))
    extra = np.asarray(roof + wall, np.float32)
    tris = np.zeros(len(field) + 4, TRIANGLE_IN_DTYPE)
    tris["v"] = np.concatenate([field, extra])
    tris["material_index"] = np.concatenate([np.zeros(len(field), np.int64), [1, 1, 2, 2]])
    mats = _materials([diffuse((0.55, 0.5, 0.45)), diffuse((0.7, 0.7, 0.7)), diffuse((0.6, 0.6, 0.6))])
    env = EnvironmentMap(sky_environment(*env_size, sun_elevation=60.0, sun_azimuth=-90.0), 1.0, 0.0)
    return Scene(camera_position=(0.0, 1.2, 2.2), camera_look_at=(0.0, -0.1, -1.0), fov_degrees=60.0,
                 width=width, height=height, samples=samples, max_bounces=max_bounces,
                 background=A.TH_BG_ENVIRONMENT, materials=mats, spheres=np.zeros(0, SPHERE_IN_DTYPE),
                 triangles=tris, environment=env, name=f"terrain_{2 * nx * nz}")


def no_attributes(n):
    return np.zeros(n, ATTR_IN_DTYPE)


def random_rotation(rng, max_degrees):
    """Unit quaternion (w, x, y, z): axis uniform on the sphere, angle U[0, max_degrees]."""
    u = rng.uniform(3).astype(np.float64)
    z = 2.0 * u[0] - 1.0
    phi = 2.0 * np.pi * u[1]
    r = np.sqrt(max(0.0, 1.0 - z * z))
    axis = np.array([r * np.cos(phi), r * np.sin(phi), z])
    ang = np.radians(max_degrees * u[2])
    return np.concatenate([[np.cos(ang / 2)], np.sin(ang / 2) * axis])


def quat_mul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def quat_matrix(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def rigid_pose(scene, rotation, translation, pivot=(0.0, 0.0, 0.0)):
    """Absolute rigid pose of the whole primitive set (transform.rs:17-82 semantics:
    v' = translation + R (v - pivot); sphere centres move the same way)."""
    R = quat_matrix(rotation)
    pv = np.asarray(pivot, np.float64)
    t = np.asarray(translation, np.float64)
    out = Scene(**{k: getattr(scene, k) for k in scene.__dataclass_fields__})
    if len(scene.triangles):
        tri = scene.triangles.copy()
        v = tri["v"].astype(np.float64) - pv
        tri["v"] = (v @ R.T + pv + t).astype(np.float32)
        out.triangles = tri
    if len(scene.spheres):
        sph = scene.spheres.copy()
        c = sph["center"].astype(np.float64) - pv
        sph["center"] = (c @ R.T + pv + t).astype(np.float32)
        out.spheres = sph
    return out


def batched_workload(n_scenes=16, n_frames=4, samples=16, max_bounces=6, seed=4,
                     tri_range=(128, 262144), resolutions=((320, 180), (640, 360), (1280, 720))):
    """Workload 4: scenes of log-uniform triangle count (random-triangle generator of
    workload 2), 0-256 extra spheres (diffuse/metal/dielectric), resolution drawn from
    `resolutions`, each with `n_frames` animation frames of a rigid pose composed per
    frame (rotation <= 10 degrees, translation within a 0.2 ball).  Returns a list of
    scenes, each a list of frames."""
    rng = Pcg32(workload_seed(seed))
    out = []
    for s in range(n_scenes):
        nt = int(round(np.exp(rng.uniform(1, np.log(tri_range[0]), np.log(tri_range[1]))[0])))
        ns = int(rng.u32(1)[0] % 257)
        res = resolutions[int(rng.u32(1)[0] % len(resolutions))]
        base = bvh_stress(nt, width=res[0], height=res[1], samples=samples, max_bounces=max_bounces)
        if ns:
            extra = np.zeros(ns, SPHERE_IN_DTYPE)
            extra["center"] = rng.uniform(3 * ns, -2.0, 2.0).reshape(ns, 3)
            extra["radius"] = np.exp(rng.uniform(ns, np.log(0.05), np.log(0.3)))
            pick = rng.uniform(ns)
            mats = np.zeros(ns, MATERIAL_IN_DTYPE)
            alb = rng.uniform(3 * ns, 0.2, 0.9).reshape(ns, 3)
            rough = rng.uniform(ns, 0.0, 0.5)
            for i in range(ns):
                if pick[i] < 0.5:
                    mats[i] = diffuse(alb[i])
                elif pick[i] < 0.8:
                    mats[i] = metal(alb[i], float(rough[i]))
                else:
                    mats[i] = dielectric(1.5)
            extra["material_index"] = len(base.materials) + np.arange(ns)
            base.materials = np.concatenate([base.materials, mats])
            base.spheres = np.concatenate([base.spheres, extra])
        base.name = f"batch{s}_{nt}t_{ns}s_{res[0]}x{res[1]}"
        frames = [base]
        q = np.array([1.0, 0.0, 0.0, 0.0])
        t = np.zeros(3)
        for f in range(1, n_frames):
            q = quat_mul(random_rotation(rng, 10.0), q)
            while True:
                d = rng.uniform(3, -0.2, 0.2).astype(np.float64)
                if d @ d <= 0.04:
                    break
            t = t + d
            fr = rigid_pose(base, q, t)
            fr.name = f"{base.name}_f{f}"
            frames.append(fr)
        out.append(frames)
    return out
