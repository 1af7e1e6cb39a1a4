"""JSON scene files and the glTF / Radiance HDR assets they reference, as far as the
reference's benchmark suite (scenes/benchmarks/001-007) uses them.

A subset of toaster-scene's loader (crates/toaster-scene/src/loader.rs:534-971):
camera, render settings, black / sky / environment background, named diffuse, metal,
dielectric and emissive materials, and sphere, triangle and mesh objects appended in
file order.  Mesh objects import a glTF file (embedded or external buffers) as
crates/toaster-assets/src/gltf_loader.rs:11-200 does: every triangle primitive of the
default scene's node tree, positions through the world transform, normals through its
inverse transpose and normalised, TEXCOORD_0, indices (or 0..n), and a material override
from the object (loader.rs:744-862).  Radiance .hdr files (flat or run-length RGBE)
decode to linear RGB as m * 2^(e - 136) per channel.
"""
import base64
import json
import os
import struct

import numpy as np

from . import _abi as A
from .scene import (ATTR_IN_DTYPE, MATERIAL_IN_DTYPE, SPHERE_IN_DTYPE, TRIANGLE_IN_DTYPE,
                     EnvironmentMap, Scene, dielectric, diffuse, emissive, metal)

_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32,
              5126: np.float32}
_WIDTH = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT4": 16}


# ------------------------------------------------------------------ glTF
def _buffers(doc, base):
    out = []
    for b in doc.get("buffers", []):
        uri = b.get("uri")
        if uri is None:
            raise ValueError("GLB binary chunks are not supported")
        if uri.startswith("data:"):
            out.append(base64.b64decode(uri.split(",", 1)[1]))
        else:
            with open(os.path.join(base, uri), "rb") as f:
                out.append(f.read())
    return out


def _accessor(doc, bufs, index):
    acc = doc["accessors"][index]
    view = doc["bufferViews"][acc["bufferView"]]
    dt = np.dtype(_COMPONENT[acc["componentType"]])
    width = _WIDTH[acc["type"]]
    count = acc["count"]
    start = view.get("byteOffset", 0) + acc.get("byteOffset", 0)
    stride = view.get("byteStride", 0) or dt.itemsize * width
    raw = bufs[view["buffer"]]
    if stride == dt.itemsize * width:
        a = np.frombuffer(raw, dt, count * width, start).reshape(count, width)
    else:
        a = np.stack([np.frombuffer(raw, dt, width, start + i * stride) for i in range(count)])
    if acc.get("normalized"):
        raise ValueError("normalized integer attributes are not supported")
    return a


def _node_matrix(node):
    if "matrix" in node:
        return np.asarray(node["matrix"], np.float64).reshape(4, 4).T  # column-major
    t = np.asarray(node.get("translation", [0, 0, 0]), np.float64)
    x, y, z, w = node.get("rotation", [0, 0, 0, 1])
    s = np.asarray(node.get("scale", [1, 1, 1]), np.float64)
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3] = r * s[None, :]
    m[:3, 3] = t
    return m


def load_gltf(path):
    """Triangles of a glTF file: list of (positions [3,3], normals [3,3] or None,
    tex_coords [3,2] or None) in node/primitive/index order."""
    with open(path) as f:
        doc = json.load(f)
    bufs = _buffers(doc, os.path.dirname(os.path.abspath(path)))
    tris = []

    def visit(ni, parent):
        node = doc["nodes"][ni]
        world = parent @ _node_matrix(node)
        if "mesh" in node:
            for prim in doc["meshes"][node["mesh"]]["primitives"]:
                if prim.get("mode", 4) != 4:
                    raise ValueError("only triangle primitives are supported")
                attr = prim["attributes"]
                pos = _accessor(doc, bufs, attr["POSITION"]).astype(np.float32)
                pos = (pos.astype(np.float64) @ world[:3, :3].T + world[:3, 3]).astype(np.float32)
                nrm = None
                if "NORMAL" in attr:
                    n = _accessor(doc, bufs, attr["NORMAL"]).astype(np.float32).astype(np.float64)
                    n = n @ np.linalg.inv(world[:3, :3])  # inverse transpose, row vectors
                    ln = np.linalg.norm(n, axis=1, keepdims=True)
                    if np.any(ln == 0) or not np.all(np.isfinite(n)):
                        raise ValueError("glTF normal is zero or not finite")
                    nrm = (n / ln).astype(np.float32)
                uv = None
                if "TEXCOORD_0" in attr:
                    uv = _accessor(doc, bufs, attr["TEXCOORD_0"]).astype(np.float32)
                if "indices" in prim:
                    idx = _accessor(doc, bufs, prim["indices"]).reshape(-1).astype(np.int64)
                else:
                    idx = np.arange(len(pos))
                if len(idx) % 3 or (len(idx) and idx.max() >= len(pos)):
                    raise ValueError("bad glTF index buffer")
                for t in idx.reshape(-1, 3):
                    tris.append((pos[t], None if nrm is None else nrm[t], None if uv is None else uv[t]))
        for c in node.get("children", []):
            visit(c, world)

    scenes = doc.get("scenes", [])
    roots = scenes[doc.get("scene", 0)]["nodes"] if scenes else range(len(doc.get("nodes", [])))
    for ni in roots:
        visit(ni, np.eye(4))
    if not tris:
        raise ValueError(f"glTF {path} contains no triangle primitives")
    return tris


# ------------------------------------------------------------------ Radiance HDR
def load_hdr(path):
    """(H, W, 3) float32 linear RGB of a Radiance RGBE file."""
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    w = h = None
    while True:
        end = data.index(b"\n", pos)
        line = data[pos:end].decode("ascii", "replace").strip()
        pos = end + 1
        if line.startswith("-Y") or line.startswith("+Y"):
            parts = line.split()
            if parts[0] != "-Y" or parts[2] != "+X":
                raise ValueError("only -Y H +X W orientation is supported")
            h, w = int(parts[1]), int(parts[3])
            break
    rgbe = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        if w >= 8 and w < 32768 and data[pos] == 2 and data[pos + 1] == 2 and (data[pos + 2] & 0x80) == 0:
            if (data[pos + 2] << 8 | data[pos + 3]) != w:
                raise ValueError("bad RLE scanline width")
            pos += 4
            for c in range(4):
                x = 0
                while x < w:
                    n = data[pos]
                    pos += 1
                    if n > 128:
                        n -= 128
                        rgbe[y, x:x + n, c] = data[pos]
                        pos += 1
                    else:
                        rgbe[y, x:x + n, c] = np.frombuffer(data, np.uint8, n, pos)
                        pos += n
                    x += n
        else:
            rgbe[y] = np.frombuffer(data, np.uint8, 4 * w, pos).reshape(w, 4)
            pos += 4 * w
    e = rgbe[..., 3].astype(np.int32)
    scale = np.where(e == 0, 0.0, np.ldexp(1.0, e - 136))
    return (rgbe[..., :3].astype(np.float64) * scale[..., None]).astype(np.float32)


# ------------------------------------------------------------------ scene JSON
def _material(m):
    t = m["type"]
    if t == "diffuse":
        return diffuse(m["albedo"])
    if t == "metal":
        return metal(m["albedo"], float(m.get("roughness", 0.0)))
    if t == "dielectric":
        return dielectric(float(m["ior"]))
    if t == "emissive":
        return emissive(m["color"], float(m["strength"]))
    raise ValueError(f"material type {t!r} is not supported by this loader")


def load_scene(path):
    """Scene of a JSON scene file (paths relative to the file)."""
    base = os.path.dirname(os.path.abspath(path))
    with open(path) as f:
        doc = json.load(f)
    names = {}
    rows = []
    for m in doc.get("materials", []):
        names[m["name"]] = len(rows)
        rows.append(_material(m))
    spheres, tris, attrs = [], [], []
    any_attr = False
    for o in doc["objects"]:
        mat = names[o["material"]] if "material" in o else None
        if o["type"] == "sphere":
            spheres.append((tuple(o["center"]), float(o["radius"]), mat))
        elif o["type"] == "triangle":
            tris.append((np.asarray(o["vertices"], np.float32), mat))
            attrs.append((0, 0, np.zeros((3, 3), np.float32), np.zeros((3, 2), np.float32)))
        elif o["type"] == "mesh":
            if mat is None:
                raise ValueError("mesh objects need a material override in this loader")
            for p, n, uv in load_gltf(os.path.join(base, o["path"])):
                tris.append((p, mat))
                attrs.append((1 if n is not None else 0, 1 if uv is not None else 0,
                              n if n is not None else np.zeros((3, 3), np.float32),
                              uv if uv is not None else np.zeros((3, 2), np.float32)))
                any_attr = any_attr or n is not None or uv is not None
        else:
            raise ValueError(f"object type {o['type']!r} is not supported by this loader")
    mats = np.zeros(len(rows), MATERIAL_IN_DTYPE)
    for i, r in enumerate(rows):
        mats[i] = r
    sph = np.zeros(len(spheres), SPHERE_IN_DTYPE)
    for i, s in enumerate(spheres):
        sph[i] = s
    tri = np.zeros(len(tris), TRIANGLE_IN_DTYPE)
    for i, (v, m) in enumerate(tris):
        tri[i] = (v, m)
    att = None
    if any_attr:
        att = np.zeros(len(tris), ATTR_IN_DTYPE)
        for i, a in enumerate(attrs):
            att[i] = a
    cam, rnd = doc["camera"], doc["render"]
    bg = rnd.get("background", "sky")
    env = None
    if isinstance(bg, dict):
        if bg.get("type") != "environment":
            raise ValueError("unknown background object")
        env = EnvironmentMap(load_hdr(os.path.join(base, bg["path"])), float(bg.get("intensity", 1.0)),
                             float(bg.get("rotation_degrees", 0.0)))
        kind = A.TH_BG_ENVIRONMENT
    else:
        kind = {"sky": A.TH_BG_SKY, "black": A.TH_BG_BLACK}[bg]
    return Scene(camera_position=tuple(cam["position"]), camera_look_at=tuple(cam["look_at"]),
                 camera_up=tuple(cam.get("up", (0.0, 1.0, 0.0))), fov_degrees=float(cam["fov_degrees"]),
                 width=int(rnd["width"]), height=int(rnd["height"]), samples=int(rnd["samples"]),
                 max_bounces=int(rnd["max_bounces"]), background=kind, materials=mats, spheres=sph,
                 triangles=tri, attributes=att, environment=env,
                 name=os.path.splitext(os.path.basename(path))[0])


SUITE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden",
                         "reference_benchmarks", "scenes", "benchmarks")

# Published medians of the reference's suite on an RTX 2080 Ti (Vulkan), post-BVH,
# 320x180, 4 spp, 4 bounces, dispatch+wait ms (README.md:135-141,
# docs/benchmarks/post-bvh/rtx-2080-ti/*.json summary.dispatch_wait_ms.median).
PUBLISHED_DISPATCH_WAIT_MS = {
    "001_triangles_128": 0.597, "002_triangles_2048": 0.968, "003_triangles_8192": 1.457,
    "004_spheres_64": 1.129, "005_spheres_512": 1.752, "006_mixed_2048t_128s": 2.478,
    "007_environment_control": 0.671,
}


def suite_paths():
    return [os.path.join(SUITE_DIR, n + ".json") for n in sorted(PUBLISHED_DISPATCH_WAIT_MS)]
