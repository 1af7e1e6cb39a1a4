"""Builds libtoaster_hip.so in-tree (hipcc, gfx950) from toaster_amd/csrc.

Device code is compiled with floating-point contraction off, correctly rounded
binary32 division and square root, and packed-FP32 instruction selection disabled,
so the kernels perform exactly the IEEE operations written in the source.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtoaster_hip.so")
ROOT = os.path.dirname(HERE)
EXAMPLE_SRC = os.path.join(HERE, "examples", "cornell_c.cpp")
EXAMPLE_BIN = os.path.join(HERE, "examples", "cornell_c")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
              "-Wno-unused-function"]

HIP_SOURCES = ["th_runtime.hip", "th_wavefront.hip", "th_batch.hip", "th_guide.hip", "th_guided.hip",
               "th_peaks.hip"]
# Per-source extra flags.  The network kernels keep MFMA results in ordinary VGPRs (no
# accumulation-register copies: one v_accvgpr_read per result element otherwise).
HIP_SOURCE_FLAGS = {"th_guide.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}
HOST_SOURCES = ["th_pack.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r


def build(verbose=False, defines=(), lib_path=LIB, jobs=None):
    """Compiles every HIP and host source (in parallel) and links libtoaster_hip.so.
    `defines` adds -D options to every compile (alternative builds go to `lib_path`)."""
    tag = "".join("_" + d.replace("=", "") for d in defines)
    objdir = os.path.join(HERE, "_build" + tag)
    os.makedirs(objdir, exist_ok=True)
    dflags = ["-D" + d for d in defines]
    cmds, objs = [], []
    for src in HOST_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        cmds.append(["g++"] + HOST_FLAGS + dflags + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        cmds.append([HIPCC] + HIP_FLAGS + HIP_SOURCE_FLAGS.get(src, []) + dflags +
                    ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    workers = jobs or min(len(cmds), max(1, (os.cpu_count() or 2) // 2), 8)
    with ThreadPoolExecutor(max_workers=workers) as pool:
        list(pool.map(_run, cmds))
    tmp = lib_path + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-lpthread"])
    os.replace(tmp, lib_path)
    if lib_path == LIB:  # the C caller example links the in-tree library
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), EXAMPLE_SRC,
              "-L", HERE, "-ltoaster_hip", "-Wl,-rpath,$ORIGIN/..", "-o", EXAMPLE_BIN])
    if verbose:
        print("built", lib_path)
    return lib_path


if __name__ == "__main__":
    build(verbose=True)
