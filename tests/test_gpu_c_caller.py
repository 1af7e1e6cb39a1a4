"""The C caller example (toaster_amd/examples/cornell_c.cpp, include/toaster_hip.h only)
produces the images of the same renders made through the Python mirror, byte for byte:
the Cornell box, then the same box with its spheres moved, refitted on the device and
rendered with the automatic schedule."""
import os
import subprocess

import numpy as np
import pytest

from toaster_amd import GpuScene, RenderOptions, _abi as A, build, scenes
from toaster_amd.renderer import display_rgba8

pytestmark = pytest.mark.gpu


def test_c_caller_matches_python_path(gpu_ctx, tmp_path):
    if not os.path.exists(build.EXAMPLE_BIN):
        pytest.fail("the C caller example is not built (run __graft_entry__.build())")
    out = tmp_path / "cornell.ppm"
    moved_out = tmp_path / "cornell_moved.ppm"
    print("case C caller cornell 320x180 4 spp", flush=True)
    r = subprocess.run([build.EXAMPLE_BIN, str(out), str(moved_out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip())
    data = out.read_bytes()
    header = b"P6\n320 180\n255\n"
    assert data.startswith(header)
    rgb_c = np.frombuffer(data[len(header):], np.uint8).reshape(180, 320, 3)
    sc = scenes.cornell()
    gs = GpuScene.from_scene(gpu_ctx, sc)
    gs.render(frame_index=0, accumulated_samples=0,
              options=RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=A.TH_SCHEDULE_OCTET))
    rgba = display_rgba8(gs.readback())
    assert np.array_equal(rgb_c, rgba[..., :3])
    # the moved frame: spheres shifted, refitted, rendered with the automatic schedule
    moved = scenes.cornell()
    moved.spheres["center"][0, 0] = moved.spheres["center"][0, 0] + np.float32(0.1)
    moved.spheres["center"][1, 1] = moved.spheres["center"][1, 1] + np.float32(0.1)
    gs.refit(moved)
    gs.render(frame_index=0, accumulated_samples=0,
              options=RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=A.TH_SCHEDULE_AUTO))
    rgba_moved = display_rgba8(gs.readback())
    gs.close()
    data = moved_out.read_bytes()
    assert data.startswith(header)
    rgb_moved = np.frombuffer(data[len(header):], np.uint8).reshape(180, 320, 3)
    assert not np.array_equal(rgb_moved, rgb_c)
    assert np.array_equal(rgb_moved, rgba_moved[..., :3])
