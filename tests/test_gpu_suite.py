"""The reference's own benchmark suite (scenes/benchmarks/001-007, loaded from the copies
in tests/golden/reference_benchmarks) on the GPU against the oracle, full size (320x180,
4 spp, 4 bounces, frame seed 0), with the image tolerance of test_gpu_parity."""
import numpy as np
import pytest

import oracle
from toaster_amd import GpuScene, RenderOptions, _abi as A
from toaster_amd.loader import load_scene, suite_paths

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("traversal,schedule", [(A.TH_TRAVERSE_REFERENCE, A.TH_SCHEDULE_STATIC),
                                                (A.TH_TRAVERSE_WIDE, A.TH_SCHEDULE_STATIC),
                                                (A.TH_TRAVERSE_WIDE, A.TH_SCHEDULE_OCTET),
                                                (A.TH_TRAVERSE_WIDE, A.TH_SCHEDULE_SPECULATIVE)])
@pytest.mark.parametrize("index", range(7))
def test_suite_scene_matches_oracle(gpu_ctx, index, traversal, schedule):
    sc = load_scene(suite_paths()[index])
    print(f"case suite {sc.name} traversal={traversal} schedule={schedule}", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    gs.render(options=RenderOptions(traversal=traversal, schedule=schedule), frame_index=0,
              accumulated_samples=0)
    img = gs.readback()
    gs.close()
    ref, _ = oracle.OracleScene(sc).render(threads=8)
    close = np.abs(img - ref) <= 1e-4 * (1.0 + np.abs(ref))
    frac = float(np.mean(np.all(close, axis=-1)))
    print(f"  {sc.name}: within tol {frac:.5f}")
    assert frac >= 0.995
