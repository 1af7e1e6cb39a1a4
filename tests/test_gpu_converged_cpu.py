"""The HIP path against the oracle's toaster_cpu mode (the CPU renderer's estimator,
restated on its own generator, traversal and sampling routines) at convergence:
relRMSE = sqrt(mean (I - R)^2 / (R^2 + 0.01)) <= 1e-2 between the GPU image at 65,536 spp
and the toaster_cpu image at 16,384 spp, on workload 1 and on workload 3 without metal
(whose fuzz differs between the reference's two renderers by design) and without its sun
disk (tests/test_oracle_cpu_mode.py says why)."""
import numpy as np
import pytest

import oracle
from toaster_amd import GpuScene, RenderOptions, _abi as A, scenes
from toaster_amd.guide import Guide, rel_rmse, render_guided

from .test_oracle_cpu_mode import workload3_without_metal_or_sun

pytestmark = pytest.mark.gpu

CASES = {
    "cornell": lambda: scenes.cornell(32, 18, samples=1),
    "workload3_no_metal_no_sun": lambda: workload3_without_metal_or_sun(32, 18),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("kernel", ["quad_per_sample", "speculative_reference_rng"])
def test_converged_gpu_matches_toaster_cpu(gpu_ctx, name, kernel):
    sc = CASES[name]()
    print(f"case converged vs toaster_cpu {name} kernel={kernel}", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    opts = (RenderOptions(rng_mode=A.TH_RNG_PER_SAMPLE, traversal=A.TH_TRAVERSE_WIDE) if kernel == "quad_per_sample"
            else RenderOptions(traversal=A.TH_TRAVERSE_WIDE, schedule=A.TH_SCHEDULE_SPECULATIVE))
    for c in range(16):
        gs.render(samples=4096, frame_index=900 + c, accumulated_samples=4096 * c, options=opts)
    img = gs.readback().copy()
    gs.close()
    oc = oracle.OracleScene(sc)
    p = oc.params
    p.samples = 16384
    ref = oc.render_cpu(params=p, threads=16)
    err = rel_rmse(img, ref)
    print(f"  relRMSE vs toaster_cpu {err:.5f}")
    assert err <= 1e-2


def test_converged_guided_matches_toaster_cpu(gpu_ctx):
    """The guided tracer (alpha 0.8, trained online, environment light by the
    continuation except at a path's last vertex) at 8,192 spp against toaster_cpu on
    workload 3 without metal or sun."""
    sc = CASES["workload3_no_metal_no_sun"]()
    print("case converged guided vs toaster_cpu", flush=True)
    gs = GpuScene.from_scene(gpu_ctx, sc)
    g = Guide(gpu_ctx, seed=3)
    gs.frame_index = 40
    render_guided(gs, g, 8192, alpha=0.8, env_nee=False, batch=1 << 12, warmup_steps=64, warmup_rate=4,
                  alpha_ramp=64, progressive=False)
    img = gs.readback().copy()
    gs.close()
    g.close()
    oc = oracle.OracleScene(sc)
    p = oc.params
    p.samples = 16384
    ref = oc.render_cpu(params=p, threads=16)
    err = rel_rmse(img, ref)
    print(f"  guided relRMSE vs toaster_cpu {err:.5f}")
    assert err <= 1e-2
