"""The device generator the render kernels use (th_rng_draws) against the known answers
computed from the reference shader's text (tests/golden/pcg_vectors.npz): pixel seeds
(wgsl:519), states and random_f32 draws (wgsl:189-199), bit for bit."""
import ctypes as C
import os

import numpy as np
import pytest

from toaster_amd import _abi as A

pytestmark = pytest.mark.gpu

V = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pcg_vectors.npz"))


def _draws(ctx, pixels, frame, mode=A.TH_RNG_REFERENCE, sample=0, n=16):
    px = np.ascontiguousarray(pixels, np.uint32)
    st = np.empty((len(px), n), np.uint32)
    va = np.empty((len(px), n), np.float32)
    A.check(A.lib().th_rng_draws(ctx.handle, px.ctypes.data, len(px), int(frame), mode, sample, n,
                                 st.ctypes.data, va.ctypes.data))
    return st, va


def test_device_streams_match_the_shader(gpu_ctx):
    print("case device pcg streams", flush=True)
    for fi, f in enumerate(V["frames"]):
        st, va = _draws(gpu_ctx, V["pixels"], f)
        assert np.array_equal(st, V["states"][fi])
        assert np.array_equal(va.view(np.uint32), V["values"][fi].view(np.uint32))


def test_per_sample_stream_zero_is_the_reference_stream(gpu_ctx):
    st0, va0 = _draws(gpu_ctx, V["pixels"], 7, A.TH_RNG_PER_SAMPLE, 0)
    ref, _ = _draws(gpu_ctx, V["pixels"], 7)
    assert np.array_equal(st0, ref)
    st1, _ = _draws(gpu_ctx, V["pixels"], 7, A.TH_RNG_PER_SAMPLE, 1)
    assert not np.array_equal(st1, ref)


def test_query_validation(gpu_ctx):
    px = np.zeros(4, np.uint32)
    st = np.empty(4, np.uint32)
    va = np.empty(4, np.float32)
    L = A.lib()
    assert L.th_rng_draws(gpu_ctx.handle, px.ctypes.data, 4, 0, 2, 0, 1, st.ctypes.data,
                          va.ctypes.data) == A.TH_E_INVALID
    assert L.th_rng_draws(gpu_ctx.handle, px.ctypes.data, 0, 0, 0, 0, 1, st.ctypes.data,
                          va.ctypes.data) == A.TH_E_INVALID
    assert L.th_rng_draws(gpu_ctx.handle, None, 4, 0, 0, 0, 1, st.ctypes.data,
                          va.ctypes.data) == A.TH_E_INVALID
