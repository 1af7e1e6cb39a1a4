"""Measured device ceilings (th_measure_copy, th_measure_mfma): the diagnostics bench.py
reports beside the published MI355X peaks.

- Every MFMA product of the all-ones operands is exact (the accumulators hold K x rounds),
  for both shapes the network kernels use, so the timed instructions really ran.
- The rates lie between a floor far below what the device reaches and the published
  peak (+5 %): a copy moves at most 8 TB/s, dense BF16 MFMA at most 2.5 PFLOP/s.
- Invalid sizes and shapes are rejected before any launch."""
import pytest

from toaster_amd import _abi as A

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("read_only", [False, True])
def test_streaming_bandwidth(gpu_ctx, read_only):
    gbs = gpu_ctx.measure_copy(256 << 20, 4, read_only=read_only)
    print(f"case streaming {'read' if read_only else 'copy'} 256 MiB: {gbs:.1f} GB/s", flush=True)
    assert 1000.0 < gbs < 8000.0 * 1.05


@pytest.mark.parametrize("shape", [0, 1])
def test_mfma_throughput_and_exactness(gpu_ctx, shape):
    for waves in (1, 4):
        tf, exact = gpu_ctx.measure_mfma(shape, waves, 2048)
        print(f"case mfma shape {shape} waves/SIMD {waves}: {tf:.1f} TFLOP/s exact {exact}", flush=True)
        assert exact
        assert 50.0 < tf < 2500.0 * 1.05


def test_invalid_arguments(gpu_ctx):
    with pytest.raises(A.ToasterError):
        gpu_ctx.measure_copy(1000, 1)  # not a multiple of 16
    with pytest.raises(A.ToasterError):
        gpu_ctx.measure_copy(1 << 20, 0)
    with pytest.raises(A.ToasterError):
        gpu_ctx.measure_mfma(2, 1, 16)
    with pytest.raises(A.ToasterError):
        gpu_ctx.measure_mfma(1, 9, 16)
    with pytest.raises(A.ToasterError):
        gpu_ctx.measure_mfma(1, 1, 1 << 20)
