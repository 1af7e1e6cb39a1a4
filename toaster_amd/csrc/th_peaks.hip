// th_peaks.hip -- measured ceilings of the device for the roofline figures: a streaming
// copy (HBM read + write bandwidth) and back-to-back BF16 MFMAs of the two shapes the
// network kernels use (dense matrix-core throughput).  Diagnostics beside the published
// peaks (8 TB/s HBM, 2.5 PFLOP/s dense BF16): the renderer does not call them.
#include <algorithm>

#include "th_internal.h"

typedef __bf16 pk_bf16x8 __attribute__((ext_vector_type(8)));
typedef float pk_f32x4 __attribute__((ext_vector_type(4)));
typedef float pk_f32x16 __attribute__((ext_vector_type(16)));

namespace {

// Streaming copy (MODE 0) or read (MODE 1) of n 16-byte words: block b takes words
// [b chunk, (b + 1) chunk), four coalesced 16-byte loads in flight per thread before their
// stores; the read XORs what it loads into one word per thread (so no load is dropped).
// Sweeps on MI355X: contiguous 64 KiB chunks per block copy at 5.3 TB/s (grid-stride
// loops 4.4-5.1, hipMemcpyAsync 4.8), 256 KiB chunks read at 6.0 TB/s.
template <int MODE>
__global__ void __launch_bounds__(256) k_stream(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                uint32_t* __restrict__ sink, uint64_t n, uint64_t chunk) {
  const uint64_t b0 = (uint64_t)blockIdx.x * chunk, b1 = b0 + chunk < n ? b0 + chunk : n;
  uint32_t x = 0u;
  for (uint64_t i = b0 + threadIdx.x; i < b1; i += 4u * 256u) {
    uint4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = i + u * 256u < b1 ? src[i + u * 256u] : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (MODE == 0) {
        if (i + u * 256u < b1) dst[i + u * 256u] = v[u];
      } else {
        x ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
      }
    }
  }
  if (MODE == 1) sink[(uint64_t)blockIdx.x * 256u + threadIdx.x] = x;
}

// `iters` rounds of four independent MFMAs per wave on all-ones BF16 operands (the four
// accumulator chains hide each instruction's latency).  Every accumulator element ends at
// K * iters exactly (K = 32 or 16 ones per product, sums below 2^24), and each lane
// writes 1 when all of its elements do, 0 otherwise.
template <int SHAPE>  // 0: v_mfma_f32_16x16x32_bf16, 1: v_mfma_f32_32x32x16_bf16
__global__ void __launch_bounds__(256) k_mfma_peak(uint32_t iters, uint32_t* __restrict__ ok) {
  uint4 ones = make_uint4(0x3f803f80u, 0x3f803f80u, 0x3f803f80u, 0x3f803f80u);  // BF16 1.0 pairs
  __asm__ volatile("" : "+v"(ones.x), "+v"(ones.y), "+v"(ones.z), "+v"(ones.w));
  const pk_bf16x8 a = __builtin_bit_cast(pk_bf16x8, ones);
  bool good = true;
  if (SHAPE == 0) {
    pk_f32x4 c0 = {}, c1 = {}, c2 = {}, c3 = {};
    for (uint32_t it = 0; it < iters; ++it) {
      c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, a, c0, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, a, c1, 0, 0, 0);
      c2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, a, c2, 0, 0, 0);
      c3 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, a, c3, 0, 0, 0);
    }
    const float want = 32.0f * (float)iters;
    for (int r = 0; r < 4; ++r) good = good && c0[r] == want && c1[r] == want && c2[r] == want && c3[r] == want;
  } else {
    pk_f32x16 c0 = {}, c1 = {}, c2 = {}, c3 = {};
    for (uint32_t it = 0; it < iters; ++it) {
      c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, a, c0, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, a, c1, 0, 0, 0);
      c2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, a, c2, 0, 0, 0);
      c3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, a, c3, 0, 0, 0);
    }
    const float want = 16.0f * (float)iters;
    for (int r = 0; r < 16; ++r) good = good && c0[r] == want && c1[r] == want && c2[r] == want && c3[r] == want;
  }
  ok[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = good ? 1u : 0u;
}

// Device milliseconds of one launch made by `launch`, after one untimed launch.
template <class F>
int32_t time_launch(th_context* ctx, F&& launch, float* ms) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) {
    launch();
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e0, ctx->stream);
  if (e == hipSuccess) {
    launch();
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (e != hipSuccess) return th_set_error(TH_E_DEVICE, std::string("timed launch: ") + hipGetErrorString(e));
  return TH_OK;
}

}  // namespace

extern "C" {

int32_t th_measure_copy(th_context* ctx, uint64_t bytes, uint32_t reps, uint32_t read_only, double* gbs) {
  if (!ctx || !gbs) return th_set_error(TH_E_INVALID, "null argument");
  if (bytes < 16u || (bytes & 15u) || reps == 0u)
    return th_set_error(TH_E_INVALID, "copy size must be a nonzero multiple of 16 bytes, reps nonzero");
  TH_CHECK(hipSetDevice(ctx->device));
  const uint64_t n = bytes / 16u;
  const uint64_t chunk = read_only ? 16384u : 4096u;  // words per block
  const uint64_t blocks = (n + chunk - 1u) / chunk;
  if (blocks > 0x7fffffffu) return th_set_error(TH_E_INVALID, "copy size too large");
  uint4 *src = nullptr, *dst = nullptr;
  uint32_t* sink = nullptr;
  hipError_t e = hipMalloc((void**)&src, bytes);
  if (e == hipSuccess) e = read_only ? hipMalloc((void**)&sink, blocks * 256u * sizeof(uint32_t))
                                     : hipMalloc((void**)&dst, bytes);
  if (e == hipSuccess) e = hipMemsetAsync(src, 0x5a, bytes, ctx->stream);
  if (e != hipSuccess) {
    (void)hipFree(src);
    (void)hipFree(dst);
    (void)hipFree(sink);
    return th_set_error(TH_E_DEVICE, std::string("copy buffers: ") + hipGetErrorString(e));
  }
  float ms = 0.0f;
  const int32_t rc = time_launch(ctx, [&] {
    for (uint32_t r = 0; r < reps; ++r) {
      if (read_only)
        hipLaunchKernelGGL(k_stream<1>, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream, src, dst, sink, n, chunk);
      else
        hipLaunchKernelGGL(k_stream<0>, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream, src, dst, sink, n, chunk);
    }
  }, &ms);
  (void)hipFree(src);
  (void)hipFree(dst);
  (void)hipFree(sink);
  if (rc) return rc;
  *gbs = ms > 0.0f ? (read_only ? 1.0 : 2.0) * (double)bytes * reps / (ms * 1e-3) / 1e9 : 0.0;
  return TH_OK;
}

int32_t th_measure_mfma(th_context* ctx, uint32_t shape, uint32_t waves_per_simd, uint32_t iters, double* tflops,
                        uint32_t* exact) {
  if (!ctx || !tflops) return th_set_error(TH_E_INVALID, "null argument");
  if (shape > 1u || waves_per_simd == 0u || waves_per_simd > 8u || iters == 0u || iters > (1u << 19))
    return th_set_error(TH_E_INVALID, "shape 0 or 1, 1..8 waves per SIMD, 1..2^19 rounds");
  TH_CHECK(hipSetDevice(ctx->device));
  // one 256-thread block per SIMD quartet and wave slot: cu_count * waves_per_simd blocks
  const uint32_t blocks = (uint32_t)std::max(1, ctx->cu_count) * waves_per_simd;
  const size_t threads = (size_t)blocks * 256u;
  uint32_t* d_ok = nullptr;
  hipError_t e = hipMalloc((void**)&d_ok, threads * sizeof(uint32_t));
  if (e != hipSuccess) return th_set_error(TH_E_DEVICE, std::string("mfma buffer: ") + hipGetErrorString(e));
  float ms = 0.0f;
  const int32_t rc = time_launch(ctx, [&] {
    if (shape == 0u)
      hipLaunchKernelGGL(k_mfma_peak<0>, dim3(blocks), dim3(256), 0, ctx->stream, iters, d_ok);
    else
      hipLaunchKernelGGL(k_mfma_peak<1>, dim3(blocks), dim3(256), 0, ctx->stream, iters, d_ok);
  }, &ms);
  std::vector<uint32_t> ok(threads, 0u);
  if (!rc) e = hipMemcpy(ok.data(), d_ok, threads * sizeof(uint32_t), hipMemcpyDeviceToHost);
  (void)hipFree(d_ok);
  if (rc) return rc;
  if (e != hipSuccess) return th_set_error(TH_E_DEVICE, std::string("mfma readback: ") + hipGetErrorString(e));
  uint64_t good = 0;
  for (uint32_t v : ok) good += v;
  if (exact) *exact = good == threads ? 1u : 0u;
  // 2 M N K FLOP per product: 16,384 (16x16x32) or 32,768 (32x32x16); four per round
  const double flop_per = shape == 0u ? 2.0 * 16 * 16 * 32 : 2.0 * 32 * 32 * 16;
  const double flops = flop_per * 4.0 * (double)iters * (double)(threads / 64u);
  *tflops = ms > 0.0f ? flops / (ms * 1e-3) / 1e12 : 0.0;
  return TH_OK;
}

}  // extern "C"
