// th_guide.hip -- the guiding network on BF16 MFMA: forward (inference and training),
// vMF-mixture loss gradient, backward (activation and weight gradients) and Adam.
//
// Every dense layer is Y[B x N] = X[B x K] * W^T with X and W in BF16 and FP32
// accumulation in v_mfma_f32_16x16x32_bf16 (gfx950).  Fragment layout of that
// instruction on a 64-lane wave (checked on the device by th_guide_mfma_selftest
// against a host product): lane l = 16 g + r holds A[r][8g..8g+7], B[8g..8g+7][r] and
// C[4g..4g+3][r].  A wave owns 16 rows of a batch; weights are kept in two padded
// BF16 layouts, out-major ([N][K], forward) and in-major ([K][N], backward), so every
// B fragment is one 16-byte load.  Inference (k_guide_infer) and a training step
// (k_train_fused) each run in one kernel with the activations in LDS; weight gradients
// dW[N x K] = dZ^T X reduce over the batch inside the MFMA K dimension (32 samples per
// instruction); each block writes one partial, a fixed-order reduction sums the
// partials, and the Adam kernel updates the FP32 master weights and refreshes both BF16
// copies.
#include <algorithm>
#include <vector>

#include "th_guide.h"
#include "th_internal.h"

typedef __bf16 tg_bf16x8 __attribute__((ext_vector_type(8)));
typedef float tg_f32x4 __attribute__((ext_vector_type(4)));

namespace {

// Every layer is stored padded to 64 x 64 (zero rows/columns beyond the true shape).
// Master parameter geometry: true in/out widths and offsets of W (out-major) and b.
struct LayerDims {
  int kin, nout, woff, boff;
};
__host__ __device__ constexpr LayerDims layer_dims(int l) {
  return l == 0 ? LayerDims{48, 64, 0, 64 * 48}
       : l == 1 ? LayerDims{64, 64, 64 * 48 + 64, 64 * 48 + 64 + 64 * 64}
       : l == 2 ? LayerDims{64, 64, 64 * 48 + 64 + 64 * 64 + 64, 64 * 48 + 64 + 2 * 64 * 64 + 64}
                : LayerDims{64, 40, 64 * 48 + 64 * 3 + 2 * 64 * 64,
                            64 * 48 + 64 * 3 + 2 * 64 * 64 + 40 * 64};
}
static_assert(layer_dims(3).boff + 40 == TG_NPARAMS, "parameter layout");
#define TG_PART (64 * 64 + 64)  // one block's partial gradient of one layer

__device__ __forceinline__ tg_bf16x8 ld8(const uint16_t* p) {
  uint4 v = *(const uint4*)p;  // 16-byte aligned
  return __builtin_bit_cast(tg_bf16x8, v);
}

__device__ __forceinline__ tg_f32x4 mfma(tg_bf16x8 a, tg_bf16x8 b, tg_f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Y = epilogue(X * M + bias) with M given as MT[N][K]; one wave per 16 rows (the
// host-facing forward, th_guide_forward).  MODE 0: ReLU, BF16 out.  MODE 1: linear, FP32 out.
template <int K, int N, int MODE>
__global__ void __launch_bounds__(256) k_gemm_rows(const uint16_t* __restrict__ X,
                                                   const uint16_t* __restrict__ MT,
                                                   const float* __restrict__ bias,
                                                   const uint32_t* __restrict__ count_ptr,
                                                   uint32_t count_static, void* __restrict__ Y) {
  const uint32_t rows = count_ptr ? *count_ptr : count_static;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t row0 = wave * 16u;
  if (row0 >= rows) return;
  const uint32_t l = threadIdx.x & 63u, r = l & 15u, g = l >> 4;
  tg_bf16x8 a[K / 32];
#pragma unroll
  for (int kb = 0; kb < K / 32; ++kb) a[kb] = ld8(X + (size_t)(row0 + r) * K + kb * 32 + g * 8);
#pragma unroll
  for (int nb = 0; nb < N / 16; ++nb) {
    tg_f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < K / 32; ++kb)
      acc = mfma(a[kb], ld8(MT + (size_t)(nb * 16 + r) * K + kb * 32 + g * 8), acc);
    const uint32_t col = nb * 16 + r;
    const float bv = bias ? bias[col] : 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t row = row0 + g * 4 + i;
      float v = acc[i] + bv;
      if (MODE == 0)
        ((uint16_t*)Y)[(size_t)row * N + col] = tg::f2bf(v > 0.0f ? v : 0.0f);
      else
        ((float*)Y)[(size_t)row * N + col] = v;
    }
  }
}

// A training step reads its sample count from device memory, so the renderer can queue
// steps on record counts it has not read back.  n = min(*count, cap) samples; fewer than
// min_count make the step a no-op (no partials, no update, the step counter unchanged).
// `step` is the device's Adam step counter (bias correction), advanced by k_weight_partial
// when the step applies and updates; `stats` (optional) accumulates samples offered,
// samples trained and steps applied.
struct StepCtl {
  const uint32_t* count;
  uint32_t cap, min_count;
  uint32_t* step;
  uint32_t* stats;
};
__device__ __forceinline__ uint32_t step_samples(const StepCtl& c) {
  const uint32_t n = min(*c.count, c.cap);
  return n < c.min_count ? 0u : n;
}
// 64-sample chunks over at most 256 blocks (one partial slot each)
__device__ __forceinline__ uint32_t step_rows_per(uint32_t n) { return ((n + 255u) / 256u + 63u) / 64u * 64u; }
__device__ __forceinline__ uint32_t step_blocks(uint32_t n) {
  return n == 0u ? 0u : (n + step_rows_per(n) - 1u) / step_rows_per(n);
}

// Loss gradient of one batch: L = -sum_b w_b log q_b(dir_b) / sum_b w_b.  The weight sum
// is reduced in a fixed order (256 block partials here, then the same 256-to-1 tree in
// every wave of the training kernel), so gradients are reproducible call to call.
// Block 0 also clears the step's scalars (sums[1] later receives the loss from k_adam),
// advances the step counter and the statistics.
__global__ void __launch_bounds__(256) k_weight_partial(const float* __restrict__ w, StepCtl c, uint32_t update,
                                                        float* __restrict__ part, float* __restrict__ sums) {
  __shared__ float red[256];
  const uint32_t n = step_samples(c);
  if (blockIdx.x == 0 && threadIdx.x < 16u) sums[threadIdx.x] = 0.0f;
  if (blockIdx.x == 0 && threadIdx.x == 0u) {
    if (n > 0u && update) *c.step += 1u;
    if (c.stats) {
      c.stats[0] += min(*c.count, c.cap);
      c.stats[1] += n;
      c.stats[2] += n > 0u ? 1u : 0u;
    }
  }
  float s = 0.0f;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) s += w[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// The batch's total weight from the 256 partials, computed by every wave on its own: the
// fixed-order tree p[t] + p[t + 128], then + [t + 64], + [t + 32], ... down to one value
// (lane t pairs with lane t ^ off, which is t + off where the tree reads it).
__device__ __forceinline__ float weight_total(const float* __restrict__ part) {
  const uint32_t t = threadIdx.x & 63u;
  float v = (part[t] + part[t + 128u]) + (part[t + 64u] + part[t + 192u]);
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return __shfl(v, 0, 64);
}

// Fused training step.  One block of 8 waves walks `rows_per` consecutive samples in
// chunks of 64 with every activation in LDS; two waves per SIMD hide each other's LDS
// and MFMA latency.  Wave w = 4h + v owns the 16 features 16v..16v+15 of every layer and
// half h of the chunk: in the forward it computes those output columns for samples
// 32h..32h+31 (A fragments from the row-major LDS tile, B fragments = its 16 rows of the
// out-major weights, loaded once into registers); the vMF-mixture loss gradient runs on
// 8 lanes per sample, one lobe each; then the wave accumulates rows 16v..16v+15, columns
// 32h..32h+31 of every layer's weight gradient (across chunks, in registers; the MFMA K
// dimension runs over 32 samples) and back-propagates the columns 16v..16v+15, samples
// 32h..32h+31 of the layer below (its 16 rows of the in-major weights, also in registers),
// gated by the ReLU of the layer's input.  Each block writes one partial gradient per
// layer for the fixed-order reduction and Adam.  Each output element is one MFMA chain
// over K = 0..31 then 32..63, whichever wave computes it.
#define TF_LD 72  // LDS row stride (BF16 elements) of a 64-wide activation tile
// The weight gradient's MFMA runs its K dimension over samples, so both its operands are
// read from transposed copies of the tiles ([feature][sample], same stride): one 16-byte
// LDS load per operand instead of eight 2-byte column reads.  The forward and the
// back-propagation write each output tile in both layouts (an MFMA output lane holds four
// consecutive rows of one column: one 8-byte store to the transposed copy).
__device__ __forceinline__ void st4t(uint16_t* t, uint32_t col, uint32_t row0, const uint16_t v[4]) {
  *(uint2*)(t + col * TF_LD + row0) =
      make_uint2((uint32_t)v[0] | ((uint32_t)v[1] << 16), (uint32_t)v[2] | ((uint32_t)v[3] << 16));
}
__device__ __forceinline__ tg_bf16x8 lds8(const uint16_t* p) {
  return __builtin_bit_cast(tg_bf16x8, *(const uint4*)p);
}
// Sums (and maxima) over the 8 lanes of a sample, lane k holding lobe k: the pairs
// (2p, 2p + 1) first, then (p0 + p1) + (p2 + p3) -- the order of a 4-lane quad sum over
// lanes that hold two lobes each.  DPP exchanges: lane ^ 1, lane ^ 2, then the other quad
// of the octet (each partial is uniform over its pair, then over its quad; + and max are
// commutative, so every lane ends with the same value).
__device__ __forceinline__ float dpp_xor1(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_xor2(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));
}
__device__ __forceinline__ float oct_sum(float v) {
  const float s = v + dpp_xor1(v);
  const float t = s + dpp_xor2(s);
  return t + oswap_f(t);
}
__device__ __forceinline__ float oct_max(float v) {
  const float m = th::fmax_(v, dpp_xor1(v));
  const float t = th::fmax_(m, dpp_xor2(m));
  return th::fmax_(t, oswap_f(t));
}

#define TF_THREADS 512
__global__ void __launch_bounds__(TF_THREADS) k_train_fused(const uint16_t* __restrict__ X,
                                                            const float* __restrict__ dirs,
                                                            const float* __restrict__ w, StepCtl ctl,
                                                            const float* __restrict__ wpart,
                                                            const uint16_t* __restrict__ woi,
                                                            const uint16_t* __restrict__ wio,
                                                            const float* __restrict__ bias,
                                                            float* __restrict__ partials,
                                                            float* __restrict__ loss_part) {
  __shared__ __align__(16) uint16_t tX[64 * TF_LD], tH1[64 * TF_LD], tH2[64 * TF_LD], tH3[64 * TF_LD];
  __shared__ __align__(16) uint16_t tDa[64 * TF_LD], tDb[64 * TF_LD];
  __shared__ __align__(16) uint16_t uX[64 * TF_LD], uH1[64 * TF_LD], uH2[64 * TF_LD], uH3[64 * TF_LD];
  __shared__ __align__(16) uint16_t uDa[64 * TF_LD], uDb[64 * TF_LD];  // transposed copies
  __shared__ __align__(16) float tO[64 * TG_OUT];
  __shared__ float lred[TF_THREADS / 64];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, wv = wave & 3u, h = wave >> 2;
  const uint32_t l = tid & 63u, r = l & 15u, g = l >> 4;
  const uint32_t count = step_samples(ctl), rows_per = step_rows_per(count);
  if (blockIdx.x >= step_blocks(count)) return;  // uniform per block (none when the step is skipped)
  const uint32_t r0 = blockIdx.x * rows_per, r1 = min(r0 + rows_per, count);
  const uint32_t col = 16u * wv + r;  // this lane's feature column in the forward / dZ passes
  const uint32_t rr = tid >> 3, lobe = tid & 7u;  // the loss gradient's sample and lobe
  // global loads first, in the order they are consumed: the weight partial, this wave's
  // weight rows, the first chunk's inputs
  tg_bf16x8 wf[4][2], wb[3][2];
#pragma unroll
  for (int L = 0; L < 4; ++L) {
    wf[L][0] = ld8(woi + L * 4096 + col * 64u + g * 8u);
    wf[L][1] = ld8(woi + L * 4096 + col * 64u + 32u + g * 8u);
  }
#pragma unroll
  for (int L = 1; L < 4; ++L) {
    wb[L - 1][0] = ld8(wio + L * 4096 + col * 64u + g * 8u);
    wb[L - 1][1] = ld8(wio + L * 4096 + col * 64u + 32u + g * 8u);
  }
  float bcol[4];
#pragma unroll
  for (int L = 0; L < 4; ++L) bcol[L] = bias[L * 64 + col];
  auto load_x = [&](uint32_t c0) {
    const uint32_t gr = c0 + (tid >> 3);
    return gr < r1 ? ((const uint4*)(X + (size_t)gr * 64u))[tid & 7u] : make_uint4(0, 0, 0, 0);
  };
  auto load_dw = [&](uint32_t c0, float dw[4]) {
    const uint32_t gr = c0 + rr;
    if (gr < r1) {
      dw[0] = dirs[3u * gr];
      dw[1] = dirs[3u * gr + 1u];
      dw[2] = dirs[3u * gr + 2u];
      dw[3] = w[gr];
    } else {
      dw[0] = dw[1] = dw[3] = 0.0f;
      dw[2] = 1.0f;
    }
  };
  uint4 xv = load_x(r0);
  float dw[4];
  load_dw(r0, dw);
  const float S = weight_total(wpart);
  const uint16_t* hin[4] = {tX, tH1, tH2, tH3};
  uint16_t* hinT[4] = {uX, uH1, uH2, uH3};
  tg_f32x4 acc[4][2];  // dW rows 16 wv .., columns 16 (2h + kk) ..
#pragma unroll
  for (int L = 0; L < 4; ++L)
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) acc[L][kk] = tg_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  float dbias[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // waves 0..3: column 16 wv + (l & 15), rows (l >> 4) * 16 ...
  float lsum = 0.0f;
  for (uint32_t c0 = r0; c0 < r1; c0 += 64u) {  // uniform per block
    {
      const uint32_t row = tid >> 3, seg = tid & 7u;
      *(uint4*)(tX + row * TF_LD + seg * 8u) = xv;
      const uint32_t w4[4] = {xv.x, xv.y, xv.z, xv.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        uX[(seg * 8u + 2u * j) * TF_LD + row] = (uint16_t)(w4[j] & 0xffffu);
        uX[(seg * 8u + 2u * j + 1u) * TF_LD + row] = (uint16_t)(w4[j] >> 16);
      }
    }
    const float dcur_w[4] = {dw[0], dw[1], dw[2], dw[3]};
    if (c0 + 64u < r1) {  // the next chunk's inputs are in flight during this one
      xv = load_x(c0 + 64u);
      load_dw(c0 + 64u, dw);
    }
    __syncthreads();
    // ---- forward: columns 16 wv .. 16 wv + 15 of every layer, samples 32 h .. 32 h + 31
#pragma unroll
    for (int L = 0; L < 4; ++L) {
      if (!(L == 3 && wv == 3u)) {  // the last layer has 40 outputs
        const uint16_t* in = hin[L];
#pragma unroll
        for (int mm = 0; mm < 2; ++mm) {
          const uint32_t mb = 2u * h + mm;
          const uint32_t arow = (16u * mb + r) * TF_LD;
          tg_f32x4 c = mfma(lds8(in + arow + g * 8u), wf[L][0], tg_f32x4{0.0f, 0.0f, 0.0f, 0.0f});
          c = mfma(lds8(in + arow + 32u + g * 8u), wf[L][1], c);
          uint16_t hv[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const uint32_t row = 16u * mb + 4u * g + i;
            const float v = c[i] + bcol[L];
            hv[i] = tg::f2bf(v > 0.0f ? v : 0.0f);
            if (L < 3)
              ((uint16_t*)hin[L + 1])[row * TF_LD + col] = hv[i];
            else if (col < (uint32_t)TG_OUT)
              tO[row * TG_OUT + col] = v;
          }
          if (L < 3) st4t(hinT[L + 1], col, 16u * mb + 4u * g, hv);
        }
      }
      __syncthreads();
    }
    // ---- loss gradient: 8 lanes per sample, lane k holds lobe k
    {
      const uint32_t gr = c0 + rr;
      const bool live = gr < r1 && S > 0.0f;
      const float* o = tO + rr * TG_OUT;
      const int k = (int)lobe;
      const float lg = o[32 + k];
      const float mx = oct_max(lg);
      const float e = tg::fexp(lg - mx);
      const float esum = oct_sum(e);
      const float wk = e / esum;
      const v3 om = live ? th::mk(dcur_w[0], dcur_w[1], dcur_w[2]) : th::mk(0.0f, 0.0f, 1.0f);
      const v3 rv = th::mk(o[3 * k], o[3 * k + 1], o[3 * k + 2]);
      const float len = th::fsqrt(th::dot(rv, rv) + 1e-8f);
      const v3 mu = th::divs(rv, len);
      const float kap = th::fmin_(tg::softplus(o[24 + k]) + 0.01f, 1000.0f);
      const float f = tg::vmf_pdf(mu, kap, om);
      const float q = th::fmax_(oct_sum(wk * f), 1e-30f);
      const float scale = live ? -dcur_w[3] / S : 0.0f;
      if (lobe == 0u && live) lsum += scale * tg::flog(q);
      uint16_t* dz = tDa + rr * TF_LD;
      auto put = [&](uint32_t c, uint16_t v) {  // both layouts of the output gradient
        dz[c] = v;
        uDa[c * TF_LD + rr] = v;
      };
      const float gk = wk * f / q;
      const float c = th::dot(mu, om);
      const v3 dr = th::muls(th::sub(om, th::muls(mu, c)), gk * kap / len);
      const float e2 = tg::fexp(-2.0f * kap);
      const float dk = gk * (1.0f / kap - 2.0f * e2 / (1.0f - e2) + (c - 1.0f));
      const float sig = kap >= 1000.0f ? 0.0f : 1.0f / (1.0f + tg::fexp(-o[24 + k]));
      put(3 * k, tg::f2bf(scale * dr.x));
      put(3 * k + 1, tg::f2bf(scale * dr.y));
      put(3 * k + 2, tg::f2bf(scale * dr.z));
      put(24 + k, tg::f2bf(scale * dk * sig));
      put(32 + k, tg::f2bf(scale * (gk - wk)));
#pragma unroll
      for (int j = 0; j < 3; ++j) put(TG_OUT + 3 * k + j, 0);
    }
    __syncthreads();
    // ---- backward, layer by layer
    uint16_t* dcur = tDa;
    uint16_t* dnext = tDb;
    uint16_t* dcurT = uDa;
    uint16_t* dnextT = uDb;
#pragma unroll
    for (int L = 3; L >= 0; --L) {
      // weight gradient: dW[n][k] += sum_rows dZ[row][n] hin[L][row][k], n = 16 wv + r,
      // k in columns 16 (2h + kk) ..; operands from the transposed tiles, K = 32 samples
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        const tg_bf16x8 A = lds8(dcurT + (16u * wv + r) * TF_LD + ks * 32 + g * 8);
#pragma unroll
        for (int kk = 0; kk < 2; ++kk)
          acc[L][kk] = mfma(A, lds8(hinT[L] + ((2u * h + kk) * 16u + r) * TF_LD + ks * 32 + g * 8), acc[L][kk]);
      }
      if (h == 0u) {  // the bias gradient over this lane's 16 samples, in sample order
        const uint4 b0 = *(const uint4*)(dcurT + (16u * wv + (l & 15u)) * TF_LD + (l >> 4) * 16);
        const uint4 b1 = *(const uint4*)(dcurT + (16u * wv + (l & 15u)) * TF_LD + (l >> 4) * 16 + 8);
        const uint32_t bw[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          dbias[L] += tg::bf2f((uint16_t)(bw[j] & 0xffffu));
          dbias[L] += tg::bf2f((uint16_t)(bw[j] >> 16));
        }
      }
      if (L > 0) {  // dZ of the layer below, columns 16 wv .. 16 wv + 15, samples 32 h ..:
                    // (dZ W) gated by the ReLU of this layer's input
#pragma unroll
        for (int mm = 0; mm < 2; ++mm) {
          const uint32_t mb = 2u * h + mm;
          const uint32_t arow = (16u * mb + r) * TF_LD;
          tg_f32x4 c = mfma(lds8(dcur + arow + g * 8u), wb[L - 1][0], tg_f32x4{0.0f, 0.0f, 0.0f, 0.0f});
          c = mfma(lds8(dcur + arow + 32u + g * 8u), wb[L - 1][1], c);
          const uint32_t row0 = 16u * mb + 4u * g;
          const uint2 gate2 = *(const uint2*)(hinT[L] + col * TF_LD + row0);
          const uint16_t gate[4] = {(uint16_t)(gate2.x & 0xffffu), (uint16_t)(gate2.x >> 16),
                                    (uint16_t)(gate2.y & 0xffffu), (uint16_t)(gate2.y >> 16)};
          uint16_t dv[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            dv[i] = tg::f2bf((gate[i] & 0x7fffu) ? c[i] : 0.0f);
            dnext[(row0 + i) * TF_LD + col] = dv[i];
          }
          st4t(dnextT, col, row0, dv);
        }
        __syncthreads();  // the whole chunk's dZ is written and the old one fully read
        uint16_t* t = dcur;
        dcur = dnext;
        dnext = t;
        t = dcurT;
        dcurT = dnextT;
        dnextT = t;
      }
    }
    __syncthreads();  // tiles are reloaded by the next chunk
  }
  // ---- per-block partials
#pragma unroll
  for (int L = 0; L < 4; ++L) {
    float* out = partials + (size_t)L * 256u * TG_PART + (size_t)blockIdx.x * TG_PART;
#pragma unroll
    for (int kk = 0; kk < 2; ++kk)
#pragma unroll
      for (int i = 0; i < 4; ++i) out[(16u * wv + 4u * g + i) * 64u + (2u * h + kk) * 16u + r] = acc[L][kk][i];
    if (h == 0u) {
      float b = dbias[L];
      b += __shfl_xor(b, 16, 64);
      b += __shfl_xor(b, 32, 64);
      if (l < 16u) out[64 * 64 + 16u * wv + l] = b;
    }
  }
  // the block's loss, waves in order (one partial per block; k_adam sums them in order)
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
  if (l == 0u) lred[wave] = lsum;
  __syncthreads();
  if (tid == 0u) {
    float t = 0.0f;
#pragma unroll
    for (int i = 0; i < TF_THREADS / 64; ++i) t += lred[i];
    loss_part[blockIdx.x] = t;
  }
}

// The gradient of a parameter is the fixed-order sum of its per-block partials: group j
// of 16 sums blocks j, j + 16, j + 32, ... in order, then the group sums are added in
// order (at most 256 blocks).
#define TG_GROUPS 16

// Adam on the FP32 master parameters (gradient = the fixed-order sum of the `nblocks`
// block partials, or only the gradient when lr == 0), then both padded BF16 copies and
// biases.  A block takes 64 consecutive parameters; its TG_ADAM_WAVES waves share the 16
// groups (wave q sums groups q * 16 / TG_ADAM_WAVES ..., 16 loads per group and lane,
// partials TG_PART floats apart), wave 0 adds the 16 group sums in order and applies the
// update.
#ifndef TG_ADAM_WAVES
#define TG_ADAM_WAVES 16u
#endif
#define TG_ADAM_GPW (TG_GROUPS / TG_ADAM_WAVES)  // groups per wave
__global__ void __launch_bounds__(64 * TG_ADAM_WAVES) k_adam(float* __restrict__ master, float* __restrict__ m1,
                       float* __restrict__ m2, const float* __restrict__ partials, StepCtl ctl, float lr,
                       float b1, float b2, float eps, float* __restrict__ grad_out,
                       uint16_t* __restrict__ woi, uint16_t* __restrict__ wio, float* __restrict__ bias,
                       const float* __restrict__ loss_part, float* __restrict__ loss) {
  __shared__ float gs[TG_GROUPS][64];
  // ctl.count null: refresh the BF16 copies and biases from the master weights only
  const bool refresh = ctl.count == nullptr;
  const uint32_t nblocks = refresh ? 0u : step_blocks(step_samples(ctl));
  if (!refresh && nblocks == 0u) return;  // skipped step: parameters, moments and gradient untouched
  // Adam's bias corrections at the step counter k_weight_partial advanced
  const float tstep = refresh ? 1.0f : (float)*ctl.step;
  const float bc1 = lr > 0.0f ? 1.0f - powf(b1, tstep) : 1.0f, bc2 = lr > 0.0f ? 1.0f - powf(b2, tstep) : 1.0f;
  const uint32_t lane = threadIdx.x & 63u, q = threadIdx.x >> 6;
  const uint32_t i = blockIdx.x * 64u + lane;
  if (loss != nullptr && blockIdx.x == 0 && q == 0u) {  // the step's loss: fixed-order sum of the block partials
    float t = 0.0f;
    for (uint32_t b = lane; b < nblocks; b += 64u) t += loss_part[b];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if (lane == 0u) *loss = t;
  }
  const bool valid = i < TG_NPARAMS;
  int l = 0;
  while (l < 3 && (int)i >= layer_dims(l + 1).woff) ++l;
  const LayerDims D = layer_dims(l);
  bool is_bias = (int)i >= D.boff;
  int n, k = 0;
  if (is_bias) {
    n = (int)i - D.boff;
  } else {
    n = ((int)i - D.woff) / D.kin;
    k = ((int)i - D.woff) % D.kin;
  }
  if (valid) {
    const float* base =
        partials + (size_t)l * 256u * TG_PART + (is_bias ? (size_t)64 * 64 + n : (size_t)n * 64 + k);
    float v[TG_ADAM_GPW][256 / TG_GROUPS];
#pragma unroll
    for (uint32_t jj = 0; jj < TG_ADAM_GPW; ++jj)
#pragma unroll
      for (uint32_t t = 0; t < 256 / TG_GROUPS; ++t) {
        const uint32_t b = TG_ADAM_GPW * q + jj + t * TG_GROUPS;
        v[jj][t] = b < nblocks ? base[(size_t)b * TG_PART] : 0.0f;
      }
#pragma unroll
    for (uint32_t jj = 0; jj < TG_ADAM_GPW; ++jj) {
      float s = 0.0f;
#pragma unroll
      for (uint32_t t = 0; t < 256 / TG_GROUPS; ++t)
        if (TG_ADAM_GPW * q + jj + t * TG_GROUPS < nblocks) s += v[jj][t];
      gs[TG_ADAM_GPW * q + jj][lane] = s;
    }
  }
  __syncthreads();
  if (!valid || q != 0u) return;
  float gsum = 0.0f;
  for (uint32_t j = 0; j < TG_GROUPS && j < nblocks; ++j) gsum += gs[j][lane];
  if (grad_out) grad_out[i] = gsum;
  float p = master[i];
  if (lr > 0.0f) {
    float a = b1 * m1[i] + (1.0f - b1) * gsum;
    float v = b2 * m2[i] + (1.0f - b2) * gsum * gsum;
    m1[i] = a;
    m2[i] = v;
    p = p - lr * (a / bc1) / (sqrtf(v / bc2) + eps);
    master[i] = p;
  }
  // padded copies: woi[l] is [out][in], wio[l] is [in][out], 64 x 64 per layer
  if (is_bias) {
    bias[l * 64 + n] = p;
  } else {
    uint16_t h = tg::f2bf(p);
    woi[l * 4096 + n * 64 + k] = h;
    wio[l * 4096 + k * 64 + n] = h;
  }
}

// BF16 encoding of host-side features: one thread per 8-element segment of a padded row
// (one 16-byte store); rows from `count` on and columns from TG_IN on are zero.
__global__ void k_encode_host(const float* __restrict__ feats, uint32_t count, uint32_t rows,
                              uint16_t* __restrict__ X) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, b = t >> 3, seg = t & 7u;
  if (b >= rows) return;
  uint32_t v[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t c = seg * 8u + 2u * j;
    const uint16_t lo = (b < count && c < TG_IN) ? tg::f2bf(feats[(size_t)b * TG_IN + c]) : 0;
    const uint16_t hi = (b < count && c + 1u < TG_IN) ? tg::f2bf(feats[(size_t)b * TG_IN + c + 1u]) : 0;
    v[j] = (uint32_t)lo | ((uint32_t)hi << 16);
  }
  *(uint4*)(X + (size_t)b * TG_INP + seg * 8u) = make_uint4(v[0], v[1], v[2], v[3]);
}

// tg::f2bf on the device (the conversion instruction) against the integer
// round-to-nearest-even over every 32-bit pattern: out[0] counts non-NaN mismatches,
// out[1] NaN inputs whose bits differ (the instruction returns a quiet NaN).
__global__ void __launch_bounds__(256) k_bf16_selftest(unsigned long long* __restrict__ out) {
  unsigned long long bad = 0, bad_nan = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (1ull << 32); i += stride) {
    const uint32_t u = (uint32_t)i;
    uint32_t ref;
    if ((u & 0x7f800000u) == 0x7f800000u)
      ref = u >> 16;
    else
      ref = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
    if (tg::f2bf(__uint_as_float(u)) != (uint16_t)ref) {
      if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) ++bad_nan;
      else ++bad;
    }
  }
  if (bad) atomicAdd(&out[0], bad);
  if (bad_nan) atomicAdd(&out[1], bad_nan);
}

__global__ void k_mfma_selftest(const float* A, const float* B, float* C) {
  // A 16x32 and B 32x16 row-major FP32 (rounded to BF16 here); C 16x16
  uint32_t l = threadIdx.x, r = l & 15u, g = l >> 4;
  uint16_t av[8], bv[8];
  for (int j = 0; j < 8; ++j) {
    av[j] = tg::f2bf(A[r * 32 + g * 8 + j]);
    bv[j] = tg::f2bf(B[(g * 8 + j) * 16 + r]);
  }
  tg_f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  acc = mfma(__builtin_bit_cast(tg_bf16x8, *(uint4*)av), __builtin_bit_cast(tg_bf16x8, *(uint4*)bv), acc);
  for (int i = 0; i < 4; ++i) C[(g * 4 + i) * 16 + r] = acc[i];
}

// One v_mfma_f32_32x32x16_bf16 (gfx950): lane l = s + 32 h holds A[s][8h .. 8h + 7] and
// B[8h .. 8h + 7][s]; accumulator element r of lane l is D[8 (r / 4) + 4 h + r % 4][s]
// (checked on the device by th_guide_mfma32_selftest against a host product).
typedef float tg_f32x16 __attribute__((ext_vector_type(16)));
__device__ __forceinline__ tg_f32x16 mfma32(tg_bf16x8 a, tg_bf16x8 b, tg_f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__global__ void k_mfma32_selftest(const float* A, const float* B, float* C) {
  // A 32x16 and B 16x32 row-major FP32 (rounded to BF16 here); C 32x32
  const uint32_t l = threadIdx.x, s = l & 31u, h = l >> 5;
  uint16_t av[8], bv[8];
  for (int e = 0; e < 8; ++e) {
    av[e] = tg::f2bf(A[s * 16 + h * 8 + e]);
    bv[e] = tg::f2bf(B[(h * 8 + e) * 32 + s]);
  }
  tg_f32x16 acc = {};
  acc = mfma32(__builtin_bit_cast(tg_bf16x8, *(uint4*)av), __builtin_bit_cast(tg_bf16x8, *(uint4*)bv), acc);
  for (int r = 0; r < 16; ++r) C[(8 * (r / 4) + 4 * h + r % 4) * 32 + s] = acc[r];
}

// Fused inference over a queue of deferred diffuse vertices.  Row i of the queue is the
// vertex (p, n, wo) stored in queue order (vp, vn, vw).  A wave takes 32 rows at a time
// on v_mfma_f32_32x32x16_bf16, lane (s, h) = s + 32 h serving row s: each layer is
// D[32 outputs x 32 rows] = W[32 x 16] X^T[16 x 32] per output block mb (two) and input
// block kb, the weights as the A operand (fragments in LDS, copied there once per block of
// TG_INFER_WAVES waves in the order the MFMA reads them: one conflict-free 16-byte load
// each) and the inputs as the B operand, so lane (s, h) ends with outputs
// 32 mb + 8 a + 4 h + c (a, c < 4) of row s.  The next layer's B fragment for input block
// kb' needs outputs 16 kb' + 8 h .. + 7 of row s: four of them are the lane's own and four
// its partner's (lane s + 32 (1 - h)), exchanged by one v_permlane32_swap per packed BF16
// pair -- no LDS round trip between layers.  Biases enter through an extra input block
// holding 1 in its first three inputs against the bias split into three BF16 parts
// (hi + mid + lo is the FP32 bias exactly).  ReLU runs on the packed BF16 pairs as 16-bit
// integer maxima with zero (a BF16 value and its bit pattern as an int16 have the same
// sign, and max(bf16(v), 0) = bf16(max(v, 0)) for every non-NaN v).  The first layer's B
// fragments are encoded by the lane itself with the training records' feature
// functions: input block 0 the sin/cos pairs of position axis h, block 1 those of the z
// axis (h = 0) or the bins of n.x and n.y (h = 1), block 2 the bins of (n.z, wo.x) or
// (wo.y, wo.z).  Loads are clamped into the queue and prefetched a tile ahead into two
// register sets used in turn; stores cover whole tiles (segments hold a multiple of 64
// rows) through a buffer resource, so the loop has no branch and the next tile's loads
// never wait for a store.  Outputs are stored as binary16 (relative rounding 2^-11, below
// the BF16 rounding of every hidden activation): the rows the shading kernel reads back
// are 80 bytes instead of 160, and the kernel's time is mostly these stores.
#ifndef TG_INFER_WAVES
#define TG_INFER_WAVES 16
#endif
#define TG_INFER_FRAGS 38  // per layer and output block: 3 or 4 input blocks + the bias block
__device__ __forceinline__ int infer_frag(int L, int mb, int kb) {  // kb == nkb: the bias block
  return L == 0 ? mb * 4 + kb : 8 + (L - 1) * 10 + mb * 5 + kb;
}

// Two floats -> packed binary16 (round to nearest even), clamped to +-65504 first so a
// large output stays finite.
__device__ __forceinline__ uint32_t f2h2(float x, float y) {
  typedef float f2 __attribute__((ext_vector_type(2)));
  typedef _Float16 hv2 __attribute__((ext_vector_type(2)));
  const f2 v = {__builtin_amdgcn_fmed3f(x, -65504.0f, 65504.0f), __builtin_amdgcn_fmed3f(y, -65504.0f, 65504.0f)};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, hv2));
}

__device__ __forceinline__ uint32_t relu_bf16x2(uint32_t v) {
  typedef short s16x2 __attribute__((ext_vector_type(2)));
  const s16x2 z = {0, 0};
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(s16x2, v), z));
}

__device__ __forceinline__ void infer_tile(const uint4* __restrict__ WF, uint32_t l, uint32_t base,
                                           uint32_t count, float4 a, float4 b, float4 c, float4 bmin,
                                           float4 bscale, __amdgpu_buffer_rsrc_t orsrc) {
  typedef unsigned v4u32 __attribute__((ext_vector_type(4)));
  const uint32_t s = l & 31u, h = l >> 5;
  const bool in = base + s < count;
  const v3 p01 = in ? th::mk((a.x - bmin.x) * bscale.x, (a.y - bmin.y) * bscale.y, (a.z - bmin.z) * bscale.z)
                    : th::mk(0.5f, 0.5f, 0.5f);
  const v3 n = in ? th::mk(b.x, b.y, b.z) : th::mk(0.0f, 1.0f, 0.0f);
  const v3 wo = in ? th::mk(c.x, c.y, c.z) : n;
  // ---- layer 0 inputs: three blocks of 16 (features 16 kb + 8 h + e)
  const v3 wz = th::mk(0.0f, 0.0f, 0.0f);
  auto sincos4 = [&](float x, v4u32& out) {  // the four octaves of axis value x
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      float sn, cs;
      tg::encode_pos_pair(2 * t, th::mk(x, 0.0f, 0.0f), &sn, &cs);
      out[t] = tg::f2bf2(sn, cs);
    }
  };
  auto bins2 = [&](float u0, float u1, v4u32& out) {  // four bins of u0, then of u1
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      out[t] = tg::f2bf2(tg::encode_feature(6 * TG_POS_OCT + 2 * t, p01, th::mk(u0, 0.0f, 0.0f), wz),
                         tg::encode_feature(6 * TG_POS_OCT + 2 * t + 1, p01, th::mk(u0, 0.0f, 0.0f), wz));
      out[2 + t] = tg::f2bf2(tg::encode_feature(6 * TG_POS_OCT + 2 * t, p01, th::mk(u1, 0.0f, 0.0f), wz),
                             tg::encode_feature(6 * TG_POS_OCT + 2 * t + 1, p01, th::mk(u1, 0.0f, 0.0f), wz));
    }
  };
  v4u32 x0, x1, x2, zs, zb;
  sincos4(h ? p01.y : p01.x, x0);
  sincos4(p01.z, zs);
  bins2(n.x, n.y, zb);
  bins2(h ? wo.y : n.z, h ? wo.z : wo.x, x2);
#pragma unroll
  for (int t = 0; t < 4; ++t) x1[t] = h ? zb[t] : zs[t];
  tg_bf16x8 B[4];
  B[0] = __builtin_bit_cast(tg_bf16x8, x0);
  B[1] = __builtin_bit_cast(tg_bf16x8, x1);
  B[2] = __builtin_bit_cast(tg_bf16x8, x2);
  B[3] = B[2];  // unused by layer 0
  // the bias block's inputs: 1, 1, 1 in the first three (lanes h = 0), zero elsewhere
  const v4u32 one3 = {h ? 0u : 0x3F803F80u, h ? 0u : 0x00003F80u, 0u, 0u};
  const tg_bf16x8 Bb = __builtin_bit_cast(tg_bf16x8, one3);
  // the weight fragments are read from LDS for every tile: an opaque copy of the lane
  // index keeps the compiler from hoisting all 38 loads out of the tile loop into
  // registers (152 VGPRs, two waves per SIMD instead of four)
  uint32_t li = l;
  __asm__ volatile("" : "+v"(li));
#pragma unroll
  for (int L = 0; L < 4; ++L) {
    const int nkb = L == 0 ? 3 : 4;
    tg_f32x16 acc[2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
      acc[mb] = mfma32(__builtin_bit_cast(tg_bf16x8, WF[infer_frag(L, mb, nkb) * 64 + li]), Bb, tg_f32x16{});
#pragma unroll
      for (int kb = 0; kb < nkb; ++kb)
        acc[mb] = mfma32(__builtin_bit_cast(tg_bf16x8, WF[infer_frag(L, mb, kb) * 64 + li]), B[kb], acc[mb]);
    }
    if (L < 3) {
      // B[kb'] (features 16 kb' + 8 h + e) from output block kb' / 2, rows a0 = 2 (kb' % 2)
      // and a0 + 1 of the accumulator groups: X = this lane's group a0, Y = its group a0 + 1;
      // swapping X of the upper half with Y of the lower gives the lower lanes (X, partner X)
      // = features 0..7 of the block and the upper lanes (partner Y, Y) = features 8..15
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        const tg_f32x16& q = acc[kb >> 1];
        const int a0 = 2 * (kb & 1);
        uint32_t X0 = relu_bf16x2(tg::f2bf2(q[4 * a0 + 0], q[4 * a0 + 1]));
        uint32_t X1 = relu_bf16x2(tg::f2bf2(q[4 * a0 + 2], q[4 * a0 + 3]));
        uint32_t Y0 = relu_bf16x2(tg::f2bf2(q[4 * a0 + 4], q[4 * a0 + 5]));
        uint32_t Y1 = relu_bf16x2(tg::f2bf2(q[4 * a0 + 6], q[4 * a0 + 7]));
        const auto s0 = __builtin_amdgcn_permlane32_swap(X0, Y0, false, false);
        const auto s1 = __builtin_amdgcn_permlane32_swap(X1, Y1, false, false);
        const v4u32 v = {s0[0], s1[0], s0[1], s1[1]};
        B[kb] = __builtin_bit_cast(tg_bf16x8, v);
      }
    } else {
      // outputs 8 a + 4 h + c of row s (block 0) and 32 + 4 h + c (block 1, a = 0), rounded
      // to binary16 (clamped to its range) and stored four at a time
      typedef unsigned v2u32 __attribute__((ext_vector_type(2)));
      const uint32_t ob = ((base + s) * TG_OUT + 4u * h) * 2u;
#pragma unroll
      for (int a = 0; a < 5; ++a) {
        const tg_f32x16& q = acc[a >> 2];
        const int g = a & 3;
        const v2u32 v = {f2h2(q[4 * g], q[4 * g + 1]), f2h2(q[4 * g + 2], q[4 * g + 3])};
        __builtin_amdgcn_raw_buffer_store_b64(v, orsrc, ob + 16u * (uint32_t)a, 0, 0);
      }
    }
  }
}

__global__ void __launch_bounds__(64 * TG_INFER_WAVES) k_guide_infer(
    const uint16_t* __restrict__ woi, const float* __restrict__ bias, const uint32_t* __restrict__ count_ptr,
    uint32_t count_stride, uint32_t nseg, uint32_t seg_cap, const float4* __restrict__ vp,
    const float4* __restrict__ vn, const float4* __restrict__ vw, float4 bmin, float4 bscale,
    uint16_t* __restrict__ O) {
  __shared__ uint4 WF[TG_INFER_FRAGS * 64];
  // block b serves queue segment b % nseg (the grid is a multiple of nseg); wave w of the
  // block takes 32-row tiles jb * TG_INFER_WAVES + w, then every nbs * TG_INFER_WAVES-th
  const uint32_t sg = blockIdx.x % nseg, jb = blockIdx.x / nseg, nbs = gridDim.x / nseg;
  const uint32_t count = count_ptr[(size_t)sg * count_stride];
  if (jb * TG_INFER_WAVES * 32u >= count) return;  // uniform per block
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < TG_INFER_FRAGS * 64u; i += 64u * TG_INFER_WAVES) {
    const uint32_t f = i >> 6, l = i & 63u, o = l & 31u, h = l >> 5;
    const uint32_t L = f < 8u ? 0u : 1u + (f - 8u) / 10u;
    const uint32_t r = f < 8u ? f : (f - 8u) % 10u;
    const uint32_t nkb = L == 0u ? 3u : 4u, mb = r / (nkb + 1u), kb = r % (nkb + 1u);
    const uint32_t row = mb * 32u + o;
    uint4 v;
    if (kb < nkb) {
      v = *(const uint4*)(woi + L * 4096u + row * 64u + kb * 16u + h * 8u);
    } else {  // bias block: hi, mid, lo BF16 parts in the first three inputs of lanes h = 0
      const float bv = bias[L * 64u + row];
      const uint16_t hi = tg::f2bf(bv);
      const float r1 = bv - tg::bf2f(hi);
      const uint16_t mid = tg::f2bf(r1);
      const uint16_t lo = tg::f2bf(r1 - tg::bf2f(mid));
      v = h ? make_uint4(0u, 0u, 0u, 0u)
            : make_uint4((uint32_t)hi | ((uint32_t)mid << 16), (uint32_t)lo, 0u, 0u);
    }
    WF[i] = v;
  }
  __syncthreads();
  const uint32_t w = t >> 6, l = t & 63u, s = l & 31u;
  const uint32_t step = nbs * TG_INFER_WAVES * 32u, last = count - 1u;
  uint32_t base = (jb * TG_INFER_WAVES + w) * 32u;
  if (base >= count) return;  // no block barrier follows
  vp += (size_t)sg * seg_cap;
  vn += (size_t)sg * seg_cap;
  vw += (size_t)sg * seg_cap;
  // this segment's rows of the output (seg_cap * TG_OUT binary16 values, below 2^31 bytes)
  const __amdgpu_buffer_rsrc_t orsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void*)(O + (size_t)sg * seg_cap * TG_OUT), (short)0, (int)(seg_cap * (uint32_t)TG_OUT * 2u), 0x00020000);
  // two register sets of vertex rows, loaded one tile ahead and used in turn (a single
  // loop-carried set would be copied at the back edge, which waits for its loads)
  uint32_t lr = std::min(base + s, last);
  float4 pa = vp[lr], na = vn[lr], wa = vw[lr];
  for (;;) {
    lr = std::min(base + step + s, last);
    const float4 pb = vp[lr], nb = vn[lr], wb = vw[lr];
    infer_tile(WF, l, base, count, pa, na, wa, bmin, bscale, orsrc);
    base += step;
    if (base >= count) break;
    lr = std::min(base + step + s, last);
    pa = vp[lr];
    na = vn[lr];
    wa = vw[lr];
    infer_tile(WF, l, base, count, pb, nb, wb, bmin, bscale, orsrc);
    base += step;
    if (base >= count) break;
  }
}

uint32_t round_up(uint32_t x, uint32_t m) { return (x + m - 1) / m * m; }

#ifndef TG_INFER_BLOCKS_PER_CU
#define TG_INFER_BLOCKS_PER_CU 1u
#endif

}  // namespace

// ====================================================================== guide object
struct th_guide {
  th_context* ctx = nullptr;
  float *d_master = nullptr, *d_m1 = nullptr, *d_m2 = nullptr, *d_bias = nullptr, *d_grad = nullptr;
  uint16_t *d_woi = nullptr, *d_wio = nullptr;
  // d_ctl[0]: the host API's sample count, [1]: Adam's step counter (steps applied)
  uint32_t* d_ctl = nullptr;
  uint32_t host_count = 0;  // source of the count's upload (stays valid until it lands)
  uint32_t step = 0;        // Adam steps applied, as far as the host knows (the guided schedule's count)
  float lr = 1e-2f, beta1 = 0.9f, beta2 = 0.99f, eps = 1e-8f;
  // batch workspace
  uint32_t cap = 0;
  uint16_t *X0 = nullptr, *H1 = nullptr, *H2 = nullptr, *H3 = nullptr;
  float *O = nullptr, *feat = nullptr, *dirs = nullptr, *w = nullptr, *partials = nullptr, *scal = nullptr;
  float *wpart = nullptr, *lpart = nullptr;
  uint32_t nblocks = 0;
  std::vector<void*> ws;
};

static const uint32_t kWgradBlocks = 256;  // partial slots per layer (k_adam sums at most 256)

static int32_t guide_reserve(th_guide* g, uint32_t rows) {
  if (g->cap >= rows) return TH_OK;
  for (void* p : g->ws) (void)hipFree(p);
  g->ws.clear();
  g->cap = 0;
  uint32_t cap = round_up(std::max(rows, 1024u), 1024u);
  auto al = [&](void** p, size_t bytes) -> int32_t {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return th_set_error(TH_E_DEVICE, "guide workspace allocation failed");
    e = hipMemset(*p, 0, bytes);
    g->ws.push_back(*p);
    return e == hipSuccess ? TH_OK : th_set_error(TH_E_DEVICE, "guide workspace clear failed");
  };
  size_t h = (size_t)cap * 64 * 2;
  int32_t rc;
  if ((rc = al((void**)&g->X0, h)) || (rc = al((void**)&g->H1, h)) || (rc = al((void**)&g->H2, h)) ||
      (rc = al((void**)&g->H3, h)) ||
      (rc = al((void**)&g->O, (size_t)cap * TG_OUTP * 4)) ||
      (rc = al((void**)&g->feat, (size_t)cap * TG_IN * 4)) || (rc = al((void**)&g->dirs, (size_t)cap * 12)) ||
      (rc = al((void**)&g->w, (size_t)cap * 4)) ||
      (rc = al((void**)&g->partials, (size_t)kWgradBlocks * 4 * TG_PART * 4)) ||
      (rc = al((void**)&g->wpart, 256 * 4)) || (rc = al((void**)&g->lpart, kWgradBlocks * 4)) ||
      (rc = al((void**)&g->scal, 64)))
    return rc;
  g->cap = cap;
  return TH_OK;
}

// Forward over `rows` rows of X (BF16, 64 wide) into the workspace H1..H3 and O.
static int32_t th_guide_forward_rows(th_guide* g, const uint16_t* X, const uint32_t* d_count,
                                     uint32_t rows, hipStream_t st) {
  uint32_t grid = (round_up(rows, 64) / 64);  // 4 waves x 16 rows per block
  if (grid == 0) return TH_OK;
  hipLaunchKernelGGL((k_gemm_rows<64, 64, 0>), dim3(grid), dim3(256), 0, st, X, g->d_woi + 0 * 4096,
                     g->d_bias + 0, d_count, rows, (void*)g->H1);
  hipLaunchKernelGGL((k_gemm_rows<64, 64, 0>), dim3(grid), dim3(256), 0, st, g->H1, g->d_woi + 1 * 4096,
                     g->d_bias + 64, d_count, rows, (void*)g->H2);
  hipLaunchKernelGGL((k_gemm_rows<64, 64, 0>), dim3(grid), dim3(256), 0, st, g->H2, g->d_woi + 2 * 4096,
                     g->d_bias + 128, d_count, rows, (void*)g->H3);
  hipLaunchKernelGGL((k_gemm_rows<64, 64, 1>), dim3(grid), dim3(256), 0, st, g->H3, g->d_woi + 3 * 4096,
                     g->d_bias + 192, d_count, rows, (void*)g->O);
  TH_CHECK(hipGetLastError());
  return TH_OK;
}

// One training step on `count` samples already on the device: BF16 features X (64 wide,
// zero beyond 48; rows up to count rounded to 32 must be finite), directions and
// weights.  lr = 0 computes the gradient only.  Workspace must hold count rows.
static int32_t guide_step(th_guide* g, const uint16_t* X, const float* dirs, const float* w, const StepCtl& c,
                          float lr) {
  hipStream_t st = g->ctx->stream;
  hipLaunchKernelGGL(k_weight_partial, dim3(256), dim3(256), 0, st, w, c, lr > 0.0f ? 1u : 0u, g->wpart, g->scal);
  // every partial slot's block is launched; those beyond the step's sample count return
  float* P = g->partials;
  hipLaunchKernelGGL(k_train_fused, dim3(kWgradBlocks), dim3(TF_THREADS), 0, st, X, dirs, w, c, g->wpart, g->d_woi,
                     g->d_wio, g->d_bias, P, g->lpart);
  TH_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_adam, dim3((TG_NPARAMS + 63) / 64), dim3(64 * TG_ADAM_WAVES), 0, st, g->d_master, g->d_m1, g->d_m2, P, c,
                     lr, g->beta1, g->beta2, g->eps, g->d_grad, g->d_woi, g->d_wio, g->d_bias, g->lpart,
                     g->scal + 1);
  TH_CHECK(hipGetLastError());
  return TH_OK;
}

// ====================================================================== renderer interface
int32_t guide_infer(th_guide* g, const uint32_t* d_count, uint32_t count_stride, uint32_t nseg,
                    uint32_t seg_cap, const float4* vp, const float4* vn, const float4* vw,
                    const float* bmin, const float* bmax, uint16_t* O, hipStream_t st) {
  if (nseg == 0u) return th_set_error(TH_E_INVALID, "guide queue without segments");
  // persistent: one block of 16 waves per compute unit (four per SIMD), so each compute
  // unit copies the network into LDS once and its waves walk many tiles (four blocks of
  // four waves, each with its own copy, measured 7 % slower); a multiple of nseg, so every
  // segment has the same number of blocks
  const uint32_t tiles = (seg_cap + 32u * TG_INFER_WAVES - 1u) / (32u * TG_INFER_WAVES);
  uint32_t per_seg = std::min<uint32_t>(
      tiles, (TG_INFER_BLOCKS_PER_CU * (uint32_t)std::max(g->ctx->cu_count, 1) + nseg - 1u) / nseg);
  const uint32_t grid = per_seg * nseg;
  if (grid == 0) return TH_OK;
  float4 lo = make_float4(bmin[0], bmin[1], bmin[2], 0.0f), sc;
  float e[3];
  for (int a = 0; a < 3; ++a) e[a] = bmax[a] - bmin[a];
  sc = make_float4(e[0] > 0.0f ? 1.0f / e[0] : 0.0f, e[1] > 0.0f ? 1.0f / e[1] : 0.0f,
                   e[2] > 0.0f ? 1.0f / e[2] : 0.0f, 0.0f);
  hipLaunchKernelGGL(k_guide_infer, dim3(grid), dim3(64 * TG_INFER_WAVES), 0, st, g->d_woi, g->d_bias, d_count, count_stride,
                     nseg, seg_cap, vp, vn, vw, lo, sc, O);
  TH_CHECK(hipGetLastError());
  return TH_OK;
}

int32_t guide_train(th_guide* g, const uint16_t* X, const float* dirs, const float* w, const uint32_t* d_count,
                    uint32_t cap, uint32_t min_count, uint32_t* d_stats) {
  int32_t rc = guide_reserve(g, cap);
  if (rc) return rc;
  const StepCtl c{d_count, cap, min_count, g->d_ctl + 1, d_stats};
  return guide_step(g, X, dirs, w, c, g->lr);
}

// The last step's loss, copied once (a copy per step would stall the stream each time).
int32_t guide_last_loss(th_guide* g, float* out) {
  TH_CHECK(hipMemcpyAsync(out, g->scal + 1, 4, hipMemcpyDeviceToHost, g->ctx->stream));
  return TH_OK;
}

uint32_t guide_steps(const th_guide* g) { return g->step; }
void guide_set_steps(th_guide* g, uint32_t steps) { g->step = steps; }
th_context* guide_context(const th_guide* g) { return g->ctx; }

// ====================================================================== C ABI
// He-uniform hidden weights (bound sqrt(6 / fan_in)) from a seeded PCG stream, zero hidden
// biases; the output layer starts at a tenth of that scale with biases that make the
// initial mixture cover the sphere evenly: lobe means on a Fibonacci sphere, kappa = 2,
// equal weights.
static void host_init_params(uint32_t seed, std::vector<float>& p) {
  p.assign(TG_NPARAMS, 0.0f);
  uint32_t s = seed * 2654435761u + 12345u;
  for (int l = 0; l < 4; ++l) {
    LayerDims D = layer_dims(l);
    float a = sqrtf(6.0f / (float)D.kin) * (l == 3 ? 0.1f : 1.0f);
    for (int i = 0; i < D.nout * D.kin; ++i) {
      s = th::pcg_hash(s);
      p[D.woff + i] = ((float)(s >> 8) * (1.0f / 16777216.0f) * 2.0f - 1.0f) * a;
    }
  }
  float* b4 = p.data() + layer_dims(3).boff;
  const float golden = 2.39996323f;  // pi (3 - sqrt 5)
  for (int k = 0; k < TG_LOBES; ++k) {
    float z = 1.0f - (2.0f * (float)k + 1.0f) / (float)TG_LOBES;
    float r = sqrtf(std::max(0.0f, 1.0f - z * z));
    b4[3 * k] = r * cosf(golden * (float)k);
    b4[3 * k + 1] = z;
    b4[3 * k + 2] = r * sinf(golden * (float)k);
    b4[24 + k] = logf(expf(2.0f - 0.01f) - 1.0f);  // softplus^-1(kappa - 0.01), kappa = 2
    b4[32 + k] = 0.0f;
  }
}

static int32_t guide_refresh(th_guide* g) {  // bf16 copies from the master weights
  const StepCtl refresh{nullptr, 0u, 0u, nullptr, nullptr};
  hipLaunchKernelGGL(k_adam, dim3((TG_NPARAMS + 63) / 64), dim3(64 * TG_ADAM_WAVES), 0, g->ctx->stream, g->d_master,
                     g->d_m1, g->d_m2, g->partials, refresh, 0.0f, g->beta1, g->beta2, g->eps, (float*)nullptr,
                     g->d_woi, g->d_wio, g->d_bias, (const float*)nullptr, (float*)nullptr);
  TH_CHECK(hipGetLastError());
  return TH_OK;
}

extern "C" {

int32_t th_guide_create(th_context* ctx, uint32_t seed, th_guide** out) {
  if (!ctx || !out) return th_set_error(TH_E_INVALID, "null argument");
  *out = nullptr;
  TH_CHECK(hipSetDevice(ctx->device));
  th_guide* g = new th_guide();
  g->ctx = ctx;
  auto al = [&](void** p, size_t bytes) -> bool {
    return hipMalloc(p, bytes) == hipSuccess && hipMemset(*p, 0, bytes) == hipSuccess;
  };
  if (!al((void**)&g->d_master, TG_NPARAMS * 4) || !al((void**)&g->d_m1, TG_NPARAMS * 4) ||
      !al((void**)&g->d_m2, TG_NPARAMS * 4) || !al((void**)&g->d_grad, TG_NPARAMS * 4) ||
      !al((void**)&g->d_bias, 4 * 64 * 4) || !al((void**)&g->d_woi, 4 * 4096 * 2) ||
      !al((void**)&g->d_wio, 4 * 4096 * 2) || !al((void**)&g->d_ctl, 16 * 4)) {
    th_guide_destroy(g);
    return th_set_error(TH_E_DEVICE, "guide allocation failed");
  }
  std::vector<float> p;
  host_init_params(seed, p);
  int32_t rc = guide_reserve(g, 1024);
  if (rc) {
    th_guide_destroy(g);
    return rc;
  }
  if (hipMemcpy(g->d_master, p.data(), TG_NPARAMS * 4, hipMemcpyHostToDevice) != hipSuccess ||
      (rc = guide_refresh(g)) || hipStreamSynchronize(ctx->stream) != hipSuccess) {
    th_guide_destroy(g);
    return rc ? rc : th_set_error(TH_E_DEVICE, "guide init failed");
  }
  *out = g;
  return TH_OK;
}

void th_guide_destroy(th_guide* g) {
  if (!g) return;
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  for (void* p : g->ws) (void)hipFree(p);
  void* ps[] = {g->d_master, g->d_m1, g->d_m2, g->d_grad, g->d_bias, g->d_woi, g->d_wio, g->d_ctl};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete g;
}

int32_t th_guide_get_params(th_guide* g, float* out, uint32_t n) {
  if (!g || !out || n != TG_NPARAMS) return th_set_error(TH_E_INVALID, "expected TG_NPARAMS floats");
  TH_CHECK(hipStreamSynchronize(g->ctx->stream));
  TH_CHECK(hipMemcpy(out, g->d_master, TG_NPARAMS * 4, hipMemcpyDeviceToHost));
  return TH_OK;
}

int32_t th_guide_set_params(th_guide* g, const float* in, uint32_t n) {
  if (!g || !in || n != TG_NPARAMS) return th_set_error(TH_E_INVALID, "expected TG_NPARAMS floats");
  TH_CHECK(hipStreamSynchronize(g->ctx->stream));
  TH_CHECK(hipMemcpy(g->d_master, in, TG_NPARAMS * 4, hipMemcpyHostToDevice));
  TH_CHECK(hipMemset(g->d_m1, 0, TG_NPARAMS * 4));
  TH_CHECK(hipMemset(g->d_m2, 0, TG_NPARAMS * 4));
  TH_CHECK(hipMemset(g->d_ctl, 0, 16 * 4));
  g->step = 0;
  int32_t rc = guide_refresh(g);
  if (rc) return rc;
  TH_CHECK(hipStreamSynchronize(g->ctx->stream));
  return TH_OK;
}

int32_t th_guide_set_hyper(th_guide* g, float lr, float beta1, float beta2, float eps) {
  if (!g || !(lr >= 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f) ||
      !(eps > 0.0f))
    return th_set_error(TH_E_INVALID, "invalid Adam hyper-parameters");
  g->lr = lr;
  g->beta1 = beta1;
  g->beta2 = beta2;
  g->eps = eps;
  return TH_OK;
}

int32_t th_guide_forward(th_guide* g, const float* feats, uint32_t count, float* out) {
  if (!g || !feats || !out || count == 0) return th_set_error(TH_E_INVALID, "null argument");
  int32_t rc = guide_reserve(g, count);
  if (rc) return rc;
  hipStream_t st = g->ctx->stream;
  uint32_t rows = round_up(count, 32);
  TH_CHECK(hipMemcpyAsync(g->feat, feats, (size_t)count * TG_IN * 4, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_encode_host, dim3((rows * 8u + 255) / 256), dim3(256), 0, st, g->feat, count, rows, g->X0);
  if ((rc = th_guide_forward_rows(g, g->X0, nullptr, rows, st))) return rc;
  std::vector<float> o((size_t)rows * TG_OUTP);
  TH_CHECK(hipMemcpyAsync(o.data(), g->O, o.size() * 4, hipMemcpyDeviceToHost, st));
  TH_CHECK(hipStreamSynchronize(st));
  for (uint32_t b = 0; b < count; ++b)
    for (int j = 0; j < TG_OUT; ++j) out[(size_t)b * TG_OUT + j] = o[(size_t)b * TG_OUTP + j];
  return TH_OK;
}

int32_t th_guide_infer_vertices(th_guide* g, const float* vertices, uint32_t count, const float bmin[3],
                                const float bmax[3], float* out) {
  if (!g || !vertices || !bmin || !bmax || !out || count == 0) return th_set_error(TH_E_INVALID, "null argument");
  if (count > (1u << 24)) return th_set_error(TH_E_INVALID, "too many vertices");
  // one queue segment holding the rows, padded to a multiple of 64 as the renderer's are
  const uint32_t cap = round_up(count, 64);
  std::vector<float4> hp(cap, make_float4(0.0f, 0.0f, 0.0f, 0.0f)), hn(hp), hw(hp);
  for (uint32_t i = 0; i < count; ++i) {
    const float* v = vertices + (size_t)i * 9;
    hp[i] = make_float4(v[0], v[1], v[2], 0.0f);
    hn[i] = make_float4(v[3], v[4], v[5], 0.0f);
    hw[i] = make_float4(v[6], v[7], v[8], 0.0f);
  }
  hipStream_t st = g->ctx->stream;
  float4 *dp = nullptr, *dn = nullptr, *dw = nullptr;
  uint16_t* dO = nullptr;
  uint32_t* dc = nullptr;
  int32_t rc = TH_OK;
  const size_t vb = (size_t)cap * sizeof(float4);
  if (hipMalloc((void**)&dp, vb) != hipSuccess || hipMalloc((void**)&dn, vb) != hipSuccess ||
      hipMalloc((void**)&dw, vb) != hipSuccess || hipMalloc((void**)&dO, (size_t)cap * TG_OUT * 2) != hipSuccess ||
      hipMalloc((void**)&dc, 4) != hipSuccess) {
    rc = th_set_error(TH_E_DEVICE, "inference workspace allocation failed");
  } else if (hipMemcpyAsync(dp, hp.data(), vb, hipMemcpyHostToDevice, st) != hipSuccess ||
             hipMemcpyAsync(dn, hn.data(), vb, hipMemcpyHostToDevice, st) != hipSuccess ||
             hipMemcpyAsync(dw, hw.data(), vb, hipMemcpyHostToDevice, st) != hipSuccess ||
             hipMemcpyAsync(dc, &count, 4, hipMemcpyHostToDevice, st) != hipSuccess) {
    rc = th_set_error(TH_E_DEVICE, "inference upload failed");
  } else if ((rc = guide_infer(g, dc, 1u, 1u, cap, dp, dn, dw, bmin, bmax, dO, st)) == TH_OK) {
    std::vector<uint16_t> ho((size_t)count * TG_OUT);
    if (hipMemcpyAsync(ho.data(), dO, ho.size() * 2, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      rc = th_set_error(TH_E_DEVICE, "inference readback failed");
    } else {
      for (size_t i = 0; i < ho.size(); ++i) out[i] = tg::h2f(ho[i]);
    }
  }
  (void)hipStreamSynchronize(st);
  for (void* p : {(void*)dp, (void*)dn, (void*)dw, (void*)dO, (void*)dc})
    if (p) (void)hipFree(p);
  return rc;
}

int32_t th_guide_train_batch(th_guide* g, const float* feats, const float* dirs, const float* weights,
                             uint32_t count, int32_t update, float* loss, float* grad) {
  if (!g || !feats || !dirs || !weights || count == 0) return th_set_error(TH_E_INVALID, "null argument");
  int32_t rc = guide_reserve(g, count);
  if (rc) return rc;
  hipStream_t st = g->ctx->stream;
  uint32_t rows = round_up(count, 32);
  TH_CHECK(hipMemcpyAsync(g->feat, feats, (size_t)count * TG_IN * 4, hipMemcpyHostToDevice, st));
  TH_CHECK(hipMemcpyAsync(g->dirs, dirs, (size_t)count * 12, hipMemcpyHostToDevice, st));
  TH_CHECK(hipMemcpyAsync(g->w, weights, (size_t)count * 4, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_encode_host, dim3((rows * 8u + 255) / 256), dim3(256), 0, st, g->feat, count, rows, g->X0);
  g->host_count = count;
  TH_CHECK(hipMemcpyAsync(g->d_ctl, &g->host_count, 4, hipMemcpyHostToDevice, st));
  const StepCtl c{g->d_ctl, count, 1u, g->d_ctl + 1, nullptr};
  if ((rc = guide_step(g, g->X0, g->dirs, g->w, c, update ? g->lr : 0.0f))) return rc;
  if (update && g->lr > 0.0f) g->step += 1;  // count >= 1: the step applies
  if (loss) TH_CHECK(hipMemcpyAsync(loss, g->scal + 1, 4, hipMemcpyDeviceToHost, st));
  if (grad) TH_CHECK(hipMemcpyAsync(grad, g->d_grad, TG_NPARAMS * 4, hipMemcpyDeviceToHost, st));
  TH_CHECK(hipStreamSynchronize(st));
  return TH_OK;
}

int32_t th_guide_mfma_selftest(th_context* ctx, const float* A, const float* B, float* C) {
  if (!ctx || !A || !B || !C) return th_set_error(TH_E_INVALID, "null argument");
  TH_CHECK(hipSetDevice(ctx->device));
  float *dA = nullptr, *dB = nullptr, *dC = nullptr;
  hipError_t e = hipMalloc((void**)&dA, 16 * 32 * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&dB, 32 * 16 * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&dC, 16 * 16 * 4);
  if (e == hipSuccess) e = hipMemcpy(dA, A, 16 * 32 * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dB, B, 32 * 16 * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, ctx->stream, dA, dB, dC);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(C, dC, 16 * 16 * 4, hipMemcpyDeviceToHost);
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  if (e != hipSuccess) return th_set_error(TH_E_DEVICE, std::string("mfma selftest: ") + hipGetErrorString(e));
  return TH_OK;
}

int32_t th_guide_mfma32_selftest(th_context* ctx, const float* A, const float* B, float* C) {
  if (!ctx || !A || !B || !C) return th_set_error(TH_E_INVALID, "null argument");
  TH_CHECK(hipSetDevice(ctx->device));
  float *dA = nullptr, *dB = nullptr, *dC = nullptr;
  hipError_t e = hipMalloc((void**)&dA, 32 * 16 * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&dB, 16 * 32 * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&dC, 32 * 32 * 4);
  if (e == hipSuccess) e = hipMemcpy(dA, A, 32 * 16 * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dB, B, 16 * 32 * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mfma32_selftest, dim3(1), dim3(64), 0, ctx->stream, dA, dB, dC);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(C, dC, 32 * 32 * 4, hipMemcpyDeviceToHost);
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  if (e != hipSuccess) return th_set_error(TH_E_DEVICE, std::string("mfma32 selftest: ") + hipGetErrorString(e));
  return TH_OK;
}

int32_t th_bf16_selftest(th_context* ctx, uint64_t* mismatches, uint64_t* nan_mismatches) {
  if (!ctx || !mismatches) return th_set_error(TH_E_INVALID, "null argument");
  TH_CHECK(hipSetDevice(ctx->device));
  unsigned long long* d = nullptr;
  unsigned long long h[2] = {0, 0};
  hipError_t e = hipMalloc((void**)&d, sizeof(h));
  if (e == hipSuccess) e = hipMemcpy(d, h, sizeof(h), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_bf16_selftest, dim3(8192), dim3(256), 0, ctx->stream, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return th_set_error(TH_E_DEVICE, std::string("bf16 selftest: ") + hipGetErrorString(e));
  *mismatches = h[0];
  if (nan_mismatches) *nan_mismatches = h[1];
  return TH_OK;
}

}  // extern "C"
