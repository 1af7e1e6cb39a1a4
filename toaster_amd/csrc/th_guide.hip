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

// Loss gradient of one batch: L = -sum_b w_b log q_b(dir_b) / sum_b w_b.  The weight sum
// is reduced in a fixed order (256 block partials here, then the same 256-to-1 tree in
// every block of the training kernel), so gradients are reproducible call to call.
// Block 0 also clears the step's scalars (sums[1]: the loss accumulator).
__global__ void __launch_bounds__(256) k_weight_partial(const float* __restrict__ w, uint32_t n,
                                                        float* __restrict__ part, float* __restrict__ sums) {
  __shared__ float red[256];
  if (blockIdx.x == 0 && threadIdx.x < 16u) sums[threadIdx.x] = 0.0f;
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

// The batch's total weight from the 256 partials (fixed-order tree; `red` holds 256
// floats of LDS).  Every thread of the block gets it.
__device__ __forceinline__ float weight_total(const float* __restrict__ part, float* red) {
  red[threadIdx.x] = part[threadIdx.x];
  __syncthreads();
  for (uint32_t off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  const float t = red[0];
  __syncthreads();
  return t;
}

// Fused training step.  One block (4 waves) walks `rows_per` consecutive samples in
// chunks of 64 (16 per wave) with every activation in LDS: forward through the four
// layers (A fragments from LDS tiles, B fragments = out-major weights from L2), the
// vMF-mixture loss gradient with 4 lanes per sample (2 lobes each, quad reductions), then
// layer by layer the chunk's weight gradient (wave w owns output rows 16w..16w+15 of
// every layer's dW, accumulated across chunks in registers; the MFMA K dimension runs
// over 32 samples) and the gated back-propagation with the in-major weights.  Each block
// writes one partial gradient per layer for the fixed-order reduction and Adam.
#define TF_LD 72  // LDS row stride (BF16 elements) of a 64-wide activation tile
__device__ __forceinline__ tg_bf16x8 lds8(const uint16_t* p) {
  return __builtin_bit_cast(tg_bf16x8, *(const uint4*)p);
}
__device__ __forceinline__ float quad_sum(float v) {
  return (qbcast_f<0>(v) + qbcast_f<1>(v)) + (qbcast_f<2>(v) + qbcast_f<3>(v));
}
__device__ __forceinline__ float quad_max(float v) {
  return th::fmax_(th::fmax_(qbcast_f<0>(v), qbcast_f<1>(v)), th::fmax_(qbcast_f<2>(v), qbcast_f<3>(v)));
}

__global__ void __launch_bounds__(256) k_train_fused(const uint16_t* __restrict__ X,
                                                     const float* __restrict__ dirs,
                                                     const float* __restrict__ w, uint32_t count,
                                                     uint32_t rows_per, const float* __restrict__ wpart,
                                                     const uint16_t* __restrict__ woi,
                                                     const uint16_t* __restrict__ wio,
                                                     const float* __restrict__ bias,
                                                     float* __restrict__ partials, float* __restrict__ loss) {
  __shared__ __align__(16) uint16_t tX[64 * TF_LD], tH1[64 * TF_LD], tH2[64 * TF_LD], tH3[64 * TF_LD];
  __shared__ __align__(16) uint16_t tDa[64 * TF_LD], tDb[64 * TF_LD];
  __shared__ __align__(16) float tO[64 * TG_OUT];
  const uint32_t tid = threadIdx.x, wv = tid >> 6, l = tid & 63u, r = l & 15u, g = l >> 4;
  const uint32_t r0 = blockIdx.x * rows_per, r1 = min(r0 + rows_per, count);
  const float S = weight_total(wpart, (float*)tDa);  // tDa is free until the backward pass
  const uint16_t* hin[4] = {tX, tH1, tH2, tH3};
  tg_f32x4 acc[4][4];
#pragma unroll
  for (int L = 0; L < 4; ++L)
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc[L][kb] = tg_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  float dbias[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // column 16 wv + (l & 15), rows (l >> 4) * 16 ...
  float lsum = 0.0f;
  for (uint32_t c0 = r0; c0 < r1; c0 += 64u) {  // uniform per block
    for (uint32_t i = tid; i < 64u * 8u; i += 256u) {
      const uint32_t row = i >> 3, seg = i & 7u, gr = c0 + row;
      uint4 v = gr < r1 ? ((const uint4*)(X + (size_t)gr * 64u))[seg] : make_uint4(0, 0, 0, 0);
      *(uint4*)(tX + row * TF_LD + seg * 8u) = v;
    }
    __syncthreads();
    // ---- forward: wave wv owns rows 16 wv .. 16 wv + 15 of the chunk
    const uint32_t arow = (16u * wv + r) * TF_LD;
#pragma unroll
    for (int L = 0; L < 4; ++L) {
      const uint16_t* in = hin[L];
      const tg_bf16x8 a0 = lds8(in + arow + g * 8u), a1 = lds8(in + arow + 32u + g * 8u);
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        if (L == 3 && nb == 3) continue;  // the last layer has 40 outputs
        const uint16_t* wr = woi + L * 4096 + (nb * 16 + r) * 64;
        tg_f32x4 c = mfma(a0, ld8(wr + g * 8u), tg_f32x4{0.0f, 0.0f, 0.0f, 0.0f});
        c = mfma(a1, ld8(wr + 32u + g * 8u), c);
        const uint32_t col = nb * 16 + r;
        const float bv = bias[L * 64 + col];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const uint32_t row = 16u * wv + 4u * g + i;
          const float v = c[i] + bv;
          if (L < 3)
            ((uint16_t*)hin[L + 1])[row * TF_LD + col] = tg::f2bf(v > 0.0f ? v : 0.0f);
          else if (col < (uint32_t)TG_OUT)
            tO[row * TG_OUT + col] = v;
        }
      }
      __syncthreads();
    }
    // ---- loss gradient: 4 lanes per sample, lobes 2p and 2p + 1
    {
      const uint32_t rr = 16u * wv + (l >> 2), p = l & 3u, gr = c0 + rr;
      const bool live = gr < r1 && S > 0.0f;
      const float* o = tO + rr * TG_OUT;
      const int k0 = 2 * (int)p;
      const float lg0 = o[32 + k0], lg1 = o[33 + k0];
      const float mx = quad_max(th::fmax_(lg0, lg1));
      const float e0 = tg::fexp(lg0 - mx), e1 = tg::fexp(lg1 - mx);
      const float esum = quad_sum(e0 + e1);
      const float wk[2] = {e0 / esum, e1 / esum};
      const v3 om = live ? th::mk(dirs[3u * gr], dirs[3u * gr + 1u], dirs[3u * gr + 2u]) : th::mk(0.0f, 0.0f, 1.0f);
      v3 mu[2];
      float len[2], kap[2], f[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int k = k0 + j;
        const v3 rv = th::mk(o[3 * k], o[3 * k + 1], o[3 * k + 2]);
        len[j] = th::fsqrt(th::dot(rv, rv) + 1e-8f);
        mu[j] = th::divs(rv, len[j]);
        kap[j] = th::fmin_(tg::softplus(o[24 + k]) + 0.01f, 1000.0f);
        f[j] = tg::vmf_pdf(mu[j], kap[j], om);
      }
      const float q = th::fmax_(quad_sum(wk[0] * f[0] + wk[1] * f[1]), 1e-30f);
      const float scale = live ? -w[gr] / S : 0.0f;
      if (p == 0u && live) lsum += scale * tg::flog(q);
      uint16_t* dz = tDa + rr * TF_LD;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int k = k0 + j;
        const float gk = wk[j] * f[j] / q;
        const float c = th::dot(mu[j], om);
        const v3 dr = th::muls(th::sub(om, th::muls(mu[j], c)), gk * kap[j] / len[j]);
        const float e2 = tg::fexp(-2.0f * kap[j]);
        const float dk = gk * (1.0f / kap[j] - 2.0f * e2 / (1.0f - e2) + (c - 1.0f));
        const float sig = kap[j] >= 1000.0f ? 0.0f : 1.0f / (1.0f + tg::fexp(-o[24 + k]));
        dz[3 * k] = tg::f2bf(scale * dr.x);
        dz[3 * k + 1] = tg::f2bf(scale * dr.y);
        dz[3 * k + 2] = tg::f2bf(scale * dr.z);
        dz[24 + k] = tg::f2bf(scale * dk * sig);
        dz[32 + k] = tg::f2bf(scale * (gk - wk[j]));
      }
#pragma unroll
      for (int j = 0; j < 6; ++j) dz[TG_OUT + 6 * p + j] = 0;
    }
    __syncthreads();
    // ---- backward, layer by layer
    uint16_t* dcur = tDa;
    uint16_t* dnext = tDb;
#pragma unroll
    for (int L = 3; L >= 0; --L) {
      // weight gradient: dW[n][k] += sum_rows dZ[row][n] hin[L][row][k], n = 16 wv + r
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        uint16_t av[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) av[j] = dcur[(ks * 32 + g * 8 + j) * TF_LD + 16u * wv + r];
        const tg_bf16x8 A = __builtin_bit_cast(tg_bf16x8, *(uint4*)av);
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
          uint16_t bv[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) bv[j] = hin[L][(ks * 32 + g * 8 + j) * TF_LD + kb * 16 + r];
          acc[L][kb] = mfma(A, __builtin_bit_cast(tg_bf16x8, *(uint4*)bv), acc[L][kb]);
        }
      }
#pragma unroll
      for (int j = 0; j < 16; ++j) dbias[L] += tg::bf2f(dcur[((l >> 4) * 16 + j) * TF_LD + 16u * wv + (l & 15u)]);
      if (L > 0) {  // dZ of the layer below: (dZ W) gated by the ReLU of this layer's input
        const tg_bf16x8 a0 = lds8(dcur + arow + g * 8u), a1 = lds8(dcur + arow + 32u + g * 8u);
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
          const uint16_t* wr = wio + L * 4096 + (nb * 16 + r) * 64;
          tg_f32x4 c = mfma(a0, ld8(wr + g * 8u), tg_f32x4{0.0f, 0.0f, 0.0f, 0.0f});
          c = mfma(a1, ld8(wr + 32u + g * 8u), c);
          const uint32_t col = nb * 16 + r;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const uint32_t row = 16u * wv + 4u * g + i;
            const uint16_t gate = hin[L][row * TF_LD + col];
            dnext[row * TF_LD + col] = tg::f2bf((gate & 0x7fffu) ? c[i] : 0.0f);
          }
        }
        __syncthreads();  // the whole chunk's dZ is written and the old one fully read
        uint16_t* t = dcur;
        dcur = dnext;
        dnext = t;
      }
    }
    __syncthreads();  // tiles are reloaded by the next chunk
  }
  // ---- per-block partials
#pragma unroll
  for (int L = 0; L < 4; ++L) {
    float* out = partials + (size_t)L * 256u * TG_PART + (size_t)blockIdx.x * TG_PART;
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
      for (int i = 0; i < 4; ++i) out[(16u * wv + 4u * g + i) * 64u + kb * 16 + r] = acc[L][kb][i];
    float b = dbias[L];
    b += __shfl_xor(b, 16, 64);
    b += __shfl_xor(b, 32, 64);
    if (l < 16u) out[64 * 64 + 16u * wv + l] = b;
  }
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
  if (l == 0u && lsum != 0.0f) atomicAdd(loss, lsum);
}

// First stage of the fixed-order sum of the per-block weight-gradient partials: group j
// of 16 sums blocks j, j + 16, j + 32, ... of every layer element.
#define TG_GROUPS 16
__global__ void __launch_bounds__(256) k_reduce_partials(const float* __restrict__ partials, uint32_t nblocks,
                                                         float* __restrict__ groups) {
  const uint32_t e = blockIdx.x * 256u + threadIdx.x, j = blockIdx.y;
  if (e >= 4u * TG_PART) return;
  const uint32_t l = e / TG_PART, off = e % TG_PART;
  const float* base = partials + (size_t)l * 256u * TG_PART + off;
  float s = 0.0f;
  for (uint32_t b = j; b < nblocks; b += TG_GROUPS) s += base[(size_t)b * TG_PART];
  groups[(size_t)j * 4u * TG_PART + e] = s;
}

// Adam on the FP32 master parameters (gradient = the sum of the TG_GROUPS group sums,
// or only the gradient when lr == 0), then both padded BF16 copies and biases.
__global__ void k_adam(float* __restrict__ master, float* __restrict__ m1, float* __restrict__ m2,
                       const float* __restrict__ groups, uint32_t ngroups, float lr, float b1,
                       float b2, float eps, float bc1, float bc2, float* __restrict__ grad_out,
                       uint16_t* __restrict__ woi, uint16_t* __restrict__ wio, float* __restrict__ bias) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= TG_NPARAMS) return;
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
  const size_t e = (size_t)l * TG_PART + (is_bias ? (size_t)64 * 64 + n : (size_t)n * 64 + k);
  float gsum = 0.0f;
  for (uint32_t j = 0; j < ngroups; ++j) gsum += groups[(size_t)j * 4u * TG_PART + e];
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

__global__ void k_encode_host(const float* __restrict__ feats, uint32_t count, uint32_t rows,
                              uint16_t* __restrict__ X) {
  uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= rows) return;
  for (int j = 0; j < TG_INP; ++j)
    X[(size_t)b * TG_INP + j] = (b < count && j < TG_IN) ? tg::f2bf(feats[(size_t)b * TG_IN + j]) : 0;
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

// Fused inference over a queue of deferred diffuse vertices, one wave per 16-row tile
// and persistent over the queue.  Row i of the queue is the vertex (p, n, wo) stored in
// queue order (vp, vn, vw).  Lane (r, g) encodes features 12g..12g+11 of row r into an
// LDS tile (each sin/cos pair once); the four layers then run on MFMA with every weight
// fragment held in registers (loaded once per wave) and activations passed through the
// same tile (BF16, row stride 72 elements = 144 bytes, 16-byte aligned rows); each row's
// 40 outputs are written in FP32.
__global__ void __launch_bounds__(64) k_guide_infer(const uint16_t* __restrict__ woi,
                                                    const float* __restrict__ bias,
                                                    const uint32_t* __restrict__ count_ptr,
                                                    const float4* __restrict__ vp,
                                                    const float4* __restrict__ vn,
                                                    const float4* __restrict__ vw, float4 bmin,
                                                    float4 bscale, float* __restrict__ O) {
  __shared__ __align__(16) uint16_t T[16 * 72];
  const uint32_t count = *count_ptr;
  if (blockIdx.x * 16u >= count) return;  // one wave per block: uniform exit
  const uint32_t l = threadIdx.x, r = l & 15u, g = l >> 4;
  tg_bf16x8 W[4][4][2];
  float bv[4][4];
#pragma unroll
  for (int L = 0; L < 4; ++L)
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
      bv[L][nb] = bias[L * 64 + nb * 16 + r];
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) W[L][nb][kb] = ld8(woi + L * 4096 + (nb * 16 + r) * 64 + kb * 32 + g * 8);
    }
  // the next tile's vertex rows are loaded while the current tile runs its layers
  const float4 zero4 = make_float4(0.5f, 0.5f, 0.5f, 0.0f);
  uint32_t base = blockIdx.x * 16u;
  float4 na = zero4, nbv = zero4, nc = zero4;
  if (base + r < count) {
    na = vp[base + r];
    nbv = vn[base + r];
    nc = vw[base + r];
  }
  for (; base < count; base += gridDim.x * 16u) {
    const uint32_t row = base + r;
    const float4 a = na, b = nbv, c = nc;
    const uint32_t nrow = row + gridDim.x * 16u;
    if (nrow < count) {
      na = vp[nrow];
      nbv = vn[nrow];
      nc = vw[nrow];
    }
    v3 p01 = th::mk(0.5f, 0.5f, 0.5f), n = th::mk(0.0f, 1.0f, 0.0f), wo = n;
    if (row < count) {
      p01 = th::mk((a.x - bmin.x) * bscale.x, (a.y - bmin.y) * bscale.y, (a.z - bmin.z) * bscale.z);
      n = th::mk(b.x, b.y, b.z);
      wo = th::mk(c.x, c.y, c.z);
    }
#pragma unroll
    for (int jj = 0; jj < 12; jj += 2) {
      const int j = 12 * (int)g + jj;
      float f0, f1;
      if (j < 6 * TG_POS_OCT) {
        tg::encode_pos_pair(j, p01, &f0, &f1);
      } else {
        f0 = tg::encode_feature(j, p01, n, wo);
        f1 = tg::encode_feature(j + 1, p01, n, wo);
      }
      *(uint32_t*)(T + r * 72 + j) = (uint32_t)tg::f2bf(f0) | ((uint32_t)tg::f2bf(f1) << 16);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) T[r * 72 + 48 + 4 * g + j] = 0;
    __syncthreads();
#pragma unroll
    for (int L = 0; L < 4; ++L) {
      const tg_bf16x8 a0 = __builtin_bit_cast(tg_bf16x8, *(const uint4*)(T + r * 72 + g * 8));
      const tg_bf16x8 a1 = __builtin_bit_cast(tg_bf16x8, *(const uint4*)(T + r * 72 + 32 + g * 8));
      const int nbs = L < 3 ? 4 : 3;  // the last layer has 40 outputs: three 16-wide tiles
      tg_f32x4 acc[4];
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        if (nb < nbs) {
          acc[nb] = mfma(a0, W[L][nb][0], tg_f32x4{0.0f, 0.0f, 0.0f, 0.0f});
          acc[nb] = mfma(a1, W[L][nb][1], acc[nb]);
        }
      }
      __syncthreads();  // every lane has read the tile before it is overwritten
      if (L < 3) {
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            float v = acc[nb][i] + bv[L][nb];
            T[(4 * g + i) * 72 + nb * 16 + r] = tg::f2bf(v > 0.0f ? v : 0.0f);
          }
        __syncthreads();
      } else {
#pragma unroll
        for (int nb = 0; nb < 3; ++nb)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const uint32_t col = nb * 16 + r, rr = base + 4 * g + i;
            if (col < (uint32_t)TG_OUT && rr < count) O[(size_t)rr * TG_OUT + col] = acc[nb][i] + bv[3][nb];
          }
      }
    }
  }
}

uint32_t round_up(uint32_t x, uint32_t m) { return (x + m - 1) / m * m; }

#ifndef TG_INFER_WAVES_PER_CU
#define TG_INFER_WAVES_PER_CU 8u
#endif

}  // namespace

// ====================================================================== guide object
struct th_guide {
  th_context* ctx = nullptr;
  float *d_master = nullptr, *d_m1 = nullptr, *d_m2 = nullptr, *d_bias = nullptr, *d_grad = nullptr;
  uint16_t *d_woi = nullptr, *d_wio = nullptr;
  uint32_t step = 0;
  float lr = 1e-2f, beta1 = 0.9f, beta2 = 0.99f, eps = 1e-8f;
  // batch workspace
  uint32_t cap = 0;
  uint16_t *X0 = nullptr, *H1 = nullptr, *H2 = nullptr, *H3 = nullptr;
  float *O = nullptr, *feat = nullptr, *dirs = nullptr, *w = nullptr, *partials = nullptr, *scal = nullptr;
  float *groups = nullptr, *wpart = nullptr;
  uint32_t nblocks = 0;
  std::vector<void*> ws;
};

static const uint32_t kWgradBlocks = 256;  // partial slots per layer (k_reduce_partials assumes 256)

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
      (rc = al((void**)&g->groups, (size_t)TG_GROUPS * 4 * TG_PART * 4)) ||
      (rc = al((void**)&g->wpart, 256 * 4)) || (rc = al((void**)&g->scal, 64)))
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
static int32_t guide_step(th_guide* g, const uint16_t* X, const float* dirs, const float* w,
                          uint32_t count, float lr) {
  hipStream_t st = g->ctx->stream;
  hipLaunchKernelGGL(k_weight_partial, dim3(256), dim3(256), 0, st, w, count, g->wpart, g->scal);
  // 64-sample chunks, at most kWgradBlocks blocks (one partial slot each)
  const uint32_t rows_per = round_up((count + kWgradBlocks - 1) / kWgradBlocks, 64);
  const uint32_t nb = (count + rows_per - 1) / rows_per;
  float* P = g->partials;
  hipLaunchKernelGGL(k_train_fused, dim3(nb), dim3(256), 0, st, X, dirs, w, count, rows_per, g->wpart, g->d_woi,
                     g->d_wio, g->d_bias, P, g->scal + 1);
  TH_CHECK(hipGetLastError());
  float bc1 = 1.0f, bc2 = 1.0f;
  if (lr > 0.0f) {
    g->step += 1;
    bc1 = 1.0f - powf(g->beta1, (float)g->step);
    bc2 = 1.0f - powf(g->beta2, (float)g->step);
  }
  // partials of layer l: kWgradBlocks slots of TG_PART floats; the first nb are summed
  hipLaunchKernelGGL(k_reduce_partials, dim3((4 * TG_PART + 255) / 256, TG_GROUPS), dim3(256), 0, st, P, nb,
                     g->groups);
  hipLaunchKernelGGL(k_adam, dim3((TG_NPARAMS + 255) / 256), dim3(256), 0, st, g->d_master,
                     g->d_m1, g->d_m2, g->groups, std::min<uint32_t>(nb, TG_GROUPS), lr, g->beta1, g->beta2,
                     g->eps, bc1, bc2, g->d_grad, g->d_woi, g->d_wio, g->d_bias);
  TH_CHECK(hipGetLastError());
  return TH_OK;
}

// ====================================================================== renderer interface
int32_t guide_infer(th_guide* g, const uint32_t* d_count, uint32_t max_rows, const float4* vp,
                    const float4* vn, const float4* vw, const float* bmin, const float* bmax, float* O,
                    hipStream_t st) {
  // persistent: as many one-wave blocks as stay resident (2 per SIMD at this kernel's
  // register count), so each wave loads the weights once and walks many tiles
  uint32_t grid = std::min<uint32_t>((max_rows + 15u) / 16u, TG_INFER_WAVES_PER_CU * (uint32_t)std::max(g->ctx->cu_count, 1));
  if (grid == 0) return TH_OK;
  float4 lo = make_float4(bmin[0], bmin[1], bmin[2], 0.0f), sc;
  float e[3];
  for (int a = 0; a < 3; ++a) e[a] = bmax[a] - bmin[a];
  sc = make_float4(e[0] > 0.0f ? 1.0f / e[0] : 0.0f, e[1] > 0.0f ? 1.0f / e[1] : 0.0f,
                   e[2] > 0.0f ? 1.0f / e[2] : 0.0f, 0.0f);
  hipLaunchKernelGGL(k_guide_infer, dim3(grid), dim3(64), 0, st, g->d_woi, g->d_bias, d_count, vp, vn, vw,
                     lo, sc, O);
  TH_CHECK(hipGetLastError());
  return TH_OK;
}

int32_t guide_train(th_guide* g, const uint16_t* X, const float* dirs, const float* w, uint32_t count) {
  int32_t rc = guide_reserve(g, count);
  if (rc) return rc;
  return guide_step(g, X, dirs, w, count, g->lr);
}

// The last step's loss, copied once (a copy per step would stall the stream each time).
int32_t guide_last_loss(th_guide* g, float* out) {
  TH_CHECK(hipMemcpyAsync(out, g->scal + 1, 4, hipMemcpyDeviceToHost, g->ctx->stream));
  return TH_OK;
}

uint32_t guide_steps(const th_guide* g) { return g->step; }
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
  hipLaunchKernelGGL(k_adam, dim3((TG_NPARAMS + 255) / 256), dim3(256), 0, g->ctx->stream, g->d_master,
                     g->d_m1, g->d_m2, g->groups, 0u, 0.0f, g->beta1, g->beta2, g->eps, 1.0f, 1.0f,
                     (float*)nullptr, g->d_woi, g->d_wio, g->d_bias);
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
      !al((void**)&g->d_wio, 4 * 4096 * 2)) {
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
  void* ps[] = {g->d_master, g->d_m1, g->d_m2, g->d_grad, g->d_bias, g->d_woi, g->d_wio};
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
  hipLaunchKernelGGL(k_encode_host, dim3((rows + 255) / 256), dim3(256), 0, st, g->feat, count, rows, g->X0);
  if ((rc = th_guide_forward_rows(g, g->X0, nullptr, rows, st))) return rc;
  std::vector<float> o((size_t)rows * TG_OUTP);
  TH_CHECK(hipMemcpyAsync(o.data(), g->O, o.size() * 4, hipMemcpyDeviceToHost, st));
  TH_CHECK(hipStreamSynchronize(st));
  for (uint32_t b = 0; b < count; ++b)
    for (int j = 0; j < TG_OUT; ++j) out[(size_t)b * TG_OUT + j] = o[(size_t)b * TG_OUTP + j];
  return TH_OK;
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
  hipLaunchKernelGGL(k_encode_host, dim3((rows + 255) / 256), dim3(256), 0, st, g->feat, count, rows, g->X0);
  if ((rc = guide_step(g, g->X0, g->dirs, g->w, count, update ? g->lr : 0.0f))) return rc;
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

}  // extern "C"
