// th_guide.h -- the learned path-guiding proposal: feature encoding, the 8-lobe
// von Mises-Fisher mixture decoded from the network output, its density and sampler,
// and the MLP parameter layout.  No counterpart in the reference (its diffuse vertices
// sample the cosine lobe only, wgsl:849-857).
//
// Network: 48 -> 64 -> 64 -> 64 -> 40, ReLU on hidden layers, BF16 weights and
// activations, FP32 accumulation (MFMA v_mfma_f32_16x16x32_bf16), FP32 master weights
// for Adam.  Output o[0..39]: lobe k has direction r_k = o[3k..3k+2] (mean = r/|r|),
// concentration kappa_k = softplus(o[24+k]) + 0.01 and mixture logit o[32+k].
//
// vMF density with the stable normaliser f(w) = kappa / (2 pi (1 - e^{-2 kappa}))
// * exp(kappa (mu.w - 1)); sampling by inversion of the cosine marginal
// w = 1 + log(xi + (1 - xi) e^{-2 kappa}) / kappa (Wenzel Jakob, "Numerically stable
// sampling of the von Mises-Fisher distribution on S^2", 2012).
#pragma once
#include <stdint.h>

#include "th_math.h"

#define TG_IN 48
#define TG_INP 64  // input padded to two MFMA K steps
#define TG_H 64
#define TG_OUT 40
#define TG_OUTP 64  // output padded to the 64-wide layer shape
#define TG_LOBES 8
// master parameter count: W1 64x48, b1 64, W2 64x64, b2, W3 64x64, b3, W4 40x64, b4 40
#define TG_NPARAMS (64 * 48 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 40 * 64 + 40)

namespace tg {
using th::v3;

// ------------------------------------------------------------------ bf16
// Round to nearest even.  The device uses the gfx950 conversion instruction
// (v_cvt_pk_bf16_f32), which gives the same result as the integer rounding below for every
// non-NaN input, denormals and infinities included (compared over all 2^32 patterns on an
// MI355X); a NaN stays a NaN there.
TH_HD uint16_t f2bf(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_bit_cast(uint16_t, (__bf16)f);
#else
  union { float f; uint32_t u; } v;
  v.f = f;
  uint32_t u = v.u;
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf/nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
#endif
}
// Two values packed low | high << 16 (one conversion instruction on the device).
TH_HD uint32_t f2bf2(float lo, float hi) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
  typedef float f32x2_t __attribute__((ext_vector_type(2)));
  const f32x2_t v = {lo, hi};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
#else
  return (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
#endif
}
// IEEE binary16 bits -> float (the network's stored outputs)
TH_HD float h2f(uint16_t h) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (float)__builtin_bit_cast(_Float16, h);
#else
  const uint32_t sgn = (uint32_t)(h & 0x8000u) << 16, ex = (h >> 10) & 0x1fu, man = h & 0x3ffu;
  union { float f; uint32_t u; } v;
  if (ex == 0x1fu) {
    v.u = sgn | 0x7f800000u | (man << 13);
  } else if (ex != 0u) {
    v.u = sgn | ((ex + 112u) << 23) | (man << 13);
  } else {  // zero or subnormal: man * 2^-24
    v.f = (float)man * 5.9604645e-8f;
    v.u |= sgn;
  }
  return v.f;
#endif
}

TH_HD float bf2f(uint16_t b) {
  union { float f; uint32_t u; } v;
  v.u = (uint32_t)b << 16;
  return v.f;
}

// ------------------------------------------------------------------ encoding
TH_HD float fexp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __expf(x);
#else
  return expf(x);
#endif
}

// Encoding layout: TG_POS_OCT frequency octaves per position axis (sin and cos each)
// and TG_DIR_BINS one-blob bins per normal / outgoing-direction component.
#ifndef TG_POS_OCT
#define TG_POS_OCT 4
#endif
#ifndef TG_DIR_BINS
#define TG_DIR_BINS 4
#endif
static_assert(6 * TG_POS_OCT + 6 * TG_DIR_BINS == TG_IN, "encoding width");

// Position features j and j + 1 (j even, j < 6 TG_POS_OCT): sin and cos of
// 2 pi 2^k x for axis a = j / (2 TG_POS_OCT), octave k.  The device uses the hardware
// sine/cosine (angles stay below 2 pi 2^TG_POS_OCT, well inside its range); training
// records and inference both encode on the device, so they see the same features.
TH_HD void encode_pos_pair(int j, v3 p01, float* s, float* c) {
  int a = j / (2 * TG_POS_OCT), k = (j % (2 * TG_POS_OCT)) >> 1;
  float x = a == 0 ? p01.x : (a == 1 ? p01.y : p01.z);
  float ang = 6.2831855f * (float)(1 << k) * x;
#if defined(__HIP_DEVICE_COMPILE__)
  __sincosf(ang, s, c);
#else
  *s = sinf(ang);
  *c = cosf(ang);
#endif
}

// Feature j (0..47) of the encoding of a vertex: scene-bounds-normalised position p01
// in [0,1]^3 -> sin/cos(2 pi 2^k x) for k < TG_POS_OCT per axis; the normal and outgoing
// direction components c -> u = (c + 1)/2 spread over TG_DIR_BINS Gaussian bins at
// (b + 1/2)/TG_DIR_BINS with sigma 1/TG_DIR_BINS.  Each feature is computed on its own,
// so lanes can share the work of one row.
TH_HD float encode_feature(int j, v3 p01, v3 n, v3 wo) {
  if (j < 6 * TG_POS_OCT) {
    float s, c;
    encode_pos_pair(j & ~1, p01, &s, &c);
    return (j & 1) ? c : s;
  }
  int jj = j - 6 * TG_POS_OCT;
  int c = jj / TG_DIR_BINS, b = jj % TG_DIR_BINS;
  float v = c == 0 ? n.x : c == 1 ? n.y : c == 2 ? n.z : c == 3 ? wo.x : c == 4 ? wo.y : wo.z;
  float u = 0.5f * (v + 1.0f);
  float d = (u - ((float)b + 0.5f) * (1.0f / (float)TG_DIR_BINS)) * (float)TG_DIR_BINS;
  return fexp(-0.5f * d * d);
}

// The whole row, each sin/cos pair computed once.
TH_HD void encode_row(v3 p01, v3 n, v3 wo, float* out /* TG_IN */) {
  for (int j = 0; j < 6 * TG_POS_OCT; j += 2) encode_pos_pair(j, p01, &out[j], &out[j + 1]);
  for (int j = 6 * TG_POS_OCT; j < TG_IN; ++j) out[j] = encode_feature(j, p01, n, wo);
}

// ------------------------------------------------------------------ mixture
struct Mixture {
  v3 mu[TG_LOBES];
  float kappa[TG_LOBES];
  float w[TG_LOBES];
  float wn[TG_LOBES];  // w_k times the lobe's normaliser kappa / (2 pi (1 - e^{-2 kappa}))
};

TH_HD float softplus(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return x > 15.0f ? x : __logf(1.0f + __expf(x));
#else
  return x > 15.0f ? x : logf(1.0f + expf(x));
#endif
}
TH_HD float flog(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __logf(x);
#else
  return logf(x);
#endif
}

// Reciprocal, reciprocal square root and square root for the guided estimator.  On the
// device these are the hardware approximations (about 1 ulp): the estimator only needs
// the sampling and the pdf to use one decoded mixture, not correctly rounded values.
TH_HD float frcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}
TH_HD float frsq(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / sqrtf(x);
#endif
}
TH_HD float fsqrt_fast(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}

TH_HD void decode(const float* o, Mixture* m) {
  float mx = -3.0e38f;
#pragma unroll
  for (int k = 0; k < TG_LOBES; ++k) mx = th::fmax_(mx, o[32 + k]);
  float sum = 0.0f;
#pragma unroll
  for (int k = 0; k < TG_LOBES; ++k) {
    v3 r = th::mk(o[3 * k], o[3 * k + 1], o[3 * k + 2]);
    m->mu[k] = th::smul(frsq(th::dot(r, r) + 1e-8f), r);
    m->kappa[k] = th::fmin_(softplus(o[24 + k]) + 0.01f, 1000.0f);
    m->w[k] = fexp(o[32 + k] - mx);
    sum += m->w[k];
  }
  const float INV_2PI = 0.15915494f;
  const float rs = frcp(sum);
#pragma unroll
  for (int k = 0; k < TG_LOBES; ++k) {
    m->w[k] = m->w[k] * rs;
    m->wn[k] = m->w[k] * (m->kappa[k] * INV_2PI * frcp(1.0f - fexp(-2.0f * m->kappa[k])));
  }
}

TH_HD float vmf_pdf(v3 mu, float kappa, v3 w) {
  const float INV_2PI = 0.15915494f;
  float norm = kappa * INV_2PI / (1.0f - fexp(-2.0f * kappa));
  return norm * fexp(kappa * (th::dot(mu, w) - 1.0f));
}

TH_HD float mixture_pdf(const Mixture& m, v3 w) {
  float p = 0.0f;
#pragma unroll
  for (int k = 0; k < TG_LOBES; ++k) p += m.wn[k] * fexp(m.kappa[k] * (th::dot(m.mu[k], w) - 1.0f));
  return p;
}

// Orthonormal basis around a unit vector (Duff et al. 2017, branchless).
TH_HD void onb(v3 n, v3* t, v3* b) {
  float s = n.z >= 0.0f ? 1.0f : -1.0f;
  float a = -frcp(s + n.z);
  float bb = n.x * n.y * a;
  *t = th::mk(1.0f + s * n.x * n.x * a, s * bb, -s * n.x);
  *b = th::mk(bb, s + n.y * n.y * a, -n.y);
}

TH_HD v3 vmf_sample(v3 mu, float kappa, float u1, float u2) {
  float w = 1.0f + flog(u1 + (1.0f - u1) * fexp(-2.0f * kappa)) * frcp(kappa);
  w = th::clampf(w, -1.0f, 1.0f);
  float r = fsqrt_fast(th::fmax_(0.0f, 1.0f - w * w));
  float s, c;
  th::sincos_(6.2831855f * u2, &s, &c);
  v3 t, b;
  onb(mu, &t, &b);
  return th::add(th::smul(w, mu), th::add(th::smul(r * c, t), th::smul(r * s, b)));
}

// Lobe k is the first whose running weight sum exceeds u0 (the last when none does).
// The index is found first, then the lobe gathered with value selects at compile-time
// positions, so the decoded mixture stays in registers (a run-time index into it, or a
// select between lobe addresses, would place it in scratch memory).
TH_HD v3 mixture_sample(const Mixture& m, float u0, float u1, float u2) {
  int ks = 0;
  float c = m.w[0];
  bool more = u0 >= c;
#pragma unroll
  for (int k = 1; k < TG_LOBES; ++k) {
    if (more) {
      ks = k;
      c += m.w[k];
      more = u0 >= c;
    }
  }
  v3 mu = m.mu[0];
  float kappa = m.kappa[0];
#pragma unroll
  for (int k = 1; k < TG_LOBES; ++k) {
    const bool s = ks == k;
    mu.x = s ? m.mu[k].x : mu.x;
    mu.y = s ? m.mu[k].y : mu.y;
    mu.z = s ? m.mu[k].z : mu.z;
    kappa = s ? m.kappa[k] : kappa;
  }
  return vmf_sample(mu, kappa, u1, u2);
}

}  // namespace tg
