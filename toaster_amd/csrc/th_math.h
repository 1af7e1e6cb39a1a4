// th_math.h -- scalar/vector math shared by the HIP kernels and the host code.
//
// Every function is a fixed sequence of IEEE-754 binary32 additions, multiplications,
// divisions, square roots and floors.  Compiled with floating-point contraction off
// (hipcc and gcc both get -ffp-contract=off) the same source yields the same bits on
// the MI355X and on the host, which is what lets a CPU restatement of the estimator
// be compared with the GPU image pixel by pixel.
//
// The built-ins follow the WGSL definitions used by the reference kernel
// (shaders/pathtrace_spheres.wgsl): dot = left-to-right sum of products,
// normalize(v) = v / length(v), mix(a, b, t) = a*(1-t) + b*t, fract(x) = x - floor(x).
// Transcendentals (sin, cos, atan2, acos) are this project's own: Cody-Waite
// reduction by pi/2 with a three-part constant and Taylor polynomials of the reduced
// argument (|r| <= pi/4), whose truncation error is below 3e-9 -- the result is
// within a few ulp of the correctly rounded value.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define TH_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define TH_HD static inline
#endif

namespace th {

// ------------------------------------------------------------------ scalars
TH_HD float fsqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_sqrtf(x);  // correctly rounded (hipcc default for HIP)
#else
  return sqrtf(x);
#endif
}
TH_HD float ffloor(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_floorf(x);
#else
  return floorf(x);
#endif
}
TH_HD float fabs_(float x) { return __builtin_fabsf(x); }
TH_HD float fmin_(float a, float b) { return b < a ? b : a; }
TH_HD float fmax_(float a, float b) { return a < b ? b : a; }
TH_HD float clampf(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
TH_HD float fract(float x) { return x - ffloor(x); }
TH_HD bool sign_negative(float x) {
  union { float f; uint32_t u; } v;
  v.f = x;
  return (v.u >> 31) != 0u;
}

// WGSL PI constant (wgsl:387) rounded to binary32.
#define TH_PI 3.14159265359f
#define TH_MIN_DISTANCE 0.001f
#define TH_MAX_DISTANCE 10000.0f
#define TH_EPSILON 1e-8f
#define TH_FLT_MAX 3.4028235e38f

// ------------------------------------------------------------------ vectors
struct v3 {
  float x, y, z;
};
TH_HD v3 mk(float x, float y, float z) {
  v3 r;
  r.x = x;
  r.y = y;
  r.z = z;
  return r;
}
TH_HD v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
TH_HD v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
TH_HD v3 mul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
TH_HD v3 muls(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
TH_HD v3 smul(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
TH_HD v3 divs(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
TH_HD v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
TH_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
TH_HD v3 cross(v3 a, v3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
TH_HD float length(v3 a) { return fsqrt(dot(a, a)); }
TH_HD v3 normalize(v3 a) { return divs(a, length(a)); }
TH_HD v3 mix(v3 a, v3 b, float t) { return add(muls(a, 1.0f - t), muls(b, t)); }
TH_HD float get(v3 a, int axis) { return axis == 0 ? a.x : (axis == 1 ? a.y : a.z); }

// ------------------------------------------------------------------ RNG
// pcg_hash of wgsl:189-193: one PCG32 LCG step (multiplier 747796405, increment
// 2891336453) followed by the RXS-M-XS output permutation (Jarzynski & Olano 2020,
// "Hash Functions for GPU Rendering").
TH_HD uint32_t pcg_hash(uint32_t input) {
  uint32_t state = input * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}
// random_f32 of wgsl:196-199: rehash then map to the inclusive unit interval.
TH_HD float rand01(uint32_t* state) {
  *state = pcg_hash(*state);
  return (float)(*state) / 4294967295.0f;
}

// ------------------------------------------------------------------ sin / cos
// Cody-Waite split of pi/2: C1 has 8 and C2 12 significant bits, so k*C1 and k*C2
// are exact for |k| < 2^11; C3 is the binary32 rounding of the remainder.
#define TH_PIO2_C1 1.5703125f
#define TH_PIO2_C2 4.8375129699707031e-4f
#define TH_PIO2_C3 7.5497901264043321e-08f
#define TH_TWO_OVER_PI 0.63661975f

TH_HD float sin_poly(float r) {  // Taylor to r^9 on |r| <= pi/4
  float r2 = r * r;
  float p = 2.7557319223985893e-06f;     //  1/9!
  p = p * r2 - 1.9841269841269841e-04f;  // -1/7!
  p = p * r2 + 8.3333333333333333e-03f;  //  1/5!
  p = p * r2 - 1.6666666666666667e-01f;  // -1/3!
  return r + (r * r2) * p;
}
TH_HD float cos_poly(float r) {  // Taylor to r^10 on |r| <= pi/4
  float r2 = r * r;
  float p = -2.7557319223985888e-07f;  // -1/10!
  p = p * r2 + 2.4801587301587302e-05f;  //  1/8!
  p = p * r2 - 1.3888888888888889e-03f;  // -1/6!
  p = p * r2 + 4.1666666666666664e-02f;  //  1/4!
  p = p * r2 - 0.5f;                     // -1/2!
  return 1.0f + r2 * p;
}
TH_HD void sincos_(float x, float* s, float* c) {
  float k = ffloor(x * TH_TWO_OVER_PI + 0.5f);
  float r = ((x - k * TH_PIO2_C1) - k * TH_PIO2_C2) - k * TH_PIO2_C3;
  int q = ((int)k) & 3;
  float sr = sin_poly(r), cr = cos_poly(r);
  if (q == 0) {
    *s = sr;
    *c = cr;
  } else if (q == 1) {
    *s = cr;
    *c = -sr;
  } else if (q == 2) {
    *s = -sr;
    *c = -cr;
  } else {
    *s = -cr;
    *c = sr;
  }
}
TH_HD float fsin(float x) {
  float s, c;
  sincos_(x, &s, &c);
  return s;
}
TH_HD float fcos(float x) {
  float s, c;
  sincos_(x, &s, &c);
  return c;
}

// ------------------------------------------------------------------ atan2 / acos
// atan on |w| <= tan(pi/8): alternating Taylor series to w^19 (truncation < 3e-9).
TH_HD float atan_poly(float w) {
  float w2 = w * w;
  float p = -1.0f / 19.0f;
  p = p * w2 + 1.0f / 17.0f;
  p = p * w2 - 1.0f / 15.0f;
  p = p * w2 + 1.0f / 13.0f;
  p = p * w2 - 1.0f / 11.0f;
  p = p * w2 + 1.0f / 9.0f;
  p = p * w2 - 1.0f / 7.0f;
  p = p * w2 + 1.0f / 5.0f;
  p = p * w2 - 1.0f / 3.0f;
  return w + (w * w2) * p;
}
#define TH_PIO4 0.78539819f
#define TH_PIO2 1.5707964f
#define TH_TAN_PIO8 0.41421357f
TH_HD float fatan2(float y, float x) {
  float ax = fabs_(x), ay = fabs_(y);
  float a;
  if (ax == 0.0f && ay == 0.0f) {
    a = sign_negative(x) ? TH_PI : 0.0f;
  } else {
    float mn = fmin_(ax, ay), mx = fmax_(ax, ay);
    float t = mn / mx;  // in [0, 1]
    if (t > TH_TAN_PIO8) {
      a = TH_PIO4 + atan_poly((t - 1.0f) / (t + 1.0f));
    } else {
      a = atan_poly(t);
    }
    if (ay > ax) a = TH_PIO2 - a;
    if (sign_negative(x)) a = TH_PI - a;
  }
  return sign_negative(y) ? -a : a;
}
// acos(x) = atan2(sqrt((1-x)(1+x)), x) for x in [-1, 1].
TH_HD float facos(float x) { return fatan2(fsqrt((1.0f - x) * (1.0f + x)), x); }

// ------------------------------------------------------- environment CDF search
// Index of the first texel whose CDF value exceeds `sel` in a row-major CDF of H rows of
// W texels (the partition point of <= that wgsl:352-362 finds with one binary search
// over all W*H entries), or W*H - 1 when none does; found row first.  rows[j] is the
// CDF value of the last texel of row j.  With a non-decreasing CDF the first row whose
// last value exceeds sel holds the first texel that does (every texel of the earlier
// rows is <= their last value <= sel), and that row's last texel bounds the search.
TH_HD uint32_t env_search_rows(const float* cdf, const float* rows, uint32_t W, uint32_t H, float sel) {
  uint32_t lo = 0, hi = H;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2u;
    if (rows[mid] <= sel)
      lo = mid + 1u;
    else
      hi = mid;
  }
  if (lo >= H) return W * H - 1u;
  const float* row = cdf + (size_t)lo * W;
  uint32_t a = 0, b = W - 1u;  // row[W - 1] = rows[lo] > sel
  while (a < b) {
    const uint32_t mid = a + (b - a) / 2u;
    if (row[mid] <= sel)
      a = mid + 1u;
    else
      b = mid;
  }
  return lo * W + a;
}

}  // namespace th
