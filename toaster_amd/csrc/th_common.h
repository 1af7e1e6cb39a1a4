// th_common.h -- helpers shared by the host packer and the device runtime.
#pragma once
#include <stdint.h>

#include <string>

#include "../../include/toaster_hip.h"
#include "th_math.h"

// Records the message returned by th_last_error() and passes the status through.
int32_t th_set_error(int32_t code, const std::string& msg);

// Traversal entry of a reference node (bit layout described in th_device.h).
TH_HD uint32_t th_entry_of(const TH_GpuBvhNode& n, uint32_t index) {
  if (n.primitive_count == 0u) return index;
  if (n.primitive_count < 63u && n.first_or_right < (1u << 25))
    return 0x80000000u | (n.primitive_count << 25) | n.first_or_right;
  return 0x80000000u | (63u << 25) | index;
}

// sRGB byte -> linear (wgsl:210-215 applied to f32(b) * (1/255)), evaluated on the
// host so the device reads the same 256 values the host computes.
void th_srgb_table(float out[256]);

// ACES fitted tone curve with the non-finite rules of color.rs:61-70.
TH_HD float th_aces(float c) {
  if (c != c) return 0.0f;
  if (c > 3.4028235e38f) return 1.0f;
  if (c < -3.4028235e38f) return 0.0f;
  c = c < 0.0f ? 0.0f : c;
  return (c * (2.51f * c + 0.03f)) / (c * (2.43f * c + 0.59f) + 0.14f);
}

// (x.clamp(0, 1) * 255).round() as u8 (image_output.rs:203-205); NaN -> 0.
TH_HD uint8_t th_u8(float x) {
  if (!(x >= 0.0f)) x = 0.0f;
  if (x > 1.0f) x = 1.0f;
  float y = x * 255.0f;
  float r = th::ffloor(y);
  if (y - r >= 0.5f) r = r + 1.0f;  // round half away from zero (y >= 0)
  return (uint8_t)r;
}

// Display byte of one tone-mapped channel v in [0, 1]: linear_to_srgb then th_u8.
// The sRGB curve is monotone, so the byte is the number of thresholds t_k <= v where
// t_k (k = 1..255) is the smallest v whose byte is >= k; the thresholds are computed
// on the host with the host's powf (th_srgb_thresholds), which keeps the device
// conversion byte-identical to the host one.
void th_srgb_thresholds(float out[256]);

TH_HD uint8_t th_encode_byte(float v, const float* thr) {
  uint32_t lo = 1, hi = 256;  // first k in [1, 255] with thr[k] > v, or 256
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (mid <= 255u && thr[mid] <= v)
      lo = mid + 1u;
    else
      hi = mid;
  }
  return (uint8_t)(lo - 1u);
}

TH_HD uint8_t th_display_channel(float c, float scale, const float* thr) {
  float v = th_aces(c * scale);
  v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
  return th_encode_byte(v, thr);
}
