// Device-side scene view, small vector helpers and the Philox generator.
// gfx950 only; compiled with -ffp-contract=off so every float operation below
// is a single IEEE operation in the written order.
#ifndef TPT_CSRC_DEV_COMMON_H
#define TPT_CSRC_DEV_COMMON_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tpt.h"

#pragma clang fp contract(off)

namespace tptd {

constexpr float kPiF = 3.14159265358979323846f;

struct v3 {
  float x, y, z;
};

__device__ __forceinline__ v3 mk(float x, float y, float z) { return v3{x, y, z}; }
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ v3 operator*(v3 a, v3 b) { return v3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ v3 operator*(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ v3 operator*(float s, v3 a) { return v3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ v3 operator/(v3 a, float s) { return v3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 cross(v3 a, v3 b) {
  return v3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ float len(v3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ v3 unit(v3 a) { return a / len(a); }
__device__ __forceinline__ float comp(v3 a, uint32_t i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Read-only scene arrays on the device (all in HBM; the hot ones are small
// enough to stay L2-resident for the benchmark scenes).
struct scene_view {
  const tpt_slot *slots;
  const uint32_t *segments;
  const tpt_entry *entries;
  const tpt_xform *xforms;
  const tpt_material *materials;
  const tpt_texture *textures;
  const uint8_t *texels;
  const tpt_perlin *perlin;
  const tpt_light *lights;
  const tpt_motion *motions;
  const uint32_t *seg_xf;  // per segment: xform | flip<<31 (TPT_SEGMENT_PLAIN), or null: none wrapped
  uint32_t n_slots, n_segments, n_lights, light_mode;
  float walk_ulp;  // 2^-22 x the largest coordinate magnitude of the scene (sphere-only walk bound)
};

// --------------------------------------------------------------- Philox ---
// Philox4x32-10 (Salmon et al., SC'11).  Render streams: draw k of bounce b of
// sample s of pixel p is word k%4 of philox(ctr = (p, s, b, k/4), key = seed).
// Each round's two 32x32 -> 64-bit products are written as 64-bit multiplies
// of zero-extended words, one v_mad_u64_u32 each instead of a separate
// v_mul_lo_u32 and v_mul_hi_u32 (same bits).
__device__ __forceinline__ uint4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                        uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
    const uint32_t lo0 = static_cast<uint32_t>(p0), hi0 = static_cast<uint32_t>(p0 >> 32);
    const uint32_t lo1 = static_cast<uint32_t>(p1), hi1 = static_cast<uint32_t>(p1 >> 32);
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  return make_uint4(c0, c1, c2, c3);
}

__device__ __forceinline__ float u01(uint32_t x) { return static_cast<float>(x >> 8) * (1.0f / 16777216.0f); }

struct stream {
  uint32_t pixel, sample, bounce, k0, k1, draw;
  uint4 cache;
  __device__ __forceinline__ void init(uint32_t p, uint32_t s, uint32_t b, uint64_t seed) {
    pixel = p;
    sample = s;
    bounce = b;
    k0 = static_cast<uint32_t>(seed);
    k1 = static_cast<uint32_t>(seed >> 32);
    draw = 0;
  }
  __device__ __forceinline__ float next() {
    if ((draw & 3u) == 0) cache = philox(pixel, sample, bounce, draw >> 2, k0, k1);
    const uint32_t w = draw & 3u;
    ++draw;
    const uint32_t x = w == 0 ? cache.x : (w == 1 ? cache.y : (w == 2 ? cache.z : cache.w));
    return u01(x);
  }
};

}  // namespace tptd

#endif  // TPT_CSRC_DEV_COMMON_H
