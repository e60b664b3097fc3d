// ORACLE — test infrastructure only.  Nothing in the product may include,
// link or call anything under oracle/.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11), restated from the paper: 10 rounds of two 32x32->64
// multiplies with M0 = 0xD2511F53, M1 = 0xCD9E8D57 and Weyl key increments
// W0 = 0x9E3779B9, W1 = 0xBB67AE85.  Pinned by the Random123 known-answer
// vectors in tests/test_oracle_rng.py.
#ifndef TPT_ORACLE_PHILOX_H
#define TPT_ORACLE_PHILOX_H

#include <cstdint>

namespace oracle {

struct philox_out {
  uint32_t v[4];
};

inline philox_out philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
  for (int round = 0; round < 10; ++round) {
    if (round > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
    const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
    const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  return philox_out{{c0, c1, c2, c3}};
}

// Uniform in [0,1) with 24 random bits: exactly representable as float.
inline float u01(uint32_t x) { return static_cast<float>(x >> 8) * (1.0f / 16777216.0f); }

// Render stream: draw k of bounce b of sample s of pixel p uses counter
// (p, s, b, k/4), word k%4, key (seed_lo, seed_hi).  Bounce 0 is the camera
// (jitter, lens); bounce d+1 is color() at depth d.
struct path_stream {
  uint32_t pixel, sample, bounce, k0, k1;
  uint32_t draw = 0;
  uint32_t cache[4] = {0, 0, 0, 0};
  double next() {
    if ((draw & 3u) == 0) {
      const philox_out o = philox4x32_10(pixel, sample, bounce, draw >> 2, k0, k1);
      for (int i = 0; i < 4; ++i) cache[i] = o.v[i];
    }
    return static_cast<double>(u01(cache[draw++ & 3u]));
  }
  void set_bounce(uint32_t b) {
    bounce = b;
    draw = 0;
  }
};

}  // namespace oracle

#endif  // TPT_ORACLE_PHILOX_H
