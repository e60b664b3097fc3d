// perlin_noise — 256-entry gradient and permutation tables
// (reference headers/perlin_noise.h:7-20, src/perlin_noise.cc:3-21).
//
// As in the reference the tables are static: every construction refills them
// (256×3 drand_r draws for the gradients, then three std::shuffle passes), so
// all Perlin textures of a scene share the tables of the last one built.  The
// reference seeds the shuffle engine from the clock; here the seed is
// shuffle_seed() (settable, default 0x5EED) so renders are reproducible.
#ifndef TPT_SCENE_PERLIN_NOISE_H
#define TPT_SCENE_PERLIN_NOISE_H

#include <array>
#include <cstdint>

#include "vec3.h"

class perlin_noise {
 public:
  perlin_noise();
  static std::array<vec3, 256> random_vec3_;
  static std::array<int, 256> permute_x_;
  static std::array<int, 256> permute_y_;
  static std::array<int, 256> permute_z_;
  static uint64_t &shuffle_seed();
};

#endif  // TPT_SCENE_PERLIN_NOISE_H
