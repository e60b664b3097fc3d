// Host utilities of the scene-building API (reference headers/utils.h).
//
// drand_r() keeps the reference's generator exactly (src/utils.cc:28-32): a
// default-seeded (5489) thread_local std::mt19937 drawn through
// std::uniform_real_distribution<double>, so scene code and the BVH build make
// the same draws as the reference.  tpt::drand_seed() reseeds the calling
// thread's generator (an addition; the reference has no way to reseed).
#ifndef TPT_SCENE_UTILS_H
#define TPT_SCENE_UTILS_H

#include <cstdint>
#include <string>

#include "vec3.h"

double drand_r(double min = 0.0, double max = 1.0);

inline float float_min(float a, float b) { return a < b ? a : b; }
inline float float_max(float a, float b) { return a > b ? a : b; }

// Loads an 8-bit image for image_texture.  The reference uses stb_image
// (src/utils.cc:236-240); this build decodes baseline JPEG and PPM (P6, P3) to
// RGB (channels = 3) and returns nullptr for other files.  The buffer is
// malloc'ed and owned by the caller.
unsigned char *load_image_texture(std::string filename, int &width, int &height, int &channels);

namespace tpt {
void drand_seed(uint32_t seed);
}  // namespace tpt

#endif  // TPT_SCENE_UTILS_H
