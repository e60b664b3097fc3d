// Named scenes: the object graph, its light list, camera and default render
// settings, built on a fresh thread with a seeded drand_r generator.
#ifndef TPT_HOST_SCENES_H
#define TPT_HOST_SCENES_H

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../scene/camera.h"

class hitable;

namespace tpt {

struct scene_spec {
  std::string name;
  hitable *world = nullptr;
  hitable *lights = nullptr;  // null: no light sampling (cosine-only)
  std::unique_ptr<camera_with_blur> cam;
  uint32_t width = 0, height = 0, spp = 0, max_depth = 50;
  float background = 0.0f;
  // the camera's construction parameters (a driver rebuilds it with its own fov / aspect)
  float lookfrom[3] = {0, 0, 0}, lookat[3] = {0, 0, -1}, vfov = 90.0f, aperture = 0.0f, focus = 1.0f;
  float t0 = 0.0f, t1 = 0.0f;  // shutter
  std::vector<unsigned char> image;  // texels owned for image_texture
};

// Comma-separated list of scene names.
const char *scene_names();

// Builds `name` with drand_r seeded by `seed` on a fresh thread.  Returns
// false with *err set for an unknown name or a construction error.
bool build_scene(const std::string &name, uint32_t seed, scene_spec *out, std::string *err);

// Seed-generated RGB8 uniform noise image (splitmix64 of seed and texel index).
std::vector<unsigned char> noise_image(int width, int height, uint64_t seed);

}  // namespace tpt

#endif  // TPT_HOST_SCENES_H
