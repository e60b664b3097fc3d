// Named scene registry over scenes.inc, compiled against this build's API.
#include "scenes.h"

#include <cmath>
#include <cstdlib>
#include <exception>
#include <thread>

#include "../scene/camera.h"
#include "../scene/hitable.h"
#include "../scene/hitable_list.h"
#include "../scene/material.h"
#include "../scene/rect_box.h"
#include "../scene/sphere.h"
#include "../scene/texture.h"
#include "../scene/utils.h"

#define TPT_NEW_PERLIN(scale, octaves) (new perlin_noise_texture((scale), (octaves)))
#include "scenes.inc"
#undef TPT_NEW_PERLIN

namespace tpt {

const char *scene_names() {
  return "random_spheres,random_spheres_textured,cornell555,random_spheres_n<N>,ref_cornell_box,"
         "ref_sphere_cornell_box,ref_two_perlin_spheres,ref_light_spheres,ref_two_checker_spheres,box_list,"
         "ref_random_scene,cornell_smoke,ref_oneweek_final,ref_earth";
}

std::vector<unsigned char> noise_image(int width, int height, uint64_t seed) {
  std::vector<unsigned char> img(static_cast<size_t>(width) * height * 3);
  uint64_t x = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  for (size_t i = 0; i < img.size(); i += 8) {
    // splitmix64 step
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    for (size_t k = 0; k < 8 && i + k < img.size(); ++k) img[i + k] = static_cast<unsigned char>(z >> (8 * k));
  }
  return img;
}

namespace {

void make_camera(scene_spec *s, vec3 from, vec3 at, float vfov, float aperture, float focus, float t0 = 0.0f,
                 float t1 = 0.0f) {
  const float aspect = static_cast<float>(s->width) / static_cast<float>(s->height);
  for (int k = 0; k < 3; ++k) s->lookfrom[k] = from[k], s->lookat[k] = at[k];
  s->vfov = vfov;
  s->aperture = aperture;
  s->focus = focus;
  s->t0 = t0;
  s->t1 = t1;
  s->cam.reset(new camera_with_blur(from, at, vec3(0, 1, 0), vfov, aspect, aperture, focus, t0, t1));
}

bool build_now(const std::string &name, uint32_t seed, scene_spec *s, std::string *err) {
  s->name = name;
  if (name == "random_spheres" || name == "random_spheres_textured" || name.rfind("random_spheres_n", 0) == 0) {
    const bool textured = name == "random_spheres_textured";
    int grid = 11;
    if (name.rfind("random_spheres_n", 0) == 0) {
      const long target = std::strtol(name.c_str() + 16, nullptr, 10);
      if (target < 1 || target > 4000000) {
        *err = "random_spheres_n<N>: N must be in [1, 4000000]";
        return false;
      }
      grid = static_cast<int>(std::ceil(std::sqrt(static_cast<double>(target)) / 2.0));
      if (grid < 1) grid = 1;
    }
    if (textured) s->image = noise_image(2048, 1024, seed);
    s->world = tpt_scene_random_spheres(grid, textured ? 1 : 0, textured ? s->image.data() : nullptr, 2048, 1024);
    s->lights = nullptr;
    s->width = s->height = textured ? 1600 : 400;
    s->spp = textured ? 300 : 16;
    s->max_depth = 50;
    s->background = 1.0f;
    const vec3 from(13, 2, 3), at(0, 0, 0);
    make_camera(s, from, at, 20.0f, 0.1f, (from - at).length());
    return true;
  }
  if (name == "cornell555") {
    s->world = tpt_scene_cornell555(&s->lights);
    s->width = s->height = 1200;
    s->spp = 2048;
    s->max_depth = 50;
    make_camera(s, vec3(278, 278, -800), vec3(278, 278, 0), 40.0f, 0.0f, 10.0f);
    return true;
  }
  if (name == "ref_cornell_box") {
    // config.ini:1-13 settings with the reference driver's camera (main.cpp:87-91).
    s->world = tpt_scene_ref_cornell_box(&s->lights);
    s->width = s->height = 400;
    s->spp = 50;
    s->max_depth = 15;
    make_camera(s, vec3(0, 0, 800), vec3(0, 0, 0), 90.0f, 0.1f, 10.0f);
    return true;
  }
  if (name == "ref_sphere_cornell_box") {
    s->world = tpt_scene_ref_sphere_cornell_box(&s->lights);
    s->width = s->height = 400;
    s->spp = 50;
    s->max_depth = 15;
    make_camera(s, vec3(0, 0, 800), vec3(0, 0, 0), 90.0f, 0.1f, 10.0f);
    return true;
  }
  if (name == "ref_two_perlin_spheres" || name == "ref_two_checker_spheres") {
    s->world = name == "ref_two_perlin_spheres" ? tpt_scene_ref_two_perlin_spheres()
                                                : tpt_scene_ref_two_checker_spheres();
    s->width = s->height = 400;
    s->spp = 50;
    s->max_depth = 15;
    s->background = 1.0f;
    make_camera(s, vec3(13, 2, 3), vec3(0, 0, 0), 20.0f, 0.0f, 10.0f);
    return true;
  }
  if (name == "ref_random_scene") {
    // the reference's random_scene with the camera its driver leaves commented
    // (main.cpp:82-84) and the shutter of config.ini's [CAM_MOTION] (0, 1)
    s->world = tpt_scene_ref_random_scene();
    s->width = s->height = 400;
    s->spp = 16;
    s->max_depth = 50;
    s->background = 1.0f;
    const vec3 from(13, 2, 3), at(0, 0, 0);
    make_camera(s, from, at, 20.0f, 0.0f, (from - at).length(), 0.0f, 1.0f);
    return true;
  }
  if (name == "ref_oneweek_final" || name == "ref_earth") {
    // the earth map from the working directory, as the reference's
    // stbi_load("earthmap.jpg") / load_image_texture find it
    int w = 0, h = 0, c = 0;
    unsigned char *px = load_image_texture("earthmap.jpg", w, h, c);
    if (!px) {
      *err = name + ": earthmap.jpg (baseline JPEG or PPM) not found or not decodable in the working directory";
      return false;
    }
    s->image.assign(px, px + static_cast<size_t>(w) * h * 3);
    std::free(px);
    if (name == "ref_earth") {
      s->world = tpt_scene_ref_earth(s->image.data(), w, h);
      s->width = s->height = 400;
      s->spp = 64;
      s->max_depth = 50;
      s->background = 1.0f;
      const vec3 from(13, 2, 3), at(0, 0, 0);
      make_camera(s, from, at, 20.0f, 0.0f, (from - at).length());
      return true;
    }
    s->world = tpt_scene_ref_oneweek_final(s->image.data(), w, h, &s->lights);
    s->width = s->height = 400;
    s->spp = 64;
    s->max_depth = 50;
    make_camera(s, vec3(478, 278, -600), vec3(278, 278, 0), 40.0f, 0.0f, 10.0f, 0.0f, 1.0f);
    return true;
  }
  if (name == "cornell_smoke") {
    s->world = tpt_scene_cornell_smoke(&s->lights);
    s->width = s->height = 400;
    s->spp = 64;
    s->max_depth = 50;
    make_camera(s, vec3(278, 278, -800), vec3(278, 278, 0), 40.0f, 0.0f, 10.0f);
    return true;
  }
  if (name == "ref_light_spheres") {
    s->world = tpt_scene_ref_light_spheres(&s->lights);
    s->width = s->height = 400;
    s->spp = 50;
    s->max_depth = 15;
    make_camera(s, vec3(13, 2, 3), vec3(0, 0, 0), 40.0f, 0.0f, 10.0f);
    return true;
  }
  if (name == "box_list") {
    s->world = tpt_scene_box_list();
    s->width = s->height = 256;
    s->spp = 16;
    s->max_depth = 20;
    s->background = 1.0f;
    make_camera(s, vec3(0, 400, -700), vec3(0, 60, 0), 50.0f, 0.0f, 10.0f);
    return true;
  }
  *err = "unknown scene '" + name + "'";
  return false;
}

}  // namespace

bool build_scene(const std::string &name, uint32_t seed, scene_spec *out, std::string *err) {
  bool ok = false;
  std::string msg;
  std::thread worker([&]() {
    try {
      drand_seed(seed);
      ok = build_now(name, seed, out, &msg);
    } catch (const std::exception &e) {
      msg = e.what();
      ok = false;
    }
  });
  worker.join();
  if (!ok && err) *err = msg;
  return ok;
}

}  // namespace tpt
