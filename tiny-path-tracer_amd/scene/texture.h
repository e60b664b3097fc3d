// Textures of the scene-building API (reference headers/texture.h:9-51).
// Data only; texture::value runs on the device (csrc/shade.h).
#ifndef TPT_SCENE_TEXTURE_H
#define TPT_SCENE_TEXTURE_H

#include "perlin_noise.h"
#include "utils.h"
#include "vec3.h"

class texture {
 public:
  virtual ~texture() = default;
};

class constant_texture : public texture {
 public:
  constant_texture() {}
  constant_texture(vec3 c) : color_(c) {}
  vec3 color_;
};

// odd where sin(10x)·sin(10y)·sin(10z) < 0, else even (src/texture.cc:4-16).
class checker_texture : public texture {
 public:
  checker_texture() {}
  checker_texture(texture *t0, texture *t1) : odd_(t0), even_(t1) {}
  texture *odd_ = nullptr;
  texture *even_ = nullptr;
};

// 0.5·(1 + sin(scale·p.z + 10·turb(p))) grey (src/texture.cc:18-25).  The
// reference fixes the turbulence depth at 5 (headers/perlin_noise.h:11);
// `octaves` is an added, defaulted parameter so existing calls are unchanged.
class perlin_noise_texture : public texture {
 public:
  perlin_noise_texture(float scale = 1.0f, int octaves = 5) : scale_(scale), octaves_(octaves) {}
  perlin_noise noise_;
  float scale_;
  int octaves_;
};

// Nearest-texel RGB8 lookup at (u·w, (1−v)·h), clamped (src/texture.cc:27-42).
// The pixel buffer is borrowed, as in the reference; the flattener copies it.
class image_texture : public texture {
 public:
  image_texture() {}
  image_texture(unsigned char *pixels, int width, int height)
      : data_(pixels), width_(width), height_(height) {}
  unsigned char *data_ = nullptr;
  int width_ = 0, height_ = 0;
};

#endif  // TPT_SCENE_TEXTURE_H
