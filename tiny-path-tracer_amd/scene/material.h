// Materials of the scene-building API (reference headers/material.h:11-80).
// Data only: scatter/scattering_pdf/emitted are evaluated by the device shade
// kernels (csrc/shade.h) from the flattened material table.
#ifndef TPT_SCENE_MATERIAL_H
#define TPT_SCENE_MATERIAL_H

#include "texture.h"
#include "vec3.h"

class material {
 public:
  virtual ~material() = default;
};

// Diffuse; albedo from a texture; sampled with cosine_pdf (value cos θ).
class lambertian : public material {
 public:
  lambertian(texture *albedo) : albedo_(albedo) {}
  texture *albedo_;
};

// Mirror reflection plus fuzz·(random point in the unit ball).
// fuzz outside [0,1) is stored as 1 (reference headers/material.h:41-46).
class metal : public material {
 public:
  metal(const vec3 &albedo, float fuzz) : albedo_(albedo), fuzz_((fuzz < 1 && fuzz >= 0) ? fuzz : 1.0f) {}
  vec3 albedo_;
  float fuzz_;
};

// Glass with Schlick reflectance.
class dielectric : public material {
 public:
  dielectric(float ri) : ref_idx_(ri) {}
  float ref_idx_;
};

// One-sided emitter: emits where dot(normal, ray dir) < 0; never scatters.
class diffuse_light : public material {
 public:
  diffuse_light(texture *a) : emit_(a) {}
  texture *emit_;
};

// Phase function of constant_medium.  Declared for source compatibility only.
class isotropic : public material {
 public:
  isotropic(texture *t) : albedo_(t) {}
  texture *albedo_;
};

#endif  // TPT_SCENE_MATERIAL_H
