// sphere / moving_sphere (reference headers/sphere.h:6-40).
#ifndef TPT_SCENE_SPHERE_H
#define TPT_SCENE_SPHERE_H

#include "hitable.h"

class sphere : public hitable {
 public:
  sphere() {}
  sphere(vec3 center, float radius, material *m) : center_(center), radius_(radius), mat_ptr_(m) {}
  bool bounding_box(float t0, float t1, AABB &box) const override;
  vec3 center_;
  float radius_ = 0.0f;
  material *mat_ptr_ = nullptr;
};

// Motion blur (headers/sphere.h:23-40): flattened to an MSPHERE slot whose
// centre the device evaluates at the ray's time with center() below.
class moving_sphere : public hitable {
 public:
  moving_sphere() {}
  moving_sphere(vec3 center0, vec3 center1, float t0, float t1, float radius, material *m)
      : center0_(center0), center1_(center1), radius_(radius), mat_ptr_(m), time0_(t0), time1_(t1) {}
  vec3 center(float time) const {
    return center0_ + (time - time0_) / (time1_ - time0_) * (center1_ - center0_);
  }
  bool bounding_box(float t0, float t1, AABB &box) const override;
  vec3 center0_, center1_;
  float radius_ = 0.0f;
  material *mat_ptr_ = nullptr;
  float time0_ = 0.0f, time1_ = 0.0f;
};

#endif  // TPT_SCENE_SPHERE_H
