// ray — origin, direction (not normalised) and time, as in the reference's
// headers/ray.h:5-18.  Host-side value type; the device keeps rays in SoA
// queues instead.
#ifndef TPT_SCENE_RAY_H
#define TPT_SCENE_RAY_H

#include "vec3.h"

class ray {
 public:
  ray() {}
  ray(const vec3 &pos, const vec3 &direction, float time = 0.0f)
      : pos_(pos), direction_(direction), time_(time) {}
  vec3 origin() const { return pos_; }
  vec3 direction() const { return direction_; }
  float time() const { return time_; }
  vec3 point_at_parameter(float t) const { return pos_ + t * direction_; }

  vec3 pos_;
  vec3 direction_;
  float time_ = 0.0f;
};

#endif  // TPT_SCENE_RAY_H
