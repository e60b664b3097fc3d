// AABB — axis-aligned box with public min_/max_ (reference headers/aabb.h:7-21).
// The per-ray slab test runs on the device only (csrc/traverse.h); the host
// keeps the box data and surrounding_box(), which the BVH build needs.
#ifndef TPT_SCENE_AABB_H
#define TPT_SCENE_AABB_H

#include "vec3.h"

class AABB {
 public:
  AABB() {}
  AABB(const vec3 &a, const vec3 &b) : min_(a), max_(b) {}
  vec3 min() const { return min_; }
  vec3 max() const { return max_; }
  vec3 min_;
  vec3 max_;
};

// Component-wise union with `a < b ? a : b` selection (reference
// src/aabb.cc:21-30 via float_min/float_max, headers/utils.h:23-24).
AABB surrounding_box(AABB box0, AABB box1);

#endif  // TPT_SCENE_AABB_H
