// hitable, bvh_node, constant_medium — the geometry interface of the
// scene-building API (reference headers/hitable.h:14-69).
//
// A hitable here is scene *data*: constructors, public members and
// bounding_box() behave as in the reference, so scene code compiles and builds
// the same object graph (same BVH topology, same boxes).  The per-ray methods
// of the reference (hit, pdf_value, random) are not host functions: the
// graph is flattened (host/flatten.cc) and intersected by the HIP kernels,
// reached through tpt_trace_closest() / tpt_render() in include/tpt.h.
#ifndef TPT_SCENE_HITABLE_H
#define TPT_SCENE_HITABLE_H

#include "aabb.h"
#include "ray.h"
#include "texture.h"

class material;

// Result record of a closest-hit query (reference headers/hitable.h:14-21).
// Filled on the host from tpt_trace_closest() output by tpt::trace().
struct hit_record {
  float t = 0.0f;
  float u = 0.0f;
  float v = 0.0f;
  vec3 point;
  vec3 normal;
  material *mat_ptr = nullptr;
};

class hitable {
 public:
  virtual ~hitable() = default;
  virtual bool bounding_box(float t0, float t1, AABB &box) const = 0;
};

// Binary BVH node.  Construction draws the split axis from drand_r(0,3),
// sorts the input by bbox.min[axis] with std::sort and splits at n/2
// (reference src/hitable.cc:10-56); n==1 stores the same child twice.
class bvh_node : public hitable {
 public:
  bvh_node() {}
  bvh_node(hitable **l, int n, float time0, float time1);
  bool bounding_box(float t0, float t1, AABB &box) const override;
  hitable *left_ = nullptr;
  hitable *right_ = nullptr;
  AABB box_;
};

// Participating medium (reference headers/hitable.h:58-69).  Declared so scene
// code that names it compiles; flattening a graph that contains one fails with
// TPT_ERR_UNSUPPORTED.
class constant_medium : public hitable {
 public:
  constant_medium(hitable *boundary, float density, texture *tex);
  bool bounding_box(float t0, float t1, AABB &box) const override {
    return boundary_->bounding_box(t0, t1, box);
  }
  hitable *boundary_;
  float density_;
  material *phase_funcion_;  // spelling kept from the reference API
};

#endif  // TPT_SCENE_HITABLE_H
