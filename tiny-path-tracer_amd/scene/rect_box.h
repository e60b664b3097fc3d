// Axis-aligned rectangles, flip_normal, box, translate, rotate_y
// (reference headers/rect_box.h:5-116, src/rect_box.cc).
#ifndef TPT_SCENE_RECT_BOX_H
#define TPT_SCENE_RECT_BOX_H

#include "hitable.h"

// z = k plane, x in [x0,x1], y in [y0,y1].
class xy_rect : public hitable {
 public:
  xy_rect() {}
  xy_rect(float x0, float x1, float y0, float y1, float k, material *mat)
      : mat_ptr_(mat), x0_(x0), x1_(x1), y0_(y0), y1_(y1), k_(k) {}
  bool bounding_box(float t0, float t1, AABB &box) const override;
  material *mat_ptr_ = nullptr;
  float x0_ = 0, x1_ = 0, y0_ = 0, y1_ = 0, k_ = 0;
};

// y = k plane, x in [x0,x1], z in [z0,z1].  The only rect with a light-sampling
// pdf in the reference (src/rect_box.cc:26-43).
class xz_rect : public hitable {
 public:
  xz_rect() {}
  xz_rect(float x0, float x1, float z0, float z1, float k, material *mat)
      : mat_ptr_(mat), x0_(x0), x1_(x1), z0_(z0), z1_(z1), k_(k) {}
  bool bounding_box(float t0, float t1, AABB &box) const override;
  material *mat_ptr_ = nullptr;
  float x0_ = 0, x1_ = 0, z0_ = 0, z1_ = 0, k_ = 0;
};

// x = k plane, y in [y0,y1], z in [z0,z1].
class yz_rect : public hitable {
 public:
  yz_rect() {}
  yz_rect(float y0, float y1, float z0, float z1, float k, material *mat)
      : mat_ptr_(mat), y0_(y0), y1_(y1), z0_(z0), z1_(z1), k_(k) {}
  bool bounding_box(float t0, float t1, AABB &box) const override;
  material *mat_ptr_ = nullptr;
  float y0_ = 0, y1_ = 0, z0_ = 0, z1_ = 0, k_ = 0;
};

// Negates the normal of whatever it wraps (reference headers/rect_box.h:44-60).
class flip_normal : public hitable {
 public:
  flip_normal(hitable *p) : ptr_(p) {}
  bool bounding_box(float t0, float t1, AABB &box) const override {
    return ptr_->bounding_box(t0, t1, box);
  }
  hitable *ptr_;
};

// Six rects in a hitable_list, in the order +z, -z(flipped), +y, -y(flipped),
// +x, -x(flipped) (reference src/rect_box.cc:93-115).
class box : public hitable {
 public:
  box() {}
  box(vec3 pmin, vec3 pmax, material *mat);
  bool bounding_box(float, float, AABB &b) const override {
    b = AABB(point_min_, point_max_);
    return true;
  }
  vec3 point_min_, point_max_;
  hitable *list_ptr_ = nullptr;
};

class translate : public hitable {
 public:
  translate(hitable *p, const vec3 &offset) : ptr_(p), offset_(offset) {}
  bool bounding_box(float t0, float t1, AABB &box) const override;
  hitable *ptr_;
  vec3 offset_;
};

// Rotation about +y by `angle` degrees.  sin/cos and the rotated bbox are fixed
// at construction exactly as the reference does (src/rect_box.cc:121-169) and
// are copied bit-for-bit to the device.
class rotate_y : public hitable {
 public:
  rotate_y(hitable *p, float angle);
  bool bounding_box(float, float, AABB &box) const override {
    box = box_;
    return has_box_;
  }
  hitable *ptr_;
  float sin_theta_ = 0.0f;
  float cos_theta_ = 1.0f;
  bool has_box_ = false;
  AABB box_;
};

#endif  // TPT_SCENE_RECT_BOX_H
