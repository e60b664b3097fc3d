// hitable_list — ordered list of hitables (reference headers/hitable_list.h:4-19).
// Intersection keeps the list's shrinking-t_max semantics on the device; used
// both as scene geometry and as the light list handed to tpt_render().
#ifndef TPT_SCENE_HITABLE_LIST_H
#define TPT_SCENE_HITABLE_LIST_H

#include "hitable.h"

class hitable_list : public hitable {
 public:
  hitable_list() {}
  hitable_list(hitable **l, int n) : list_(l), list_size_(n) {}
  bool bounding_box(float t0, float t1, AABB &box) const override;
  hitable **list_ = nullptr;
  int list_size_ = 0;
};

#endif  // TPT_SCENE_HITABLE_LIST_H
