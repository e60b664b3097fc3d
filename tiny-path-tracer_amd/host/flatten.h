// Flattener: turns the scene API's object graph (world + light list) into the
// device arrays of tpt_flat_scene (include/tpt.h).
//
// It walks the graph the way the reference's hitable::hit recursion does
// (bvh_node: left then right, src/hitable.cc:63-90; hitable_list: in order,
// src/hitable_list.cc:38-51; box -> its six-rect list, src/rect_box.cc:117-119;
// translate / rotate_y / flip_normal wrappers, headers/rect_box.h:50-95,
// src/rect_box.cc:171-195) and records, for every primitive occurrence, the
// data the device needs to repeat that recursion's arithmetic exactly.
// Boxes, sines/cosines and primitive parameters are copied, never recomputed.
#ifndef TPT_HOST_FLATTEN_H
#define TPT_HOST_FLATTEN_H

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/tpt.h"

class hitable;
class camera_with_blur;

namespace tpt {

struct flat_scene {
  std::vector<tpt_slot> slots;
  std::vector<uint32_t> segments;
  std::vector<tpt_entry> entries;
  std::vector<tpt_xform> xforms;
  std::vector<tpt_material> materials;
  std::vector<tpt_texture> textures;
  std::vector<uint8_t> texels;
  std::vector<tpt_perlin> perlin;  // 0 or 1 element
  std::vector<tpt_light> lights;
  std::vector<tpt_motion> motions;
  std::vector<uint32_t> segment_xforms;  // per segment; empty when no segment is wrapped
  uint32_t light_mode = 0;
  uint32_t n_leaves = 0;
  uint32_t max_depth = 0;
  float bounds[6] = {0, 0, 0, 1, 1, 1};

  tpt_flat_scene view() const;
};

// Flattens `world` and the light list `lights` (may be null: no light
// sampling).  Returns TPT_OK or a negative tpt_status with *err set.
int flatten(const hitable *world, const hitable *lights, flat_scene *out, std::string *err);

tpt_camera camera_fields(const camera_with_blur &cam);

}  // namespace tpt

#endif  // TPT_HOST_FLATTEN_H
