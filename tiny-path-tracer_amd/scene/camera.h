// Thin-lens camera (reference headers/camera.h:6-23, src/camera.cc:2-21).
// The constructor computes the basis on the host with the reference's
// arithmetic; ray generation (get_ray) runs in the device ray-gen kernel,
// which receives these fields through tpt_camera (include/tpt.h).
#ifndef TPT_SCENE_CAMERA_H
#define TPT_SCENE_CAMERA_H

#include "ray.h"
#include "utils.h"

class camera_with_blur {
 public:
  camera_with_blur(vec3 lookfrom, vec3 lookat, vec3 vup, float vfov, float aspect, float aperture,
                   float focus_dist, float t0, float t1);
  vec3 origin_;
  vec3 lower_left_corner_;
  vec3 vertical_;
  vec3 horizontal_;
  vec3 u_, v_, w_;
  float lens_radius_;
  float time0, time1;
};

using camera = camera_with_blur;

#endif  // TPT_SCENE_CAMERA_H
