// Host half of the C ABI (include/tpt.h): error reporting and the scene
// library.  The device half lives in csrc/tpt_device.hip.
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/tpt.h"
#include "errors.h"
#include "flatten.h"
#include "image_load.h"
#include "scenes.h"

namespace {
thread_local std::string g_last_error;
}

int tpt::fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

struct tpt_host_scene {
  tpt::scene_spec spec;
  tpt::flat_scene flat;
};

extern "C" {

int tpt_abi_version(void) { return TPT_ABI_VERSION; }

const char *tpt_last_error(void) { return g_last_error.c_str(); }

const char *tpt_scene_names(void) { return tpt::scene_names(); }

int tpt_host_scene_build(const char *name, uint32_t seed, tpt_host_scene **out) {
  if (!name || !out) return tpt::fail(TPT_ERR_INVALID, "tpt_host_scene_build: null argument");
  *out = nullptr;
  auto *hs = new (std::nothrow) tpt_host_scene();
  if (!hs) return tpt::fail(TPT_ERR_NOMEM, "tpt_host_scene_build: out of memory");
  std::string err;
  if (!tpt::build_scene(name, seed, &hs->spec, &err)) {
    delete hs;
    return tpt::fail(TPT_ERR_INVALID, err);
  }
  const int rc = tpt::flatten(hs->spec.world, hs->spec.lights, &hs->flat, &err);
  if (rc != TPT_OK) {
    delete hs;
    return tpt::fail(rc, err);
  }
  *out = hs;
  return TPT_OK;
}

int tpt_host_scene_flat(const tpt_host_scene *hs, tpt_flat_scene *flat) {
  if (!hs || !flat) return tpt::fail(TPT_ERR_INVALID, "tpt_host_scene_flat: null argument");
  *flat = hs->flat.view();
  return TPT_OK;
}

int tpt_host_scene_camera(const tpt_host_scene *hs, tpt_camera *camera) {
  if (!hs || !camera) return tpt::fail(TPT_ERR_INVALID, "tpt_host_scene_camera: null argument");
  *camera = tpt::camera_fields(*hs->spec.cam);
  return TPT_OK;
}

int tpt_host_scene_info(const tpt_host_scene *hs, tpt_scene_info *info) {
  if (!hs || !info) return tpt::fail(TPT_ERR_INVALID, "tpt_host_scene_info: null argument");
  info->width = hs->spec.width;
  info->height = hs->spec.height;
  info->spp = hs->spec.spp;
  info->max_depth = hs->spec.max_depth;
  info->background = hs->spec.background;
  info->n_leaves = hs->flat.n_leaves;
  for (int k = 0; k < 3; ++k) info->lookfrom[k] = hs->spec.lookfrom[k], info->lookat[k] = hs->spec.lookat[k];
  info->vfov = hs->spec.vfov;
  info->aperture = hs->spec.aperture;
  info->focus = hs->spec.focus;
  info->time0 = hs->spec.t0;
  info->time1 = hs->spec.t1;
  return TPT_OK;
}

int tpt_camera_make(const float lookfrom[3], const float lookat[3], const float vup[3], float vfov, float aspect,
                    float aperture, float focus_dist, float t0, float t1, tpt_camera *out) {
  if (!lookfrom || !lookat || !vup || !out) return tpt::fail(TPT_ERR_INVALID, "tpt_camera_make: null argument");
  if (!(aspect > 0.0f)) return tpt::fail(TPT_ERR_INVALID, "tpt_camera_make: aspect must be positive");
  const camera_with_blur cam(vec3(lookfrom[0], lookfrom[1], lookfrom[2]), vec3(lookat[0], lookat[1], lookat[2]),
                             vec3(vup[0], vup[1], vup[2]), vfov, aspect, aperture, focus_dist, t0, t1);
  *out = tpt::camera_fields(cam);
  return TPT_OK;
}

int tpt_image_load(const char *path, uint8_t **out, int *width, int *height) {
  if (!path || !out || !width || !height) return tpt::fail(TPT_ERR_INVALID, "tpt_image_load: null argument");
  *out = nullptr;
  std::vector<uint8_t> rgb;
  std::string err;
  if (!tpt::load_image_rgb8(path, &rgb, width, height, &err)) return tpt::fail(TPT_ERR_INVALID, err);
  if (rgb.size() != static_cast<size_t>(*width) * static_cast<size_t>(*height) * 3)
    return tpt::fail(TPT_ERR_INVALID, "tpt_image_load: decoded size does not match the header");
  *out = static_cast<uint8_t *>(std::malloc(rgb.size()));
  if (!*out) return tpt::fail(TPT_ERR_NOMEM, "tpt_image_load: out of memory");
  std::memcpy(*out, rgb.data(), rgb.size());
  return TPT_OK;
}

void tpt_image_free(uint8_t *pixels) { std::free(pixels); }

int tpt_host_scene_destroy(tpt_host_scene *hs) {
  // The object graph uses raw new and is never freed, like the reference's
  // factories (src/utils.cc:96-414); the flat arrays and texels are released.
  delete hs;
  return TPT_OK;
}

}  // extern "C"
