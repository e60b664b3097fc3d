// ORACLE — test infrastructure only.  Nothing in the product may include,
// link or call anything under oracle/; only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg use it, and only as the checker.
//
// A CPU restatement of the reference's render path (BlurryLight/
// tiny-path-tracer), evaluated directly on the object graph built by the
// scene API (tiny-path-tracer_amd/scene), by recursive virtual-style
// dispatch, with the reference's exact float/double arithmetic:
//   hit()      bvh_node / AABB / hitable_list / sphere / moving_sphere / rects /
//              flip_normal / box / translate / rotate_y / constant_medium
//              (src/hitable.cc:63-128,
//              src/aabb.cc:3-19, src/hitable_list.cc:38-51, src/sphere.cc:13-75,
//              src/rect_box.cc:8-24,51-68,75-91,117-119,171-195,
//              headers/rect_box.h:50-57,87-95)
//   color()    src/utils.cc:58-94 with materials (src/material.cc), pdfs
//              (headers/utils.h:72-98, src/utils.cc:427-456,
//              headers/rect_box.h:118-131, src/rect_box.cc:26-43,
//              src/sphere.cc:93-120, src/hitable_list.cc:24-36), textures
//              (src/texture.cc, src/utils.cc:160-225), camera
//              (src/camera.cc:23-31) and the sample loop (main.cpp:115-142).
// Compiled with g++ -O2 -ffp-contract=off (no fast-math), like the reference.
//
// Pinned by: the reference's own code compiled from /root/reference
// (oracle/_ref, bit-exact hits, camera fields, BVH topology, Perlin noise
// with the reference's tables installed), closed-form
// and special-case tests, and brute-force cross-checks (tests/test_oracle_*.py).
#ifndef TPT_ORACLE_ORACLE_H
#define TPT_ORACLE_ORACLE_H

#include <cstdint>
#include <map>
#include <vector>

#include "../tiny-path-tracer_amd/scene/camera.h"
#include "../tiny-path-tracer_amd/scene/hitable.h"
#include "philox.h"

class material;
class texture;

namespace oracle {

// Random source for color(): drand_r() in the reference.
struct rng {
  virtual ~rng() = default;
  virtual double next() = 0;
  virtual void bounce(uint32_t) {}
  // Draw k (0: free path, 1-3: normal) of constant_medium `leaf` in the
  // current query.  The reference draws them from the same sequential
  // generator (src/hitable.cc:114, 121); the Philox streams key them by leaf.
  virtual double medium(uint32_t leaf, int k) {
    (void)leaf;
    (void)k;
    return next();
  }
};

struct graph_info {
  std::map<const hitable *, uint32_t> leaf_count;    // leaf ids under a node
  std::map<const material *, uint32_t> material_id;  // first-seen DFS order
};

// Leaf ids and material ids in the flattener's numbering (depth-first order,
// an n==1 bvh_node's duplicate child counted once).
void index_graph(const hitable *world, graph_info *info);

// hitable::hit restated; *leaf receives the leaf id of the winning primitive.
// R supplies constant_medium's draws (scenes without media may pass null).
bool hit(const hitable *h, const graph_info &g, const ray &r, float t_min, float t_max, hit_record &rec,
         uint32_t base, uint32_t *leaf, rng *R = nullptr);

// AABB::hit (src/aabb.cc:3-19).
bool slab_hit(const AABB &box, const ray &r, float tmin, float tmax);

struct render_opts {
  int max_depth = 50;
  float background = 0.0f;  // sky gradient scale for misses (0 = black)
  bool media_scatter = false;  // isotropic scatters (the book's intent) instead of ending the path
};

// color() (src/utils.cc:58-94).  lights may be null (cosine-only extension).
vec3 color(const ray &r, const hitable *world, const graph_info &g, const hitable *lights, int depth,
           const render_opts &o, rng &R);

// camera_with_blur::get_ray (src/camera.cc:23-31).
ray camera_ray(const camera_with_blur &cam, float s, float t, rng &R);

// One bounce of color() without the recursion: the ray the material scatters
// for hit `rec` (mixture/cosine sampling for lambertian, as color() draws it).
// Returns false when the path would end there.  Used to build secondary rays.
bool scatter_once(const ray &r, const hit_record &rec, const graph_info &g, const hitable *lights, rng &R,
                  ray *out);

// perlin_noise::noise and perlin_noise::turb (src/utils.cc:160-225) over the
// scene API's static tables, and texture::value for any texture
// (src/texture.cc).
float perlin_noise_at(const vec3 &p);
float perlin_turb(const vec3 &p, int depth);
vec3 texture_value(const texture *t, float u, float v, const vec3 &p);

}  // namespace oracle

#endif  // TPT_ORACLE_ORACLE_H
