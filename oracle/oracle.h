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
// (oracle/_ref): bit-exact hits, camera fields, BVH topology, Perlin noise
// with the reference's tables installed, the samplers, light pdf_value /
// random of spheres, rects and lists, checker and image textures,
// metal::scatter, and color() sample by sample on whole scenes, with the
// reference's own mt19937 draws in the reference build's argument order
// (rng::args_right_to_left); closed-form and special-case tests (slab, roots,
// rects, the sky and cosine-only extensions, media free paths) and
// brute-force cross-checks (tests/test_oracle_*.py).  The guided estimator
// (oracle_guide.cc) has no reference counterpart: its network equals
// oracle/guide_oracle.py's, its sampler draws from the density it reports,
// and it keeps the pinned plain estimator's expectation
// (tests/test_guide_oracle_render.py).
#ifndef TPT_ORACLE_ORACLE_H
#define TPT_ORACLE_ORACLE_H

#include <cstdint>
#include <map>
#include <vector>

#include "../tiny-path-tracer_amd/scene/camera.h"
#include "../tiny-path-tracer_amd/scene/hitable.h"
#include "philox.h"

class material;
class metal;
class texture;

namespace oracle {

// Random source for color(): drand_r() in the reference.
struct rng {
  virtual ~rng() = default;
  virtual double next() = 0;
  // The reference draws the coordinates of some vectors inside one
  // constructor call, vec3(drand_r(), drand_r(), ...) (src/utils.cc:16, 24;
  // src/rect_box.cc:40-41), whose argument order C++ leaves unspecified.  A
  // source that replays the reference build's own draws says which order that
  // build used (true: the last coordinate takes the first draw); the Philox
  // streams of the device take them first coordinate first.
  virtual bool args_right_to_left() const { return false; }
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
  // Guided estimator (oracle_guide.cc; no reference counterpart): network
  // parameters, or null for the reference estimator; the first guide_verts
  // lambertian vertices of a path are guided; box = the scene box that maps
  // hit points to the network's [0,1]^3 input.
  const float *guide = nullptr;
  int guide_verts = 2;
  float box_lo[3] = {0, 0, 0}, box_hi[3] = {1, 1, 1};
};

// color() (src/utils.cc:58-94).  lights may be null (cosine-only extension).
// ndiff: lambertian vertices of the path so far (guided mode).
vec3 color(const ray &r, const hitable *world, const graph_info &g, const hitable *lights, int depth,
           const render_opts &o, rng &R, int ndiff = 0);

// ---- the guided proposal (oracle_guide.cc) ----
struct guide_mix {
  float mu[8][3], kappa[8], w[8], wn[8];  // wn: w times the lobe's normaliser
};
float bf16_round(float f);
// Network outputs (40) for input (p^, n, wo), float64 sums of bf16 operands.
void guide_forward(const float *params, const float in9[9], double out[40]);
// Forward, outputs rounded to bf16, decoded; false when an output is not finite.
bool guide_decode(const float *params, const float in9[9], guide_mix *m);
float guide_alpha(const float *params);  // sigmoid of the mix logit
float guide_mix_pdf(const guide_mix &m, const vec3 &unit_dir);
vec3 guide_mix_sample(const guide_mix &m, float u0, float u1, float u2);

// camera_with_blur::get_ray (src/camera.cc:23-31).
ray camera_ray(const camera_with_blur &cam, float s, float t, rng &R);

// One bounce of color() without the recursion: the ray the material scatters
// for hit `rec` (mixture/cosine sampling for lambertian, as color() draws it).
// Returns false when the path would end there.  Used to build secondary rays.
bool scatter_once(const ray &r, const hit_record &rec, const graph_info &g, const hitable *lights, rng &R,
                  ray *out);

// The samplers and light densities color() uses, for direct comparison with
// the reference's: random_in_unit_sphere / random_in_unit_disk
// (src/utils.cc:13-27), hitable::pdf_value / random of a light shape or list
// (src/hitable_list.cc:24-36, src/rect_box.cc:26-43, src/sphere.cc:93-120),
// metal::scatter (src/material.cc:88-98: false when the path is absorbed).
vec3 unit_sphere_sample(rng &R);
vec3 unit_disk_sample(rng &R);
float light_pdf_value(const hitable *lights, const graph_info &g, const vec3 &o, const vec3 &d);
vec3 light_sample(const hitable *lights, const vec3 &o, rng &R);
bool metal_scatter(const metal *m, const ray &r, const hit_record &rec, rng &R, ray *out);

// perlin_noise::noise and perlin_noise::turb (src/utils.cc:160-225) over the
// scene API's static tables, and texture::value for any texture
// (src/texture.cc).
float perlin_noise_at(const vec3 &p);
float perlin_turb(const vec3 &p, int depth);
vec3 texture_value(const texture *t, float u, float v, const vec3 &p);

}  // namespace oracle

#endif  // TPT_ORACLE_ORACLE_H
