// ORACLE — test infrastructure only (see oracle.h).
//
// Harness linked with the reference's own, unmodified sources
// (/root/reference/src/*.cc, compiled where they lie by oracle/Makefile) into
// oracle/_ref/libtpt_ref.so.  It builds scenes with the REFERENCE's classes —
// either through the shared scene text tiny-path-tracer_amd/host/scenes.inc or
// through the reference's own factories (cornell_box() etc. in
// src/utils.cc) — and answers hit queries with the reference's hitable::hit
// and renders with its color() in main.cpp's thread-per-row loop
// (main.cpp:109-159).  Used to pin the oracle and as the CPU baseline.
//
// Include paths resolve to the reference's headers (-I .../headers).
#include <atomic>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "camera.h"
#include "hitable.h"
#include "hitable_list.h"
#include "material.h"
#include "rect_box.h"
#include "sphere.h"
#include "texture.h"
#include "utils.h"

#define TPT_NEW_PERLIN(scale, octaves) (new perlin_noise_texture((scale)))
#include "../tiny-path-tracer_amd/host/scenes.inc"
#undef TPT_NEW_PERLIN

hitable *two_checker_spheres();  // defined in src/utils.cc, not declared in its header

namespace {

// Material stand-in that tags each primitive occurrence with its leaf id
// (hit_record has no id field; the proxy rides in rec.mat_ptr).
struct proxy_material : material {
  int leaf = -1;
  int mat_index = -1;
};

struct ref_scene_t {
  hitable *world = nullptr;   // materials replaced by proxies
  hitable *render_world = nullptr;  // original materials
  hitable *lights = nullptr;
  std::vector<unsigned char> image;
};

struct tagger {
  int next_leaf = 0;
  std::map<material *, int> mat_index;
  int index_of(material *m) {
    auto it = mat_index.find(m);
    if (it != mat_index.end()) return it->second;
    const int id = static_cast<int>(mat_index.size());
    mat_index[m] = id;
    return id;
  }
  void tag(material **slot) {
    auto *p = new proxy_material();
    p->leaf = next_leaf++;
    p->mat_index = index_of(*slot);
    *slot = p;
  }
  // Depth-first, left before right, an n==1 node's duplicate child once.
  void walk(hitable *h) {
    if (auto *b = dynamic_cast<bvh_node *>(h)) {
      walk(b->left_);
      if (b->right_ != b->left_) walk(b->right_);
    } else if (auto *l = dynamic_cast<hitable_list *>(h)) {
      for (int i = 0; i < l->list_size_; ++i) walk(l->list_[i]);
    } else if (auto *bx = dynamic_cast<box *>(h)) {
      walk(bx->list_ptr_);
    } else if (auto *f = dynamic_cast<flip_normal *>(h)) {
      walk(f->ptr_);
    } else if (auto *t = dynamic_cast<translate *>(h)) {
      walk(t->ptr_);
    } else if (auto *r = dynamic_cast<rotate_y *>(h)) {
      walk(r->ptr_);
    } else if (auto *s = dynamic_cast<sphere *>(h)) {
      tag(&s->mat_ptr_);
    } else if (auto *ms = dynamic_cast<moving_sphere *>(h)) {
      tag(&ms->mat_ptr_);
    } else if (auto *cm = dynamic_cast<constant_medium *>(h)) {
      tag(&cm->phase_funcion_);  // the medium is one leaf; its boundary is not walked
    } else if (auto *q = dynamic_cast<xy_rect *>(h)) {
      tag(&q->mat_ptr_);
    } else if (auto *q2 = dynamic_cast<xz_rect *>(h)) {
      tag(&q2->mat_ptr_);
    } else if (auto *q3 = dynamic_cast<yz_rect *>(h)) {
      tag(&q3->mat_ptr_);
    }
  }
};

std::vector<unsigned char> noise_image(int width, int height, uint64_t seed) {
  // Same generator as tiny-path-tracer_amd/host/scenes.cc (splitmix64).
  std::vector<unsigned char> img(static_cast<size_t>(width) * height * 3);
  uint64_t x = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  for (size_t i = 0; i < img.size(); i += 8) {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    for (size_t k = 0; k < 8 && i + k < img.size(); ++k) img[i + k] = static_cast<unsigned char>(z >> (8 * k));
  }
  return img;
}

bool build(const std::string &name, ref_scene_t *s, hitable **world) {
  hitable *lights = nullptr;
  if (name == "cornell555") {
    *world = tpt_scene_cornell555(&lights);
  } else if (name == "random_spheres") {
    *world = tpt_scene_random_spheres(11, 0, nullptr, 0, 0);
  } else if (name == "random_spheres_textured") {
    if (s->image.empty()) s->image = noise_image(2048, 1024, 5489);
    *world = tpt_scene_random_spheres(11, 1, s->image.data(), 2048, 1024);
  } else if (name.rfind("random_spheres_n", 0) == 0) {
    const long target = std::strtol(name.c_str() + 16, nullptr, 10);
    int grid = static_cast<int>(std::ceil(std::sqrt(static_cast<double>(target)) / 2.0));
    if (grid < 1) grid = 1;
    *world = tpt_scene_random_spheres(grid, 0, nullptr, 0, 0);
  } else if (name == "ref_cornell_box") {
    *world = tpt_scene_ref_cornell_box(&lights);
  } else if (name == "ref_sphere_cornell_box") {
    *world = tpt_scene_ref_sphere_cornell_box(&lights);
  } else if (name == "ref_two_perlin_spheres") {
    *world = tpt_scene_ref_two_perlin_spheres();
  } else if (name == "ref_light_spheres") {
    *world = tpt_scene_ref_light_spheres(&lights);
  } else if (name == "ref_two_checker_spheres") {
    *world = tpt_scene_ref_two_checker_spheres();
  } else if (name == "box_list") {
    *world = tpt_scene_box_list();
  } else if (name == "ref_random_scene") {
    *world = tpt_scene_ref_random_scene();
  } else if (name == "cornell_smoke") {
    *world = tpt_scene_cornell_smoke(&lights);
  } else if (name == "ref_oneweek_final" || name == "ref_earth") {
    // stb_image through the reference's own loader, from the working directory
    int w = 0, h = 0, c = 0;
    unsigned char *px = load_image_texture("earthmap.jpg", w, h, c);
    if (!px || c != 3) return false;
    s->image.assign(px, px + static_cast<size_t>(w) * h * 3);
    free(px);
    *world = name == "ref_earth" ? tpt_scene_ref_earth(s->image.data(), w, h)
                                 : tpt_scene_ref_oneweek_final(s->image.data(), w, h, &lights);
  } else if (name == "native:cornell_box") {
    // The reference's own factory and its driver's light list (main.cpp:99-106).
    *world = cornell_box();
    hitable **l = new hitable *[2];
    l[0] = new xz_rect(-100, 100, -150, -50, 298, nullptr);
    l[1] = new sphere(vec3(120, -50, 40), 120, nullptr);
    lights = new hitable_list(l, 2);
  } else if (name == "native:sphere_cornell_box") {
    *world = sphere_cornell_box();
    lights = new xz_rect(-150, 150, -150, 150, 300, nullptr);
  } else if (name == "native:two_perlin_spheres") {
    *world = two_perlin_spheres();
  } else if (name == "native:light_spheres") {
    *world = light_spheres();
  } else if (name == "native:two_checker_spheres") {
    *world = two_checker_spheres();
  } else {
    return false;
  }
  if (!s->lights) s->lights = lights;
  return true;
}

}  // namespace

extern "C" {

// Builds `name` twice on fresh threads (drand_r at the default 5489 state):
// one copy for hit queries (materials tagged) and one for rendering.
int ref_scene_build(const char *name, void **out) {
  auto *s = new ref_scene_t();
  bool ok = true;
  std::thread t1([&] { ok = build(name, s, &s->world); });
  t1.join();
  if (!ok) {
    delete s;
    return -1;
  }
  std::thread t2([&] { build(name, s, &s->render_world); });
  t2.join();
  tagger tg;
  tg.walk(s->world);
  *out = s;
  return 0;
}

int ref_leaf_count(void *sp) {
  tagger tg;
  // count without retagging: walk a fresh build would be costly; count proxies
  auto *s = static_cast<ref_scene_t *>(sp);
  struct counter {
    int n = 0;
    void walk(hitable *h) {
      if (auto *b = dynamic_cast<bvh_node *>(h)) {
        walk(b->left_);
        if (b->right_ != b->left_) walk(b->right_);
      } else if (auto *l = dynamic_cast<hitable_list *>(h)) {
        for (int i = 0; i < l->list_size_; ++i) walk(l->list_[i]);
      } else if (auto *bx = dynamic_cast<box *>(h)) {
        walk(bx->list_ptr_);
      } else if (auto *f = dynamic_cast<flip_normal *>(h)) {
        walk(f->ptr_);
      } else if (auto *t = dynamic_cast<translate *>(h)) {
        walk(t->ptr_);
      } else if (auto *r = dynamic_cast<rotate_y *>(h)) {
        walk(r->ptr_);
      } else {
        ++n;
      }
    }
  } c;
  c.walk(s->world);
  (void)tg;
  return c.n;
}

int ref_trace(void *sp, const float *rays, uint64_t n, float tmin, float tmax, int32_t *leaf, float *t, float *rec,
              int32_t *mat, int nthreads, const float *time) {
  auto *s = static_cast<ref_scene_t *>(sp);
  std::atomic<uint64_t> next{0};
  auto worker = [&]() {
    for (;;) {
      const uint64_t b = next.fetch_add(1024);
      if (b >= n) return;
      const uint64_t e = std::min<uint64_t>(n, b + 1024);
      for (uint64_t i = b; i < e; ++i) {
        const ray r(vec3(rays[i], rays[n + i], rays[2 * n + i]),
                    vec3(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]), time ? time[i] : 0.0f);
        hit_record hr;
        if (s->world->hit(r, tmin, tmax, hr)) {
          auto *p = static_cast<proxy_material *>(hr.mat_ptr);
          leaf[i] = p->leaf;
          t[i] = hr.t;
          if (mat) mat[i] = p->mat_index;
          if (rec) {
            const float v[8] = {hr.u, hr.v, hr.point.x(), hr.point.y(), hr.point.z(),
                                hr.normal.x(), hr.normal.y(), hr.normal.z()};
            for (int k = 0; k < 8; ++k) rec[k * n + i] = v[k];
          }
        } else {
          leaf[i] = -1;
          t[i] = 0.0f;
          if (mat) mat[i] = -1;
          if (rec)
            for (int k = 0; k < 8; ++k) rec[k * n + i] = 0.0f;
        }
      }
    }
  };
  std::vector<std::thread> ts;
  for (int k = 1; k < nthreads; ++k) ts.emplace_back(worker);
  worker();
  for (auto &th : ts) th.join();
  return 0;
}

// The reference's load_image_texture (src/utils.cc:236-240, stb_image):
// copies at most cap bytes of the decoded image into out.
int ref_load_image(const char *path, unsigned char *out, long cap, int *w, int *h, int *c) {
  unsigned char *data = load_image_texture(std::string(path), *w, *h, *c);
  if (!data) return -1;
  const long n = static_cast<long>(*w) * *h * *c;
  std::memcpy(out, data, static_cast<size_t>(n < cap ? n : cap));
  free(data);
  return 0;
}

// params: lookfrom[3], lookat[3], vup[3], vfov, aspect, aperture, focus, t0, t1.
// out: 24 floats in the order of oracle_camera_fields.
int ref_camera_fields(const float *p, float *out) {
  camera cam(vec3(p[0], p[1], p[2]), vec3(p[3], p[4], p[5]), vec3(p[6], p[7], p[8]), p[9], p[10], p[11], p[12],
             p[13], p[14]);
  const vec3 *vs[7] = {&cam.origin_, &cam.lower_left_corner_, &cam.horizontal_, &cam.vertical_,
                       &cam.u_,      &cam.v_,                 &cam.w_};
  for (int i = 0; i < 7; ++i)
    for (int j = 0; j < 3; ++j) out[3 * i + j] = vs[i]->e[j];
  out[21] = cam.lens_radius_;
  out[22] = cam.time0;
  out[23] = cam.time1;
  return 0;
}

// Pre-order dump of bvh_node boxes: 6 floats per node.  Returns the node count
// (at most cap nodes are written).
int ref_bvh_dump(void *sp, float *out, int cap) {
  auto *s = static_cast<ref_scene_t *>(sp);
  int n = 0;
  struct dumper {
    float *out;
    int cap;
    int *n;
    void walk(hitable *h) {
      auto *b = dynamic_cast<bvh_node *>(h);
      if (!b) return;
      if (*n < cap)
        for (int k = 0; k < 3; ++k) {
          out[6 * *n + k] = b->box_.min_[k];
          out[6 * *n + 3 + k] = b->box_.max_[k];
        }
      ++*n;
      walk(b->left_);
      if (b->right_ != b->left_) walk(b->right_);
    }
  } d{out, cap, &n};
  d.walk(s->world);
  return n;
}

// The reference's render loop (main.cpp:109-159) on this scene: one
// std::async task per row, each with a fresh thread_local generator, summing
// de_nan(color()) per pixel into sum (3·w·h floats, row 0 = bottom).  Camera
// from params as in ref_camera_fields.  Returns 0, or -1 without a light list.
// decorrelate_rows: see below (statistical tests only; timing uses 0).
int ref_render(void *sp, const float *camp, uint32_t w, uint32_t h, uint32_t spp, uint32_t max_depth,
               float *sum, float *sumsq, int decorrelate_rows) {
  auto *s = static_cast<ref_scene_t *>(sp);
  if (!s->lights) return -1;
  camera cam(vec3(camp[0], camp[1], camp[2]), vec3(camp[3], camp[4], camp[5]), vec3(camp[6], camp[7], camp[8]),
             camp[9], camp[10], camp[11], camp[12], camp[13], camp[14]);
  auto row = [&](int j) {
    // Each row thread starts from the default-seeded generator (the
    // reference's behaviour).  With decorrelate_rows, row j first consumes
    // j * 2^20 draws of its own stream, so rows sample independently.
    if (decorrelate_rows)
      for (uint64_t k = 0; k < (static_cast<uint64_t>(j) << 20); ++k) (void)drand_r();
    for (uint32_t i = 0; i < w; i++) {
      vec3 col(0, 0, 0), col2(0, 0, 0);
      for (uint32_t k = 0; k < spp; k++) {
        const float u = ((float)i + drand_r()) / (float)w;
        const float v = ((float)j + drand_r()) / (float)h;
        ray r = cam.get_ray(u, v);
        const vec3 c = de_nan(color(r, s->render_world, s->lights, 0, static_cast<int>(max_depth)));
        col += c;
        col2 += c * c;
      }
      for (int c = 0; c < 3; ++c) {
        sum[(static_cast<size_t>(j) * w + i) * 3 + c] = col[c];
        if (sumsq) sumsq[(static_cast<size_t>(j) * w + i) * 3 + c] = col2[c];
      }
    }
  };
  std::vector<std::future<void>> futures;
  for (int j = static_cast<int>(h) - 1; j >= 0; j--)
    futures.emplace_back(std::async(std::launch::deferred | std::launch::async, row, j));
  for (auto &f : futures) f.wait();
  return 0;
}

// The reference's render loop with its progressive "bonus" snapshots
// (main.cpp:109-133): every pixel's running sum of de_nan(color()) is kept
// after each `every` samples; snaps holds spp / every images (3·w·h floats
// each, row 0 = bottom), the last one the full render.  Each row on a fresh
// thread, as std::async tasks (max_threads <= 0: all at once, as the driver
// launches them) or at most max_threads threads at a time (1: one thread).
// Returns 0, or -1 without a light list or when every does not divide spp.
int ref_render_snapshots(void *sp, const float *camp, uint32_t w, uint32_t h, uint32_t spp, uint32_t every,
                         uint32_t max_depth, int max_threads, float *snaps) {
  auto *s = static_cast<ref_scene_t *>(sp);
  if (!s->lights || every == 0 || spp % every != 0) return -1;
  camera cam(vec3(camp[0], camp[1], camp[2]), vec3(camp[3], camp[4], camp[5]), vec3(camp[6], camp[7], camp[8]),
             camp[9], camp[10], camp[11], camp[12], camp[13], camp[14]);
  const size_t image = static_cast<size_t>(w) * h * 3;
  auto row = [&](int j) {
    for (uint32_t i = 0; i < w; i++) {
      vec3 col(0, 0, 0);
      for (uint32_t k = 0; k < spp; k++) {
        const float u = ((float)i + drand_r()) / (float)w;
        const float v = ((float)j + drand_r()) / (float)h;
        ray r = cam.get_ray(u, v);
        col += de_nan(color(r, s->render_world, s->lights, 0, static_cast<int>(max_depth)));
        if ((k + 1) % every == 0) {
          float *out = snaps + ((k + 1) / every - 1) * image + (static_cast<size_t>(j) * w + i) * 3;
          for (int c = 0; c < 3; ++c) out[c] = col[c];
        }
      }
    }
  };
  if (max_threads <= 0) {
    std::vector<std::future<void>> futures;
    for (int j = static_cast<int>(h) - 1; j >= 0; j--)
      futures.emplace_back(std::async(std::launch::deferred | std::launch::async, row, j));
    for (auto &f : futures) f.wait();
  } else {
    for (int j0 = static_cast<int>(h) - 1; j0 >= 0; j0 -= max_threads) {
      std::vector<std::thread> ts;
      for (int j = j0; j > j0 - max_threads && j >= 0; --j) ts.emplace_back(row, j);
      for (auto &t : ts) t.join();
    }
  }
  return 0;
}

}  // extern "C"

// ---- the reference's Perlin noise (src/perlin_noise.cc, src/utils.cc:155-225) ----
extern "C" {

// The reference's current static tables: grad[256][3], perm[3][256].
void ref_perlin_tables(float *grad, int32_t *perm) {
  for (int i = 0; i < 256; ++i) {
    for (int j = 0; j < 3; ++j) grad[3 * i + j] = perlin_noise::random_vec3_[i][j];
    perm[i] = perlin_noise::permute_x_[i];
    perm[256 + i] = perlin_noise::permute_y_[i];
    perm[512 + i] = perlin_noise::permute_z_[i];
  }
}

// Constructs the reference's perlin_noise_texture(scale) (its constructor
// reshuffles the permutations from the clock), installs the given tables in
// its static arrays, then evaluates noise(p), turb(p) and value(0, 0, p).x()
// with the reference's own member functions at n points pts[n][3].
int ref_perlin_eval(const float *grad, const int32_t *perm, const float *pts, uint64_t n, float scale, float *noise,
                    float *turb, float *tex) {
  std::thread t([&] {
    perlin_noise_texture texture(scale);
    for (int i = 0; i < 256; ++i) {
      perlin_noise::random_vec3_[i] = vec3(grad[3 * i], grad[3 * i + 1], grad[3 * i + 2]);
      perlin_noise::permute_x_[i] = perm[i];
      perlin_noise::permute_y_[i] = perm[256 + i];
      perlin_noise::permute_z_[i] = perm[512 + i];
    }
    for (uint64_t k = 0; k < n; ++k) {
      const vec3 p(pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]);
      noise[k] = texture.noise_.noise(p);
      turb[k] = texture.noise_.turb(p);
      tex[k] = texture.value(0.0f, 0.0f, p).x();
    }
  });
  t.join();
  return 0;
}

}  // extern "C"

// ---- single functions of the reference, for bit-exact pins of the oracle ----
// Every function that draws runs on a fresh thread, so drand_r starts from its
// default state (a std::mt19937 seeded 5489), as on each row thread of the
// reference's driver.
namespace {

// Light shapes: kind 0 sphere (cx, cy, cz, r); kind 1 xz_rect (x0, x1, z0, z1,
// k); kind 2 the reference driver's light list (main.cpp:99-106).
hitable *make_light(int kind, const float *p) {
  if (kind == 0) return new sphere(vec3(p[0], p[1], p[2]), p[3], nullptr);
  if (kind == 1) return new xz_rect(p[0], p[1], p[2], p[3], p[4], nullptr);
  if (kind == 2) {
    hitable **l = new hitable *[2];
    l[0] = new xz_rect(-100, 100, -150, -50, 298, nullptr);
    l[1] = new sphere(vec3(120, -50, 40), 120, nullptr);
    return new hitable_list(l, 2);
  }
  return nullptr;
}

template <class F>
void on_fresh_thread(F &&fn) {
  std::thread t(fn);
  t.join();
}

}  // namespace

extern "C" {

// random_in_unit_sphere() (kind 0) or random_in_unit_disk() (kind 1), n
// successive samples, 3 floats each.
int ref_sampler(int kind, uint64_t n, float *out) {
  on_fresh_thread([&] {
    for (uint64_t i = 0; i < n; ++i) {
      const vec3 v = kind == 0 ? random_in_unit_sphere() : random_in_unit_disk();
      for (int k = 0; k < 3; ++k) out[3 * i + k] = v[k];
    }
  });
  return 0;
}

// hitable::pdf_value(o[i], d[i]) of a light shape (make_light).
int ref_light_pdf(int kind, const float *p, const float *o, const float *d, uint64_t n, float *out) {
  hitable *h = make_light(kind, p);
  if (!h) return -1;
  for (uint64_t i = 0; i < n; ++i)
    out[i] = h->pdf_value(vec3(o[3 * i], o[3 * i + 1], o[3 * i + 2]), vec3(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
  return 0;
}

// hitable::random(o[i]) of a light shape, successive draws of one generator.
int ref_light_random(int kind, const float *p, const float *o, uint64_t n, float *out) {
  hitable *h = make_light(kind, p);
  if (!h) return -1;
  on_fresh_thread([&] {
    for (uint64_t i = 0; i < n; ++i) {
      const vec3 v = h->random(vec3(o[3 * i], o[3 * i + 1], o[3 * i + 2]));
      for (int k = 0; k < 3; ++k) out[3 * i + k] = v[k];
    }
  });
  return 0;
}

// texture::value(u, v, p) for queries q[i] = (u, v, px, py, pz): kind 0
// checker_texture(constant_texture(c[0..2]), constant_texture(c[3..5])) (odd,
// even); kind 1 image_texture(texels, w, h).
int ref_texture_value(int kind, const float *c, const unsigned char *texels, int w, int h, const float *q, uint64_t n,
                      float *out) {
  texture *t = nullptr;
  if (kind == 0)
    t = new checker_texture(new constant_texture(vec3(c[0], c[1], c[2])), new constant_texture(vec3(c[3], c[4], c[5])));
  if (kind == 1) t = new image_texture(const_cast<unsigned char *>(texels), w, h);
  if (!t) return -1;
  for (uint64_t i = 0; i < n; ++i) {
    const float *x = q + 5 * i;
    const vec3 v = t->value(x[0], x[1], vec3(x[2], x[3], x[4]));
    for (int k = 0; k < 3; ++k) out[3 * i + k] = v[k];
  }
  return 0;
}

// metal(albedo, fuzz)::scatter for queries q[i] = (ray o, ray d, time, rec
// point, rec normal): the scattered direction and the return value,
// successive draws of one generator.
int ref_metal_scatter(const float *albedo, float fuzz, const float *q, uint64_t n, float *dir, int32_t *ok) {
  metal m(vec3(albedo[0], albedo[1], albedo[2]), fuzz);
  on_fresh_thread([&] {
    for (uint64_t i = 0; i < n; ++i) {
      const float *x = q + 13 * i;
      const ray r(vec3(x[0], x[1], x[2]), vec3(x[3], x[4], x[5]), x[6]);
      hit_record rec;
      rec.t = 1.0f;
      rec.point = vec3(x[7], x[8], x[9]);
      rec.normal = vec3(x[10], x[11], x[12]);
      rec.mat_ptr = &m;
      scatter_record sr;
      ok[i] = m.scatter(r, rec, sr) ? 1 : 0;
      for (int k = 0; k < 3; ++k) dir[3 * i + k] = sr.specular_ray.direction()[k];
    }
  });
  return 0;
}

// The reference driver's sample loop (main.cpp:115-126) for listed pixels
// pix[2k] = i, pix[2k + 1] = j: u, v, get_ray, de_nan(color()), spp samples
// each, every pixel on a fresh thread (a fresh generator).  out: npix * spp
// * 3 floats.  Camera from params as in ref_camera_fields.
int ref_pixel_samples(void *sp, const float *camp, uint32_t w, uint32_t h, const uint32_t *pix, uint64_t npix,
                      uint32_t spp, uint32_t max_depth, int nthreads, float *out) {
  auto *s = static_cast<ref_scene_t *>(sp);
  if (!s->lights) return -1;
  camera cam(vec3(camp[0], camp[1], camp[2]), vec3(camp[3], camp[4], camp[5]), vec3(camp[6], camp[7], camp[8]),
             camp[9], camp[10], camp[11], camp[12], camp[13], camp[14]);
  auto pixel = [&](uint64_t k) {
    const uint32_t i = pix[2 * k], j = pix[2 * k + 1];
    for (uint32_t smp = 0; smp < spp; ++smp) {
      const float u = ((float)i + drand_r()) / (float)w;
      const float v = ((float)j + drand_r()) / (float)h;
      ray r = cam.get_ray(u, v);
      const vec3 c = de_nan(color(r, s->render_world, s->lights, 0, static_cast<int>(max_depth)));
      for (int ch = 0; ch < 3; ++ch) out[(k * spp + smp) * 3 + ch] = c[ch];
    }
  };
  if (nthreads < 1) nthreads = 1;
  for (uint64_t b = 0; b < npix; b += static_cast<uint64_t>(nthreads)) {
    std::vector<std::thread> ts;
    for (uint64_t k = b; k < std::min<uint64_t>(npix, b + nthreads); ++k) ts.emplace_back(pixel, k);
    for (auto &t : ts) t.join();
  }
  return 0;
}

// Replaces the reference's static Perlin tables (its constructor shuffles
// them from the clock), so that both sides evaluate the same noise.
void ref_perlin_install(const float *grad, const int32_t *perm) {
  for (int i = 0; i < 256; ++i) {
    perlin_noise::random_vec3_[i] = vec3(grad[3 * i], grad[3 * i + 1], grad[3 * i + 2]);
    perlin_noise::permute_x_[i] = perm[i];
    perlin_noise::permute_y_[i] = perm[256 + i];
    perlin_noise::permute_z_[i] = perm[512 + i];
  }
}

}  // extern "C"
