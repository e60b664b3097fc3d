// ORACLE — test infrastructure only (see oracle.h).
// C entry points of liboracle.so for the tests: builds scenes with the
// product's scene code, then intersects / renders them with the restatement.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../tiny-path-tracer_amd/host/scenes.h"
#include "../tiny-path-tracer_amd/scene/hitable_list.h"
#include "../tiny-path-tracer_amd/scene/material.h"
#include "../tiny-path-tracer_amd/scene/rect_box.h"
#include "../tiny-path-tracer_amd/scene/sphere.h"
#include "../tiny-path-tracer_amd/scene/texture.h"
#include "oracle.h"

namespace {

thread_local std::string g_err;

struct philox_rng : oracle::rng {
  oracle::path_stream s;
  double next() override { return s.next(); }
  void bounce(uint32_t b) override { s.set_bounce(b); }
  // counter (pixel, sample, bounce, 2^20 + leaf), as the device draws them
  double medium(uint32_t leaf, int k) override {
    const oracle::philox_out o = oracle::philox4x32_10(s.pixel, s.sample, s.bounce, (1u << 20) + leaf, s.k0, s.k1);
    return static_cast<double>(oracle::u01(o.v[k]));
  }
};

struct mt_rng : oracle::rng {
  std::mt19937 gen;  // default seed 5489, as a fresh row thread of the reference
  std::uniform_real_distribution<double> dist{0.0, 1.0};
  bool rtl = false;  // the reference build's argument order (rng::args_right_to_left)
  double next() override { return dist(gen); }
  bool args_right_to_left() const override { return rtl; }
};

template <class F>
void parallel_for(uint64_t n, int nthreads, F &&fn) {
  if (nthreads < 1) nthreads = 1;
  std::atomic<uint64_t> next{0};
  // chunks of up to 1024 indices, at least 8 per thread (image rows are few)
  const uint64_t chunk =
      std::max<uint64_t>(1, std::min<uint64_t>(1024, n / (8 * static_cast<uint64_t>(nthreads))));
  auto worker = [&]() {
    for (;;) {
      const uint64_t b = next.fetch_add(chunk);
      if (b >= n) return;
      const uint64_t e = std::min(n, b + chunk);
      for (uint64_t i = b; i < e; ++i) fn(i);
    }
  };
  std::vector<std::thread> ts;
  for (int t = 1; t < nthreads; ++t) ts.emplace_back(worker);
  worker();
  for (auto &t : ts) t.join();
}

// World-space primitives and boxes used to aim adversarial rays.
struct aim_targets {
  std::vector<AABB> boxes;
  struct rect {
    int axis;
    float k, a0, a1, b0, b1;
  };
  std::vector<rect> rects;
  std::vector<std::pair<vec3, float>> spheres;
};

void gather_targets(const hitable *h, vec3 offset, bool translated_only, aim_targets *t) {
  if (auto *b = dynamic_cast<const bvh_node *>(h)) {
    t->boxes.push_back(b->box_);
    gather_targets(b->left_, offset, translated_only, t);
    if (b->right_ != b->left_) gather_targets(b->right_, offset, translated_only, t);
  } else if (auto *l = dynamic_cast<const hitable_list *>(h)) {
    for (int i = 0; i < l->list_size_; ++i) gather_targets(l->list_[i], offset, translated_only, t);
  } else if (auto *bx = dynamic_cast<const box *>(h)) {
    gather_targets(bx->list_ptr_, offset, translated_only, t);
  } else if (auto *f = dynamic_cast<const flip_normal *>(h)) {
    gather_targets(f->ptr_, offset, translated_only, t);
  } else if (auto *tr = dynamic_cast<const translate *>(h)) {
    gather_targets(tr->ptr_, offset + tr->offset_, translated_only, t);
  } else if (dynamic_cast<const rotate_y *>(h)) {
    // rotated content: aim only at its box
    AABB bb;
    if (h->bounding_box(0, 0, bb)) t->boxes.push_back(AABB(bb.min_ + offset, bb.max_ + offset));
  } else if (auto *s = dynamic_cast<const sphere *>(h)) {
    t->spheres.push_back({s->center_ + offset, s->radius_});
  } else if (auto *q = dynamic_cast<const xy_rect *>(h)) {
    t->rects.push_back({2, q->k_ + offset.z(), q->x0_ + offset.x(), q->x1_ + offset.x(), q->y0_ + offset.y(),
                        q->y1_ + offset.y()});
  } else if (auto *q2 = dynamic_cast<const xz_rect *>(h)) {
    t->rects.push_back({1, q2->k_ + offset.y(), q2->x0_ + offset.x(), q2->x1_ + offset.x(),
                        q2->z0_ + offset.z(), q2->z1_ + offset.z()});
  } else if (auto *q3 = dynamic_cast<const yz_rect *>(h)) {
    t->rects.push_back({0, q3->k_ + offset.x(), q3->y0_ + offset.y(), q3->y1_ + offset.y(),
                        q3->z0_ + offset.z(), q3->z1_ + offset.z()});
  }
  (void)translated_only;
}

}  // namespace

struct oracle_scene {
  tpt::scene_spec spec;
  oracle::graph_info g;
  AABB world_box;
};

extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

int oracle_scene_build(const char *name, uint32_t seed, oracle_scene **out) {
  auto *s = new oracle_scene();
  std::string err;
  if (!tpt::build_scene(name, seed, &s->spec, &err)) {
    g_err = err;
    delete s;
    return -1;
  }
  oracle::index_graph(s->spec.world, &s->g);
  if (!s->spec.world->bounding_box(0, 1, s->world_box)) s->world_box = AABB(vec3(-1, -1, -1), vec3(1, 1, 1));
  *out = s;
  return 0;
}

void oracle_scene_destroy(oracle_scene *s) { delete s; }

uint32_t oracle_leaf_count(const oracle_scene *s) { return s->g.leaf_count.at(s->spec.world); }

int oracle_scene_settings(const oracle_scene *s, uint32_t *w, uint32_t *h, uint32_t *spp, uint32_t *depth,
                          float *bg) {
  *w = s->spec.width;
  *h = s->spec.height;
  *spp = s->spec.spp;
  *depth = s->spec.max_depth;
  *bg = s->spec.background;
  return 0;
}

// 24 floats: origin, lower_left, horizontal, vertical, u, v, w, then
// lens_radius, time0, time1.
int oracle_camera_fields(const oracle_scene *s, float *out) {
  const camera_with_blur &c = *s->spec.cam;
  const vec3 *vs[7] = {&c.origin_, &c.lower_left_corner_, &c.horizontal_, &c.vertical_, &c.u_, &c.v_, &c.w_};
  for (int i = 0; i < 7; ++i)
    for (int j = 0; j < 3; ++j) out[3 * i + j] = vs[i]->e[j];
  out[21] = c.lens_radius_;
  out[22] = c.time0;
  out[23] = c.time1;
  return 0;
}

// rays: 6 arrays of n floats (ox, oy, oz, dx, dy, dz).  rec: 8 arrays of n
// (u, v, px, py, pz, nx, ny, nz); rec and mat may be null.  time: n ray times
// (ray::time()), or null for time 0.
int oracle_trace(const oracle_scene *s, const float *rays, uint64_t n, float tmin, float tmax, int32_t *leaf,
                 float *t, float *rec, int32_t *mat, int nthreads, const float *time) {
  parallel_for(n, nthreads, [&](uint64_t i) {
    const ray r(vec3(rays[i], rays[n + i], rays[2 * n + i]), vec3(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]),
                time ? time[i] : 0.0f);
    // media draws of ray i: counter (i mod 2^32, i / 2^32, 0, 2^20 + leaf), key 0
    philox_rng M;
    M.s.pixel = static_cast<uint32_t>(i);
    M.s.sample = static_cast<uint32_t>(i >> 32);
    M.s.k0 = M.s.k1 = 0;
    M.s.set_bounce(0);
    hit_record hr;
    uint32_t lf = 0;
    if (oracle::hit(s->spec.world, s->g, r, tmin, tmax, hr, 0, &lf, &M)) {
      leaf[i] = static_cast<int32_t>(lf);
      t[i] = hr.t;
      if (rec) {
        const float vals[8] = {hr.u, hr.v, hr.point.x(), hr.point.y(), hr.point.z(),
                               hr.normal.x(), hr.normal.y(), hr.normal.z()};
        for (int k = 0; k < 8; ++k) rec[k * n + i] = vals[k];
      }
      if (mat) mat[i] = static_cast<int32_t>(s->g.material_id.at(hr.mat_ptr));
    } else {
      leaf[i] = -1;
      t[i] = 0.0f;
      if (rec)
        for (int k = 0; k < 8; ++k) rec[k * n + i] = 0.0f;
      if (mat) mat[i] = -1;
    }
  });
  return 0;
}

// kind 0: camera primaries; 1: one-bounce secondaries scattered by the
// oracle; 2: origins uniform in the (10% grown) world box, directions uniform
// on the sphere times a random length in [0.25, 4); 3: adversarial rays
// (axis-aligned, in-plane, at box corners/edges, sphere tangents, extreme
// magnitudes).  Deterministic in (kind, seed, i).
int oracle_make_rays(const oracle_scene *s, int kind, uint64_t n, uint64_t seed, float *rays, int nthreads) {
  const camera_with_blur &cam = *s->spec.cam;
  const uint32_t W = s->spec.width, H = s->spec.height;
  aim_targets tg;
  if (kind == 3) gather_targets(s->spec.world, vec3(0, 0, 0), true, &tg);
  const vec3 lo = s->world_box.min_, hi = s->world_box.max_;
  const vec3 ext = hi - lo;
  parallel_for(n, nthreads, [&](uint64_t i) {
    philox_rng R;
    R.s.pixel = static_cast<uint32_t>(i);
    R.s.sample = static_cast<uint32_t>(i >> 32) ^ 0xA5A5u;
    R.s.k0 = static_cast<uint32_t>(seed);
    R.s.k1 = static_cast<uint32_t>(seed >> 32) ^ static_cast<uint32_t>(kind * 0x1234567);
    R.s.set_bounce(0);
    ray out;
    auto primary = [&]() {
      const uint32_t px = static_cast<uint32_t>(R.next() * W);
      const uint32_t py = static_cast<uint32_t>(R.next() * H);
      const float u = static_cast<float>((static_cast<float>(px) + R.next()) / static_cast<float>(W));
      const float v = static_cast<float>((static_cast<float>(py) + R.next()) / static_cast<float>(H));
      return oracle::camera_ray(cam, u, v, R);
    };
    if (kind == 0) {
      out = primary();
    } else if (kind == 1) {
      out = primary();
      for (int tries = 0; tries < 8; ++tries) {
        hit_record rec;
        uint32_t lf;
        if (oracle::hit(s->spec.world, s->g, out, 0.001f, std::numeric_limits<float>::max(), rec, 0, &lf, &R)) {
          ray next;
          if (oracle::scatter_once(out, rec, s->g, s->spec.lights, R, &next)) {
            out = next;
            break;
          }
        }
        out = primary();
      }
    } else if (kind == 2 || (kind == 3 && tg.boxes.empty() && tg.rects.empty() && tg.spheres.empty())) {
      const vec3 o(lo.x() - 0.05f * ext.x() + static_cast<float>(R.next()) * 1.1f * ext.x(),
                   lo.y() - 0.05f * ext.y() + static_cast<float>(R.next()) * 1.1f * ext.y(),
                   lo.z() - 0.05f * ext.z() + static_cast<float>(R.next()) * 1.1f * ext.z());
      const float z = static_cast<float>(1 - 2 * R.next());
      const float phi = static_cast<float>(2 * M_PI * R.next());
      const float rr = std::sqrt(std::max(0.0f, 1 - z * z));
      const float len = static_cast<float>(std::exp2(-2.0 + 4.0 * R.next()));
      out = ray(o, len * vec3(rr * std::cos(phi), rr * std::sin(phi), z));
    } else {
      const int sub = static_cast<int>(R.next() * 8);
      auto rnd_in = [&](float a, float b) { return a + static_cast<float>(R.next()) * (b - a); };
      auto rnd_origin = [&]() {
        return vec3(rnd_in(lo.x() - 0.2f * ext.x(), hi.x() + 0.2f * ext.x()),
                    rnd_in(lo.y() - 0.2f * ext.y(), hi.y() + 0.2f * ext.y()),
                    rnd_in(lo.z() - 0.2f * ext.z(), hi.z() + 0.2f * ext.z()));
      };
      vec3 o = rnd_origin();
      vec3 d;
      if (sub == 0) {  // axis-aligned, signed zeros in the other components
        const int ax = static_cast<int>(R.next() * 3);
        d = vec3(R.next() < 0.5 ? 0.0f : -0.0f, R.next() < 0.5 ? 0.0f : -0.0f, R.next() < 0.5 ? 0.0f : -0.0f);
        d[ax] = R.next() < 0.5 ? 1.0f : -1.0f;
      } else if (sub == 1 && !tg.rects.empty()) {  // in the plane of a rect (NaN t)
        const auto &q = tg.rects[static_cast<size_t>(R.next() * tg.rects.size())];
        o[q.axis] = q.k;
        d = vec3(static_cast<float>(R.next() - 0.5), static_cast<float>(R.next() - 0.5),
                 static_cast<float>(R.next() - 0.5));
        d[q.axis] = R.next() < 0.5 ? 0.0f : -0.0f;
      } else if (sub == 2 && !tg.boxes.empty()) {  // at a box corner
        const AABB &b = tg.boxes[static_cast<size_t>(R.next() * tg.boxes.size())];
        const vec3 c(R.next() < 0.5 ? b.min_.x() : b.max_.x(), R.next() < 0.5 ? b.min_.y() : b.max_.y(),
                     R.next() < 0.5 ? b.min_.z() : b.max_.z());
        d = c - o;
      } else if (sub == 3 && !tg.rects.empty()) {  // at a rect edge or corner
        const auto &q = tg.rects[static_cast<size_t>(R.next() * tg.rects.size())];
        vec3 p;
        const int i0 = q.axis == 0 ? 1 : 0, i1 = q.axis == 2 ? 1 : 2;
        p[q.axis] = q.k;
        p[i0] = R.next() < 0.5 ? q.a0 : q.a1;
        p[i1] = R.next() < 0.3 ? (R.next() < 0.5 ? q.b0 : q.b1) : rnd_in(q.b0, q.b1);
        d = p - o;
      } else if (sub == 4 && !tg.spheres.empty()) {  // tangent to a sphere
        const auto &sp = tg.spheres[static_cast<size_t>(R.next() * tg.spheres.size())];
        const vec3 to = sp.first - o;
        const vec3 perp = unit_vector(cross(to, vec3(0.3f, 1.0f, 0.7f)));
        const vec3 touch = sp.first + sp.second * perp;
        d = touch - o;
      } else if (sub == 5 && !tg.rects.empty()) {  // origin exactly on a rect plane
        const auto &q = tg.rects[static_cast<size_t>(R.next() * tg.rects.size())];
        o[q.axis] = q.k;
        d = vec3(static_cast<float>(R.next() - 0.5), static_cast<float>(R.next() - 0.5),
                 static_cast<float>(R.next() - 0.5));
      } else if (sub == 6) {  // extreme magnitudes
        const float scale = R.next() < 0.5 ? 1e-30f : 1e30f;
        d = scale * vec3(static_cast<float>(R.next() - 0.5), static_cast<float>(R.next() - 0.5),
                         static_cast<float>(R.next() - 0.5));
      } else {  // far origin aimed at the scene
        o = 1e5f * vec3(static_cast<float>(R.next() - 0.5), static_cast<float>(R.next() - 0.5),
                        static_cast<float>(R.next() - 0.5));
        d = rnd_origin() - o;
      }
      out = ray(o, d);
    }
    for (int k = 0; k < 3; ++k) {
      rays[k * n + i] = out.origin()[k];
      rays[(3 + k) * n + i] = out.direction()[k];
    }
  });
  return 0;
}

// Renders spp samples per pixel into sum/sumsq (3·w·h floats each, row 0 =
// bottom, accumulated with +=).  rng_mode 0: the Philox path streams the
// device uses; 1: a fresh default-seeded mt19937 per row (the reference's
// thread-per-row generator state, main.cpp:157-159 + src/utils.cc:28-32);
// 2: as 1, constructor arguments drawn last first (rng::args_right_to_left).
int oracle_render(const oracle_scene *s, uint32_t w, uint32_t h, uint32_t spp, uint32_t max_depth, uint64_t seed,
                  uint32_t sample_offset, float background, int rng_mode, int nthreads, float *sum,
                  float *sumsq, int media_scatter) {
  oracle::render_opts o;
  o.max_depth = static_cast<int>(max_depth);
  o.background = background;
  o.media_scatter = media_scatter != 0;
  parallel_for(h, nthreads, [&](uint64_t j) {
    mt_rng mt;
    philox_rng ph;
    ph.s.k0 = static_cast<uint32_t>(seed);
    ph.s.k1 = static_cast<uint32_t>(seed >> 32);
    for (uint32_t i = 0; i < w; ++i) {
      const uint64_t px = j * w + i;
      for (uint32_t k = 0; k < spp; ++k) {
        oracle::rng *R = &mt;
        mt.rtl = rng_mode == 2;
        if (rng_mode == 0) {
          ph.s.pixel = static_cast<uint32_t>(px);
          ph.s.sample = sample_offset + k;
          ph.s.set_bounce(0);
          R = &ph;
        }
        const float u = static_cast<float>((static_cast<float>(i) + R->next()) / static_cast<float>(w));
        const float v = static_cast<float>((static_cast<float>(j) + R->next()) / static_cast<float>(h));
        const ray r = oracle::camera_ray(*s->spec.cam, u, v, *R);
        const vec3 c = oracle::color(r, s->spec.world, s->g, s->spec.lights, 0, o, *R);
        for (int ch = 0; ch < 3; ++ch) {
          const float x = std::isnan(c[ch]) ? 0.0f : c[ch];
          sum[px * 3 + ch] += x;
          if (sumsq) sumsq[px * 3 + ch] += x * x;
        }
      }
    }
  });
  return 0;
}

}  // extern "C"

// ---- primitive-level access for closed-form tests --------------------------
extern "C" {

void oracle_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t *out) {
  const oracle::philox_out o = oracle::philox4x32_10(c0, c1, c2, c3, k0, k1);
  for (int i = 0; i < 4; ++i) out[i] = o.v[i];
}

// AABB::hit on box (lo, hi) for ray (o, d).
int oracle_slab(const float *lo, const float *hi, const float *o, const float *d, float tmin, float tmax) {
  return oracle::slab_hit(AABB(vec3(lo[0], lo[1], lo[2]), vec3(hi[0], hi[1], hi[2])),
                          ray(vec3(o[0], o[1], o[2]), vec3(d[0], d[1], d[2])), tmin, tmax)
             ? 1
             : 0;
}

// sphere::hit / rect hits on a single primitive; rec as in oracle_trace.
// kind 0: sphere p = (cx, cy, cz, r); kind 1/2/3: xy/xz/yz rect p = (a0, a1, b0, b1, k).
int oracle_prim_hit(int kind, const float *p, const float *o, const float *d, float tmin, float tmax, float *t,
                    float *rec) {
  hitable *h = nullptr;
  if (kind == 0) h = new sphere(vec3(p[0], p[1], p[2]), p[3], nullptr);
  if (kind == 1) h = new xy_rect(p[0], p[1], p[2], p[3], p[4], nullptr);
  if (kind == 2) h = new xz_rect(p[0], p[1], p[2], p[3], p[4], nullptr);
  if (kind == 3) h = new yz_rect(p[0], p[1], p[2], p[3], p[4], nullptr);
  if (!h) return -1;
  oracle::graph_info g;
  oracle::index_graph(h, &g);
  hit_record hr;
  uint32_t leaf;
  const bool ok = oracle::hit(h, g, ray(vec3(o[0], o[1], o[2]), vec3(d[0], d[1], d[2])), tmin, tmax, hr, 0, &leaf);
  delete h;
  if (!ok) return 0;
  *t = hr.t;
  const float v[8] = {hr.u, hr.v, hr.point.x(), hr.point.y(), hr.point.z(), hr.normal.x(), hr.normal.y(), hr.normal.z()};
  for (int k = 0; k < 8; ++k) rec[k] = v[k];
  return 1;
}

// The product's camera constructor with explicit parameters (for comparison
// with the reference's): params as ref_camera_fields.
int oracle_camera_from_params(const float *p, float *out) {
  camera_with_blur cam(vec3(p[0], p[1], p[2]), vec3(p[3], p[4], p[5]), vec3(p[6], p[7], p[8]), p[9], p[10], p[11],
                       p[12], p[13], p[14]);
  const vec3 *vs[7] = {&cam.origin_, &cam.lower_left_corner_, &cam.horizontal_, &cam.vertical_,
                       &cam.u_,      &cam.v_,                 &cam.w_};
  for (int i = 0; i < 7; ++i)
    for (int j = 0; j < 3; ++j) out[3 * i + j] = vs[i]->e[j];
  out[21] = cam.lens_radius_;
  out[22] = cam.time0;
  out[23] = cam.time1;
  return 0;
}

// Pre-order bvh_node boxes of the scene's world, as ref_bvh_dump.
int oracle_bvh_dump(const oracle_scene *s, float *out, int cap) {
  int n = 0;
  struct dumper {
    float *out;
    int cap;
    int *n;
    void walk(const hitable *h) {
      auto *b = dynamic_cast<const bvh_node *>(h);
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
  d.walk(s->spec.world);
  return n;
}

}  // extern "C"

// ---- Perlin tables and evaluation, for the comparison with the reference ----
extern "C" {

// The scene API's current tables: grad[256][3], perm[3][256] (x, y, z).
void oracle_perlin_tables(float *grad, int32_t *perm) {
  for (int i = 0; i < 256; ++i) {
    for (int j = 0; j < 3; ++j) grad[3 * i + j] = perlin_noise::random_vec3_[i].e[j];
    perm[i] = perlin_noise::permute_x_[i];
    perm[256 + i] = perlin_noise::permute_y_[i];
    perm[512 + i] = perlin_noise::permute_z_[i];
  }
}

// Installs the given tables, then evaluates noise(p), turb(p, octaves) and
// perlin_noise_texture(scale, octaves).value(0, 0, p).x() at n points
// pts[n][3].  Runs on a thread of its own: the texture's constructor draws
// from that thread's drand_r before the tables are overwritten.
int oracle_perlin_eval(const float *grad, const int32_t *perm, const float *pts, uint64_t n, float scale, int octaves,
                       float *noise, float *turb, float *tex) {
  std::thread t([&] {
    perlin_noise_texture texture(scale, octaves);
    for (int i = 0; i < 256; ++i) {
      perlin_noise::random_vec3_[i] = vec3(grad[3 * i], grad[3 * i + 1], grad[3 * i + 2]);
      perlin_noise::permute_x_[i] = perm[i];
      perlin_noise::permute_y_[i] = perm[256 + i];
      perlin_noise::permute_z_[i] = perm[512 + i];
    }
    for (uint64_t k = 0; k < n; ++k) {
      const vec3 p(pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]);
      noise[k] = oracle::perlin_noise_at(p);
      turb[k] = oracle::perlin_turb(p, octaves);
      tex[k] = oracle::texture_value(&texture, 0.0f, 0.0f, p).x();
    }
  });
  t.join();
  return 0;
}

}  // extern "C"

// ---- single functions, for bit-exact comparison with the reference's ----
// Same arguments as the ref_* functions of oracle/ref_harness.cc; `rtl` picks
// the argument order of multi-draw constructors (rng::args_right_to_left).
namespace {

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

}  // namespace

extern "C" {

int oracle_sampler(int kind, uint64_t n, int rtl, float *out) {
  mt_rng R;
  R.rtl = rtl != 0;
  for (uint64_t i = 0; i < n; ++i) {
    const vec3 v = kind == 0 ? oracle::unit_sphere_sample(R) : oracle::unit_disk_sample(R);
    for (int k = 0; k < 3; ++k) out[3 * i + k] = v[k];
  }
  return 0;
}

int oracle_light_pdf(int kind, const float *p, const float *o, const float *d, uint64_t n, float *out) {
  hitable *h = make_light(kind, p);
  if (!h) return -1;
  oracle::graph_info g;
  oracle::index_graph(h, &g);
  for (uint64_t i = 0; i < n; ++i)
    out[i] = oracle::light_pdf_value(h, g, vec3(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
                                     vec3(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
  return 0;
}

int oracle_light_random(int kind, const float *p, const float *o, uint64_t n, int rtl, float *out) {
  hitable *h = make_light(kind, p);
  if (!h) return -1;
  mt_rng R;
  R.rtl = rtl != 0;
  for (uint64_t i = 0; i < n; ++i) {
    const vec3 v = oracle::light_sample(h, vec3(o[3 * i], o[3 * i + 1], o[3 * i + 2]), R);
    for (int k = 0; k < 3; ++k) out[3 * i + k] = v[k];
  }
  return 0;
}

int oracle_texture_value(int kind, const float *c, const unsigned char *texels, int w, int h, const float *q,
                         uint64_t n, float *out) {
  texture *t = nullptr;
  if (kind == 0)
    t = new checker_texture(new constant_texture(vec3(c[0], c[1], c[2])), new constant_texture(vec3(c[3], c[4], c[5])));
  if (kind == 1) t = new image_texture(const_cast<unsigned char *>(texels), w, h);
  if (!t) return -1;
  for (uint64_t i = 0; i < n; ++i) {
    const float *x = q + 5 * i;
    const vec3 v = oracle::texture_value(t, x[0], x[1], vec3(x[2], x[3], x[4]));
    for (int k = 0; k < 3; ++k) out[3 * i + k] = v[k];
  }
  return 0;
}

int oracle_metal_scatter(const float *albedo, float fuzz, const float *q, uint64_t n, int rtl, float *dir,
                         int32_t *ok) {
  const metal m(vec3(albedo[0], albedo[1], albedo[2]), fuzz);
  mt_rng R;
  R.rtl = rtl != 0;
  for (uint64_t i = 0; i < n; ++i) {
    const float *x = q + 13 * i;
    const ray r(vec3(x[0], x[1], x[2]), vec3(x[3], x[4], x[5]), x[6]);
    hit_record rec;
    rec.t = 1.0f;
    rec.point = vec3(x[7], x[8], x[9]);
    rec.normal = vec3(x[10], x[11], x[12]);
    ray out;
    ok[i] = oracle::metal_scatter(&m, r, rec, R, &out) ? 1 : 0;
    for (int k = 0; k < 3; ++k) dir[3 * i + k] = out.direction()[k];
  }
  return 0;
}

// The sample loop of ref_pixel_samples with the restated color(): each
// listed pixel from a fresh mt19937.
int oracle_pixel_samples(const oracle_scene *s, const float *camp, uint32_t w, uint32_t h, const uint32_t *pix,
                         uint64_t npix, uint32_t spp, uint32_t max_depth, int rtl, int nthreads, float *out) {
  if (!s->spec.lights) return -1;
  const camera_with_blur cam(vec3(camp[0], camp[1], camp[2]), vec3(camp[3], camp[4], camp[5]),
                             vec3(camp[6], camp[7], camp[8]), camp[9], camp[10], camp[11], camp[12], camp[13],
                             camp[14]);
  oracle::render_opts o;
  o.max_depth = static_cast<int>(max_depth);
  parallel_for(npix, nthreads, [&](uint64_t k) {
    mt_rng R;
    R.rtl = rtl != 0;
    const uint32_t i = pix[2 * k], j = pix[2 * k + 1];
    for (uint32_t smp = 0; smp < spp; ++smp) {
      const float u = static_cast<float>((static_cast<float>(i) + R.next()) / static_cast<float>(w));
      const float v = static_cast<float>((static_cast<float>(j) + R.next()) / static_cast<float>(h));
      const ray r = oracle::camera_ray(cam, u, v, R);
      const vec3 c = oracle::color(r, s->spec.world, s->g, s->spec.lights, 0, o, R);
      for (int ch = 0; ch < 3; ++ch) out[(k * spp + smp) * 3 + ch] = std::isnan(c[ch]) ? 0.0f : c[ch];
    }
  });
  return 0;
}

// color() of explicit rays (6 x n planes, as oracle_trace) with the Philox
// streams of the device (pixel = ray index, sample = `sample`), misses
// returning the sky gradient times `background`.  out: 3 x n.
int oracle_color_rays(const oracle_scene *s, const float *rays, uint64_t n, uint32_t max_depth, float background,
                      uint64_t seed, uint32_t sample, int nthreads, float *out) {
  oracle::render_opts o;
  o.max_depth = static_cast<int>(max_depth);
  o.background = background;
  parallel_for(n, nthreads, [&](uint64_t i) {
    philox_rng R;
    R.s.k0 = static_cast<uint32_t>(seed);
    R.s.k1 = static_cast<uint32_t>(seed >> 32);
    R.s.pixel = static_cast<uint32_t>(i);
    R.s.sample = sample;
    R.s.set_bounce(0);
    const ray r(vec3(rays[i], rays[n + i], rays[2 * n + i]), vec3(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]));
    const vec3 c = oracle::color(r, s->spec.world, s->g, s->spec.lights, 0, o, R);
    for (int ch = 0; ch < 3; ++ch) out[ch * n + i] = c[ch];
  });
  return 0;
}

}  // extern "C"

// ---- the guided estimator (oracle_guide.cc) ----
extern "C" {

// Network outputs (n x 40, before the bf16 rounding) for inputs in9 (n x 9).
int oracle_guide_forward(const float *params, const float *in9, uint64_t n, float *out) {
  for (uint64_t i = 0; i < n; ++i) {
    double o[40];
    oracle::guide_forward(params, in9 + 9 * i, o);
    for (int k = 0; k < 40; ++k) out[40 * i + k] = static_cast<float>(o[k]);
  }
  return 0;
}

// The decoded mixture of input in9: n directions sampled with draws u (n x 3)
// and the mixture density at each (of the unit direction); -1 if the
// network output is not finite.
int oracle_guide_sample(const float *params, const float *in9, const float *u, uint64_t n, float *dirs, float *pdf) {
  oracle::guide_mix m;
  if (!oracle::guide_decode(params, in9, &m)) return -1;
  for (uint64_t i = 0; i < n; ++i) {
    const vec3 d = oracle::guide_mix_sample(m, u[3 * i], u[3 * i + 1], u[3 * i + 2]);
    for (int k = 0; k < 3; ++k) dirs[3 * i + k] = d[k];
    pdf[i] = oracle::guide_mix_pdf(m, unit_vector(d));
  }
  return 0;
}

// Mixture density of input in9 at n unit directions.
int oracle_guide_pdf(const float *params, const float *in9, const float *dirs, uint64_t n, float *pdf) {
  oracle::guide_mix m;
  if (!oracle::guide_decode(params, in9, &m)) return -1;
  for (uint64_t i = 0; i < n; ++i) pdf[i] = oracle::guide_mix_pdf(m, vec3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
  return 0;
}

// oracle_render (Philox streams) with the guided estimator: the first
// guide_verts lambertian vertices of every path sample from the proposal of
// `params` (padded layout + mix logit), the scene box mapping hit points to
// the network input.
int oracle_render_guided(const oracle_scene *s, uint32_t w, uint32_t h, uint32_t spp, uint32_t max_depth,
                         uint64_t seed, uint32_t sample_offset, const float *params, int guide_verts, int nthreads,
                         float *sum, float *sumsq) {
  oracle::render_opts o;
  o.max_depth = static_cast<int>(max_depth);
  o.guide = params;
  o.guide_verts = guide_verts;
  for (int k = 0; k < 3; ++k) o.box_lo[k] = s->world_box.min_[k], o.box_hi[k] = s->world_box.max_[k];
  float bg = 0.0f;
  {
    uint32_t a, b, c, d;
    oracle_scene_settings(s, &a, &b, &c, &d, &bg);
  }
  o.background = bg;
  parallel_for(h, nthreads, [&](uint64_t j) {
    philox_rng ph;
    ph.s.k0 = static_cast<uint32_t>(seed);
    ph.s.k1 = static_cast<uint32_t>(seed >> 32);
    for (uint32_t i = 0; i < w; ++i) {
      const uint64_t px = j * w + i;
      for (uint32_t k = 0; k < spp; ++k) {
        ph.s.pixel = static_cast<uint32_t>(px);
        ph.s.sample = sample_offset + k;
        ph.s.set_bounce(0);
        const float u = static_cast<float>((static_cast<float>(i) + ph.next()) / static_cast<float>(w));
        const float v = static_cast<float>((static_cast<float>(j) + ph.next()) / static_cast<float>(h));
        const ray r = oracle::camera_ray(*s->spec.cam, u, v, ph);
        const vec3 c = oracle::color(r, s->spec.world, s->g, s->spec.lights, 0, o, ph);
        for (int ch = 0; ch < 3; ++ch) {
          const float x = std::isnan(c[ch]) ? 0.0f : c[ch];
          sum[px * 3 + ch] += x;
          if (sumsq) sumsq[px * 3 + ch] += x * x;
        }
      }
    }
  });
  return 0;
}

}  // extern "C"

