// Host implementation of the scene-building API: constructors, bounding boxes,
// the BVH build, drand_r, the camera basis and the Perlin tables.
//
// Every function here must reproduce the reference's host arithmetic bit for
// bit, because its results (node boxes, BVH topology, rotate_y sin/cos,
// camera vectors, Perlin gradients) are copied verbatim to the device and the
// closest-hit kernel is checked bit-exactly against the reference's
// hitable::hit.  Compiled by g++ with -ffp-contract=off and no fast-math.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <random>
#include <stdexcept>

#include "../scene/camera.h"
#include "../scene/hitable.h"
#include "../scene/hitable_list.h"
#include "../scene/material.h"
#include "../scene/perlin_noise.h"
#include "../scene/rect_box.h"
#include "../scene/sphere.h"
#include "../scene/texture.h"
#include "../scene/utils.h"
#include "image_load.h"

// ---------------------------------------------------------------- drand_r ---
namespace {
std::mt19937 &scene_generator() {
  static thread_local std::mt19937 gen;  // default seed 5489, per thread
  return gen;
}
}  // namespace

// Reference src/utils.cc:28-32: one uniform_real_distribution<double> draw
// (two 32-bit mt19937 words in libstdc++) from the thread's generator.
double drand_r(double min, double max) {
  std::uniform_real_distribution<double> dist(min, max);
  return dist(scene_generator());
}

void tpt::drand_seed(uint32_t seed) { scene_generator().seed(seed); }

// ------------------------------------------------------------------- AABB ---
AABB surrounding_box(AABB box0, AABB box1) {
  vec3 lo, hi;
  for (int i = 0; i < 3; ++i) {
    lo.e[i] = float_min(box0.min_.e[i], box1.min_.e[i]);
    hi.e[i] = float_max(box0.max_.e[i], box1.max_.e[i]);
  }
  return AABB(lo, hi);
}

// ---------------------------------------------------------------- bvh_node ---
namespace {
// Orders hitables by their box's lower corner on one axis.  Same predicate as
// the reference's box_axis_compare (src/hitable.cc:10-30): "a before b" unless
// isgreaterequal(a.min[axis], b.min[axis]); boxes are queried at time (0,0).
struct lower_corner_before {
  int axis;
  bool operator()(const hitable *a, const hitable *b) const {
    AABB ba, bb;
    const bool ok_a = a->bounding_box(0, 0, ba);
    const bool ok_b = b->bounding_box(0, 0, bb);
    if (!ok_a || !ok_b) throw std::runtime_error("bvh_node: a child has no bounding box");
    return !std::isgreaterequal(ba.min_.e[axis], bb.min_.e[axis]);
  }
};
}  // namespace

bvh_node::bvh_node(hitable **l, int n, float time0, float time1) {
  // The axis draw precedes the sort and the child builds (pre-order draws).
  const int axis = static_cast<int>(drand_r(0, 3.0));
  std::sort(l, l + n, lower_corner_before{axis < 0 ? 0 : (axis > 2 ? 2 : axis)});
  if (n == 1) {
    left_ = right_ = l[0];
  } else if (n == 2) {
    left_ = l[0];
    right_ = l[1];
  } else {
    const int half = n / 2;
    left_ = new bvh_node(l, half, time0, time1);
    right_ = new bvh_node(l + half, n - half, time0, time1);
  }
  AABB lb, rb;
  const bool ok = left_->bounding_box(time0, time1, lb) && right_->bounding_box(time0, time1, rb);
  if (!ok) std::fprintf(stderr, "bvh_node: child without bounding box\n");
  box_ = surrounding_box(lb, rb);
}

bool bvh_node::bounding_box(float, float, AABB &box) const {
  box = box_;
  return true;
}

// ----------------------------------------------------------- hitable_list ---
bool hitable_list::bounding_box(float t0, float t1, AABB &box) const {
  if (list_size_ < 1) return false;
  AABB acc, cur;
  if (!list_[0]->bounding_box(t0, t1, acc)) return false;
  for (int i = 1; i < list_size_; ++i) {
    if (!list_[i]->bounding_box(t0, t1, cur)) return false;
    acc = surrounding_box(acc, cur);
  }
  box = acc;
  return true;
}

// ----------------------------------------------------------------- sphere ---
bool sphere::bounding_box(float, float, AABB &box) const {
  const vec3 r(radius_, radius_, radius_);
  box = AABB(center_ - r, center_ + r);
  return true;
}

bool moving_sphere::bounding_box(float t0, float t1, AABB &box) const {
  const vec3 r(radius_, radius_, radius_);
  const AABB a(center(t0) - r, center(t0) + r);
  const AABB b(center(t1) - r, center(t1) + r);
  box = surrounding_box(a, b);
  return true;
}

// ------------------------------------------------------------------ rects ---
// The ±0.0001 slab padding is applied in double and rounded to float, as the
// reference's `k_ - 0.0001` does (src/rect_box.cc:3-6, 45-49, 70-73).
namespace {
inline float pad_lo(float k) { return static_cast<float>(static_cast<double>(k) - 0.0001); }
inline float pad_hi(float k) { return static_cast<float>(static_cast<double>(k) + 0.0001); }
}  // namespace

bool xy_rect::bounding_box(float, float, AABB &box) const {
  box = AABB(vec3(x0_, y0_, pad_lo(k_)), vec3(x1_, y1_, pad_hi(k_)));
  return true;
}
bool xz_rect::bounding_box(float, float, AABB &box) const {
  box = AABB(vec3(x0_, pad_lo(k_), z0_), vec3(x1_, pad_hi(k_), z1_));
  return true;
}
bool yz_rect::bounding_box(float, float, AABB &box) const {
  box = AABB(vec3(pad_lo(k_), y0_, z0_), vec3(pad_hi(k_), y1_, z1_));
  return true;
}

// -------------------------------------------------------------------- box ---
box::box(vec3 pmin, vec3 pmax, material *mat) : point_min_(pmin), point_max_(pmax) {
  hitable **faces = new hitable *[6];
  faces[0] = new xy_rect(pmin.x(), pmax.x(), pmin.y(), pmax.y(), pmax.z(), mat);
  faces[1] = new flip_normal(new xy_rect(pmin.x(), pmax.x(), pmin.y(), pmax.y(), pmin.z(), mat));
  faces[2] = new xz_rect(pmin.x(), pmax.x(), pmin.z(), pmax.z(), pmax.y(), mat);
  faces[3] = new flip_normal(new xz_rect(pmin.x(), pmax.x(), pmin.z(), pmax.z(), pmin.y(), mat));
  faces[4] = new yz_rect(pmin.y(), pmax.y(), pmin.z(), pmax.z(), pmax.x(), mat);
  faces[5] = new flip_normal(new yz_rect(pmin.y(), pmax.y(), pmin.z(), pmax.z(), pmin.x(), mat));
  list_ptr_ = new hitable_list(faces, 6);
}

// -------------------------------------------------------------- translate ---
bool translate::bounding_box(float t0, float t1, AABB &box) const {
  if (!ptr_->bounding_box(t0, t1, box)) return false;
  box = AABB(box.min_ + offset_, box.max_ + offset_);
  return true;
}

// --------------------------------------------------------------- rotate_y ---
// Reference src/rect_box.cc:121-169: radians rounded to float, sinf/cosf, and
// the rotated box by the per-column min/max method.
rotate_y::rotate_y(hitable *p, float angle) : ptr_(p) {
  const float radians = static_cast<float>(static_cast<double>(angle / 180.0f) * M_PI);
  sin_theta_ = std::sin(radians);
  cos_theta_ = std::cos(radians);
  has_box_ = ptr_->bounding_box(0, 0, box_);
  const vec3 col_x(cos_theta_, 0, -sin_theta_);
  const vec3 col_y(0, 1, 0);
  const vec3 col_z(sin_theta_, 0, cos_theta_);
  const vec3 xa = col_x * box_.max_.x(), xb = col_x * box_.min_.x();
  const vec3 ya = col_y * box_.max_.y(), yb = col_y * box_.min_.y();
  const vec3 za = col_z * box_.max_.z(), zb = col_z * box_.min_.z();
  const vec3 lo = vec_min(xa, xb) + vec_min(ya, yb) + vec_min(za, zb);
  const vec3 hi = vec_max(xa, xb) + vec_max(ya, yb) + vec_max(za, zb);
  box_ = AABB(lo, hi);
}

// -------------------------------------------------------- constant_medium ---
constant_medium::constant_medium(hitable *boundary, float density, texture *tex)
    : boundary_(boundary), density_(density), phase_funcion_(new isotropic(tex)) {}

// ----------------------------------------------------------------- camera ---
// Reference src/camera.cc:2-21.  theta and tan() are evaluated in double and
// rounded to float (the unqualified tan() there resolves to ::tan(double)).
camera_with_blur::camera_with_blur(vec3 lookfrom, vec3 lookat, vec3 vup, float vfov, float aspect,
                                   float aperture, float focus_dist, float t0, float t1) {
  time0 = t0;
  time1 = t1;
  lens_radius_ = aperture / 2;
  const float theta = static_cast<float>(static_cast<double>(vfov) * M_PI / 180);
  const float half_height = static_cast<float>(::tan(static_cast<double>(theta / 2)));
  const float half_width = aspect * half_height;
  origin_ = lookfrom;
  w_ = unit_vector(lookfrom - lookat);
  u_ = unit_vector(cross(vup, w_));
  v_ = unit_vector(cross(w_, u_));
  lower_left_corner_ =
      origin_ - (half_width * focus_dist) * u_ - (half_height * focus_dist) * v_ - w_ * focus_dist;
  horizontal_ = ((2 * half_width) * u_) * focus_dist;
  vertical_ = ((2 * half_height) * v_) * focus_dist;
}

// ----------------------------------------------------------- perlin_noise ---
std::array<vec3, 256> perlin_noise::random_vec3_;
std::array<int, 256> perlin_noise::permute_x_;
std::array<int, 256> perlin_noise::permute_y_;
std::array<int, 256> perlin_noise::permute_z_;

uint64_t &perlin_noise::shuffle_seed() {
  static uint64_t seed = 0x5EED;
  return seed;
}

// Reference src/perlin_noise.cc:3-21: 256 unit gradients from three
// sequenced drand_r draws each, then three shuffles of the identity with one
// default_random_engine (seeded from shuffle_seed() instead of the clock).
perlin_noise::perlin_noise() {
  for (vec3 &g : random_vec3_) {
    const float a = static_cast<float>(2 * drand_r() - 1);
    const float b = static_cast<float>(2 * drand_r() - 1);
    const float c = static_cast<float>(2 * drand_r() - 1);
    g = unit_vector(vec3(a, b, c));
  }
  for (int i = 0; i < 256; ++i) permute_x_[i] = permute_y_[i] = permute_z_[i] = i;
  std::default_random_engine engine(static_cast<std::default_random_engine::result_type>(shuffle_seed()));
  std::shuffle(permute_x_.begin(), permute_x_.end(), engine);
  std::shuffle(permute_y_.begin(), permute_y_.end(), engine);
  std::shuffle(permute_z_.begin(), permute_z_.end(), engine);
}

// ------------------------------------------------------ load_image_texture ---
// src/utils.cc:236-240 returns stbi_load's buffer (rows top first, `channels`
// bytes per pixel); image_texture reads 3 bytes per pixel.  This build's own
// decoder (host/image_load.cc) always returns 3 channels (gray replicated).
unsigned char *load_image_texture(std::string filename, int &width, int &height, int &channels) {
  std::vector<uint8_t> rgb;
  std::string err;
  int w = 0, h = 0;
  if (!tpt::load_image_rgb8(filename, &rgb, &w, &h, &err)) return nullptr;
  auto *data = static_cast<unsigned char *>(std::malloc(rgb.size()));
  if (!data) return nullptr;
  std::memcpy(data, rgb.data(), rgb.size());
  width = w;
  height = h;
  channels = 3;
  return data;
}
