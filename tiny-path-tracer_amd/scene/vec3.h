// vec3 — three-float vector of the scene-building API.
//
// Same layout and interface as the reference's vec3 (headers/vec3.h:6-159):
// a public `float e[3]`, x()/y()/z() and r()/g()/b() accessors, the
// arithmetic operators, dot/cross/unit_vector and vec_min/vec_max, so scene
// code written against the reference compiles unchanged.  Host-side only:
// scene construction, bounding boxes and the flattener use it.  Every
// operation is a single IEEE float op in the written order (the library is
// compiled with -ffp-contract=off), because bounding boxes, rotate_y
// sines/cosines and camera vectors are copied bit-for-bit to the device.
#ifndef TPT_SCENE_VEC3_H
#define TPT_SCENE_VEC3_H

#include <algorithm>
#include <cmath>
#include <iostream>

struct vec3 {
  float e[3];

  vec3() : e{0.0f, 0.0f, 0.0f} {}
  vec3(float a, float b, float c) : e{a, b, c} {}
  vec3(float s) : e{s, s, s} {}  // implicit on purpose, as in the reference

  float x() const { return e[0]; }
  float y() const { return e[1]; }
  float z() const { return e[2]; }
  float r() const { return e[0]; }
  float g() const { return e[1]; }
  float b() const { return e[2]; }

  const vec3 &operator+() const { return *this; }
  vec3 operator-() const { return vec3(-e[0], -e[1], -e[2]); }
  float operator[](int i) const { return e[i]; }
  float &operator[](int i) { return e[i]; }

  vec3 &operator+=(const vec3 &o) {
    for (int i = 0; i < 3; ++i) e[i] += o.e[i];
    return *this;
  }
  vec3 &operator-=(const vec3 &o) {
    for (int i = 0; i < 3; ++i) e[i] -= o.e[i];
    return *this;
  }
  vec3 &operator*=(const vec3 &o) {
    for (int i = 0; i < 3; ++i) e[i] *= o.e[i];
    return *this;
  }
  vec3 &operator/=(const vec3 &o) {
    for (int i = 0; i < 3; ++i) e[i] /= o.e[i];
    return *this;
  }
  vec3 &operator*=(float s) {
    for (int i = 0; i < 3; ++i) e[i] *= s;
    return *this;
  }
  vec3 &operator/=(float s) {
    for (int i = 0; i < 3; ++i) e[i] /= s;
    return *this;
  }

  float squared_length() const { return e[0] * e[0] + e[1] * e[1] + e[2] * e[2]; }
  float length() const { return std::sqrt(squared_length()); }
  // Reciprocal taken in double, then scaled in float (headers/vec3.h:95-100).
  void make_unit_vector() {
    const float k = static_cast<float>(1.0 / static_cast<double>(length()));
    *this *= k;
  }
  void normalize() { make_unit_vector(); }
};

inline vec3 operator+(const vec3 &a, const vec3 &b) {
  return vec3(a.e[0] + b.e[0], a.e[1] + b.e[1], a.e[2] + b.e[2]);
}
inline vec3 operator-(const vec3 &a, const vec3 &b) {
  return vec3(a.e[0] - b.e[0], a.e[1] - b.e[1], a.e[2] - b.e[2]);
}
inline vec3 operator*(const vec3 &a, const vec3 &b) {
  return vec3(a.e[0] * b.e[0], a.e[1] * b.e[1], a.e[2] * b.e[2]);
}
inline vec3 operator/(const vec3 &a, const vec3 &b) {
  return vec3(a.e[0] / b.e[0], a.e[1] / b.e[1], a.e[2] / b.e[2]);
}
inline vec3 operator*(const vec3 &a, float s) { return vec3(a.e[0] * s, a.e[1] * s, a.e[2] * s); }
inline vec3 operator*(float s, const vec3 &a) { return a * s; }
inline vec3 operator/(const vec3 &a, float s) { return vec3(a.e[0] / s, a.e[1] / s, a.e[2] / s); }

inline float dot(const vec3 &a, const vec3 &b) {
  return a.e[0] * b.e[0] + a.e[1] * b.e[1] + a.e[2] * b.e[2];
}
inline vec3 cross(const vec3 &a, const vec3 &b) {
  return vec3(a.e[1] * b.e[2] - b.e[1] * a.e[2], a.e[2] * b.e[0] - b.e[2] * a.e[0],
              a.e[0] * b.e[1] - b.e[0] * a.e[1]);
}
inline vec3 normalize(vec3 v) { return v / v.length(); }
inline vec3 unit_vector(vec3 v) { return v / v.length(); }
inline vec3 vec_min(const vec3 &a, const vec3 &b) {
  return vec3(std::min(a.e[0], b.e[0]), std::min(a.e[1], b.e[1]), std::min(a.e[2], b.e[2]));
}
inline vec3 vec_max(const vec3 &a, const vec3 &b) {
  return vec3(std::max(a.e[0], b.e[0]), std::max(a.e[1], b.e[1]), std::max(a.e[2], b.e[2]));
}

inline std::istream &operator>>(std::istream &is, vec3 &v) { return is >> v.e[0] >> v.e[1] >> v.e[2]; }
inline std::ostream &operator<<(std::ostream &os, const vec3 &v) {
  return os << v.e[0] << ' ' << v.e[1] << ' ' << v.e[2];
}

#endif  // TPT_SCENE_VEC3_H
