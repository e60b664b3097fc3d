// Device shading: textures, light-list pdf/sampling, material scatter.
//
// Follows the reference's per-ray code (cited per function) in float, with
// the same draw order as the CPU oracle so both consume one Philox stream
// per (pixel, sample, bounce).  Shading needs statistical parity only (the
// reference itself mixes glibc float/double transcendentals), so sin/cos/
// pow use the device's implementations.
#ifndef TPT_CSRC_SHADE_H
#define TPT_CSRC_SHADE_H

#include "traverse.h"

#pragma clang fp contract(off)

namespace tptd {

constexpr float kFltMax = 3.402823466e38f;
constexpr float kPi = 3.14159265358979323846f;

// ------------------------------------------------------------- samplers ---
// random_in_unit_disk / random_in_unit_sphere (src/utils.cc:13-27): rejection,
// draws in component order.
__device__ __forceinline__ v3 in_unit_disk(stream &R) {
  v3 p;
  do {
    const float a = R.next();
    const float b = R.next();
    p = v3{2.0f * a - 1.0f, 2.0f * b - 1.0f, 0.0f};
  } while (p.x * p.x + p.y * p.y + p.z * p.z >= 1.0f);
  return p;
}

__device__ __forceinline__ v3 in_unit_sphere(stream &R) {
  v3 p;
  do {
    const float a = R.next();
    const float b = R.next();
    const float c = R.next();
    p = v3{2.0f * a - 1.0f, 2.0f * b - 1.0f, 2.0f * c - 1.0f};
  } while (len(p) >= 1.0f);
  return p;
}

// random_on_hemisphere (src/utils.cc:427-435): cosine-weighted, local frame.
__device__ __forceinline__ v3 on_hemisphere(stream &R) {
  const float r1 = R.next();
  const float r2 = R.next();
  const float phi = static_cast<float>(2.0 * M_PI * static_cast<double>(r1));
  float s, c;
  sincosf(phi, &s, &c);
  const float sr = sqrtf(r2);
  return v3{c * sr, s * sr, sqrtf(1.0f - r2)};
}

// onb::build_from_w (src/utils.cc:437-450); w is the normal as given.
struct onb {
  v3 u, v, w;
  __device__ __forceinline__ explicit onb(v3 n) {
    w = n;
    const v3 helper = fabsf(n.x) > 0.9f ? v3{0.0f, 1.0f, 0.0f} : v3{1.0f, 0.0f, 0.0f};
    v = unit(cross(n, helper));
    u = cross(v, w);
  }
  __device__ __forceinline__ v3 local(v3 a) const { return a.x * u + a.y * v + a.z * w; }
};

__device__ __forceinline__ v3 reflect(v3 v, v3 n) { return v - (2.0f * dot(v, n)) * n; }

// refract (src/utils.cc:34-44).
__device__ __forceinline__ bool refract(v3 v, v3 n, float ni_over_nt, v3 *out) {
  const v3 uv = unit(v);
  const float dt = dot(uv, n);
  const float disc = 1.0f - ni_over_nt * ni_over_nt * (1.0f - dt * dt);
  if (disc > 0.0f) {
    *out = ni_over_nt * (uv - n * dt) - n * sqrtf(disc);
    return true;
  }
  return false;
}

// schlick (src/utils.cc:46-51).
__device__ __forceinline__ float schlick(float cosine, float ior) {
  float r0 = (1.0f - ior) / (1.0f + ior);
  r0 = r0 * r0;
  const float x = 1.0f - cosine;
  const float x2 = x * x;
  return r0 + (1.0f - r0) * (x2 * x2 * x);
}

// -------------------------------------------------------------- textures ---
// perlin_noise::noise / perlin_interpolate / turb (src/utils.cc:160-225): the
// quintic fade is evaluated in double as the reference's pow() calls are, and
// the gradient offsets use the faded weights, as there.
__device__ __forceinline__ float fade(float x) {
  const double d = x, d3 = d * d * d, d4 = d3 * d, d5 = d4 * d;
  return static_cast<float>(6.0 * d5 - 15.0 * d4 + 10.0 * d3);
}

__device__ float perlin_noise(const tpt_perlin *P, v3 p) {
  const float fx = floorf(p.x), fy = floorf(p.y), fz = floorf(p.z);
  const float u = fade(p.x - fx), v = fade(p.y - fy), w = fade(p.z - fz);
  const int i = static_cast<int>(fx) & 255, j = static_cast<int>(fy) & 255, k = static_cast<int>(fz) & 255;
  float acc = 0.0f;
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int c = 0; c < 2; c++) {
        const int h = P->perm[0][(i + a) & 255] ^ P->perm[1][(j + b) & 255] ^ P->perm[2][(k + c) & 255];
        const float4 g = *reinterpret_cast<const float4 *>(P->grad[h]);
        const float d = g.x * (u - a) + g.y * (v - b) + g.z * (w - c);
        acc += (a * u + (1 - a) * (1.0f - u)) * (b * v + (1 - b) * (1.0f - v)) * (c * w + (1 - c) * (1.0f - w)) * d;
      }
  return fabsf(acc);
}

__device__ float turb(const tpt_perlin *P, v3 p, uint32_t depth) {
  float acc = 0.0f, weight = 1.0f;
  for (uint32_t i = 0; i < depth; i++) {
    acc += weight * perlin_noise(P, p);
    weight *= 0.5f;
    p = p * 2.0f;
  }
  return fabsf(acc);
}

// SIMPLE (template flag of the shading code): every texture of the scene is a
// constant_texture and the light root is absent or one xz_rect, so the
// Perlin, image, checker, sphere-light and light-list code is left out
// (fewer registers, more waves).

// texture::value for the four textures (headers/texture.h:15-24,
// src/texture.cc:4-42).  Checker nesting is followed iteratively.
template <bool SIMPLE = false>
__device__ v3 tex_value(const scene_view &S, uint32_t tex, float u, float v, v3 p) {
  for (int guard = 0; guard < 32; ++guard) {
    const tpt_texture T = S.textures[tex];
    if (SIMPLE || T.kind == TPT_TEX_CONSTANT) return v3{T.p[0], T.p[1], T.p[2]};
    if (T.kind == TPT_TEX_CHECKER) {
      const float s = sinf(10.0f * p.x) * sinf(10.0f * p.y) * sinf(10.0f * p.z);
      tex = s < 0.0f ? T.a : T.b;
      continue;
    }
    if (T.kind == TPT_TEX_PERLIN) {
      const float s = 1.0f + sinf(T.p[0] * p.z + 10.0f * turb(S.perlin + T.b, p, T.a));  // b: the table of its scene
      const float g = 0.5f * s;
      return v3{g, g, g};
    }
    // image: nearest texel, clamped.
    const int w = static_cast<int>(T.b), h = static_cast<int>(T.c);
    int i = static_cast<int>(u * static_cast<float>(w));
    int j = static_cast<int>((1.0f - v) * static_cast<float>(h));
    i = i < 0 ? 0 : (i > w - 1 ? w - 1 : i);
    j = j < 0 ? 0 : (j > h - 1 ? h - 1 : j);
    const uint8_t *px = S.texels + T.a + 3 * (static_cast<size_t>(i) + static_cast<size_t>(w) * j);
    return v3{px[0] / 255.0f, px[1] / 255.0f, px[2] / 255.0f};
  }
  return v3{0.0f, 0.0f, 0.0f};
}

// ------------------------------------------------------------ light list ---
// pdf_value of one light (src/rect_box.cc:26-37, src/sphere.cc:93-106,
// headers/hitable.h:35-37 for other hitables).
template <bool SIMPLE = false>
__device__ __forceinline__ float one_light_pdf(const tpt_light &L, v3 o, v3 d) {
  if (SIMPLE || L.kind == TPT_LIGHT_XZ_RECT) {
    const float f[5] = {L.p[4], L.p[0], L.p[1], L.p[2], L.p[3]};
    float t;
    if (rect_t(f, 1, o, d, 0.0001f, kFltMax, &t)) {
      const float area = fabsf((L.p[1] - L.p[0]) * (L.p[3] - L.p[2]));
      const v3 td = t * d;
      const float dist2 = dot(td, td);
      const float cosine = fabsf(d.y / len(d));
      return dist2 / (cosine * area);
    }
    return 0.0f;
  }
  if (L.kind == TPT_LIGHT_SPHERE) {
    float t;
    if (sphere_t(L.p[0], L.p[1], L.p[2], L.p[3], o, d, 0.001f, kFltMax, &t)) {
      const v3 c{L.p[0], L.p[1], L.p[2]};
      const v3 co = c - o;
      const float tmp = (L.p[3] * L.p[3]) / dot(co, co);
      const float cmax = sqrtf(1.0f - tmp);
      const float solid = 2.0f * kPi * (1.0f - cmax);
      if (solid != solid) return 0.0f;
      return 1.0f / solid;
    }
    return 0.0f;
  }
  return 0.0f;
}

// hitable_pdf::value on the light root (src/utils.cc:452-454): a bare shape's
// pdf, or hitable_list's weighted sum (src/hitable_list.cc:24-31).
template <bool SIMPLE = false>
__device__ __forceinline__ float light_pdf(const scene_view &S, v3 o, v3 d) {
  if (SIMPLE || S.light_mode == 0) return one_light_pdf<SIMPLE>(S.lights[0], o, d);
  float sum = 0.0f;
  for (uint32_t i = 0; i < S.n_lights; ++i) {
    const tpt_light L = S.lights[i];
    sum += L.weight * one_light_pdf(L, o, d);
  }
  return sum;
}

// hitable_pdf::generate (src/utils.cc:456): pick (list: int(u·n)), then the
// shape's random(): a uniform point on the rect (src/rect_box.cc:39-43) or a
// cone direction toward the sphere (src/sphere.cc:108-120).
template <bool SIMPLE = false>
__device__ __forceinline__ v3 light_random(const scene_view &S, v3 o, stream &R) {
  uint32_t idx = 0;
  if (SIMPLE) {
  } else if (S.light_mode == 1) {
    idx = static_cast<uint32_t>(static_cast<int>(static_cast<double>(R.next()) * S.n_lights));
    if (idx >= S.n_lights) idx = S.n_lights - 1;
  } else if (S.light_mode == 2) {
    const float u = R.next();
    float acc = 0.0f;
    idx = S.n_lights - 1;
    for (uint32_t i = 0; i < S.n_lights; ++i) {
      acc += S.lights[i].weight;
      if (u < acc) {
        idx = i;
        break;
      }
    }
  }
  const tpt_light L = S.lights[idx];
  if (SIMPLE || L.kind == TPT_LIGHT_XZ_RECT) {
    const float a = R.next();
    const float b = R.next();
    const v3 pt{L.p[0] + a * (L.p[1] - L.p[0]), L.p[4], L.p[2] + b * (L.p[3] - L.p[2])};
    return pt - o;
  }
  if (L.kind == TPT_LIGHT_SPHERE) {
    const v3 c{L.p[0], L.p[1], L.p[2]};
    const v3 dir = c - o;
    const onb uvw(unit(dir));
    const float tmp = (L.p[3] * L.p[3]) / dot(dir, dir);
    const float cmax = sqrtf(1.0f - tmp);
    const float r1 = R.next();
    const float r2 = R.next();
    const float z = 1.0f + r2 * (cmax - 1.0f);
    float s, co;
    sincosf(2.0f * kPi * r1, &s, &co);
    const float sz = sqrtf(1.0f - z * z);
    return uvw.local(v3{co * sz, s * sz, z});
  }
  return v3{1.0f, 0.0f, 0.0f};
}

__device__ __forceinline__ v3 sky(v3 d, float scale) {
  const v3 u = unit(d);
  const float t = 0.5f * (u.y + 1.0f);
  return scale * ((1.0f - t) * v3{1.0f, 1.0f, 1.0f} + t * v3{0.5f, 0.7f, 1.0f});
}

}  // namespace tptd

#endif  // TPT_CSRC_SHADE_H
