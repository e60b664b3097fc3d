// Closest-hit traversal on the flattened scene, bit-exact with the
// reference's recursive hitable::hit.
//
// How the reference's semantics map onto the flat layout (include/tpt.h):
//  * bvh_node::hit (src/hitable.cc:63-90) tests its box with the caller's
//    (t_min, t_max), then both children with the SAME t_max, and keeps the
//    left record only if strictly closer.  The set of boxes tested never
//    depends on hits found, so a pre-order walk with skip links (box miss ->
//    jump past the subtree) visits exactly the reference's nodes in its
//    order, with no stack.  For non-NaN t the nested "left if strictly
//    closer" combines equal "last visited among the minima", i.e. replace on
//    t <= best.  A NaN t (an in-plane ray on a rect: 0/0, accepted by the
//    reference's `t > t_max || t < t_min` test) makes the combine depend on
//    the tree shape, so a segment that produced one is re-evaluated by
//    segment_exact(), which replays the recursion with an explicit stack.
//  * AABB::hit (src/aabb.cc:3-19): 1.0f/d per axis (precomputed once per ray,
//    same bits), swap iff inv < 0, NaN-ignoring ternaries.  No early exit is
//    needed: t_min only grows and t_max only shrinks, so the final test equals
//    the per-axis one.
//  * sphere::hit (src/sphere.cc:13-43): float a, b, c and discriminant; roots
//    in FP64 ((double)-b -+ sqrt(disc)) / (double)(2a), rounded to float.
//  * rect hit (src/rect_box.cc:8-91): t = (k - o_axis) / d_axis (IEEE
//    division), bounds inclusive.
//  * hitable_list (src/hitable_list.cc:38-51), box, translate, rotate_y,
//    flip_normal: a LIST leaf holds its primitives in evaluation order with
//    their wrapper chains; the list's t_max shrinks to each accepted t.
#ifndef TPT_CSRC_TRAVERSE_H
#define TPT_CSRC_TRAVERSE_H

#include "dev_common.h"

#pragma clang fp contract(off)

namespace tptd {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr float kFltMaxT = 3.402823466e38f;  // std::numeric_limits<float>::max()
// Sphere-only walk: a box whose subtree holds a sphere of radius above
// kWalkBigRadius has kBoxUnbounded set in its skip word (device copy only;
// skip links are below 2^28) and is always tested against t_max.
constexpr float kWalkBigRadius = 2.0f;
constexpr uint32_t kBoxUnbounded = 0x80000000u;

struct hit_t {
  float t;
  uint32_t slot;   // slot index of the winning leaf
  uint32_t entry;  // entry index for LIST leaves, kNone otherwise
  uint32_t mat;    // material word (material | axis<<28 | flip<<30)
};

struct slot_t {
  float f[6];
  uint32_t a, b;
};

// Per-query context: the ray's time (moving spheres) and the Philox counter
// words and key of its media draws (a medium with leaf id L draws from
// counter (c0, c1, c2, 2^20 + L)).
struct ray_ctx {
  float time;
  uint32_t c0, c1, c2, k0, k1;
};

__device__ __forceinline__ slot_t load_slot(const tpt_slot *p) {
  const float4 *q = reinterpret_cast<const float4 *>(p);
  const float4 x = q[0], y = q[1];
  slot_t s;
  s.f[0] = x.x, s.f[1] = x.y, s.f[2] = x.z, s.f[3] = x.w, s.f[4] = y.x, s.f[5] = y.y;
  s.a = __float_as_uint(y.z);
  s.b = __float_as_uint(y.w);
  return s;
}

__device__ __forceinline__ bool slab(const slot_t &s, v3 o, v3 inv, float tmin, float tmax) {
  float t0 = (s.f[0] - o.x) * inv.x, t1 = (s.f[3] - o.x) * inv.x;
  if (inv.x < 0.0f) {
    const float k = t0;
    t0 = t1;
    t1 = k;
  }
  tmin = t0 > tmin ? t0 : tmin;
  tmax = t1 < tmax ? t1 : tmax;
  t0 = (s.f[1] - o.y) * inv.y, t1 = (s.f[4] - o.y) * inv.y;
  if (inv.y < 0.0f) {
    const float k = t0;
    t0 = t1;
    t1 = k;
  }
  tmin = t0 > tmin ? t0 : tmin;
  tmax = t1 < tmax ? t1 : tmax;
  t0 = (s.f[2] - o.z) * inv.z, t1 = (s.f[5] - o.z) * inv.z;
  if (inv.z < 0.0f) {
    const float k = t0;
    t0 = t1;
    t1 = k;
  }
  tmin = t0 > tmin ? t0 : tmin;
  tmax = t1 < tmax ? t1 : tmax;
  return !(tmin > tmax);
}

// The same test when t_min and t_max are not NaN.  The reference's updates
// `t0 > tmin ? t0 : tmin` and `t1 < tmax ? t1 : tmax` then equal IEEE maxNum
// and minNum: a NaN t0 or t1 leaves the bound unchanged in both forms, and the
// bounds stay non-NaN; where the forms pick a zero of the other sign the value
// still compares equal, so every later comparison and the result agree.  The
// swap becomes a select on the ray's fixed direction signs.
__device__ __forceinline__ bool slab_nn(const slot_t &s, v3 o, v3 inv, float tmin, float tmax) {
  const float x0 = (s.f[0] - o.x) * inv.x, x1 = (s.f[3] - o.x) * inv.x;
  const float y0 = (s.f[1] - o.y) * inv.y, y1 = (s.f[4] - o.y) * inv.y;
  const float z0 = (s.f[2] - o.z) * inv.z, z1 = (s.f[5] - o.z) * inv.z;
  const bool nx = inv.x < 0.0f, ny = inv.y < 0.0f, nz = inv.z < 0.0f;
  tmin = fmaxf(nx ? x1 : x0, tmin);
  tmax = fminf(nx ? x0 : x1, tmax);
  tmin = fmaxf(ny ? y1 : y0, tmin);
  tmax = fminf(ny ? y0 : y1, tmax);
  tmin = fmaxf(nz ? z1 : z0, tmin);
  tmax = fminf(nz ? z0 : z1, tmax);
  return !(tmin > tmax);
}

__device__ __forceinline__ bool sphere_t(float cx, float cy, float cz, float r, v3 o, v3 d, float tmin,
                                         float tmax, float *t) {
  const float ox = o.x - cx, oy = o.y - cy, oz = o.z - cz;
  const float a = d.x * d.x + d.y * d.y + d.z * d.z;
  const float b = 2.0f * (d.x * ox + d.y * oy + d.z * oz);  // == (float)(2.0 * (double)dot), exact scaling
  const float c = (ox * ox + oy * oy + oz * oz) - r * r;
  const float disc = b * b - (4.0f * a) * c;
  if (disc > 0.0f) {
    const double sq = __builtin_sqrt(static_cast<double>(disc));
    const double den = static_cast<double>(2.0f * a);
    const double nb = static_cast<double>(-b);
    const float t0 = static_cast<float>((nb - sq) / den);
    if (t0 < tmax && t0 > tmin) {
      *t = t0;
      return true;
    }
    const float t1 = static_cast<float>((nb + sq) / den);
    if (t1 < tmax && t1 > tmin) {
      *t = t1;
      return true;
    }
  }
  return false;
}

// moving_sphere::center (headers/sphere.h:28-31), in float.
__device__ __forceinline__ v3 motion_center(const tpt_motion &m, float time) {
  const float s = (time - m.time0) / (m.time1 - m.time0);
  return v3{m.center0[0] + s * (m.center1[0] - m.center0[0]), m.center0[1] + s * (m.center1[1] - m.center0[1]),
            m.center0[2] + s * (m.center1[2] - m.center0[2])};
}

// moving_sphere::hit (src/sphere.cc:45-75): sphere::hit's arithmetic about
// the centre at the ray's time.
__device__ __forceinline__ bool msphere_t(const scene_view &S, float idx_bits, v3 o, v3 d, float time, float tmin,
                                          float tmax, float *t) {
  const tpt_motion m = S.motions[__float_as_uint(idx_bits)];
  const v3 c = motion_center(m, time);
  return sphere_t(c.x, c.y, c.z, m.radius, o, d, tmin, tmax, t);
}

// The draws of medium `leaf` for this query: word 0 the free path, 1-3 the
// arbitrary normal of a hit.
__device__ __forceinline__ uint4 medium_draws(const ray_ctx &rc, uint32_t leaf) {
  return philox(rc.c0, rc.c1, rc.c2, (1u << 20) + leaf, rc.k0, rc.k1);
}

// f = (k, a0, a1, b0, b1); axis = normal axis; in-plane axes (i0, i1) are
// (1,2) for yz, (0,2) for xz, (0,1) for xy.
__device__ __forceinline__ bool rect_t(const float *f, uint32_t axis, v3 o, v3 d, float tmin, float tmax,
                                       float *t) {
  const float oa = axis == 0 ? o.x : (axis == 1 ? o.y : o.z);
  const float da = axis == 0 ? d.x : (axis == 1 ? d.y : d.z);
  const float tt = (f[0] - oa) / da;
  if (tt > tmax || tt < tmin) return false;
  const float o0 = axis == 0 ? o.y : o.x, d0 = axis == 0 ? d.y : d.x;
  const float o1 = axis == 2 ? o.y : o.z, d1 = axis == 2 ? d.y : d.z;
  const float x = o0 + tt * d0;
  const float y = o1 + tt * d1;
  if (x < f[1] || x > f[2] || y < f[3] || y > f[4]) return false;
  *t = tt;
  return true;
}

// Ray in the frame of a wrapper chain (outermost op first), as translate::hit
// and rotate_y::hit build it (headers/rect_box.h:87-95, src/rect_box.cc:171-180).
__device__ __forceinline__ void to_local(const tpt_xform &x, v3 *o, v3 *d) {
  for (uint32_t i = 0; i < x.nops; ++i) {
    const uint32_t kind = (x.kinds >> (2 * i)) & 3u;
    if (kind == TPT_XFORM_TRANSLATE) {
      *o = v3{o->x - x.p[i][0], o->y - x.p[i][1], o->z - x.p[i][2]};
    } else {
      const float s = x.p[i][0], c = x.p[i][1];
      *o = v3{c * o->x - s * o->z, o->y, s * o->x + c * o->z};
      *d = v3{c * d->x - s * d->z, d->y, s * d->x + c * d->z};
    }
  }
}

// Inverse mapping of a hit point / normal (innermost op first).
__device__ __forceinline__ void to_world(const tpt_xform &x, v3 *p, v3 *n) {
  for (int i = static_cast<int>(x.nops) - 1; i >= 0; --i) {
    const uint32_t kind = (x.kinds >> (2 * i)) & 3u;
    if (kind == TPT_XFORM_TRANSLATE) {
      *p = v3{p->x + x.p[i][0], p->y + x.p[i][1], p->z + x.p[i][2]};
    } else {
      const float s = x.p[i][0], c = x.p[i][1];
      *p = v3{c * p->x + s * p->z, p->y, -s * p->x + c * p->z};
      *n = v3{c * n->x + s * n->z, n->y, -s * n->x + c * n->z};
    }
  }
}

// Operations of a walk by kind, for the counted VALU work of a render (the
// counting variant of k_trace).  Every walk function takes an optional pointer;
// the other kernels pass a constant null, and with the calls inlined the
// counting code is removed from them.
struct op_counts {
  uint32_t box = 0;     // slab tests (box slots visited)
  uint32_t sphere = 0;  // sphere and moving-sphere tests
  uint32_t rect = 0;    // rect tests
  uint32_t xform = 0;   // wrapper chains applied to a ray (translate / rotate_y)
};

template <bool MEDIA>
__device__ __forceinline__ bool list_eval(const scene_view &S, uint32_t first, uint32_t count, uint32_t leaf0, v3 o,
                                          v3 d, const ray_ctx &rc, float tmin, float tmax, float *t, uint32_t *entry,
                                          uint32_t *mat, op_counts *oc = nullptr);

// constant_medium::hit (src/hitable.cc:98-128): the boundary's first hit over
// the whole line, its next hit from there + 0.0001, clipped to (t_min, t_max);
// the medium stops the ray at a free path of -ln(u)/density along it.
__device__ __forceinline__ bool medium_t(const scene_view &S, const float *f, uint32_t leaf, v3 o, v3 d,
                                         const ray_ctx &rc, float tmin, float tmax, float *t) {
  const uint32_t first = __float_as_uint(f[0]), count = __float_as_uint(f[1]);
  float t1, t2;
  uint32_t e, m;
  if (!list_eval<false>(S, first, count, 0, o, d, rc, -kFltMaxT, kFltMaxT, &t1, &e, &m)) return false;
  const float tmin2 = static_cast<float>(static_cast<double>(t1) + 0.0001);
  if (!list_eval<false>(S, first, count, 0, o, d, rc, tmin2, kFltMaxT, &t2, &e, &m)) return false;
  if (t1 < tmin) t1 = tmin;
  if (t2 > tmax) t2 = tmax;
  if (t1 >= t2) return false;
  const float len = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
  const float inside = (t2 - t1) * len;
  const uint4 w = medium_draws(rc, leaf);
  const float hit_distance =
      static_cast<float>(static_cast<double>(-1.0f / f[2]) * log(static_cast<double>(u01(w.x))));
  if (hit_distance < inside) {
    *t = t1 + hit_distance / len;
    return true;
  }
  return false;
}

// LIST leaf: primitives in order, t_max shrinking to each accepted t (a
// medium's boundary is evaluated the same way, MEDIA = false).
template <bool MEDIA>
__device__ __forceinline__ bool list_eval(const scene_view &S, uint32_t first, uint32_t count, uint32_t leaf0, v3 o,
                                          v3 d, const ray_ctx &rc, float tmin, float tmax, float *t, uint32_t *entry,
                                          uint32_t *mat, op_counts *oc) {
  const float time = rc.time;
  float cur = tmax;
  bool found = false;
  uint32_t cached = kNone;
  v3 lo = o, ld = d;
  for (uint32_t e = first; e < first + count; ++e) {
    const tpt_entry en = S.entries[e];
    if (en.xform != cached) {
      lo = o;
      ld = d;
      if (en.xform != kNone) to_local(S.xforms[en.xform], &lo, &ld);
      if (oc && en.xform != kNone) ++oc->xform;
      cached = en.xform;
    }
    if (oc) {
      if (en.b == TPT_SLOT_SPHERE || en.b == TPT_SLOT_MSPHERE) ++oc->sphere;
      if (en.b == TPT_SLOT_RECT) ++oc->rect;
    }
    float th;
    bool h;
    if (MEDIA && en.b == TPT_SLOT_MEDIUM)
      h = medium_t(S, en.f, leaf0 + (e - first), lo, ld, rc, tmin, cur, &th);
    else
      h = en.b == TPT_SLOT_SPHERE    ? sphere_t(en.f[0], en.f[1], en.f[2], en.f[3], lo, ld, tmin, cur, &th)
          : en.b == TPT_SLOT_MSPHERE ? msphere_t(S, en.f[0], lo, ld, time, tmin, cur, &th)
                                     : rect_t(en.f, (en.a >> 28) & 3u, lo, ld, tmin, cur, &th);
    if (h) {
      cur = th;
      found = true;
      *entry = e;
      *mat = en.a;
    }
  }
  *t = cur;
  return found;
}

// SPH: the scene's leaves are all spheres (no rect or list code is compiled).
// MEDIA: the scene may hold constant_media (their code is compiled only then).
template <bool SPH = false, bool MEDIA = false>
__device__ __forceinline__ bool leaf_t(const scene_view &S, const slot_t &s, uint32_t kind, v3 o, v3 d,
                                       const ray_ctx &rc, float tmin, float tmax, float *t, uint32_t *entry,
                                       uint32_t *mat, op_counts *oc = nullptr) {
  if (oc) {
    if (SPH || kind == TPT_SLOT_SPHERE || kind == TPT_SLOT_MSPHERE) ++oc->sphere;
    if (!SPH && kind == TPT_SLOT_RECT) ++oc->rect;
  }
  if (SPH || kind == TPT_SLOT_SPHERE) {
    *entry = kNone;
    *mat = s.a;
    return sphere_t(s.f[0], s.f[1], s.f[2], s.f[3], o, d, tmin, tmax, t);
  }
  if (kind == TPT_SLOT_RECT) {
    *entry = kNone;
    *mat = s.a;
    return rect_t(s.f, (s.a >> 28) & 3u, o, d, tmin, tmax, t);
  }
  if (kind == TPT_SLOT_MSPHERE) {
    *entry = kNone;
    *mat = s.a;
    return msphere_t(S, s.f[0], o, d, rc.time, tmin, tmax, t);
  }
  if (MEDIA && kind == TPT_SLOT_MEDIUM) {
    *entry = kNone;
    *mat = s.a;
    return medium_t(S, s.f, s.b & TPT_PAYLOAD_MASK, o, d, rc, tmin, tmax, t);
  }
  return list_eval<MEDIA>(S, s.a, __float_as_uint(s.f[1]), __float_as_uint(s.f[0]), o, d, rc, tmin, tmax, t, entry,
                         mat, oc);
}

// Exact replay of the bvh_node recursion for one segment: evaluates each node
// with a fresh result, combining (left, right) as `left.t < right.t ? left :
// right`.  Stack entries: a pending right child, or a finished left result.
template <bool SPH = false, bool MEDIA = false>
__device__ __forceinline__ bool segment_exact(const scene_view &S, uint32_t start, v3 o, v3 d, v3 inv,
                                                        const ray_ctx &rc, float tmin, float tmax, hit_t *out) {
  struct frame {
    uint32_t tag;  // bit31 set: finished-left result; else pending right slot
    uint32_t slot, entry, mat;
    float t;
  };
  frame stk[64];
  int sp = 0;
  bool have = false;
  hit_t acc{0.0f, kNone, kNone, 0};
  uint32_t i = start;
  for (;;) {
    // Evaluate node i; acc holds its result when this block ends.
    for (;;) {
      const slot_t s = load_slot(&S.slots[i]);
      const uint32_t kind = s.b >> 28;
      if (kind == TPT_SLOT_BOX) {
        if (slab(s, o, inv, tmin, tmax)) {
          const uint32_t right = s.b & TPT_PAYLOAD_MASK;
          if (right != TPT_NO_RIGHT && sp < 64) stk[sp++] = frame{right, 0, 0, 0, 0.0f};
          i = i + 1;
          continue;
        }
        have = false;
      } else {
        float t;
        uint32_t entry, mat;
        have = leaf_t<SPH, MEDIA>(S, s, kind, o, d, rc, tmin, tmax, &t, &entry, &mat);
        acc = hit_t{t, i, entry, mat};
      }
      break;
    }
    // Unwind: combine finished results, or start a pending right child.
    bool resumed = false;
    while (sp > 0) {
      const frame f = stk[--sp];
      if (f.tag & 0x80000000u) {
        const bool lhave = (f.tag & 1u) != 0;
        if (lhave && have) {
          if (f.t < acc.t) acc = hit_t{f.t, f.slot, f.entry, f.mat};
        } else if (lhave) {
          acc = hit_t{f.t, f.slot, f.entry, f.mat};
          have = true;
        }
      } else {
        stk[sp++] = frame{0x80000000u | (have ? 1u : 0u), acc.slot, acc.entry, acc.mat, acc.t};
        have = false;
        i = f.tag;
        resumed = true;
        break;
      }
    }
    if (!resumed) break;
  }
  if (have) *out = acc;
  return have;
}

// Box bound of the sphere-only walk once a hit at t is held: no sphere in a
// box entered beyond it has a computed t at or below t.  In units of t (the
// ray's |d| as the unit of length):
//  * a sphere's exact entry lies in its exact box [c - r, c + r]; the box
//    holds c -+ r rounded to float (within 2^-24 M of exact, M the scene's
//    largest coordinate magnitude; bvh boxes are exact unions), so the
//    computed box entry exceeds the sphere's exact entry by at most
//    2^-24 M max|1/d_i| plus two roundings of the slab arithmetic;
//  * sphere_t's float coefficients a, b, c move a root by at most
//    8e-4 (|t| + r/|d|) (of order sqrt(eps) D/|d|, D = |o - c|; worst at
//    tangency, where the rounding of b^2 - 4ac passes through the square root;
//    5.7e-4 the worst measured over 10^6 grazing and general rays).
// The bound adds 2^-6 (|t| + |t_min|), 2^-6 kWalkBigRadius / |d| and
// 2^-22 M max|1/d_i| (reach, per ray): each at least 4x its term, 19x for the
// root error.  Spheres larger than kWalkBigRadius are never bounded (their
// boxes are flagged).  So a skipped box holds only spheres with t strictly
// above the held one, and skipping it changes neither the winner nor a tie.
// A non-finite reach (a zero direction component, an infinite coordinate)
// leaves t_max.
__device__ __forceinline__ float sphere_walk_bound(float t, float tmin, float tmax, float reach) {
  return fminf(tmax, t + (fabsf(t) + fabsf(tmin)) * 0.015625f + reach);
}

// Fast pre-order walk of one segment (t_min, t_max not NaN).  *saw_nan is set
// when a leaf returned a NaN t (the caller then uses segment_exact).
//
// SPH (every leaf a sphere, so no NaN t): boxes are tested against
// sphere_walk_bound of the hit held so far instead of t_max.  The reference
// tests every box the segment crosses (bvh_node::hit passes the same t_max to
// both children); the skipped boxes hold no sphere that could win, so the
// result is the reference's bit for bit.  Leaves keep the caller's bounds.
template <bool SPH = false, bool MEDIA = false>
__device__ __forceinline__ bool segment_fast(const scene_view &S, uint32_t start, uint32_t end, v3 o, v3 d, v3 inv,
                                             const ray_ctx &rc, float tmin, float tmax, hit_t *out, bool *saw_nan,
                                             uint32_t *nodes, op_counts *oc = nullptr) {
  bool have = false;
  hit_t best{0.0f, kNone, kNone, 0};
  uint32_t i = start;
  uint32_t visited = 0;
  float bound = tmax;  // box bound (SPH: shrinks with the held hit)
  float reach = 0.0f;
  if (SPH)
    reach = 0.015625f * kWalkBigRadius * rsqrtf(d.x * d.x + d.y * d.y + d.z * d.z) +
            S.walk_ulp * fmaxf(fmaxf(fabsf(inv.x), fabsf(inv.y)), fabsf(inv.z));
  while (i < end) {
    const slot_t s = load_slot(&S.slots[i]);
    const uint32_t kind = s.b >> 28;
    ++visited;
    uint32_t ni = i + 1;
    if (kind == TPT_SLOT_BOX) {
      if (oc) ++oc->box;
      const bool unbounded = SPH && (s.a & kBoxUnbounded);
      if (!slab_nn(s, o, inv, tmin, unbounded ? tmax : bound)) ni = s.a & TPT_PAYLOAD_MASK;
    } else {
      float t;
      uint32_t entry, mat;
      if (leaf_t<SPH, MEDIA>(S, s, kind, o, d, rc, tmin, tmax, &t, &entry, &mat, oc)) {
        if (t != t) *saw_nan = true;
        if (!have || t <= best.t) {
          best = hit_t{t, i, entry, mat};
          have = true;
          if (SPH) bound = sphere_walk_bound(t, tmin, tmax, reach);
        }
      }
    }
    i = ni;
  }
  *nodes += visited;
  if (have) *out = best;
  return have;
}

// Smallest value over the active lanes of the wave (the device library's
// wavefront reduction, DPP moves instead of LDS permutes).
__device__ __forceinline__ uint32_t wave_min(uint32_t v) { return __ockl_wfred_min_u32(v); }

// Wave-uniform sweep of one segment, for small scenes: the wave steps through
// the slots together, always at the smallest slot some lane still has to
// visit; a lane acts only on the slot equal to its own next pre-order slot.
// Each lane therefore evaluates exactly the slots of its own pre-order walk,
// in the same order (same results as segment_fast), while the slot index is
// wave-uniform: slot data comes through scalar loads and the branch on the
// slot kind does not diverge.  Lanes with `live` false take no part.
// Must be called by all lanes of the wave together.  SPH: a sphere-only
// scene, each lane's boxes bounded by its own held hit as in segment_fast
// (the sweep then serves large trees for rays whose walks mostly coincide,
// such as one wave's camera rays).
template <bool SPH = false>
__device__ __forceinline__ bool segment_sweep(const scene_view &S, uint32_t start, uint32_t end, bool live, v3 o, v3 d,
                                              v3 inv, const ray_ctx &rc, float tmin, float tmax, hit_t *out,
                                              bool *saw_nan, uint32_t *nodes, op_counts *oc = nullptr) {
  bool have = false;
  hit_t best{0.0f, kNone, kNone, 0};
  uint32_t next = live ? start : end;
  uint32_t visited = 0;
  float bound = tmax;
  float reach = 0.0f;
  if (SPH)
    reach = 0.015625f * kWalkBigRadius * rsqrtf(d.x * d.x + d.y * d.y + d.z * d.z) +
            S.walk_ulp * fmaxf(fmaxf(fabsf(inv.x), fabsf(inv.y)), fabsf(inv.z));
  uint32_t i = __builtin_amdgcn_readfirstlane(wave_min(next));
  while (i < end) {
    const slot_t s = load_slot(&S.slots[i]);
    const uint32_t kind = s.b >> 28;
    const bool mine = next == i;
    if (mine) {
      ++visited;
      if (kind == TPT_SLOT_BOX) {
        if (oc) ++oc->box;
        const bool unbounded = !SPH || (s.a & kBoxUnbounded);
        next = slab_nn(s, o, inv, tmin, unbounded ? tmax : bound) ? i + 1 : (s.a & TPT_PAYLOAD_MASK);
      } else {
        float t;
        uint32_t entry, mat;
        if (leaf_t<SPH>(S, s, kind, o, d, rc, tmin, tmax, &t, &entry, &mat, oc)) {
          if (t != t) *saw_nan = true;
          if (!have || t <= best.t) {
            best = hit_t{t, i, entry, mat};
            have = true;
            if (SPH) bound = sphere_walk_bound(t, tmin, tmax, reach);
          }
        }
        next = i + 1;
      }
    }
    // Next slot: i + 1 when any lane goes there (the common case), else the
    // smallest pending skip target.
    const bool step = __ballot(next == i + 1) != 0ull;
    i = step ? i + 1 : __builtin_amdgcn_readfirstlane(wave_min(next));
  }
  *nodes += visited;
  if (have) *out = best;
  return have;
}

// World = segments evaluated like a hitable_list (t_max shrinks between them).
// SWEEP selects segment_sweep (all lanes of the wave must call together).  A
// NaN bound (a caller's t_min, or t_max after a segment whose winner had a
// NaN t) sends the segment to segment_exact, which keeps the reference's slab
// updates verbatim.
template <bool SWEEP = false, bool SPH = false, bool MEDIA = false>
__device__ __forceinline__ bool trace_world(const scene_view &S, v3 o, v3 d, const ray_ctx &rc, float tmin, float tmax,
                                            hit_t *out, uint32_t *nodes, bool live = true, op_counts *oc = nullptr) {
  const v3 inv0{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
  const v3 o0 = o, d0 = d;
  float cur = tmax;
  bool any = false;
  for (uint32_t g = 0; g < S.n_segments; ++g) {
    const uint32_t start = S.segments[2 * g], end = S.segments[2 * g + 1];
    hit_t h;
    bool saw_nan = false;
    if (start >= end) continue;
    // a bvh_node seen through translate / rotate_y: its walk takes the ray
    // those wrappers hand down (headers/rect_box.h:87-95, src/rect_box.cc:171-180);
    // scenes with such segments use neither the sweep nor the sphere-only walk
    o = o0, d = d0;
    v3 inv = inv0;
    const uint32_t sx = !SWEEP && !SPH && S.seg_xf ? S.seg_xf[g] : TPT_SEGMENT_PLAIN;
    if ((sx & TPT_SEGMENT_PLAIN) != TPT_SEGMENT_PLAIN) {
      to_local(S.xforms[sx & TPT_SEGMENT_PLAIN], &o, &d);
      if (oc) ++oc->xform;
      inv = v3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    }
    const bool bounded = tmin == tmin && cur == cur;
    bool got = false;
    if (SWEEP) {
      got = segment_sweep<SPH>(S, start, end, live && bounded, o, d, inv, rc, tmin, cur, &h, &saw_nan, nodes, oc);
    } else {
      if (!live) continue;
      if (bounded) got = segment_fast<SPH, MEDIA>(S, start, end, o, d, inv, rc, tmin, cur, &h, &saw_nan, nodes, oc);
    }
    if (live && (saw_nan || !bounded)) got = segment_exact<SPH, MEDIA>(S, start, o, d, inv, rc, tmin, cur, &h);
    if (got) {
      *out = h;
      cur = h.t;
      any = true;
    }
  }
  return any;
}

// Leaf id reported by tpt_trace_closest.
__device__ __forceinline__ uint32_t leaf_id(const scene_view &S, const hit_t &h) {
  const slot_t s = load_slot(&S.slots[h.slot]);
  if ((s.b >> 28) == TPT_SLOT_LIST) return __float_as_uint(s.f[0]) + (h.entry - s.a);
  return s.b & TPT_PAYLOAD_MASK;
}

// Full hit_record of the winner: point, normal (flip applied), u, v.
// MEDIUM_NORMAL: draw a medium hit's arbitrary normal as constant_medium::hit
// does; the shading kernels skip it (the isotropic phase function ignores the
// normal) and report (0, 1, 0).
template <bool MEDIUM_NORMAL = false>
__device__ __forceinline__ void hit_record(const scene_view &S, const hit_t &h, v3 o, v3 d, const ray_ctx &rc, v3 *p,
                                           v3 *n, float *u, float *v) {
  const slot_t s = load_slot(&S.slots[h.slot]);
  uint32_t kind = s.b >> 28;
  float f[5] = {s.f[0], s.f[1], s.f[2], s.f[3], s.f[4]};
  uint32_t word = s.a;
  v3 lo = o, ld = d;
  uint32_t xf = kNone;
  // the wrappers of the winner's segment, if any (outermost, applied first)
  uint32_t sx = TPT_SEGMENT_PLAIN;
  if (S.seg_xf)
    for (uint32_t g = 0; g < S.n_segments; ++g)
      if (h.slot >= S.segments[2 * g] && h.slot < S.segments[2 * g + 1]) sx = S.seg_xf[g];
  if ((sx & TPT_SEGMENT_PLAIN) != TPT_SEGMENT_PLAIN) to_local(S.xforms[sx & TPT_SEGMENT_PLAIN], &lo, &ld);
  if (kind == TPT_SLOT_LIST) {
    const tpt_entry en = S.entries[h.entry];
    for (int k = 0; k < 5; ++k) f[k] = en.f[k];
    word = en.a;
    kind = en.b;
    xf = en.xform;
    if (xf != kNone) to_local(S.xforms[xf], &lo, &ld);
  }
  const float t = h.t;
  v3 pt{lo.x + t * ld.x, lo.y + t * ld.y, lo.z + t * ld.z};
  v3 nn;
  if (kind == TPT_SLOT_MEDIUM) {
    // constant_medium::hit: an arbitrary normal vec3(u, u, u).normalize()
    // (src/hitable.cc:121-122); u and v are not set there (0 here).
    nn = v3{0.0f, 1.0f, 0.0f};
    if (MEDIUM_NORMAL) {
      const uint4 w = medium_draws(rc, leaf_id(S, h));
      nn = v3{u01(w.y), u01(w.z), u01(w.w)};
      const float k = static_cast<float>(1.0 / static_cast<double>(sqrtf(nn.x * nn.x + nn.y * nn.y + nn.z * nn.z)));
      nn = v3{nn.x * k, nn.y * k, nn.z * k};
    }
    *u = 0.0f;
    *v = 0.0f;
  } else if (kind == TPT_SLOT_MSPHERE) {
    // moving_sphere::hit (src/sphere.cc:53-57): normal about the centre at
    // the ray's time, uv about center0.
    const tpt_motion m = S.motions[__float_as_uint(f[0])];
    const v3 c = motion_center(m, rc.time);
    const float r = m.radius;
    const v3 q{(pt.x - m.center0[0]) / r, (pt.y - m.center0[1]) / r, (pt.z - m.center0[2]) / r};
    *u = static_cast<float>(static_cast<double>(atan2f(q.z, q.x)) / (2 * M_PI)) + 0.5f;
    *v = static_cast<float>(static_cast<double>(asinf(q.y)) / M_PI) + 0.5f;
    nn = v3{(pt.x - c.x) / r, (pt.y - c.y) / r, (pt.z - c.z) / r};
  } else if (kind == TPT_SLOT_SPHERE) {
    const float r = f[3];
    const v3 q{(pt.x - f[0]) / r, (pt.y - f[1]) / r, (pt.z - f[2]) / r};
    // get_uv_map (headers/utils.h:38-43); atan2f/asinf are the device's.
    *u = static_cast<float>(static_cast<double>(atan2f(q.z, q.x)) / (2 * M_PI)) + 0.5f;
    *v = static_cast<float>(static_cast<double>(asinf(q.y)) / M_PI) + 0.5f;
    nn = q;
  } else {
    const uint32_t axis = (word >> 28) & 3u;
    const float x = axis == 0 ? pt.y : pt.x;
    const float y = axis == 2 ? pt.y : pt.z;
    *u = (x - f[1]) / (f[2] - f[1]);
    *v = (y - f[3]) / (f[4] - f[3]);
    nn = v3{axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f, axis == 2 ? 1.0f : 0.0f};
  }
  if ((word >> 30) & 1u) nn = -nn;
  if (xf != kNone) to_world(S.xforms[xf], &pt, &nn);
  if (sx & TPT_SEGMENT_FLIP) nn = -nn;
  if ((sx & TPT_SEGMENT_PLAIN) != TPT_SEGMENT_PLAIN) to_world(S.xforms[sx & TPT_SEGMENT_PLAIN], &pt, &nn);
  *p = pt;
  *n = nn;
}

}  // namespace tptd

#endif  // TPT_CSRC_TRAVERSE_H
