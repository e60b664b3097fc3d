// Device half of the C ABI (include/tpt.h): scene upload, the closest-hit
// kernel and the wavefront renderer.
//
// Renderer (replaces color() + the sample loop, src/utils.cc:58-94 and
// main.cpp:115-134): a pool of path slots lives in HBM as SoA arrays.  Each
// slot owns one pixel at a time and runs its samples one after another, so
// accumulation needs no atomics and is deterministic.  Per bounce two kernels
// run on ping-pong queues of slot ids:
//   k_trace  closest hit for every queued ray, then sorts the slot into one
//            of three material queues (diffuse / specular / terminal) with a
//            64-lane ballot + prefix count and one atomic per wave;
//   k_shade  walks the three queues back to back (waves see one material
//            class), scatters or finishes each path, regenerates camera rays
//            for finished slots, and appends continuing rays to the next ray
//            queue (same wave-level compaction).
// Queue counters are double-buffered and reset by the kernel that precedes
// their next use, so the bounce loop needs no host round trip; the host only
// polls the live-ray count every few bounces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../host/errors.h"
#include "shade.h"

using namespace tptd;

namespace {

// ------------------------------------------------------------ wave utils ---
// Appends one element per lane with `pred` to a queue: one atomic per wave.
// Must be called by all lanes of the wave together.
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool pred) {
  const unsigned long long mask = __ballot(pred);
  const uint32_t lane = __lane_id();
  const uint32_t prefix = __popcll(mask & ((1ull << lane) - 1ull));
  const uint32_t total = __popcll(mask);
  uint32_t base = 0;
  if (total) {
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    if (static_cast<int>(lane) == leader) base = atomicAdd(counter, total);
    base = __shfl(base, leader);
  }
  return base + prefix;
}

__device__ __forceinline__ void wave_count(unsigned long long *ctr, uint32_t v) {
  // Sum over the wave, one atomic.
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if (__lane_id() == 0 && v) atomicAdd(ctr, static_cast<unsigned long long>(v));
}

// -------------------------------------------------------- closest hit API ---
__global__ void __launch_bounds__(256) k_trace_closest(scene_view S, const float *__restrict__ rays, uint64_t n,
                                                       float tmin, float tmax, int32_t *__restrict__ leaf,
                                                       float *__restrict__ tout, float *__restrict__ rec,
                                                       int32_t *__restrict__ mat) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const v3 o{rays[i], rays[n + i], rays[2 * n + i]};
  const v3 d{rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]};
  hit_t h;
  uint32_t nodes = 0;
  if (trace_world(S, o, d, tmin, tmax, &h, &nodes)) {
    leaf[i] = static_cast<int32_t>(leaf_id(S, h));
    tout[i] = h.t;
    if (rec) {
      v3 p, nn;
      float u, v;
      hit_record(S, h, o, d, &p, &nn, &u, &v);
      const float vals[8] = {u, v, p.x, p.y, p.z, nn.x, nn.y, nn.z};
#pragma unroll
      for (int k = 0; k < 8; ++k) rec[k * n + i] = vals[k];
    }
    if (mat) mat[i] = static_cast<int32_t>(h.mat & TPT_PAYLOAD_MASK);
  } else {
    leaf[i] = -1;
    tout[i] = 0.0f;
    if (rec)
#pragma unroll
      for (int k = 0; k < 8; ++k) rec[k * n + i] = 0.0f;
    if (mat) mat[i] = -1;
  }
}

// --------------------------------------------------------------- renderer ---
enum { Q_DIFFUSE = 0, Q_SPECULAR = 1, Q_TERMINAL = 2 };
// Queue counters (uint32): [0,1] ray queues by parity, [2 + 3*parity + k]
// material queues.  Stats (uint64): segments, light queries, nan samples,
// paths, nodes visited.
enum { ST_SEGMENTS = 0, ST_LIGHT = 1, ST_NAN = 2, ST_PATHS = 3, ST_NODES = 4, ST_COUNT = 8 };

struct paths {
  float *ox, *oy, *oz, *dx, *dy, *dz;
  float *tr, *tg, *tb;  // throughput
  float *rr, *rg, *rb;  // radiance of the current sample
  uint32_t *pixel, *sample, *depth;
  float *hit_t;
  uint32_t *hit_slot, *hit_entry, *hit_mat;
  uint32_t *q_ray[2];
  uint32_t *q_mat[2][3];
  uint32_t *qcount;
  unsigned long long *stats;
};

struct render_params {
  tpt_camera cam;
  uint32_t width, height, n_pixels, capacity;
  uint32_t spp_begin, spp_end, max_depth;
  uint64_t seed;
  float background;
  float *sum, *sumsq;
};

__device__ __forceinline__ v3 cam_vec(const float *a) { return v3{a[0], a[1], a[2]}; }

// Camera ray for (pixel, sample): jitter (main.cpp:121-122, double adds),
// then camera_with_blur::get_ray (src/camera.cc:23-31).  Stream bounce 0.
__device__ __forceinline__ void raygen(const render_params &P, const paths &Q, uint32_t slot) {
  const uint32_t px = Q.pixel[slot], s = Q.sample[slot];
  stream R;
  R.init(px, s, 0, P.seed);
  const uint32_t i = px % P.width, j = px / P.width;
  const float u = static_cast<float>((static_cast<double>(static_cast<float>(i)) + static_cast<double>(R.next())) /
                                     static_cast<double>(static_cast<float>(P.width)));
  const float v = static_cast<float>((static_cast<double>(static_cast<float>(j)) + static_cast<double>(R.next())) /
                                     static_cast<double>(static_cast<float>(P.height)));
  const tpt_camera &C = P.cam;
  const v3 disk = in_unit_disk(R);
  const v3 rd = disk * C.lens_radius;
  const v3 off = cam_vec(C.u) * rd.x + cam_vec(C.v) * rd.y;
  (void)R.next();  // ray time draw (time0 + xi*(time1 - time0)); static scenes ignore it
  const v3 org = cam_vec(C.origin);
  const v3 o = org + off;
  const v3 d = cam_vec(C.lower_left) + u * cam_vec(C.horizontal) + v * cam_vec(C.vertical) - org - off;
  Q.ox[slot] = o.x, Q.oy[slot] = o.y, Q.oz[slot] = o.z;
  Q.dx[slot] = d.x, Q.dy[slot] = d.y, Q.dz[slot] = d.z;
  Q.tr[slot] = 1.0f, Q.tg[slot] = 1.0f, Q.tb[slot] = 1.0f;
  Q.rr[slot] = 0.0f, Q.rg[slot] = 0.0f, Q.rb[slot] = 0.0f;
  Q.depth[slot] = 0;
}

__global__ void k_reset(paths Q) {
  const uint32_t t = threadIdx.x;
  if (t < 8) Q.qcount[t] = 0;
  if (t < ST_COUNT) Q.stats[t] = 0;
}

// Gives every slot its first pixel and camera ray; queues it for bounce 0.
__global__ void __launch_bounds__(256) k_init(render_params P, paths Q) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < P.capacity; base += stride) {
    const uint32_t slot = base + threadIdx.x;
    const bool live = slot < P.capacity && slot < P.n_pixels && P.spp_begin < P.spp_end;
    if (live) {
      Q.pixel[slot] = slot;
      Q.sample[slot] = P.spp_begin;
      raygen(P, Q, slot);
    }
    const uint32_t at = wave_append(&Q.qcount[0], live);
    if (live) Q.q_ray[0][at] = slot;
  }
}

__global__ void __launch_bounds__(256) k_trace(scene_view S, render_params P, paths Q, uint32_t parity) {
  const uint32_t count = Q.qcount[parity];
  if (blockIdx.x == 0 && threadIdx.x == 0) Q.qcount[parity ^ 1u] = 0;  // next ray queue, filled by k_shade
  const uint32_t *queue = Q.q_ray[parity];
  uint32_t *mc = &Q.qcount[2 + 3 * parity];
  uint32_t nodes = 0, segs = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride) {
    const uint32_t qi = base + threadIdx.x;
    const bool valid = qi < count;
    uint32_t cls = 3, slot = 0;
    if (valid) {
      slot = queue[qi];
      const v3 o{Q.ox[slot], Q.oy[slot], Q.oz[slot]};
      const v3 d{Q.dx[slot], Q.dy[slot], Q.dz[slot]};
      hit_t h;
      ++segs;
      if (trace_world(S, o, d, 0.001f, kFltMax, &h, &nodes)) {
        Q.hit_t[slot] = h.t;
        Q.hit_slot[slot] = h.slot;
        Q.hit_entry[slot] = h.entry;
        const uint32_t m = h.mat & TPT_PAYLOAD_MASK;
        Q.hit_mat[slot] = m;
        const uint32_t kind = S.materials[m].kind;
        if (kind == TPT_MAT_DIFFUSE_LIGHT || Q.depth[slot] >= P.max_depth) {
          cls = Q_TERMINAL;
        } else {
          cls = kind == TPT_MAT_LAMBERTIAN ? Q_DIFFUSE : Q_SPECULAR;
        }
      } else {
        Q.hit_slot[slot] = kNone;
        cls = Q_TERMINAL;
      }
    }
#pragma unroll
    for (uint32_t k = 0; k < 3; ++k) {
      const uint32_t at = wave_append(&mc[k], cls == k);
      if (cls == k) Q.q_mat[parity][k][at] = slot;
    }
  }
  wave_count(&Q.stats[ST_SEGMENTS], segs);
  wave_count(&Q.stats[ST_NODES], nodes);
}

// de_nan (headers/utils.h:100-109) and accumulation of the finished sample;
// then the slot's next sample (or next pixel) is generated.  Returns whether
// the slot has a new camera ray.
__device__ __forceinline__ bool finish(const render_params &P, const paths &Q, uint32_t slot, uint32_t *nan_count) {
  const uint32_t px = Q.pixel[slot];
  float c[3] = {Q.rr[slot], Q.rg[slot], Q.rb[slot]};
  bool any_nan = false;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (c[k] != c[k]) {
      c[k] = 0.0f;
      any_nan = true;
    }
  }
  *nan_count += any_nan ? 1u : 0u;
  float *sum = P.sum + 3ull * px;
#pragma unroll
  for (int k = 0; k < 3; ++k) sum[k] += c[k];
  if (P.sumsq) {
    float *sq = P.sumsq + 3ull * px;
#pragma unroll
    for (int k = 0; k < 3; ++k) sq[k] += c[k] * c[k];
  }
  uint32_t s = Q.sample[slot] + 1;
  uint32_t p = px;
  if (s >= P.spp_end) {
    p = px + P.capacity;
    s = P.spp_begin;
    if (p >= P.n_pixels) return false;
  }
  Q.pixel[slot] = p;
  Q.sample[slot] = s;
  raygen(P, Q, slot);
  return true;
}

__global__ void __launch_bounds__(256) k_shade(scene_view S, render_params P, paths Q, uint32_t parity) {
  const uint32_t *mc = &Q.qcount[2 + 3 * parity];
  const uint32_t n0 = mc[0], n1 = mc[1], n2 = mc[2];
  const uint32_t total = n0 + n1 + n2;
  if (blockIdx.x == 0 && threadIdx.x < 3) Q.qcount[2 + 3 * (parity ^ 1u) + threadIdx.x] = 0;
  uint32_t *next_count = &Q.qcount[parity ^ 1u];
  uint32_t *next_queue = Q.q_ray[parity ^ 1u];
  uint32_t nan_count = 0, paths_done = 0, light_q = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += stride) {
    const uint32_t qi = base + threadIdx.x;
    bool emit = false;
    uint32_t slot = 0;
    if (qi < total) {
      const uint32_t cls = qi < n0 ? Q_DIFFUSE : (qi < n0 + n1 ? Q_SPECULAR : Q_TERMINAL);
      const uint32_t off = cls == Q_DIFFUSE ? qi : (cls == Q_SPECULAR ? qi - n0 : qi - n0 - n1);
      slot = Q.q_mat[parity][cls][off];
      const v3 o{Q.ox[slot], Q.oy[slot], Q.oz[slot]};
      const v3 d{Q.dx[slot], Q.dy[slot], Q.dz[slot]};
      v3 thr{Q.tr[slot], Q.tg[slot], Q.tb[slot]};
      v3 rad{Q.rr[slot], Q.rg[slot], Q.rb[slot]};
      const uint32_t hslot = Q.hit_slot[slot];
      bool cont = false;
      if (hslot == kNone) {
        // miss: black (src/utils.cc:85-86) or the sky extension
        const v3 bg = P.background > 0.0f ? sky(d, P.background) : v3{0.0f, 0.0f, 0.0f};
        rad = rad + thr * bg;
      } else {
        const hit_t h{Q.hit_t[slot], hslot, Q.hit_entry[slot], 0};
        v3 p, n;
        float u, v;
        hit_record(S, h, o, d, &p, &n, &u, &v);
        const tpt_material M = S.materials[Q.hit_mat[slot]];
        const uint32_t depth = Q.depth[slot];
        stream R;
        R.init(Q.pixel[slot], Q.sample[slot], depth + 1, P.seed);
        v3 nd{0.0f, 0.0f, 0.0f};
        if (cls == Q_TERMINAL) {
          // diffuse_light::emitted is one-sided (src/material.cc:79-86);
          // every other material emits nothing.
          v3 e{0.0f, 0.0f, 0.0f};
          if (M.kind == TPT_MAT_DIFFUSE_LIGHT && dot(n, d) < 0.0f) e = tex_value(S, M.tex, u, v, p);
          rad = rad + thr * e;
        } else if (M.kind == TPT_MAT_METAL) {
          // metal::scatter (src/material.cc:88-98); specular: no emitted term.
          const v3 refl = reflect(unit(d), n);
          nd = refl + M.p[3] * in_unit_sphere(R);
          if (dot(nd, n) > 0.0f) {
            thr = thr * v3{M.p[0], M.p[1], M.p[2]};
            cont = true;
          } else {
            rad = rad + thr * v3{0.0f, 0.0f, 0.0f};
          }
        } else if (M.kind == TPT_MAT_DIELECTRIC) {
          // dielectric::scatter (src/material.cc:19-70).
          const float ior = M.p[0];
          const v3 reflected = reflect(d, n);
          v3 outward;
          float ni, cosine;
          if (dot(d, n) > 0.0f) {
            outward = -n;
            ni = ior;
            cosine = dot(d, n) / len(d);
            cosine = sqrtf(1.0f - ior * ior * (1.0f - cosine * cosine));
          } else {
            outward = n;
            ni = 1.0f / ior;
            cosine = -dot(d, n) / len(d);
          }
          v3 refracted{0.0f, 0.0f, 0.0f};
          const float prob = refract(d, outward, ni, &refracted) ? schlick(cosine, ior) : 1.0f;
          nd = R.next() < prob ? reflected : refracted;
          cont = true;
        } else {
          // lambertian + mixture_pdf(hitable_pdf(lights), cosine_pdf)
          // (src/material.cc:3-17, src/utils.cc:73-81); cosine_pdf::value
          // is cos θ without 1/π, as in the reference.
          const v3 atten = tex_value(S, M.tex, u, v, p);
          const onb uvw(n);
          float pdf;
          if (S.n_lights > 0) {
            if (R.next() < 0.5f) {
              nd = light_random(S, p, R);
            } else {
              nd = uvw.local(on_hemisphere(R));
            }
            const float pl = light_pdf(S, p, nd);
            ++light_q;
            float pc = dot(unit(nd), uvw.w);
            pc = pc > 0.0f ? pc : 0.0f;
            pdf = 0.5f * pl + 0.5f * pc;
          } else {
            nd = uvw.local(on_hemisphere(R));
            float pc = dot(unit(nd), uvw.w);
            pdf = pc > 0.0f ? pc : 0.0f;
          }
          float spdf = dot(n, unit(nd));
          spdf = spdf < 0.0f ? 0.0f : spdf / kPi;
          const v3 w = (atten * spdf) / pdf;
          rad = rad + thr * v3{0.0f, 0.0f, 0.0f};  // carries a NaN throughput into the sample
          thr = thr * w;
          cont = true;
        }
        if (cont) {
          Q.ox[slot] = p.x, Q.oy[slot] = p.y, Q.oz[slot] = p.z;
          Q.dx[slot] = nd.x, Q.dy[slot] = nd.y, Q.dz[slot] = nd.z;
          Q.tr[slot] = thr.x, Q.tg[slot] = thr.y, Q.tb[slot] = thr.z;
          Q.rr[slot] = rad.x, Q.rg[slot] = rad.y, Q.rb[slot] = rad.z;
          Q.depth[slot] = depth + 1;
          emit = true;
        }
      }
      if (!emit) {
        Q.rr[slot] = rad.x, Q.rg[slot] = rad.y, Q.rb[slot] = rad.z;
        ++paths_done;
        emit = finish(P, Q, slot, &nan_count);
      }
    }
    const uint32_t at = wave_append(next_count, emit);
    if (emit) next_queue[at] = slot;
  }
  wave_count(&Q.stats[ST_NAN], nan_count);
  wave_count(&Q.stats[ST_PATHS], paths_done);
  wave_count(&Q.stats[ST_LIGHT], light_q);
}

// ------------------------------------------------------------ host side ---
#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) return tpt::fail(TPT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int upload(T **dst, const T *src, size_t n) {
  *dst = nullptr;
  if (n == 0) return TPT_OK;
  HIP_TRY(hipMalloc(reinterpret_cast<void **>(dst), n * sizeof(T)));
  HIP_TRY(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return TPT_OK;
}

bool is_gfx950(int dev) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return false;
  return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

}  // namespace

struct tpt_scene {
  int device = 0;
  int n_cu = 256;
  tpt_slot *slots = nullptr;
  uint32_t *segments = nullptr;
  tpt_entry *entries = nullptr;
  tpt_xform *xforms = nullptr;
  tpt_material *materials = nullptr;
  tpt_texture *textures = nullptr;
  uint8_t *texels = nullptr;
  tpt_perlin *perlin = nullptr;
  tpt_light *lights = nullptr;
  uint32_t n_slots = 0, n_segments = 0, n_lights = 0, light_mode = 0, n_leaves = 0, max_depth = 0;
  uint32_t n_materials = 0;
  // render state, grown on demand
  void *state = nullptr;
  size_t state_bytes = 0;
  uint32_t capacity = 0;
  paths Q{};
  uint32_t *host_count = nullptr;  // pinned
  std::vector<hipEvent_t> events;  // TPT_RENDER_PROFILE timing

  scene_view view() const {
    scene_view v;
    v.slots = slots;
    v.segments = segments;
    v.entries = entries;
    v.xforms = xforms;
    v.materials = materials;
    v.textures = textures;
    v.texels = texels;
    v.perlin = perlin;
    v.lights = lights;
    v.n_slots = n_slots;
    v.n_segments = n_segments;
    v.n_lights = n_lights;
    v.light_mode = light_mode;
    return v;
  }
};

namespace {

void free_scene(tpt_scene *s) {
  (void)hipFree(s->slots);
  (void)hipFree(s->segments);
  (void)hipFree(s->entries);
  (void)hipFree(s->xforms);
  (void)hipFree(s->materials);
  (void)hipFree(s->textures);
  (void)hipFree(s->texels);
  (void)hipFree(s->perlin);
  (void)hipFree(s->lights);
  (void)hipFree(s->state);
  for (auto e : s->events) (void)hipEventDestroy(e);
  if (s->host_count) (void)hipHostFree(s->host_count);
}

int ensure_state(tpt_scene *s, uint32_t capacity) {
  if (capacity <= s->capacity && s->state) return TPT_OK;
  (void)hipFree(s->state);
  s->state = nullptr;
  s->capacity = 0;
  const size_t n = capacity;
  // 15 float/uint arrays + 4 hit arrays + 2 ray queues + 6 material queues
  const size_t words = n * (15 + 4 + 2 + 6) + 16 + 2 * ST_COUNT;
  const size_t bytes = words * 4 + 256;
  HIP_TRY(hipMalloc(&s->state, bytes));
  s->state_bytes = bytes;
  uint32_t *w = static_cast<uint32_t *>(s->state);
  auto take = [&](size_t k) {
    uint32_t *p = w;
    w += k;
    return p;
  };
  paths &Q = s->Q;
  Q.stats = reinterpret_cast<unsigned long long *>(take(2 * ST_COUNT));
  Q.qcount = take(16);
  Q.ox = reinterpret_cast<float *>(take(n));
  Q.oy = reinterpret_cast<float *>(take(n));
  Q.oz = reinterpret_cast<float *>(take(n));
  Q.dx = reinterpret_cast<float *>(take(n));
  Q.dy = reinterpret_cast<float *>(take(n));
  Q.dz = reinterpret_cast<float *>(take(n));
  Q.tr = reinterpret_cast<float *>(take(n));
  Q.tg = reinterpret_cast<float *>(take(n));
  Q.tb = reinterpret_cast<float *>(take(n));
  Q.rr = reinterpret_cast<float *>(take(n));
  Q.rg = reinterpret_cast<float *>(take(n));
  Q.rb = reinterpret_cast<float *>(take(n));
  Q.pixel = take(n);
  Q.sample = take(n);
  Q.depth = take(n);
  Q.hit_t = reinterpret_cast<float *>(take(n));
  Q.hit_slot = take(n);
  Q.hit_entry = take(n);
  Q.hit_mat = take(n);
  Q.q_ray[0] = take(n);
  Q.q_ray[1] = take(n);
  for (int p = 0; p < 2; ++p)
    for (int k = 0; k < 3; ++k) Q.q_mat[p][k] = take(n);
  s->capacity = capacity;
  return TPT_OK;
}

}  // namespace

extern "C" {

int tpt_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  int good = 0;
  for (int d = 0; d < n; ++d) good += is_gfx950(d) ? 1 : 0;
  return good;
}

int tpt_scene_create(const tpt_flat_scene *f, tpt_scene **out) {
  if (!f || !out) return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: null argument");
  *out = nullptr;
  if (f->n_segments > 0 && (!f->slots || !f->segments))
    return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: missing slots/segments");
  if (f->max_depth > 60) return tpt::fail(TPT_ERR_UNSUPPORTED, "tpt_scene_create: bvh deeper than 60 levels");
  for (uint32_t g = 0; g < f->n_segments; ++g) {
    if (f->segments[2 * g] > f->segments[2 * g + 1] || f->segments[2 * g + 1] > f->n_slots)
      return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: segment out of range");
  }
  // Every index the kernels follow must stay inside its array.
  for (uint32_t i = 0; i < f->n_slots; ++i) {
    const tpt_slot &s = f->slots[i];
    const uint32_t kind = s.b >> 28;
    if (kind == TPT_SLOT_BOX) {
      const uint32_t right = s.b & TPT_PAYLOAD_MASK;
      if (s.a < i + 2 || s.a > f->n_slots || (right != TPT_NO_RIGHT && (right <= i + 1 || right >= s.a)))
        return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: bad box links");
    } else if (kind == TPT_SLOT_LIST) {
      uint32_t cnt;
      std::memcpy(&cnt, &s.f[1], 4);
      if (static_cast<uint64_t>(s.a) + cnt > f->n_entries)
        return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: list entries out of range");
    } else if (kind == TPT_SLOT_SPHERE || kind == TPT_SLOT_RECT) {
      if ((s.a & TPT_PAYLOAD_MASK) >= f->n_materials)
        return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: material index out of range");
    } else {
      return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: unknown slot kind");
    }
  }
  for (uint32_t e = 0; e < f->n_entries; ++e) {
    const tpt_entry &en = f->entries[e];
    if ((en.a & TPT_PAYLOAD_MASK) >= f->n_materials || (en.xform != 0xFFFFFFFFu && en.xform >= f->n_xforms) ||
        (en.b != TPT_SLOT_SPHERE && en.b != TPT_SLOT_RECT))
      return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: bad list entry");
  }
  for (uint32_t x = 0; x < f->n_xforms; ++x)
    if (f->xforms[x].nops > TPT_XFORM_MAX_OPS) return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: bad xform");
  for (uint32_t m = 0; m < f->n_materials; ++m) {
    const tpt_material &M = f->materials[m];
    if ((M.kind == TPT_MAT_LAMBERTIAN || M.kind == TPT_MAT_DIFFUSE_LIGHT) && M.tex >= f->n_textures)
      return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: texture index out of range");
    if (M.kind > TPT_MAT_DIFFUSE_LIGHT) return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: bad material");
  }
  for (uint32_t t = 0; t < f->n_textures; ++t) {
    const tpt_texture &T = f->textures[t];
    if (T.kind == TPT_TEX_CHECKER && (T.a >= f->n_textures || T.b >= f->n_textures))
      return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: checker child out of range");
    if (T.kind == TPT_TEX_PERLIN && !f->perlin) return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: perlin tables missing");
    if (T.kind == TPT_TEX_IMAGE &&
        (T.b == 0 || T.c == 0 || static_cast<uint64_t>(T.a) + 3ull * T.b * T.c > f->n_texels))
      return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: image texels out of range");
    if (T.kind == TPT_TEX_PERLIN && f->perlin) {
      for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 256; ++i)
          if (f->perlin->perm[k][i] < 0 || f->perlin->perm[k][i] > 255)
            return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: perlin permutation out of range");
    }
  }
  if (f->n_lights > 0 && f->light_mode > 2) return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: bad light mode");
  if (tpt_device_count() < 1) return tpt::fail(TPT_ERR_NODEVICE, "tpt_scene_create: no gfx950 device");
  auto *s = new tpt_scene();
  int rc = TPT_OK;
  HIP_TRY(hipGetDevice(&s->device));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, s->device) == hipSuccess) s->n_cu = prop.multiProcessorCount;
  if ((rc = upload(&s->slots, f->slots, f->n_slots)) || (rc = upload(&s->segments, f->segments, 2ull * f->n_segments)) ||
      (rc = upload(&s->entries, f->entries, f->n_entries)) || (rc = upload(&s->xforms, f->xforms, f->n_xforms)) ||
      (rc = upload(&s->materials, f->materials, f->n_materials)) ||
      (rc = upload(&s->textures, f->textures, f->n_textures)) || (rc = upload(&s->texels, f->texels, f->n_texels)) ||
      (rc = upload(&s->perlin, f->perlin, f->perlin ? 1 : 0)) || (rc = upload(&s->lights, f->lights, f->n_lights))) {
    free_scene(s);
    delete s;
    return rc;
  }
  s->n_slots = f->n_slots;
  s->n_segments = f->n_segments;
  s->n_lights = f->n_lights;
  s->light_mode = f->light_mode;
  s->n_leaves = f->n_leaves;
  s->max_depth = f->max_depth;
  s->n_materials = f->n_materials;
  if (hipHostMalloc(reinterpret_cast<void **>(&s->host_count), 64 * sizeof(uint32_t), hipHostMallocDefault) !=
      hipSuccess) {
    free_scene(s);
    delete s;
    return tpt::fail(TPT_ERR_HIP, "tpt_scene_create: hipHostMalloc failed");
  }
  *out = s;
  return TPT_OK;
}

int tpt_scene_destroy(tpt_scene *s) {
  if (!s) return TPT_OK;
  (void)hipDeviceSynchronize();
  free_scene(s);
  delete s;
  return TPT_OK;
}

int tpt_trace_closest_record(tpt_scene *s, const float *d_rays, uint64_t n, float t_min, float t_max,
                             int32_t *d_leaf, float *d_t, float *d_rec, int32_t *d_mat, void *stream) {
  if (!s || (n && (!d_rays || !d_leaf || !d_t))) return tpt::fail(TPT_ERR_INVALID, "tpt_trace_closest: null argument");
  if (n == 0) return TPT_OK;
  const uint64_t blocks = (n + 255) / 256;
  if (blocks > 0x7FFFFFFFull) return tpt::fail(TPT_ERR_INVALID, "tpt_trace_closest: too many rays");
  hipLaunchKernelGGL(k_trace_closest, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), s->view(), d_rays, n, t_min, t_max, d_leaf, d_t, d_rec, d_mat);
  HIP_TRY(hipGetLastError());
  return TPT_OK;
}

int tpt_trace_closest(tpt_scene *s, const float *d_rays, uint64_t n, float t_min, float t_max, int32_t *d_leaf,
                      float *d_t, void *stream) {
  return tpt_trace_closest_record(s, d_rays, n, t_min, t_max, d_leaf, d_t, nullptr, nullptr, stream);
}

int tpt_render(tpt_scene *s, const tpt_camera *camera, const tpt_render_desc *desc, float *d_sum, float *d_sumsq,
               tpt_render_stats *stats, void *stream_ptr) {
  if (!s || !camera || !desc || !d_sum) return tpt::fail(TPT_ERR_INVALID, "tpt_render: null argument");
  if (desc->width == 0 || desc->height == 0) return tpt::fail(TPT_ERR_INVALID, "tpt_render: empty image");
  if (desc->flags & ~TPT_RENDER_PROFILE) return tpt::fail(TPT_ERR_INVALID, "tpt_render: unknown flags");
  const bool profile = (desc->flags & TPT_RENDER_PROFILE) != 0;
  if (desc->guiding != 0) return tpt::fail(TPT_ERR_UNSUPPORTED, "tpt_render: guiding needs tpt_guide_* (not built)");
  const uint64_t n_pixels = static_cast<uint64_t>(desc->width) * desc->height;
  if (n_pixels > 0x7FFFFFFFull) return tpt::fail(TPT_ERR_INVALID, "tpt_render: image too large");
  if (static_cast<uint64_t>(desc->sample_offset) + desc->spp > 0xFFFFFFFFull)
    return tpt::fail(TPT_ERR_INVALID, "tpt_render: sample range overflows");
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (desc->spp == 0) return TPT_OK;
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
  uint32_t cap = desc->queue_capacity ? desc->queue_capacity : (1u << 21);
  if (cap > n_pixels) cap = static_cast<uint32_t>(n_pixels);
  if (cap > (1u << 26)) return tpt::fail(TPT_ERR_INVALID, "tpt_render: queue_capacity above 2^26");
  int rc = ensure_state(s, cap);
  if (rc != TPT_OK) return rc;

  render_params P{};
  P.cam = *camera;
  P.width = desc->width;
  P.height = desc->height;
  P.n_pixels = static_cast<uint32_t>(n_pixels);
  P.capacity = cap;
  P.spp_begin = desc->sample_offset;
  P.spp_end = desc->sample_offset + desc->spp;
  P.max_depth = desc->max_depth;
  P.seed = desc->seed;
  P.background = desc->background;
  P.sum = d_sum;
  P.sumsq = d_sumsq;
  const scene_view S = s->view();
  const paths &Q = s->Q;

  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, stream));
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, stream, Q);
  const uint32_t grid_init = std::min<uint32_t>((cap + 255) / 256, static_cast<uint32_t>(s->n_cu) * 8);
  hipLaunchKernelGGL(k_init, dim3(grid_init), dim3(256), 0, stream, P, Q);
  const uint32_t grid = static_cast<uint32_t>(s->n_cu) * 8;
  // Upper bound on bounces: every slot runs (pixels/cap) pixels of spp paths
  // of at most max_depth + 1 segments.
  const uint64_t rounds = (n_pixels + cap - 1) / cap;
  const uint64_t max_iter = rounds * static_cast<uint64_t>(desc->spp) * (desc->max_depth + 2ull) + 8;
  const int poll = 8;
  if (profile && s->events.empty()) {
    s->events.resize(2 * poll + 1);
    for (auto &e : s->events) HIP_TRY(hipEventCreate(&e));
  }
  double trace_ms = 0.0, shade_ms = 0.0;
  uint64_t n_trace = 0, n_shade = 0;
  auto harvest = [&](int iters) {
    for (int k = 0; k < iters; ++k) {
      float a = 0.0f, b = 0.0f;
      if (hipEventElapsedTime(&a, s->events[2 * k], s->events[2 * k + 1]) == hipSuccess) trace_ms += a;
      if (hipEventElapsedTime(&b, s->events[2 * k + 1], s->events[2 * k + 2]) == hipSuccess) shade_ms += b;
    }
    n_trace += iters;
    n_shade += iters;
  };
  uint64_t it = 0;
  int batch = 0;
  for (; it < max_iter; ++it) {
    const uint32_t parity = static_cast<uint32_t>(it & 1u);
    if (profile) HIP_TRY(hipEventRecord(s->events[2 * batch], stream));
    hipLaunchKernelGGL(k_trace, dim3(grid), dim3(256), 0, stream, S, P, Q, parity);
    if (profile) HIP_TRY(hipEventRecord(s->events[2 * batch + 1], stream));
    hipLaunchKernelGGL(k_shade, dim3(grid), dim3(256), 0, stream, S, P, Q, parity);
    if (profile) HIP_TRY(hipEventRecord(s->events[2 * batch + 2], stream));
    ++batch;
    if ((it % poll) == poll - 1) {
      HIP_TRY(hipMemcpyAsync(s->host_count, Q.qcount, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipStreamSynchronize(stream));
      if (profile) harvest(batch);
      batch = 0;
      if (s->host_count[parity ^ 1u] == 0) break;
    }
  }
  if (profile && batch > 0) {
    HIP_TRY(hipStreamSynchronize(stream));
    harvest(batch);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(e1, stream));
  unsigned long long st[ST_COUNT];
  HIP_TRY(hipMemcpyAsync(st, Q.stats, sizeof st, hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (it >= max_iter) return tpt::fail(TPT_ERR_HIP, "tpt_render: bounce loop did not drain");
  if (stats) {
    stats->paths = st[ST_PATHS];
    stats->segments = st[ST_SEGMENTS];
    stats->light_queries = st[ST_LIGHT];
    stats->nan_samples = st[ST_NAN];
    stats->kernel_ms = ms;
    stats->nodes_visited = st[ST_NODES];
    stats->trace_ms = trace_ms;
    stats->shade_ms = shade_ms;
    stats->trace_launches = n_trace;
    stats->shade_launches = n_shade;
  }
  return TPT_OK;
}

}  // extern "C"
