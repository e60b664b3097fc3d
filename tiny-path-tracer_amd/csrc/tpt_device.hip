// Device half of the C ABI (include/tpt.h): scene upload, the closest-hit
// kernel and the host loop of the wavefront renderer (kernels in render.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../host/errors.h"
#include "guide_train.h"
#include "render.h"

#include <random>

using namespace tptd;

namespace {

// -------------------------------------------------------- closest hit API ---
template <bool SWEEP>
__global__ void __launch_bounds__(256) k_trace_closest(scene_view S, const tpt_slot *__restrict__ slots,
                                                       const tpt_entry *__restrict__ entries,
                                                       const float *__restrict__ rays, uint64_t n, float tmin,
                                                       float tmax, int32_t *__restrict__ leaf,
                                                       float *__restrict__ tout, float *__restrict__ rec,
                                                       int32_t *__restrict__ mat) {
  S.slots = slots;
  S.entries = entries;
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool live = i < n;
  v3 o{0.0f, 0.0f, 0.0f}, d{0.0f, 0.0f, 0.0f};
  if (live) {
    o = v3{rays[i], rays[n + i], rays[2 * n + i]};
    d = v3{rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]};
  }
  hit_t h;
  uint32_t nodes = 0;
  const bool got = trace_world<SWEEP>(S, o, d, tmin, tmax, &h, &nodes, live);
  if (!live) return;
  if (got) {
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
  bool sweep = false;  // wave-uniform traversal (small scenes)
  float bounds[6] = {0, 0, 0, 1, 1, 1};
  // render state, grown on demand
  void *state = nullptr;
  size_t state_bytes = 0;
  uint32_t capacity = 0;   // path records in flight
  uint64_t res_items = 0;  // items the res buffer holds
  uint32_t stat_blocks = 0;
  render_state Q{};
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

int ensure_state(tpt_scene *s, uint32_t capacity, uint64_t res_items, uint32_t stat_blocks) {
  if (s->state && capacity <= s->capacity && res_items <= s->res_items && stat_blocks <= s->stat_blocks)
    return TPT_OK;
  capacity = std::max(capacity, s->capacity);
  res_items = std::max(res_items, s->res_items);
  stat_blocks = std::max(stat_blocks, s->stat_blocks);
  (void)hipFree(s->state);
  s->state = nullptr;
  s->capacity = 0;
  s->res_items = 0;
  s->stat_blocks = 0;
  // Region size: block b appends to region b % NSUB.  Live records never
  // exceed `capacity`; each block takes grid-stride chunks of 256, so one
  // block of a grid of G handles at most t / G + 256 of t entries and the
  // ceil(G / NSUB) blocks of a residue at most t / NSUB + t / G +
  // ceil(G / NSUB) * 256.  Up to two kernels append to the same ray queue,
  // with stat_blocks / 4 <= G <= stat_blocks, and their t add up to at most
  // `capacity`.
  const size_t per_residue = (stat_blocks + NSUB - 1) / NSUB;
  const size_t region = capacity / NSUB + 2 * (4 * static_cast<size_t>(capacity) / stat_blocks + per_residue * 256 + 256);
  const size_t n = region * NSUB;
  // 2 ray queues x 3 planes + 3 material queues x 4 planes, 16 B per entry
  const size_t qbytes = n * 16 * (2 * 3 + 3 * 4) + 40 * 256;  // + rounding of every take()
  const size_t bytes = 256 + COUNTER_WORDS * sizeof(uint32_t) + stat_blocks * ST_COUNT * sizeof(unsigned long long) +
                       qbytes + res_items * 3 * sizeof(float) + 256;
  HIP_TRY(hipMalloc(&s->state, bytes));
  s->state_bytes = bytes;
  char *w = static_cast<char *>(s->state);
  auto take = [&](size_t k) {
    char *p = w;
    w += (k + 255) & ~size_t(255);
    return p;
  };
  render_state &Q = s->Q;
  Q.region = static_cast<uint32_t>(region);
  Q.counters = reinterpret_cast<uint32_t *>(take(COUNTER_WORDS * sizeof(uint32_t)));
  Q.stats = reinterpret_cast<unsigned long long *>(take(stat_blocks * ST_COUNT * sizeof(unsigned long long)));
  for (int p = 0; p < 2; ++p)
    for (int k = 0; k < 3; ++k) Q.rq[p][k] = reinterpret_cast<float4 *>(take(n * 16));
  for (int c = 0; c < 3; ++c)
    for (int k = 0; k < 4; ++k) Q.mq[c][k] = reinterpret_cast<float4 *>(take(n * 16));
  Q.res = reinterpret_cast<float *>(take(res_items * 3 * sizeof(float)));
  s->capacity = capacity;
  s->res_items = res_items;
  s->stat_blocks = stat_blocks;
  return TPT_OK;
}

}  // namespace

struct tpt_guide {
  int device = 0;
  int n_cu = 256;
  float *params = nullptr;   // G_NPARAM_PAD + 1 (mix logit)
  float *adam_m = nullptr, *adam_v = nullptr;
  uint32_t *xtiles = nullptr;  // inference tiles, GX_TILES x 256 words
  uint32_t *rtiles = nullptr;  // training tiles, GR_TILES x 256 words
  float *gacc = nullptr;       // gradient accumulator (G_NPARAM_PAD + 2), cleared by k_guide_adam
  float *partial = nullptr;    // train blocks x G_PARTIAL
  float *scratch = nullptr;    // gradient / loss staging (G_NPARAM_PAD + 2)
  int train_blocks = 0;
  int step = 0;
  float4 *rec_side = nullptr, *rec_rows = nullptr;
  uint64_t rec_cap = 0;  // recorded paths the buffers hold
  uint64_t n_rows = 0;
};

namespace {

int guide_pack(tpt_guide *g, hipStream_t stream) {
  hipLaunchKernelGGL(k_guide_pack, dim3((G_B1 + 255) / 256), dim3(256), 0, stream, g->params,
                     reinterpret_cast<uint16_t *>(g->xtiles), reinterpret_cast<uint16_t *>(g->rtiles));
  HIP_TRY(hipGetLastError());
  return TPT_OK;
}

// One fused forward/backward pass over rows [row0, row0 + n) (cyclic in
// n_total), partials reduced by k_guide_reduce (mode 0: gradient, 1: Adam).
int guide_step(tpt_guide *g, const float *rows, uint64_t n, uint64_t row0, uint64_t n_total, int mode, float lr,
               float *grad, float *loss_dev, hipStream_t stream) {
  const int blocks = static_cast<int>(std::min<uint64_t>((n + 63) / 64, static_cast<uint64_t>(g->train_blocks)));
  const size_t lds = k_guide_train_lds();
  hipLaunchKernelGGL(k_guide_train, dim3(blocks), dim3(256), lds, stream, g->rtiles, g->params, rows, n, row0,
                     n_total, g->partial);
  HIP_TRY(hipGetLastError());
  if (mode == 1) ++g->step;
  const dim3 rgrid((G_NPARAM_PAD + 2 + 255) / 256, (blocks + PART_CHUNK - 1) / PART_CHUNK);
  hipLaunchKernelGGL(k_guide_reduce, rgrid, dim3(256), 0, stream, g->partial, blocks, g->gacc);
  hipLaunchKernelGGL(k_guide_adam, dim3((G_NPARAM_PAD + 2 + 255) / 256), dim3(256), 0, stream, g->gacc, g->params,
                     g->adam_m, g->adam_v, grad, loss_dev, mode, lr, g->step);
  HIP_TRY(hipGetLastError());
  if (mode == 1) return guide_pack(g, stream);
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
  // Wave-uniform sweeps pay for the union of the lanes' visits; that stays
  // close to each lane's own count only when the whole tree is small.
  s->sweep = f->n_slots <= 64;
  for (int k = 0; k < 6; ++k) s->bounds[k] = f->bounds[k];
  if (hipHostMalloc(reinterpret_cast<void **>(&s->host_count), NSUB * CSTRIDE * sizeof(uint32_t), hipHostMallocDefault) !=
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
  const scene_view S = s->view();
  if (s->sweep)
    hipLaunchKernelGGL(k_trace_closest<true>, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), S, S.slots, S.entries, d_rays, n, t_min, t_max, d_leaf, d_t,
                       d_rec, d_mat);
  else
    hipLaunchKernelGGL(k_trace_closest<false>, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), S, S.slots, S.entries, d_rays, n, t_min, t_max, d_leaf, d_t,
                       d_rec, d_mat);
  HIP_TRY(hipGetLastError());
  return TPT_OK;
}

int tpt_trace_closest(tpt_scene *s, const float *d_rays, uint64_t n, float t_min, float t_max, int32_t *d_leaf,
                      float *d_t, void *stream) {
  return tpt_trace_closest_record(s, d_rays, n, t_min, t_max, d_leaf, d_t, nullptr, nullptr, stream);
}

}  // extern "C"

namespace {

int render_impl(tpt_scene *s, tpt_guide *guide, const tpt_camera *camera, const tpt_render_desc *desc, float *d_sum,
                float *d_sumsq, tpt_render_stats *stats, void *stream_ptr) {
  if (!s || !camera || !desc || !d_sum) return tpt::fail(TPT_ERR_INVALID, "tpt_render: null argument");
  if (desc->width == 0 || desc->height == 0) return tpt::fail(TPT_ERR_INVALID, "tpt_render: empty image");
  if (desc->flags & ~TPT_RENDER_PROFILE) return tpt::fail(TPT_ERR_INVALID, "tpt_render: unknown flags");
  const bool profile = (desc->flags & TPT_RENDER_PROFILE) != 0;
  if (!guide && desc->guiding != 0) return tpt::fail(TPT_ERR_INVALID, "tpt_render: guiding needs tpt_render_guided");
  const uint64_t n_pix = static_cast<uint64_t>(desc->width) * desc->height;
  if (n_pix > (1ull << 26)) return tpt::fail(TPT_ERR_INVALID, "tpt_render: more than 2^26 pixels");
  if (static_cast<uint64_t>(desc->sample_offset) + desc->spp > 0xFFFFFFFFull)
    return tpt::fail(TPT_ERR_INVALID, "tpt_render: sample range overflows");
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (desc->spp == 0) return TPT_OK;
  hipStream_t stream = static_cast<hipStream_t>(stream_ptr);

  // Pass size: at most 2^26 items (768 MB of per-sample results).
  const uint64_t max_items = 1ull << 26;
  const uint32_t pass_spp = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(desc->spp, max_items / n_pix)));
  const uint64_t pass_items = n_pix * pass_spp;
  uint32_t cap = desc->queue_capacity ? desc->queue_capacity : (1u << 21);
  if (cap > (1u << 26)) return tpt::fail(TPT_ERR_INVALID, "tpt_render: queue_capacity above 2^26");
  if (cap > pass_items) cap = static_cast<uint32_t>(pass_items);
  const uint32_t grid = static_cast<uint32_t>(s->n_cu) * 8;
  // k_guide_shade holds 51 KB of LDS (weights + output table): 3 blocks per CU
  const uint32_t guide_grid = static_cast<uint32_t>(s->n_cu) * 3;
  int rc = ensure_state(s, cap, pass_items, grid);
  if (rc != TPT_OK) return rc;

  render_params P{};
  P.cam = *camera;
  P.width = desc->width;
  P.height = desc->height;
  P.n_pix = static_cast<uint32_t>(n_pix);
  P.max_depth = desc->max_depth;
  P.seed = desc->seed;
  P.background = desc->background;
  P.sum = d_sum;
  P.sumsq = d_sumsq;
  const scene_view S = s->view();
  const render_state &Q = s->Q;
  P.guide_verts = guide ? (desc->guide_vertices ? desc->guide_vertices : 2u) : ~0u;
  if (guide) {
    P.g_tiles = guide->xtiles;
    P.g_params = guide->params;
    for (int k = 0; k < 6; ++k) P.bounds[k] = s->bounds[k];
    P.rec_stride = desc->guiding;
    if (P.rec_stride) {
      const uint64_t need = pass_items / P.rec_stride + 1;
      if (need > guide->rec_cap) {
        (void)hipFree(guide->rec_side);
        (void)hipFree(guide->rec_rows);
        guide->rec_side = guide->rec_rows = nullptr;
        guide->rec_cap = 0;
        HIP_TRY(hipMalloc(&guide->rec_side, need * REC_VERTS * 5 * sizeof(float4)));
        HIP_TRY(hipMalloc(&guide->rec_rows, need * REC_VERTS * 4 * sizeof(float4)));
        guide->rec_cap = need;
      }
      P.rec_side = guide->rec_side;
      P.rec_rows = guide->rec_rows;
      guide->n_rows = (std::min<uint64_t>(pass_items, static_cast<uint64_t>(n_pix) * desc->spp) + P.rec_stride - 1) /
                      P.rec_stride * REC_VERTS;
    }
  }

  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, stream));
  hipLaunchKernelGGL(k_zero_stats, dim3(64), dim3(256), 0, stream, Q.stats, grid * ST_COUNT);
  const int poll = 16;
  if (profile && s->events.size() < static_cast<size_t>(2 * poll + 1)) {
    for (auto e : s->events) (void)hipEventDestroy(e);
    s->events.assign(2 * poll + 1, nullptr);
    for (auto &e : s->events) HIP_TRY(hipEventCreate(&e));
  }
  double trace_ms = 0.0, shade_ms = 0.0;
  uint64_t n_launch = 0;
  auto harvest = [&](int iters) {
    for (int k = 0; k < iters; ++k) {
      float a = 0.0f, b = 0.0f;
      if (hipEventElapsedTime(&a, s->events[2 * k], s->events[2 * k + 1]) == hipSuccess) trace_ms += a;
      if (hipEventElapsedTime(&b, s->events[2 * k + 1], s->events[2 * k + 2]) == hipSuccess) shade_ms += b;
    }
    n_launch += iters;
  };
  bool drained = true;
  for (uint32_t done = 0; done < desc->spp && drained; done += P.pass_spp) {
    P.s0 = desc->sample_offset + done;
    P.pass_spp = std::min(pass_spp, desc->spp - done);
    P.n_items = static_cast<uint32_t>(n_pix * P.pass_spp);
    P.n_first = std::min(cap, P.n_items);
    hipLaunchKernelGGL(k_init, dim3(std::min<uint32_t>((P.n_first + 255) / 256 + 1, grid)), dim3(256), 0, stream, P,
                       Q);
    // A path has at most max_depth + 1 segments and each bounce retires at
    // least one path per live record, so the loop is bounded.
    const uint64_t max_iter = (static_cast<uint64_t>(P.n_items) / cap + 2) * (desc->max_depth + 2ull) + 64;
    drained = false;
    int batch = 0;
    for (uint64_t it = 0; it < max_iter; ++it) {
      const uint32_t parity = static_cast<uint32_t>(it & 1u);
      if (profile) HIP_TRY(hipEventRecord(s->events[2 * batch], stream));
      if (s->sweep)
        hipLaunchKernelGGL(k_trace<true>, dim3(grid), dim3(256), 0, stream, S, S.slots, S.entries, P, Q, parity);
      else
        hipLaunchKernelGGL(k_trace<false>, dim3(grid), dim3(256), 0, stream, S, S.slots, S.entries, P, Q, parity);
      if (profile) HIP_TRY(hipEventRecord(s->events[2 * batch + 1], stream));
      if (guide) {
        hipLaunchKernelGGL(k_shade<true>, dim3(grid), dim3(256), 0, stream, S, P, Q, parity);
        hipLaunchKernelGGL(k_guide_shade, dim3(guide_grid), dim3(256), 0, stream, S, P, Q, parity);
      } else {
        hipLaunchKernelGGL(k_shade<false>, dim3(grid), dim3(256), 0, stream, S, P, Q, parity);
      }
      if (profile) HIP_TRY(hipEventRecord(s->events[2 * batch + 2], stream));
      ++batch;
      if ((it % poll) == poll - 1) {
        HIP_TRY(hipMemcpyAsync(s->host_count, Q.counters + rc_at(parity ^ 1u, 0), NSUB * CSTRIDE * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        if (profile) harvest(batch);
        batch = 0;
        uint64_t live = 0;
        for (uint32_t k = 0; k < NSUB; ++k) live += s->host_count[k * CSTRIDE];
        if (live == 0) {
          drained = true;
          break;
        }
      }
    }
    if (profile && batch > 0) {
      HIP_TRY(hipStreamSynchronize(stream));
      harvest(batch);
    }
    if (!drained) break;
    hipLaunchKernelGGL(k_accumulate, dim3(std::min<uint32_t>((P.n_pix + 255) / 256, grid)), dim3(256), 0, stream, P,
                       static_cast<const float *>(Q.res));
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(e1, stream));
  std::vector<unsigned long long> st(static_cast<size_t>(grid) * ST_COUNT);
  HIP_TRY(hipMemcpyAsync(st.data(), Q.stats, st.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!drained) return tpt::fail(TPT_ERR_HIP, "tpt_render: bounce loop did not drain");
  if (stats) {
    unsigned long long tot[ST_COUNT] = {0};
    for (uint32_t b = 0; b < grid; ++b)
      for (int k = 0; k < ST_COUNT; ++k) tot[k] += st[static_cast<size_t>(b) * ST_COUNT + k];
    stats->paths = tot[ST_PATHS];
    stats->segments = tot[ST_SEGMENTS];
    stats->light_queries = tot[ST_LIGHT];
    stats->nan_samples = tot[ST_NAN];
    stats->nodes_visited = tot[ST_NODES];
    stats->kernel_ms = ms;
    stats->trace_ms = trace_ms;
    stats->shade_ms = shade_ms;
    stats->trace_launches = n_launch;
    stats->shade_launches = n_launch;
  }
  return TPT_OK;
}

}  // namespace

extern "C" {

int tpt_render(tpt_scene *s, const tpt_camera *camera, const tpt_render_desc *desc, float *d_sum, float *d_sumsq,
               tpt_render_stats *stats, void *stream) {
  return render_impl(s, nullptr, camera, desc, d_sum, d_sumsq, stats, stream);
}

int tpt_render_guided(tpt_scene *s, tpt_guide *guide, const tpt_camera *camera, const tpt_render_desc *desc,
                      float *d_sum, float *d_sumsq, tpt_render_stats *stats, void *stream) {
  if (!guide) return tpt::fail(TPT_ERR_INVALID, "tpt_render_guided: null guide");
  return render_impl(s, guide, camera, desc, d_sum, d_sumsq, stats, stream);
}

int tpt_guide_create(uint64_t seed, tpt_guide **out) {
  if (!out) return tpt::fail(TPT_ERR_INVALID, "tpt_guide_create: null argument");
  *out = nullptr;
  if (tpt_device_count() < 1) return tpt::fail(TPT_ERR_NODEVICE, "tpt_guide_create: no gfx950 device");
  // He-normal weights (std sqrt(2 / fan_in)), zero biases, mix logit 0 (a = 0.5).
  std::vector<float> h(G_NPARAM_PAD + 1, 0.0f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  for (int L = 0; L < 4; ++L) {
    int off, rows, cols;
    layer_geom(L, &off, &rows, &cols);
    const float fan_in = L == 0 ? static_cast<float>(G_IN) : 64.0f;
    const float sd = std::sqrt(2.0f / fan_in) * (L == 3 ? 0.1f : 1.0f);
    for (int o = 0; o < rows; ++o)
      for (int i = 0; i < cols; ++i) {
        const int p = off + o * cols + i;
        const float v = nd(rng) * sd;
        h[p] = param_live(p) ? v : 0.0f;
      }
  }
  auto *g = new tpt_guide();
  HIP_TRY(hipGetDevice(&g->device));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, g->device) == hipSuccess) g->n_cu = prop.multiProcessorCount;
  g->train_blocks = g->n_cu;
  const size_t np = G_NPARAM_PAD + 1;
  HIP_TRY(hipMalloc(&g->params, np * sizeof(float)));
  HIP_TRY(hipMalloc(&g->adam_m, np * sizeof(float)));
  HIP_TRY(hipMalloc(&g->adam_v, np * sizeof(float)));
  HIP_TRY(hipMalloc(&g->xtiles, GX_TILES * GX_TILE_WORDS * sizeof(uint32_t)));
  HIP_TRY(hipMemset(g->xtiles, 0, GX_TILES * GX_TILE_WORDS * sizeof(uint32_t)));
  HIP_TRY(hipMalloc(&g->rtiles, GR_TILES * 256 * sizeof(uint32_t)));
  HIP_TRY(hipMemset(g->rtiles, 0, GR_TILES * 256 * sizeof(uint32_t)));
  HIP_TRY(hipMalloc(&g->gacc, (G_NPARAM_PAD + 2) * sizeof(float)));
  HIP_TRY(hipMemset(g->gacc, 0, (G_NPARAM_PAD + 2) * sizeof(float)));
  HIP_TRY(hipMalloc(&g->partial, static_cast<size_t>(g->train_blocks) * G_PARTIAL * sizeof(float)));
  HIP_TRY(hipMalloc(&g->scratch, (G_NPARAM_PAD + 2) * sizeof(float)));
  HIP_TRY(hipMemcpy(g->params, h.data(), np * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(g->adam_m, 0, np * sizeof(float)));
  HIP_TRY(hipMemset(g->adam_v, 0, np * sizeof(float)));
  HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_guide_train),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(k_guide_train_lds())));
  const int rc = guide_pack(g, nullptr);
  if (rc != TPT_OK) return rc;
  HIP_TRY(hipDeviceSynchronize());
  *out = g;
  return TPT_OK;
}

int tpt_guide_destroy(tpt_guide *g) {
  if (!g) return TPT_OK;
  (void)hipDeviceSynchronize();
  (void)hipFree(g->params);
  (void)hipFree(g->adam_m);
  (void)hipFree(g->adam_v);
  (void)hipFree(g->xtiles);
  (void)hipFree(g->rtiles);
  (void)hipFree(g->gacc);
  (void)hipFree(g->partial);
  (void)hipFree(g->scratch);
  (void)hipFree(g->rec_side);
  (void)hipFree(g->rec_rows);
  delete g;
  return TPT_OK;
}

int tpt_guide_get_params(tpt_guide *g, float *host) {
  if (!g || !host) return tpt::fail(TPT_ERR_INVALID, "tpt_guide_get_params: null argument");
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(host, g->params, (G_NPARAM_PAD + 1) * sizeof(float), hipMemcpyDeviceToHost));
  return TPT_OK;
}

int tpt_guide_set_params(tpt_guide *g, const float *host) {
  if (!g || !host) return tpt::fail(TPT_ERR_INVALID, "tpt_guide_set_params: null argument");
  std::vector<float> h(host, host + G_NPARAM_PAD + 1);
  for (int p = 0; p < G_NPARAM_PAD; ++p)
    if (!param_live(p)) h[p] = 0.0f;
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(g->params, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(g->adam_m, 0, h.size() * sizeof(float)));
  HIP_TRY(hipMemset(g->adam_v, 0, h.size() * sizeof(float)));
  g->step = 0;
  const int rc = guide_pack(g, nullptr);
  if (rc != TPT_OK) return rc;
  HIP_TRY(hipDeviceSynchronize());
  return TPT_OK;
}

int tpt_guide_infer(tpt_guide *g, const float *d_in, uint64_t n, float *d_out, void *stream) {
  if (!g || (n && (!d_in || !d_out))) return tpt::fail(TPT_ERR_INVALID, "tpt_guide_infer: null argument");
  if (n == 0) return TPT_OK;
  const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((n + 63) / 64, static_cast<uint64_t>(g->n_cu) * 8));
  hipLaunchKernelGGL(k_guide_infer, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream), g->xtiles,
                     g->params, d_in, n, d_out);
  HIP_TRY(hipGetLastError());
  return TPT_OK;
}

int tpt_guide_gradient(tpt_guide *g, const float *d_rows, uint64_t n, float *d_grad, float *loss, void *stream) {
  if (!g || !d_rows || !d_grad || n == 0) return tpt::fail(TPT_ERR_INVALID, "tpt_guide_gradient: bad argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int rc = guide_step(g, d_rows, n, 0, n, 0, 0.0f, d_grad, g->scratch, st);
  if (rc != TPT_OK) return rc;
  float l = 0.0f;
  HIP_TRY(hipMemcpyAsync(&l, g->scratch, sizeof(float), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (loss) *loss = l / static_cast<float>(n);
  return TPT_OK;
}

int tpt_guide_train(tpt_guide *g, const float *d_rows, uint64_t n, uint64_t n_batch, uint32_t iters, float lr,
                    float *loss, void *stream) {
  if (g && !d_rows) {  // the rows recorded by the last tpt_render_guided
    d_rows = reinterpret_cast<const float *>(g->rec_rows);
    n = g->rec_rows ? g->n_rows : 0;
  }
  if (!g || !d_rows || n == 0) return tpt::fail(TPT_ERR_INVALID, "tpt_guide_train: no rows");
  if (n_batch == 0 || n_batch > n) n_batch = n;
  hipStream_t st = static_cast<hipStream_t>(stream);
  uint64_t row0 = 0;
  for (uint32_t it = 0; it < iters; ++it) {
    const int rc = guide_step(g, d_rows, n_batch, row0, n, 1, lr, g->scratch, g->scratch + G_NPARAM_PAD + 1, st);
    if (rc != TPT_OK) return rc;
    row0 = (row0 + n_batch) % n;
  }
  if (loss) {
    float l = 0.0f;
    HIP_TRY(hipMemcpyAsync(&l, g->scratch + G_NPARAM_PAD + 1, sizeof(float), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    *loss = l / static_cast<float>(n_batch);
  }
  return TPT_OK;
}

int tpt_guide_batch(tpt_guide *g, const float **d_rows, uint64_t *n_rows) {
  if (!g || !d_rows || !n_rows) return tpt::fail(TPT_ERR_INVALID, "tpt_guide_batch: null argument");
  *d_rows = reinterpret_cast<const float *>(g->rec_rows);
  *n_rows = g->rec_rows ? g->n_rows : 0;
  return TPT_OK;
}

}  // extern "C"
