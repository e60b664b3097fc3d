// Device half of the C ABI (include/tpt.h): scene upload, the closest-hit
// kernel and the host loop of the wavefront renderer (kernels in render.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../host/errors.h"
#include "guide_train.h"
#include "render.h"

#include <random>

using namespace tptd;

namespace {

// -------------------------------------------------------- closest hit API ---
// SPH: a sphere-only scene takes the render's sphere-only walk.
template <bool SWEEP, bool MEDIA, bool SPH = false>
__global__ void __launch_bounds__(256) k_trace_closest(scene_view S, const tpt_slot *__restrict__ slots,
                                                       const tpt_entry *__restrict__ entries,
                                                       const float *__restrict__ rays,
                                                       const float *__restrict__ times, uint64_t n, float tmin,
                                                       float tmax, int32_t *__restrict__ leaf,
                                                       float *__restrict__ tout, float *__restrict__ rec,
                                                       int32_t *__restrict__ mat) {
  S.slots = slots;
  S.entries = entries;
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool live = i < n;
  v3 o{0.0f, 0.0f, 0.0f}, d{0.0f, 0.0f, 0.0f};
  // media draws: counter (i mod 2^32, i / 2^32, 0, 2^20 + leaf), key 0
  ray_ctx rc{0.0f, static_cast<uint32_t>(i), static_cast<uint32_t>(i >> 32), 0u, 0u, 0u};
  if (live) {
    o = v3{rays[i], rays[n + i], rays[2 * n + i]};
    d = v3{rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]};
    if (times) rc.time = times[i];
  }
  hit_t h;
  uint32_t nodes = 0;
  const bool got = trace_world<SWEEP, SPH, MEDIA>(S, o, d, rc, tmin, tmax, &h, &nodes, live);
  if (!live) return;
  if (got) {
    leaf[i] = static_cast<int32_t>(leaf_id(S, h));
    tout[i] = h.t;
    if (rec) {
      v3 p, nn;
      float u, v;
      hit_record<MEDIA>(S, h, o, d, rc, &p, &nn, &u, &v);
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

// A scene handle: one or more flattened scenes packed into one pool (indices
// rebased), plus the render state grown on demand.
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
  tpt_motion *motions = nullptr;
  uint32_t *seg_xf = nullptr;  // per pooled segment; null when no scene has a wrapped segment
  uint32_t n_slots = 0, n_segments = 0, n_lights = 0;
  // the pool's scenes
  std::vector<scene_dev> scn;
  std::vector<uint8_t> sweep;     // per scene: wave-uniform traversal (small scenes)
  std::vector<uint8_t> sph;       // per scene: every leaf is a sphere
  std::vector<uint8_t> simple;    // per scene: constant textures, no light or one xz_rect light
  std::vector<uint8_t> motion;    // per scene: has moving spheres (rays need their time)
  std::vector<uint8_t> media;     // per scene: has constant_medium (queries need media draws)
  std::vector<uint32_t> n_slots_of;
  std::vector<uint32_t> n_leaves_of;
  scene_dev *d_scenes = nullptr;
  // render state, grown on demand
  void *state = nullptr;
  size_t state_bytes = 0;
  uint32_t capacity = 0;   // path records in flight
  uint64_t res_items = 0;  // items the res buffer holds
  uint32_t stat_blocks = 0, region_grid = 0, region_guide_grid = 0;
  render_state Q{};
  uint32_t *host_count = nullptr;  // pinned
  std::vector<hipEvent_t> events;  // TPT_RENDER_PROFILE timing
  // job and pass tables (device) and the pinned staging of the pass table
  job_dev *d_jobs = nullptr;
  uint32_t jobs_cap = 0;
  pass_hdr *d_hdr = nullptr, *h_hdr = nullptr;
  pass_ent *d_pass = nullptr, *h_pass = nullptr;
  uint32_t pass_cap = 0;
  // captured bounce loop (hipGraph), replayed in every pass of a render call
  hipStream_t gstream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<uint8_t> gkey;

  // The pooled arrays seen as scene k (its segments and lights).
  scene_view view(uint32_t k = 0) const {
    scene_view v;
    v.slots = slots;
    v.segments = segments + 2ull * scn[k].seg_begin;
    v.entries = entries;
    v.xforms = xforms;
    v.materials = materials;
    v.textures = textures;
    v.texels = texels;
    v.perlin = perlin;
    v.lights = lights + scn[k].light_begin;
    v.motions = motions;
    v.seg_xf = seg_xf ? seg_xf + scn[k].seg_begin : nullptr;
    v.n_slots = n_slots;
    v.n_segments = scn[k].n_seg;
    v.n_lights = scn[k].n_lights;
    v.light_mode = scn[k].light_mode;
    v.walk_ulp = scn[k].walk_ulp;
    return v;
  }
  // The pooled arrays with every segment and light (multi-job launches).
  scene_view pool_view() const {
    scene_view v = view(0);
    v.segments = segments;
    v.seg_xf = seg_xf;
    v.n_segments = n_segments;
    v.lights = lights;
    v.n_lights = n_lights;
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
  (void)hipFree(s->motions);
  (void)hipFree(s->seg_xf);
  (void)hipFree(s->d_scenes);
  (void)hipFree(s->state);
  (void)hipFree(s->d_jobs);
  (void)hipFree(s->d_hdr);
  (void)hipFree(s->d_pass);
  if (s->h_hdr) (void)hipHostFree(s->h_hdr);
  if (s->h_pass) (void)hipHostFree(s->h_pass);
  for (auto e : s->events) (void)hipEventDestroy(e);
  if (s->host_count) (void)hipHostFree(s->host_count);
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->ev_in) (void)hipEventDestroy(s->ev_in);
  if (s->ev_out) (void)hipEventDestroy(s->ev_out);
  if (s->gstream) (void)hipStreamDestroy(s->gstream);
}

// Region size: block b of a grid of G appends to region b % NSUB.  Live
// records never exceed `capacity`; each block takes grid-stride chunks of its
// block size B, so it handles at most t / G + B of t entries and the
// ceil(G / NSUB) blocks of a residue at most t / NSUB + t / G + ceil(G / NSUB) * B.
// Two kernels (grid G1 of 256-thread blocks, grid G2 of k_guide_shade's
// blocks, at most 512 threads) may append to the same ray queue, their t adding
// up to at most `capacity`.  G2 = 0: no guided shading (plain renders), only
// G1 kernels append.
size_t region_size(uint32_t capacity, uint32_t g1, uint32_t g2) {
  const size_t r1 = (g1 + NSUB - 1) / NSUB, r2 = (g2 + NSUB - 1) / NSUB;
  return capacity / NSUB + capacity / g1 + (g2 ? capacity / g2 : 0) + r1 * 256 + r2 * 512 + 256;
}

// Device bytes of a render state (as ensure_state lays it out) for a capacity,
// res_items samples and the launch grids.
size_t state_bytes_for(uint32_t capacity, uint64_t res_items, uint32_t grid, uint32_t guide_grid) {
  const size_t n = region_size(capacity, grid, guide_grid) * NSUB;
  return 256 + COUNTER_WORDS * sizeof(uint32_t) + std::max(grid, guide_grid) * ST_COUNT * sizeof(unsigned long long) +
         n * 16 * (2 * 3 + 3 * 4) + 40 * 256 + res_items * 3 * sizeof(float) + 256;
}

int ensure_state(tpt_scene *s, uint32_t capacity, uint64_t res_items, uint32_t grid, uint32_t guide_grid) {
  const uint32_t stat_blocks = std::max(grid, guide_grid);
  if (s->state && capacity <= s->capacity && res_items <= s->res_items && stat_blocks <= s->stat_blocks &&
      region_size(capacity, grid, guide_grid) <= s->Q.region)
    return TPT_OK;
  capacity = std::max(capacity, s->capacity);
  res_items = std::max(res_items, s->res_items);
  const uint32_t sb = std::max(stat_blocks, s->stat_blocks);
  const size_t region = std::max<size_t>(region_size(capacity, grid, guide_grid), s->Q.region);
  if (region * NSUB > 0xFFFFFFFFull) return tpt::fail(TPT_ERR_INVALID, "tpt_render: queue capacity too large");
  (void)hipFree(s->state);
  s->state = nullptr;
  s->capacity = 0;
  s->res_items = 0;
  s->stat_blocks = 0;
  s->Q.region = 0;
  if (s->gexec) {  // the captured loop holds the old queue pointers
    (void)hipGraphExecDestroy(s->gexec);
    s->gexec = nullptr;
    s->gkey.clear();
  }
  const size_t n = region * NSUB;
  // 2 ray queues x 3 planes + 3 material queues x 4 planes, 16 B per entry
  const size_t qbytes = n * 16 * (2 * 3 + 3 * 4) + 40 * 256;  // + rounding of every take()
  const size_t bytes = 256 + COUNTER_WORDS * sizeof(uint32_t) + sb * ST_COUNT * sizeof(unsigned long long) + qbytes +
                       res_items * 3 * sizeof(float) + 256;
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
  Q.stats = reinterpret_cast<unsigned long long *>(take(sb * ST_COUNT * sizeof(unsigned long long)));
  for (int p = 0; p < 2; ++p)
    for (int k = 0; k < 3; ++k) Q.rq[p][k] = reinterpret_cast<float4 *>(take(n * 16));
  for (int c = 0; c < 3; ++c)
    for (int k = 0; k < 4; ++k) Q.mq[c][k] = reinterpret_cast<float4 *>(take(n * 16));
  Q.res = reinterpret_cast<float *>(take(res_items * 3 * sizeof(float)));
  s->capacity = capacity;
  s->res_items = res_items;
  s->stat_blocks = sb;
  return TPT_OK;
}

// Every index the kernels follow must stay inside its array.
// A medium's boundary: entries in range, none of them a medium (the device
// evaluates a boundary without media).
bool medium_ok(const tpt_flat_scene *f, const float *w) {
  uint32_t first, count;
  std::memcpy(&first, &w[0], 4);
  std::memcpy(&count, &w[1], 4);
  if (count == 0 || static_cast<uint64_t>(first) + count > f->n_entries) return false;
  for (uint32_t e = first; e < first + count; ++e)
    if (f->entries[e].b == TPT_SLOT_MEDIUM) return false;
  return true;
}

int validate_flat(const tpt_flat_scene *f) {
  if (!f) return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: null scene");
  if (f->n_segments > 0 && (!f->slots || !f->segments))
    return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: missing slots/segments");
  if (f->max_depth > 60) return tpt::fail(TPT_ERR_UNSUPPORTED, "tpt_scene_create: bvh deeper than 60 levels");
  for (uint32_t g = 0; g < f->n_segments; ++g) {
    if (f->segments[2 * g] > f->segments[2 * g + 1] || f->segments[2 * g + 1] > f->n_slots)
      return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: segment out of range");
    if (f->segment_xforms) {
      const uint32_t x = f->segment_xforms[g] & TPT_SEGMENT_PLAIN;
      if (x != TPT_SEGMENT_PLAIN && x >= f->n_xforms)
        return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: segment xform out of range");
    }
  }
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
    } else if (kind == TPT_SLOT_SPHERE || kind == TPT_SLOT_RECT || kind == TPT_SLOT_MSPHERE ||
               kind == TPT_SLOT_MEDIUM) {
      if ((s.a & TPT_PAYLOAD_MASK) >= f->n_materials)
        return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: material index out of range");
      uint32_t mi;
      std::memcpy(&mi, &s.f[0], 4);
      if (kind == TPT_SLOT_MSPHERE && (!f->motions || mi >= f->n_motions))
        return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: motion index out of range");
      if (kind == TPT_SLOT_MEDIUM && !medium_ok(f, s.f)) return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: bad medium");
    } else {
      return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: unknown slot kind");
    }
  }
  for (uint32_t e = 0; e < f->n_entries; ++e) {
    const tpt_entry &en = f->entries[e];
    uint32_t mi;
    std::memcpy(&mi, &en.f[0], 4);
    if ((en.a & TPT_PAYLOAD_MASK) >= f->n_materials || (en.xform != 0xFFFFFFFFu && en.xform >= f->n_xforms) ||
        (en.b != TPT_SLOT_SPHERE && en.b != TPT_SLOT_RECT && en.b != TPT_SLOT_MSPHERE && en.b != TPT_SLOT_MEDIUM) ||
        (en.b == TPT_SLOT_MSPHERE && (!f->motions || mi >= f->n_motions)) ||
        (en.b == TPT_SLOT_MEDIUM && (en.xform != 0xFFFFFFFFu || !medium_ok(f, en.f))))
      return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: bad list entry");
  }
  for (uint32_t x = 0; x < f->n_xforms; ++x)
    if (f->xforms[x].nops > TPT_XFORM_MAX_OPS) return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: bad xform");
  for (uint32_t m = 0; m < f->n_materials; ++m) {
    const tpt_material &M = f->materials[m];
    if ((M.kind == TPT_MAT_LAMBERTIAN || M.kind == TPT_MAT_DIFFUSE_LIGHT) && M.tex >= f->n_textures)
      return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: texture index out of range");
    if ((M.kind == TPT_MAT_ISOTROPIC) && M.tex >= f->n_textures)
      return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: texture index out of range");
    if (M.kind > TPT_MAT_ISOTROPIC) return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: bad material");
  }
  for (uint32_t t = 0; t < f->n_textures; ++t) {
    const tpt_texture &T = f->textures[t];
    if (T.kind == TPT_TEX_CHECKER && (T.a >= f->n_textures || T.b >= f->n_textures))
      return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: checker child out of range");
    if (T.kind == TPT_TEX_PERLIN && (!f->perlin || T.b != 0))
      return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: perlin tables missing");
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
  return TPT_OK;
}

// Index fields that point into another array, rebased by that array's offset
// in the pool (the material word keeps its axis/flip bits).
uint32_t rebase_word(uint32_t word, uint32_t base) {
  return (word & ~TPT_PAYLOAD_MASK) | ((word & TPT_PAYLOAD_MASK) + base);
}

// An index kept as the bits of a float field (the motion index of MSPHERE).
float rebase_bits(float f, uint32_t base) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += base;
  std::memcpy(&f, &u, 4);
  return f;
}

int create_pool(const tpt_flat_scene *scenes, uint32_t n, tpt_scene **out) {
  if (!scenes || !out || n == 0) return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: no scenes");
  *out = nullptr;
  for (uint32_t k = 0; k < n; ++k) {
    const int rc = validate_flat(&scenes[k]);
    if (rc != TPT_OK) return rc;
  }
  std::vector<tpt_slot> slots;
  std::vector<uint32_t> segments;
  std::vector<tpt_entry> entries;
  std::vector<tpt_xform> xforms;
  std::vector<tpt_material> materials;
  std::vector<tpt_texture> textures;
  std::vector<uint8_t> texels;
  std::vector<tpt_perlin> perlin;
  std::vector<tpt_light> lights;
  std::vector<tpt_motion> motions;
  std::vector<uint32_t> seg_xf;
  bool any_wrapped = false;
  std::vector<scene_dev> scn(n);
  std::vector<uint8_t> sweep(n), sph(n), simple(n), motion(n), media(n);
  std::vector<uint32_t> nslots(n), nleaves(n);
  for (uint32_t k = 0; k < n; ++k) {
    const tpt_flat_scene &f = scenes[k];
    const uint64_t sb = slots.size(), eb = entries.size(), xb = xforms.size(), mb = materials.size(),
                   tb = textures.size(), bb = texels.size(), pb = perlin.size(), lb = lights.size(),
                   gb = segments.size() / 2, mob = motions.size();
    if (sb + f.n_slots >= TPT_NO_RIGHT || mb + f.n_materials > TPT_PAYLOAD_MASK || bb + f.n_texels > 0xFFFFFFFFull ||
        eb + f.n_entries > 0xFFFFFFF0ull)
      return tpt::fail(TPT_ERR_INVALID, "tpt_scene_create: pool too large for 32-bit indices");
    sph[k] = 1;
    for (uint32_t i = 0; i < f.n_slots; ++i) {
      tpt_slot s = f.slots[i];
      const uint32_t kind = s.b >> 28;
      if (kind != TPT_SLOT_BOX && kind != TPT_SLOT_SPHERE) sph[k] = 0;
      if (kind == TPT_SLOT_BOX) {
        s.a += static_cast<uint32_t>(sb);
        const uint32_t right = s.b & TPT_PAYLOAD_MASK;
        if (right != TPT_NO_RIGHT) s.b = (s.b & ~TPT_PAYLOAD_MASK) | (right + static_cast<uint32_t>(sb));
      } else if (kind == TPT_SLOT_LIST) {
        s.a += static_cast<uint32_t>(eb);
      } else {
        s.a = rebase_word(s.a, static_cast<uint32_t>(mb));
        if (kind == TPT_SLOT_MSPHERE) s.f[0] = rebase_bits(s.f[0], static_cast<uint32_t>(mob));
        if (kind == TPT_SLOT_MEDIUM) s.f[0] = rebase_bits(s.f[0], static_cast<uint32_t>(eb));
      }
      slots.push_back(s);
    }
    for (uint32_t g = 0; g < 2 * f.n_segments; ++g) segments.push_back(f.segments[g] + static_cast<uint32_t>(sb));
    for (uint32_t g = 0; g < f.n_segments; ++g) {
      uint32_t w = f.segment_xforms ? f.segment_xforms[g] : TPT_SEGMENT_PLAIN;
      if ((w & TPT_SEGMENT_PLAIN) != TPT_SEGMENT_PLAIN) w = (w & TPT_SEGMENT_FLIP) | ((w & TPT_SEGMENT_PLAIN) + static_cast<uint32_t>(xb));
      any_wrapped = any_wrapped || w != TPT_SEGMENT_PLAIN;
      seg_xf.push_back(w);
    }
    for (uint32_t e = 0; e < f.n_entries; ++e) {
      tpt_entry en = f.entries[e];
      if (en.xform != 0xFFFFFFFFu) en.xform += static_cast<uint32_t>(xb);
      en.a = rebase_word(en.a, static_cast<uint32_t>(mb));
      if (en.b == TPT_SLOT_MSPHERE) en.f[0] = rebase_bits(en.f[0], static_cast<uint32_t>(mob));
      if (en.b == TPT_SLOT_MEDIUM) en.f[0] = rebase_bits(en.f[0], static_cast<uint32_t>(eb));
      entries.push_back(en);
    }
    xforms.insert(xforms.end(), f.xforms, f.xforms + f.n_xforms);
    for (uint32_t m = 0; m < f.n_materials; ++m) {
      tpt_material M = f.materials[m];
      if (M.kind == TPT_MAT_LAMBERTIAN || M.kind == TPT_MAT_DIFFUSE_LIGHT || M.kind == TPT_MAT_ISOTROPIC)
        M.tex += static_cast<uint32_t>(tb);
      materials.push_back(M);
    }
    for (uint32_t t = 0; t < f.n_textures; ++t) {
      tpt_texture T = f.textures[t];
      if (T.kind == TPT_TEX_CHECKER) T.a += static_cast<uint32_t>(tb), T.b += static_cast<uint32_t>(tb);
      if (T.kind == TPT_TEX_IMAGE) T.a += static_cast<uint32_t>(bb);
      if (T.kind == TPT_TEX_PERLIN) T.b = static_cast<uint32_t>(pb);
      textures.push_back(T);
    }
    texels.insert(texels.end(), f.texels, f.texels + f.n_texels);
    if (f.perlin) perlin.push_back(*f.perlin);
    lights.insert(lights.end(), f.lights, f.lights + f.n_lights);
    if (f.n_motions) motions.insert(motions.end(), f.motions, f.motions + f.n_motions);
    scene_dev &d = scn[k];
    std::memset(&d, 0, sizeof(d));
    d.seg_begin = static_cast<uint32_t>(gb);
    d.n_seg = f.n_segments;
    d.light_begin = static_cast<uint32_t>(lb);
    d.n_lights = f.n_lights;
    d.light_mode = f.n_lights ? f.light_mode : 0;
    for (int q = 0; q < 6; ++q) d.bounds[q] = f.bounds[q];
    // Sphere-only walk (traverse.h, sphere_walk_bound): the scene's largest
    // coordinate magnitude, and the boxes holding a sphere of radius above
    // kWalkBigRadius flagged in the device copy of their skip word.
    if (sph[k]) {
      float mag = 0.0f;
      bool finite = true;
      std::vector<uint32_t> big(f.n_slots + 1, 0u);
      for (uint32_t i = 0; i < f.n_slots; ++i) {
        const tpt_slot &s = f.slots[i];
        const bool box = (s.b >> 28) == TPT_SLOT_BOX;
        for (int q = 0; q < (box ? 6 : 4); ++q) finite = finite && std::fabs(s.f[q]) < INFINITY;
        for (int q = 0; q < (box ? 6 : 3); ++q) mag = std::fmax(mag, std::fabs(s.f[q]) + (box ? 0.0f : std::fabs(s.f[3])));
        big[i + 1] = big[i] + (!box && !(std::fabs(s.f[3]) <= kWalkBigRadius) ? 1u : 0u);
      }
      // a non-finite coordinate leaves every box at t_max (an infinite reach)
      d.walk_ulp = finite ? std::ldexp(mag, -22) : INFINITY;
      for (uint32_t i = 0; i < f.n_slots; ++i)
        if ((f.slots[i].b >> 28) == TPT_SLOT_BOX && big[f.slots[i].a] != big[i]) slots[sb + i].a |= kBoxUnbounded;
    }
    motion[k] = f.n_motions > 0;
    media[k] = 0;
    for (uint32_t i = 0; i < f.n_slots; ++i) media[k] |= (f.slots[i].b >> 28) == TPT_SLOT_MEDIUM;
    for (uint32_t e = 0; e < f.n_entries; ++e) media[k] |= f.entries[e].b == TPT_SLOT_MEDIUM;
    // Wave-uniform sweeps pay for the union of the lanes' visits; that stays
    // close to each lane's own count only when the whole tree is small.  (The
    // sweep kernel is built without the ray-time replay and the media draws.)
    bool wrapped = false;
    for (uint32_t g = 0; f.segment_xforms && g < f.n_segments; ++g) wrapped |= f.segment_xforms[g] != TPT_SEGMENT_PLAIN;
    sweep[k] = f.n_slots <= 64 && !motion[k] && !media[k] && !wrapped;
    if (wrapped) sph[k] = 0;  // the sphere-only walk is built without segment transforms
    simple[k] = !motion[k] && !media[k] &&
                (f.n_lights == 0 || (f.light_mode == 0 && f.lights[0].kind == TPT_LIGHT_XZ_RECT));
    for (uint32_t t = 0; t < f.n_textures; ++t) simple[k] = simple[k] && f.textures[t].kind == TPT_TEX_CONSTANT;
    nslots[k] = f.n_slots;
    nleaves[k] = f.n_leaves;
  }
  if (tpt_device_count() < 1) return tpt::fail(TPT_ERR_NODEVICE, "tpt_scene_create: no gfx950 device");
  auto *s = new tpt_scene();
  int rc = TPT_OK;
  HIP_TRY(hipGetDevice(&s->device));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, s->device) == hipSuccess) s->n_cu = prop.multiProcessorCount;
  if ((rc = upload(&s->slots, slots.data(), slots.size())) ||
      (rc = upload(&s->segments, segments.data(), segments.size())) ||
      (rc = upload(&s->entries, entries.data(), entries.size())) ||
      (rc = upload(&s->xforms, xforms.data(), xforms.size())) ||
      (rc = upload(&s->materials, materials.data(), materials.size())) ||
      (rc = upload(&s->textures, textures.data(), textures.size())) ||
      (rc = upload(&s->texels, texels.data(), texels.size())) ||
      (rc = upload(&s->perlin, perlin.data(), perlin.size())) ||
      (rc = upload(&s->lights, lights.data(), lights.size())) ||
      (rc = upload(&s->motions, motions.data(), motions.size())) ||
      (rc = upload(&s->seg_xf, seg_xf.data(), any_wrapped ? seg_xf.size() : 0)) ||
      (rc = upload(&s->d_scenes, scn.data(), scn.size()))) {
    free_scene(s);
    delete s;
    return rc;
  }
  s->n_slots = static_cast<uint32_t>(slots.size());
  s->n_segments = static_cast<uint32_t>(segments.size() / 2);
  s->n_lights = static_cast<uint32_t>(lights.size());
  s->scn = scn;
  s->sweep = sweep;
  s->sph = sph;
  s->simple = simple;
  s->motion = motion;
  s->media = media;
  s->n_slots_of = nslots;
  s->n_leaves_of = nleaves;
  if (hipHostMalloc(reinterpret_cast<void **>(&s->host_count), NSUB * CSTRIDE * sizeof(uint32_t), hipHostMallocDefault) !=
      hipSuccess) {
    free_scene(s);
    delete s;
    return tpt::fail(TPT_ERR_HIP, "tpt_scene_create: hipHostMalloc failed");
  }
  *out = s;
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
  const int blocks =
      static_cast<int>(std::min<uint64_t>((n + GT_ROWS - 1) / GT_ROWS, static_cast<uint64_t>(g->train_blocks)));
  const size_t lds = k_guide_train_lds();
  hipLaunchKernelGGL(k_guide_train, dim3(blocks), dim3(GT_THREADS), lds, stream, g->rtiles, g->params, rows, n, row0,
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
  return create_pool(f, 1, out);
}

int tpt_scene_create_pool(const tpt_flat_scene *scenes, uint32_t n_scenes, tpt_scene **out) {
  return create_pool(scenes, n_scenes, out);
}

int tpt_scene_destroy(tpt_scene *s) {
  if (!s) return TPT_OK;
  (void)hipDeviceSynchronize();
  free_scene(s);
  delete s;
  return TPT_OK;
}

int tpt_trace_closest_timed(tpt_scene *s, const float *d_rays, const float *d_time, uint64_t n, float t_min,
                            float t_max, int32_t *d_leaf, float *d_t, float *d_rec, int32_t *d_mat, void *stream) {
  if (!s || (n && (!d_rays || !d_leaf || !d_t))) return tpt::fail(TPT_ERR_INVALID, "tpt_trace_closest: null argument");
  if (n == 0) return TPT_OK;
  const uint64_t blocks = (n + 255) / 256;
  if (blocks > 0x7FFFFFFFull) return tpt::fail(TPT_ERR_INVALID, "tpt_trace_closest: too many rays");
  const scene_view S = s->view(0);
  if (s->sweep[0])
    hipLaunchKernelGGL((k_trace_closest<true, false>), dim3(static_cast<uint32_t>(blocks)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), S, S.slots, S.entries, d_rays, d_time, n, t_min, t_max,
                       d_leaf, d_t, d_rec, d_mat);
  else if (s->media[0])
    hipLaunchKernelGGL((k_trace_closest<false, true>), dim3(static_cast<uint32_t>(blocks)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), S, S.slots, S.entries, d_rays, d_time, n, t_min, t_max,
                       d_leaf, d_t, d_rec, d_mat);
  else if (s->sph[0])
    hipLaunchKernelGGL((k_trace_closest<false, false, true>), dim3(static_cast<uint32_t>(blocks)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), S, S.slots, S.entries, d_rays, d_time, n, t_min, t_max,
                       d_leaf, d_t, d_rec, d_mat);
  else
    hipLaunchKernelGGL((k_trace_closest<false, false>), dim3(static_cast<uint32_t>(blocks)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), S, S.slots, S.entries, d_rays, d_time, n, t_min, t_max,
                       d_leaf, d_t, d_rec, d_mat);
  HIP_TRY(hipGetLastError());
  return TPT_OK;
}

int tpt_trace_closest_record(tpt_scene *s, const float *d_rays, uint64_t n, float t_min, float t_max,
                             int32_t *d_leaf, float *d_t, float *d_rec, int32_t *d_mat, void *stream) {
  return tpt_trace_closest_timed(s, d_rays, nullptr, n, t_min, t_max, d_leaf, d_t, d_rec, d_mat, stream);
}

int tpt_trace_closest(tpt_scene *s, const float *d_rays, uint64_t n, float t_min, float t_max, int32_t *d_leaf,
                      float *d_t, void *stream) {
  return tpt_trace_closest_record(s, d_rays, n, t_min, t_max, d_leaf, d_t, nullptr, nullptr, stream);
}

}  // extern "C"

namespace {

// Items per pass.  A pass ends by draining the queues (its last bounces run
// at low occupancy) and starts with a fill, so passes are long: 2^31 items
// hold 25.8 GB of per-sample results (item indices stay below 2^31 + 64, the
// overshoot of a shard cursor).  With training rows recorded a pass is 2^26
// items (the rows of one pass are kept until the next).
constexpr uint64_t kPassItems = 1ull << 31;
constexpr uint64_t kPassItemsRecording = 1ull << 26;
constexpr uint32_t kMaxEnt = 0xFFFFu;        // pass entry index shares a word with the depth

struct pass_plan {
  std::vector<pass_ent> ents;
  uint32_t n_items = 0, n_acc = 0;
};

// The jobs' sample layers, job after job, packed into passes of at most
// pass_items items and kMaxEnt entries (a job has one entry per pass).
std::vector<pass_plan> plan_passes(const tpt_job *jobs, uint32_t n_jobs, uint64_t pass_items) {
  std::vector<pass_plan> out;
  pass_plan cur;
  for (uint32_t j = 0; j < n_jobs; ++j) {
    const uint32_t npix = jobs[j].width * jobs[j].height;
    uint32_t s = jobs[j].sample_offset;
    const uint32_t s_end = jobs[j].sample_offset + jobs[j].spp;
    while (s < s_end) {
      const uint64_t room = (pass_items - cur.n_items) / npix;
      if (room == 0 || cur.ents.size() >= kMaxEnt) {
        out.push_back(cur);
        cur = pass_plan();
        continue;
      }
      const uint32_t take = static_cast<uint32_t>(std::min<uint64_t>(room, s_end - s));
      pass_ent e{};
      e.item_begin = cur.n_items;
      e.job = j;
      e.s_begin = s;
      e.n_layers = take;
      e.acc_begin = cur.n_acc;
      cur.ents.push_back(e);
      cur.n_items += take * npix;
      cur.n_acc += npix;
      s += take;
    }
  }
  if (!cur.ents.empty()) out.push_back(cur);
  return out;
}

template <class T>
int grow_device(T **p, uint32_t *cap, uint32_t need) {
  if (need <= *cap) return TPT_OK;
  (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  HIP_TRY(hipMalloc(reinterpret_cast<void **>(p), static_cast<size_t>(need) * sizeof(T)));
  *cap = need;
  return TPT_OK;
}

// Live-count polls of the bounce loop (see render_core).
constexpr uint64_t kPoll = 16, kPollFirst = 8, kPollTail = 4;

struct bounce_launch {
  scene_view S;
  render_params P;
  render_state Q;
  uint32_t grid, guide_grid;
  bool sweep, multi, guided;
  bool sph;     // every leaf of the launch's scenes is a sphere
  bool simple;  // their textures are constant and their light root absent or one xz_rect
  bool tail;    // passes end with a fused k_tail launch (simple scenes, sweep or sphere-only walk)
  bool count;   // TPT_RENDER_COUNT_OPS: the counting variant of k_trace, no fused tail
};

// The fused tail on ray queue `parity` (k_tail), `blocks` blocks.
void launch_tail(const bounce_launch &B, hipStream_t st, uint32_t parity, uint32_t blocks) {
  const dim3 g(blocks), b(256);
  if (B.sweep) {
    if (B.guided)
      hipLaunchKernelGGL((k_tail<true, false, false, true>), g, b, 0, st, B.S, B.S.slots, B.S.entries, B.P, B.Q,
                         parity);
    else
      hipLaunchKernelGGL((k_tail<true, false, false, false>), g, b, 0, st, B.S, B.S.slots, B.S.entries, B.P, B.Q,
                         parity);
  } else if (B.multi) {
    if (B.guided)
      hipLaunchKernelGGL((k_tail<false, true, true, true>), g, b, 0, st, B.S, B.S.slots, B.S.entries, B.P, B.Q,
                         parity);
    else
      hipLaunchKernelGGL((k_tail<false, true, true, false>), g, b, 0, st, B.S, B.S.slots, B.S.entries, B.P, B.Q,
                         parity);
  } else {
    if (B.guided)
      hipLaunchKernelGGL((k_tail<false, false, true, true>), g, b, 0, st, B.S, B.S.slots, B.S.entries, B.P, B.Q,
                         parity);
    else
      hipLaunchKernelGGL((k_tail<false, false, true, false>), g, b, 0, st, B.S, B.S.slots, B.S.entries, B.P, B.Q,
                         parity);
  }
}

template <bool COUNT>
void launch_trace_t(const bounce_launch &B, hipStream_t st, uint32_t parity) {
  const dim3 g(B.grid), b(TRACE_THREADS);
  if (B.sweep)
    hipLaunchKernelGGL((k_trace<true, false, false, false, COUNT>), g, b, 0, st, B.S, B.S.slots, B.S.entries, B.P,
                       B.Q, parity);
  else if (B.multi && B.P.media)
    hipLaunchKernelGGL((k_trace<false, true, false, true, COUNT>), g, b, 0, st, B.S, B.S.slots, B.S.entries, B.P,
                       B.Q, parity);
  else if (B.P.media)
    hipLaunchKernelGGL((k_trace<false, false, false, true, COUNT>), g, b, 0, st, B.S, B.S.slots, B.S.entries, B.P,
                       B.Q, parity);
  else if (B.multi && B.sph)
    hipLaunchKernelGGL((k_trace<false, true, true, false, COUNT>), g, b, 0, st, B.S, B.S.slots, B.S.entries, B.P,
                       B.Q, parity);
  else if (B.multi)
    hipLaunchKernelGGL((k_trace<false, true, false, false, COUNT>), g, b, 0, st, B.S, B.S.slots, B.S.entries, B.P,
                       B.Q, parity);
  else if (B.sph)
    hipLaunchKernelGGL((k_trace<false, false, true, false, COUNT>), g, b, 0, st, B.S, B.S.slots, B.S.entries, B.P,
                       B.Q, parity);
  else
    hipLaunchKernelGGL((k_trace<false, false, false, false, COUNT>), g, b, 0, st, B.S, B.S.slots, B.S.entries, B.P,
                       B.Q, parity);
}

void launch_trace(const bounce_launch &B, hipStream_t st, uint32_t parity) {
  if (B.count)
    launch_trace_t<true>(B, st, parity);
  else
    launch_trace_t<false>(B, st, parity);
}

void launch_shade(const bounce_launch &B, hipStream_t st, uint32_t parity) {
  const dim3 g(B.grid), b(256);
  if (B.guided) {
    const dim3 gg(B.guide_grid), gbs(guide_shade_threads(true)), gbg(guide_shade_threads(false));
    if (B.multi && B.simple) {
      hipLaunchKernelGGL((k_shade<true, true, true>), g, b, 0, st, B.S, B.P, B.Q, parity);
      hipLaunchKernelGGL((k_guide_shade<true, true>), gg, gbs, 0, st, B.S, B.P, B.Q, parity);
    } else if (B.multi) {
      hipLaunchKernelGGL((k_shade<true, true>), g, b, 0, st, B.S, B.P, B.Q, parity);
      hipLaunchKernelGGL((k_guide_shade<true>), gg, gbg, 0, st, B.S, B.P, B.Q, parity);
    } else if (B.simple) {
      hipLaunchKernelGGL((k_shade<true, false, true>), g, b, 0, st, B.S, B.P, B.Q, parity);
      hipLaunchKernelGGL((k_guide_shade<false, true>), gg, gbs, 0, st, B.S, B.P, B.Q, parity);
    } else {
      hipLaunchKernelGGL((k_shade<true, false>), g, b, 0, st, B.S, B.P, B.Q, parity);
      hipLaunchKernelGGL((k_guide_shade<false>), gg, gbg, 0, st, B.S, B.P, B.Q, parity);
    }
  } else if (B.multi && B.simple) {
    hipLaunchKernelGGL((k_shade<false, true, true>), g, b, 0, st, B.S, B.P, B.Q, parity);
  } else if (B.multi) {
    hipLaunchKernelGGL((k_shade<false, true>), g, b, 0, st, B.S, B.P, B.Q, parity);
  } else if (B.simple) {
    hipLaunchKernelGGL((k_shade<false, false, true>), g, b, 0, st, B.S, B.P, B.Q, parity);
  } else {
    hipLaunchKernelGGL((k_shade<false, false>), g, b, 0, st, B.S, B.P, B.Q, parity);
  }
}

// Renders every job; the render loop described in render.h.
int render_core(tpt_scene *s, const tpt_job *jobs, uint32_t n_jobs, const tpt_launch &L, tpt_guide *guide,
                tpt_render_stats *stats, void *stream_ptr) {
  if (!s || (n_jobs && !jobs)) return tpt::fail(TPT_ERR_INVALID, "tpt_render: null argument");
  if (L.flags & ~(TPT_RENDER_PROFILE | TPT_RENDER_MEDIA_SCATTER | TPT_RENDER_NO_TAIL | TPT_RENDER_COUNT_OPS))
    return tpt::fail(TPT_ERR_INVALID, "tpt_render: unknown flags");
  if (!guide && L.guiding != 0) return tpt::fail(TPT_ERR_INVALID, "tpt_render: guiding needs tpt_render_guided");
  if (L.queue_capacity > (1u << 29)) return tpt::fail(TPT_ERR_INVALID, "tpt_render: queue_capacity above 2^29");
  if (L.blocks_per_cu > 64) return tpt::fail(TPT_ERR_INVALID, "tpt_render: blocks_per_cu above 64");
  if (L.graph_iters > 256) return tpt::fail(TPT_ERR_INVALID, "tpt_render: graph_iters above 256");
  uint32_t max_depth = 0;
  uint64_t total = 0;
  for (uint32_t j = 0; j < n_jobs; ++j) {
    const tpt_job &J = jobs[j];
    if (J.scene >= s->scn.size()) return tpt::fail(TPT_ERR_INVALID, "tpt_render: job scene out of range");
    if (J.width == 0 || J.height == 0) return tpt::fail(TPT_ERR_INVALID, "tpt_render: empty image");
    if (static_cast<uint64_t>(J.width) * J.height > kPassItemsRecording)
      return tpt::fail(TPT_ERR_INVALID, "tpt_render: more than 2^26 pixels");
    if (!J.d_sum) return tpt::fail(TPT_ERR_INVALID, "tpt_render: null d_sum");
    if (J.max_depth > 0xFFFEu) return tpt::fail(TPT_ERR_INVALID, "tpt_render: max_depth above 65534");
    if (static_cast<uint64_t>(J.sample_offset) + J.spp > 0xFFFFFFFFull)
      return tpt::fail(TPT_ERR_INVALID, "tpt_render: sample range overflows");
    max_depth = std::max(max_depth, J.max_depth);
    total += static_cast<uint64_t>(J.spp) * J.width * J.height;
  }
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (total == 0) return TPT_OK;
  const bool profile = (L.flags & TPT_RENDER_PROFILE) != 0;
  const bool multi = n_jobs > 1;
  const bool use_graph = L.graph_iters > 0 && !profile;
  hipStream_t cs = static_cast<hipStream_t>(stream_ptr);

  // job table
  std::vector<job_dev> jd(n_jobs);
  for (uint32_t j = 0; j < n_jobs; ++j) {
    const tpt_job &J = jobs[j];
    job_dev &d = jd[j];
    std::memset(&d, 0, sizeof(d));
    d.cam = J.camera;
    d.width = J.width;
    d.height = J.height;
    d.n_pix = J.width * J.height;
    d.max_depth = J.max_depth;
    d.seed = J.seed;
    d.background = J.background;
    d.scene = J.scene;
    d.sum = J.d_sum;
    d.sumsq = J.d_sumsq;
  }
  const std::vector<pass_plan> plan =
      plan_passes(jobs, n_jobs, guide && L.guiding ? kPassItemsRecording : kPassItems);
  uint32_t max_items = 0, max_ent = 0;
  for (const auto &pp : plan) {
    max_items = std::max(max_items, pp.n_items);
    max_ent = std::max<uint32_t>(max_ent, static_cast<uint32_t>(pp.ents.size()));
  }
  int rc = grow_device(&s->d_jobs, &s->jobs_cap, n_jobs);
  if (rc != TPT_OK) return rc;
  if (max_ent > s->pass_cap) {
    if (s->h_pass) (void)hipHostFree(s->h_pass);
    s->h_pass = nullptr;
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&s->h_pass), max_ent * sizeof(pass_ent), hipHostMallocDefault));
    if ((rc = grow_device(&s->d_pass, &s->pass_cap, max_ent)) != TPT_OK) return rc;
  }
  if (!s->d_hdr) {
    HIP_TRY(hipMalloc(&s->d_hdr, sizeof(pass_hdr)));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&s->h_hdr), sizeof(pass_hdr), hipHostMallocDefault));
  }
  if (use_graph && !s->gstream) {
    HIP_TRY(hipStreamCreateWithFlags(&s->gstream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&s->ev_in, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&s->ev_out, hipEventDisableTiming));
  }
  hipStream_t ws = use_graph ? s->gstream : cs;
  if (use_graph) {
    HIP_TRY(hipEventRecord(s->ev_in, cs));
    HIP_TRY(hipStreamWaitEvent(ws, s->ev_in, 0));
  }
  HIP_TRY(hipMemcpyAsync(s->d_jobs, jd.data(), n_jobs * sizeof(job_dev), hipMemcpyHostToDevice, ws));

  // Default grid: 64 blocks per CU for passes of 2^25 items or more (the
  // launch's last blocks are short, so its end is balanced; +2% on the Cornell
  // 1200^2 frame against 32), 32 below that, where the extra blocks' prologue
  // costs more than the balance returns.
  const uint32_t bpc = L.blocks_per_cu ? L.blocks_per_cu : (max_items >= (1u << 25) ? 64u : 32u);
  const uint32_t grid = static_cast<uint32_t>(s->n_cu) * bpc;
  // k_guide_shade's LDS (weights + output table) allows 2 blocks of 512 per CU
  // (simple scenes) or 3 of 256
  bool all_simple = true;
  for (uint32_t j = 0; j < n_jobs; ++j) all_simple = all_simple && s->simple[jobs[j].scene];
  const uint32_t guide_grid = static_cast<uint32_t>(s->n_cu) * guide_shade_blocks_per_cu(all_simple);
  // queue regions and statistics rows for k_guide_shade's grid only when it runs
  const uint32_t state_guide_grid = guide ? guide_grid : 0u;
  // Records in flight by default: 2^29 for single-scene renders that sweep
  // their tree (small scenes), half the largest pass rounded down to a power
  // of two (at least 2^27, at most 2^29) for the others, at most the pass, and
  // while the state fits in 70% of the device memory free now (with what this
  // scene's state already holds; never below 2^20).  More records make each
  // bounce launch longer, so its ramp-up and tail cost less; for per-lane
  // walks a fill that covers most of the pass costs more than that.  Measured
  // (Msamples/s at 2^27 / 2^28 / 2^29 records): the Cornell 1200^2 frame
  // (passes of 2^30 items) 3538 / 3613 / 3667, its guided frame (one pass of
  // 368M items) 2669 / 2689 / 2718, the textured spheres' 1600^2 frame (one
  // pass of 768M items, per-lane walks) 1736 / 1724 / 1672.
  const bool sweep_scene = !multi && s->sweep[jobs[0].scene];
  uint32_t cap = L.queue_capacity;
  if (!cap) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    const size_t budget = (free_b + s->state_bytes) / 10 * 7;
    cap = 1u << 29;
    while (!sweep_scene && cap > (1u << 27) && static_cast<uint64_t>(cap) * 2 > max_items) cap >>= 1;
    while (cap > (1u << 20) &&
           state_bytes_for(std::min(cap, max_items), max_items, grid, state_guide_grid) > budget)
      cap >>= 1;
  }
  cap = std::min(cap, max_items);
  if ((rc = ensure_state(s, cap, max_items, grid, state_guide_grid)) != TPT_OK) return rc;

  bounce_launch B;
  std::memset(&B, 0, sizeof(B));  // the captured-graph key compares its bytes
  B.S = multi ? s->pool_view() : s->view(jobs[0].scene);
  B.Q = s->Q;
  B.grid = grid;
  B.guide_grid = guide_grid;
  B.sweep = sweep_scene;
  B.sph = true;
  B.simple = true;
  for (uint32_t j = 0; j < n_jobs; ++j) {
    B.sph = B.sph && s->sph[jobs[j].scene];
    B.simple = B.simple && s->simple[jobs[j].scene];
    if (s->motion[jobs[j].scene]) B.P.motion = 1;
    if (s->media[jobs[j].scene]) B.P.media = 1;
  }
  B.multi = multi;
  B.guided = guide != nullptr;
  B.count = (L.flags & TPT_RENDER_COUNT_OPS) != 0;
  B.tail = !(L.flags & TPT_RENDER_NO_TAIL) && !B.count && B.simple && (B.sweep || B.sph);
  render_params &P = B.P;
  P.jobs = s->d_jobs;
  P.scenes = s->d_scenes;
  P.hdr = s->d_hdr;
  P.pass = s->d_pass;
  P.guide_verts = ~0u;
  P.media_scatter = (L.flags & TPT_RENDER_MEDIA_SCATTER) ? 1u : 0u;
  if (guide) {
    P.g_tiles = guide->xtiles;
    P.g_params = guide->params;
    P.guide_verts = L.guide_vertices ? L.guide_vertices : 2u;
    P.rec_stride = L.guiding;
    if (P.rec_stride) {
      const uint64_t need = max_items / P.rec_stride + 1;
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
      guide->n_rows = (static_cast<uint64_t>(max_items) + P.rec_stride - 1) / P.rec_stride * REC_VERTS;
    }
  }

  // The captured bounce loop: graph_iters iterations (even, so a replay
  // starts and ends at parity 0), kept while nothing it captured changes.
  const uint32_t K = use_graph ? (L.graph_iters + 1) / 2 * 2 : 0;
  if (use_graph) {
    std::vector<uint8_t> key(sizeof(bounce_launch) + sizeof(K));
    std::memcpy(key.data(), &B, sizeof(bounce_launch));
    std::memcpy(key.data() + sizeof(bounce_launch), &K, sizeof(K));
    if (!s->gexec || key != s->gkey) {
      if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
      s->gexec = nullptr;
      s->gkey.clear();
      hipGraph_t graph = nullptr;
      HIP_TRY(hipStreamBeginCapture(ws, hipStreamCaptureModeThreadLocal));
      for (uint32_t it = 0; it < K; ++it) {
        launch_trace(B, ws, it & 1u);
        launch_shade(B, ws, it & 1u);
      }
      HIP_TRY(hipStreamEndCapture(ws, &graph));
      const hipError_t ie = hipGraphInstantiate(&s->gexec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) {
        s->gexec = nullptr;
        return tpt::fail(TPT_ERR_HIP, std::string("tpt_render: hipGraphInstantiate: ") + hipGetErrorString(ie));
      }
      s->gkey = key;
    }
  }

  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, ws));
  const uint32_t n_stat = std::max(grid, state_guide_grid);  // rows written by block_stats
  hipLaunchKernelGGL(k_zero_stats, dim3(64), dim3(256), 0, ws, s->Q.stats, n_stat * ST_COUNT);
  const int poll = static_cast<int>(kPoll);
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
  // live records of the next input queue (parity p)
  auto live_count = [&](uint32_t p, uint64_t *live) -> int {
    HIP_TRY(hipMemcpyAsync(s->host_count, s->Q.counters + rc_at(p, 0), NSUB * CSTRIDE * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, ws));
    HIP_TRY(hipStreamSynchronize(ws));
    *live = 0;
    for (uint32_t k = 0; k < NSUB; ++k) *live += s->host_count[k * CSTRIDE];
    return TPT_OK;
  };
  // The fused tail (k_tail) takes over a pass once it has handed out every
  // item (fewer live records than the capacity) and at most tail_live paths
  // are left: from there on a bounce launch pair costs mostly latency.  Polls
  // come every kPoll iterations, the first after kPollFirst, and every
  // kPollTail once the items are out, so the switch comes early.
  const uint64_t tail_live = std::min<uint64_t>(cap / 4, 1u << 18);
  double tail_ms = 0.0;
  uint64_t n_tail = 0;
  auto run_tail = [&](uint32_t q, uint64_t live) -> int {
    // at least NSUB blocks (the hand-back appends stay within their regions)
    const uint32_t blocks =
        static_cast<uint32_t>(std::max<uint64_t>(NSUB, std::min<uint64_t>(grid, (live + 255) / 256)));
    if (profile) HIP_TRY(hipEventRecord(s->events[0], ws));
    launch_tail(B, ws, q, blocks);
    if (profile) {
      HIP_TRY(hipEventRecord(s->events[1], ws));
      HIP_TRY(hipEventSynchronize(s->events[1]));
      float a = 0.0f;
      if (hipEventElapsedTime(&a, s->events[0], s->events[1]) == hipSuccess) tail_ms += a;
    }
    ++n_tail;
    return TPT_OK;
  };
  bool drained = true;
  for (const pass_plan &pp : plan) {
    if (!drained) break;
    // The staging copies of the previous pass completed at its last poll.
    s->h_hdr->n_items = pp.n_items;
    s->h_hdr->n_first = std::min(cap, pp.n_items);
    s->h_hdr->n_ent = static_cast<uint32_t>(pp.ents.size());
    s->h_hdr->n_acc = pp.n_acc;
    std::memcpy(s->h_pass, pp.ents.data(), pp.ents.size() * sizeof(pass_ent));
    HIP_TRY(hipMemcpyAsync(s->d_hdr, s->h_hdr, sizeof(pass_hdr), hipMemcpyHostToDevice, ws));
    HIP_TRY(hipMemcpyAsync(s->d_pass, s->h_pass, pp.ents.size() * sizeof(pass_ent), hipMemcpyHostToDevice, ws));
    // counters only: the first k_trace generates the first items' rays
    hipLaunchKernelGGL(k_init, dim3((COUNTER_WORDS / CSTRIDE + 255) / 256), dim3(256), 0, ws, P, s->Q);
    // A path has at most max_depth + 1 segments and each bounce retires at
    // least one path per live record, so the loop is bounded.
    const uint64_t max_iter = (static_cast<uint64_t>(pp.n_items) / cap + 2) * (max_depth + 2ull) + 64;
    drained = false;
    bool exhausted = pp.n_items <= cap;
    if (use_graph) {
      for (uint64_t it = 0; it < max_iter; it += K) {
        HIP_TRY(hipGraphLaunch(s->gexec, ws));
        uint64_t live = 0;
        if ((rc = live_count(0, &live)) != TPT_OK) return rc;
        if (live == 0) {
          drained = true;
          break;
        }
        exhausted = exhausted || live < cap;
        if (B.tail && exhausted && live <= tail_live) {
          if ((rc = run_tail(0, live)) != TPT_OK) return rc;
          if (!B.guided) {
            drained = true;
            break;
          }
          // handed-back guided vertices: shade them and bring their paths
          // back to queue 0, where the next replay starts
          launch_shade(B, ws, 0);
          launch_trace(B, ws, 1);
          launch_shade(B, ws, 1);
          if ((rc = live_count(0, &live)) != TPT_OK) return rc;
          if (live == 0) {
            drained = true;
            break;
          }
        }
      }
    } else {
      int batch = 0;
      uint64_t next_poll = (B.tail ? kPollFirst : static_cast<uint64_t>(poll)) - 1;
      for (uint64_t it = 0; it < max_iter; ++it) {
        const uint32_t parity = static_cast<uint32_t>(it & 1u);
        if (profile) HIP_TRY(hipEventRecord(s->events[2 * batch], ws));
        launch_trace(B, ws, parity);
        if (profile) HIP_TRY(hipEventRecord(s->events[2 * batch + 1], ws));
        launch_shade(B, ws, parity);
        if (profile) HIP_TRY(hipEventRecord(s->events[2 * batch + 2], ws));
        ++batch;
        if (it == next_poll) {
          uint64_t live = 0;
          if ((rc = live_count(parity ^ 1u, &live)) != TPT_OK) return rc;
          if (profile) harvest(batch);
          batch = 0;
          if (live == 0) {
            drained = true;
            break;
          }
          exhausted = exhausted || live < cap;
          if (B.tail && exhausted && live <= tail_live) {
            if ((rc = run_tail(parity ^ 1u, live)) != TPT_OK) return rc;
            if (!B.guided) {
              drained = true;
              break;
            }
            // handed-back guided vertices: shade them; their paths continue
            // in queue `parity`, the input of iteration it + 2
            launch_shade(B, ws, parity ^ 1u);
            ++it;
            if ((rc = live_count(parity, &live)) != TPT_OK) return rc;
            if (live == 0) {
              drained = true;
              break;
            }
          }
          next_poll = it + (B.tail && exhausted ? kPollTail : static_cast<uint64_t>(poll));
        }
      }
      if (profile && batch > 0) {
        HIP_TRY(hipStreamSynchronize(ws));
        harvest(batch);
      }
    }
    if (!drained) break;
    const dim3 ag(std::min<uint32_t>((pp.n_acc + 255) / 256, grid));
    if (multi)
      hipLaunchKernelGGL(k_accumulate<true>, ag, dim3(256), 0, ws, P, static_cast<const float *>(s->Q.res));
    else
      hipLaunchKernelGGL(k_accumulate<false>, ag, dim3(256), 0, ws, P, static_cast<const float *>(s->Q.res));
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(e1, ws));
  std::vector<unsigned long long> st(static_cast<size_t>(n_stat) * ST_COUNT);
  HIP_TRY(hipMemcpyAsync(st.data(), s->Q.stats, st.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, ws));
  HIP_TRY(hipStreamSynchronize(ws));
  if (use_graph) {
    HIP_TRY(hipEventRecord(s->ev_out, ws));
    HIP_TRY(hipStreamWaitEvent(cs, s->ev_out, 0));
  }
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!drained) return tpt::fail(TPT_ERR_HIP, "tpt_render: bounce loop did not drain");
  if (stats) {
    unsigned long long tot[ST_COUNT] = {0};
    for (uint32_t b = 0; b < n_stat; ++b)
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
    stats->tail_segments = tot[ST_TAIL];
    stats->tail_ms = tail_ms;
    stats->tail_launches = n_tail;
    stats->box_tests = tot[ST_BOX];
    stats->sphere_tests = tot[ST_SPHERE];
    stats->rect_tests = tot[ST_RECT];
    stats->xforms = tot[ST_XFORM];
  }
  return TPT_OK;
}

int render_single(tpt_scene *s, tpt_guide *guide, const tpt_camera *camera, const tpt_render_desc *desc, float *d_sum,
                  float *d_sumsq, tpt_render_stats *stats, void *stream) {
  if (!s || !camera || !desc || !d_sum) return tpt::fail(TPT_ERR_INVALID, "tpt_render: null argument");
  tpt_job J;
  std::memset(&J, 0, sizeof(J));
  J.scene = 0;
  J.width = desc->width;
  J.height = desc->height;
  J.spp = desc->spp;
  J.max_depth = desc->max_depth;
  J.sample_offset = desc->sample_offset;
  J.seed = desc->seed;
  J.background = desc->background;
  J.camera = *camera;
  J.d_sum = d_sum;
  J.d_sumsq = d_sumsq;
  tpt_launch L;
  std::memset(&L, 0, sizeof(L));
  L.queue_capacity = desc->queue_capacity;
  L.flags = desc->flags;
  L.guiding = desc->guiding;
  L.guide_vertices = desc->guide_vertices;
  return render_core(s, &J, 1, L, guide, stats, stream);
}

// ---------------------------------------------- render without allocation ---
// All render state of tpt_render_noalloc lives in `scratch`.  In host code the
// name stands for a host-side shadow, not device memory: the array's device
// address on each GPU comes from hipGetSymbolAddress (noalloc_base), and only
// that address is passed to kernels.
constexpr int kNoallocMaxDevices = 64;
std::mutex g_noalloc_mu;
char *g_noalloc_base[kNoallocMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[TPT_NOALLOC_SCRATCH_BYTES];

namespace {

// Device address of `scratch` on device `dev`, which must be current; taken
// once per device and kept.
int noalloc_base(int dev, char **out) {
  if (dev < 0 || dev >= kNoallocMaxDevices)
    return tpt::fail(TPT_ERR_INVALID, "tpt_render_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_base[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return tpt::fail(TPT_ERR_HIP, std::string("tpt_render_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_noalloc_base[dev] = static_cast<char *>(p);
  }
  *out = g_noalloc_base[dev];
  return TPT_OK;
}

// The job, the pass header and the one pass-table entry of a pass, as kernel
// arguments: the host needs no staging buffer.
__global__ void k_noalloc_pass(job_dev *jobs, job_dev job, pass_hdr *hdr, pass_hdr h, pass_ent *pass, pass_ent e) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    jobs[0] = job;
    *hdr = h;
    pass[0] = e;
  }
}

// Column sums of `rows` per-block statistics rows into out[0..ST_COUNT).
__global__ void __launch_bounds__(256) k_noalloc_stats(const unsigned long long *stats, uint32_t rows,
                                                       unsigned long long *out) {
  __shared__ unsigned long long part[256];
  constexpr uint32_t kLanes = 256 / ST_COUNT;  // partial sums per counter
  const uint32_t k = threadIdx.x % ST_COUNT, lane = threadIdx.x / ST_COUNT;
  unsigned long long v = 0;
  for (uint32_t r = lane; r < rows; r += kLanes) v += stats[static_cast<size_t>(r) * ST_COUNT + k];
  part[threadIdx.x] = v;
  __syncthreads();
  if (threadIdx.x < ST_COUNT) {
    unsigned long long t = 0;
    for (uint32_t l = 0; l < kLanes; ++l) t += part[l * ST_COUNT + threadIdx.x];
    out[threadIdx.x] = t;
  }
}

// Grid of the no-allocation render: 4 blocks of 256 per CU, at most 1024
// blocks (the statistics rows and the region bound are sized for that), at
// least NSUB (the fused tail launches at least NSUB blocks).
constexpr uint32_t kNoallocMaxGrid = 1024;

// Byte offsets of the render state inside `scratch`, in 256-byte pieces.
struct noalloc_layout {
  size_t hdr, pass, jobs, counters, stats, totals, rq[2][3], mq[3][4], res, end;
  uint32_t region;
};

noalloc_layout noalloc_plan(uint32_t grid) {
  noalloc_layout L{};
  // Plain renders: only the grid-G trace and shade launches of 256-thread
  // blocks append (k_trace to the material queues, k_shade to the next ray
  // queue), at most one entry per live record, so a region holds at most
  // t/NSUB + t/G + ceil(G/NSUB)*256 entries (region_size without the guided
  // shading term); the first items' region counts that k_init sets (runs of
  // 64 entries) fit in the +256.
  const uint32_t cap = TPT_NOALLOC_PATHS;
  L.region = cap / NSUB + cap / grid + (grid + NSUB - 1) / NSUB * 256 + 256;
  const size_t n = static_cast<size_t>(L.region) * NSUB;
  size_t w = 0;
  auto take = [&](size_t bytes) {
    const size_t at = w;
    w += (bytes + 255) & ~size_t(255);
    return at;
  };
  L.hdr = take(sizeof(pass_hdr));
  L.pass = take(sizeof(pass_ent));
  L.jobs = take(sizeof(job_dev));
  L.counters = take(COUNTER_WORDS * sizeof(uint32_t));
  L.stats = take(static_cast<size_t>(kNoallocMaxGrid) * ST_COUNT * sizeof(unsigned long long));
  L.totals = take(ST_COUNT * sizeof(unsigned long long));
  for (int q = 0; q < 2; ++q)
    for (int k = 0; k < 3; ++k) L.rq[q][k] = take(n * 16);
  for (int c = 0; c < 3; ++c)
    for (int k = 0; k < 4; ++k) L.mq[c][k] = take(n * 16);
  L.res = take(static_cast<size_t>(TPT_NOALLOC_PASS_ITEMS) * 3 * sizeof(float));
  L.end = w;
  return L;
}

int render_noalloc(tpt_scene *s, const tpt_camera *camera, const tpt_render_desc *desc, float *d_sum,
                   float *d_sumsq, tpt_render_stats *stats, void *stream_ptr) {
  if (!s || !camera || !desc || !d_sum) return tpt::fail(TPT_ERR_INVALID, "tpt_render_noalloc: null argument");
  if (desc->flags & ~(TPT_RENDER_MEDIA_SCATTER | TPT_RENDER_NO_TAIL))
    return tpt::fail(TPT_ERR_INVALID, "tpt_render_noalloc: unsupported flags (profiling needs events)");
  if (desc->guiding) return tpt::fail(TPT_ERR_INVALID, "tpt_render_noalloc: guiding needs tpt_render_guided");
  if (desc->queue_capacity > TPT_NOALLOC_PATHS)
    return tpt::fail(TPT_ERR_INVALID, "tpt_render_noalloc: queue_capacity above TPT_NOALLOC_PATHS (2^17)");
  if (desc->width == 0 || desc->height == 0) return tpt::fail(TPT_ERR_INVALID, "tpt_render_noalloc: empty image");
  const uint64_t npix = static_cast<uint64_t>(desc->width) * desc->height;
  if (npix > TPT_NOALLOC_PASS_ITEMS)
    return tpt::fail(TPT_ERR_INVALID, "tpt_render_noalloc: image of " + std::to_string(npix) +
                                          " pixels above TPT_NOALLOC_PASS_ITEMS (" +
                                          std::to_string(TPT_NOALLOC_PASS_ITEMS) + ")");
  if (desc->max_depth > 0xFFFEu) return tpt::fail(TPT_ERR_INVALID, "tpt_render_noalloc: max_depth above 65534");
  if (static_cast<uint64_t>(desc->sample_offset) + desc->spp > 0xFFFFFFFFull)
    return tpt::fail(TPT_ERR_INVALID, "tpt_render_noalloc: sample range overflows");
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (desc->spp == 0) return TPT_OK;
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev != s->device) return tpt::fail(TPT_ERR_INVALID, "tpt_render_noalloc: the scene's device is not current");
  char *base = nullptr;
  int rc = noalloc_base(dev, &base);
  if (rc != TPT_OK) return rc;
  const uint32_t grid = std::max<uint32_t>(NSUB, 4u * static_cast<uint32_t>(std::min(s->n_cu, 256)));
  const noalloc_layout Lo = noalloc_plan(grid);
  if (Lo.end > TPT_NOALLOC_SCRATCH_BYTES)
    return tpt::fail(TPT_ERR_INVALID, "tpt_render_noalloc: render state exceeds the scratch array");
  const auto t_begin = std::chrono::steady_clock::now();
  hipStream_t ws = static_cast<hipStream_t>(stream_ptr);

  render_state Q{};
  Q.region = Lo.region;
  Q.counters = reinterpret_cast<uint32_t *>(base + Lo.counters);
  Q.stats = reinterpret_cast<unsigned long long *>(base + Lo.stats);
  for (int q = 0; q < 2; ++q)
    for (int k = 0; k < 3; ++k) Q.rq[q][k] = reinterpret_cast<float4 *>(base + Lo.rq[q][k]);
  for (int c = 0; c < 3; ++c)
    for (int k = 0; k < 4; ++k) Q.mq[c][k] = reinterpret_cast<float4 *>(base + Lo.mq[c][k]);
  Q.res = reinterpret_cast<float *>(base + Lo.res);
  job_dev *d_jobs = reinterpret_cast<job_dev *>(base + Lo.jobs);
  pass_hdr *d_hdr = reinterpret_cast<pass_hdr *>(base + Lo.hdr);
  pass_ent *d_pass = reinterpret_cast<pass_ent *>(base + Lo.pass);
  unsigned long long *d_totals = reinterpret_cast<unsigned long long *>(base + Lo.totals);

  job_dev jd;
  std::memset(&jd, 0, sizeof(jd));
  jd.cam = *camera;
  jd.width = desc->width;
  jd.height = desc->height;
  jd.n_pix = static_cast<uint32_t>(npix);
  jd.max_depth = desc->max_depth;
  jd.seed = desc->seed;
  jd.background = desc->background;
  jd.scene = 0;
  jd.sum = d_sum;
  jd.sumsq = d_sumsq;

  bounce_launch B;
  std::memset(&B, 0, sizeof(B));
  B.S = s->view(0);
  B.Q = Q;
  B.grid = grid;
  B.guide_grid = grid;
  B.sweep = s->sweep[0];
  B.sph = s->sph[0];
  B.simple = s->simple[0];
  B.multi = false;
  B.guided = false;
  B.tail = !(desc->flags & TPT_RENDER_NO_TAIL) && B.simple && (B.sweep || B.sph);
  render_params &P = B.P;
  P.jobs = d_jobs;
  P.scenes = s->d_scenes;
  P.hdr = d_hdr;
  P.pass = d_pass;
  P.guide_verts = ~0u;
  P.motion = s->motion[0];
  P.media = s->media[0];
  P.media_scatter = (desc->flags & TPT_RENDER_MEDIA_SCATTER) ? 1u : 0u;

  const uint32_t layers = static_cast<uint32_t>(TPT_NOALLOC_PASS_ITEMS / npix);
  const uint32_t first_items = static_cast<uint32_t>(std::min<uint64_t>(layers, desc->spp) * npix);
  uint32_t cap = desc->queue_capacity ? desc->queue_capacity : TPT_NOALLOC_PATHS;
  cap = std::min(cap, first_items);
  const uint64_t tail_live = std::min<uint64_t>(cap / 4, 1u << 18);

  // live records of ray queue p (the scene's pinned count buffer)
  auto live_count = [&](uint32_t p, uint64_t *live) -> int {
    HIP_TRY(hipMemcpyAsync(s->host_count, Q.counters + rc_at(p, 0), NSUB * CSTRIDE * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, ws));
    HIP_TRY(hipStreamSynchronize(ws));
    *live = 0;
    for (uint32_t k = 0; k < NSUB; ++k) *live += s->host_count[k * CSTRIDE];
    return TPT_OK;
  };
  uint64_t n_tail = 0;
  hipLaunchKernelGGL(k_zero_stats, dim3(64), dim3(256), 0, ws, Q.stats, kNoallocMaxGrid * ST_COUNT);
  bool drained = true;
  const uint32_t s_end = desc->sample_offset + desc->spp;
  for (uint32_t s0 = desc->sample_offset; s0 < s_end && drained;) {
    const uint32_t take = std::min(layers, s_end - s0);
    pass_hdr h{};
    h.n_items = static_cast<uint32_t>(take * npix);
    h.n_first = std::min(cap, h.n_items);
    h.n_ent = 1;
    h.n_acc = static_cast<uint32_t>(npix);
    pass_ent e{};
    e.item_begin = 0;
    e.job = 0;
    e.s_begin = s0;
    e.n_layers = take;
    e.acc_begin = 0;
    hipLaunchKernelGGL(k_noalloc_pass, dim3(1), dim3(64), 0, ws, d_jobs, jd, d_hdr, h, d_pass, e);
    hipLaunchKernelGGL(k_init, dim3((COUNTER_WORDS / CSTRIDE + 255) / 256), dim3(256), 0, ws, P, Q);
    const uint64_t max_iter = (static_cast<uint64_t>(h.n_items) / cap + 2) * (desc->max_depth + 2ull) + 64;
    drained = false;
    bool exhausted = h.n_items <= cap;
    uint64_t next_poll = (B.tail ? kPollFirst : kPoll) - 1;
    for (uint64_t it = 0; it < max_iter; ++it) {
      const uint32_t parity = static_cast<uint32_t>(it & 1u);
      launch_trace(B, ws, parity);
      launch_shade(B, ws, parity);
      if (it != next_poll) continue;
      uint64_t live = 0;
      if ((rc = live_count(parity ^ 1u, &live)) != TPT_OK) return rc;
      if (live == 0) {
        drained = true;
        break;
      }
      exhausted = exhausted || live < cap;
      if (B.tail && exhausted && live <= tail_live) {
        const uint32_t blocks =
            static_cast<uint32_t>(std::max<uint64_t>(NSUB, std::min<uint64_t>(grid, (live + 255) / 256)));
        launch_tail(B, ws, parity ^ 1u, blocks);
        ++n_tail;
        drained = true;
        break;
      }
      next_poll = it + (B.tail && exhausted ? kPollTail : kPoll);
    }
    if (!drained) break;
    hipLaunchKernelGGL(k_accumulate<false>, dim3(std::min<uint32_t>((h.n_acc + 255) / 256, grid)), dim3(256), 0, ws,
                       P, static_cast<const float *>(Q.res));
    s0 += take;
  }
  hipLaunchKernelGGL(k_noalloc_stats, dim3(1), dim3(256), 0, ws, static_cast<const unsigned long long *>(Q.stats),
                     grid, d_totals);
  HIP_TRY(hipGetLastError());
  // the totals come back through the scene's pinned count buffer (8 KB)
  static_assert(ST_COUNT * sizeof(unsigned long long) <= NSUB * CSTRIDE * sizeof(uint32_t), "pinned buffer size");
  HIP_TRY(hipMemcpyAsync(s->host_count, d_totals, ST_COUNT * sizeof(unsigned long long), hipMemcpyDeviceToHost, ws));
  HIP_TRY(hipStreamSynchronize(ws));
  if (!drained) return tpt::fail(TPT_ERR_HIP, "tpt_render_noalloc: bounce loop did not drain");
  if (stats) {
    unsigned long long tot[ST_COUNT];
    std::memcpy(tot, s->host_count, sizeof(tot));
    stats->paths = tot[ST_PATHS];
    stats->segments = tot[ST_SEGMENTS];
    stats->light_queries = tot[ST_LIGHT];
    stats->nan_samples = tot[ST_NAN];
    stats->nodes_visited = tot[ST_NODES];
    stats->tail_segments = tot[ST_TAIL];
    stats->tail_launches = n_tail;
    stats->kernel_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_begin).count();
  }
  return TPT_OK;
}

}  // namespace

extern "C" {

int tpt_render_noalloc(tpt_scene *s, const tpt_camera *camera, const tpt_render_desc *desc, float *d_sum,
                       float *d_sumsq, tpt_render_stats *stats, void *stream) {
  return render_noalloc(s, camera, desc, d_sum, d_sumsq, stats, stream);
}

int tpt_render(tpt_scene *s, const tpt_camera *camera, const tpt_render_desc *desc, float *d_sum, float *d_sumsq,
               tpt_render_stats *stats, void *stream) {
  return render_single(s, nullptr, camera, desc, d_sum, d_sumsq, stats, stream);
}

int tpt_render_guided(tpt_scene *s, tpt_guide *guide, const tpt_camera *camera, const tpt_render_desc *desc,
                      float *d_sum, float *d_sumsq, tpt_render_stats *stats, void *stream) {
  if (!guide) return tpt::fail(TPT_ERR_INVALID, "tpt_render_guided: null guide");
  return render_single(s, guide, camera, desc, d_sum, d_sumsq, stats, stream);
}

int tpt_render_jobs(tpt_scene *s, const tpt_job *jobs, uint32_t n_jobs, const tpt_launch *launch, tpt_guide *guide,
                    tpt_render_stats *stats, void *stream) {
  tpt_launch L;
  std::memset(&L, 0, sizeof(L));
  if (launch) L = *launch;
  return render_core(s, jobs, n_jobs, L, guide, stats, stream);
}

int tpt_autotune_plan(const tpt_job *jobs, uint32_t n_jobs, const tpt_launch *base, uint32_t calib_spp,
                      uint32_t caps[TPT_AUTOTUNE_MAX_CAPS], uint32_t *n_caps, uint32_t *calib_spp_out) {
  if (!jobs || n_jobs == 0 || !caps || !n_caps || !calib_spp_out)
    return tpt::fail(TPT_ERR_INVALID, "tpt_autotune_plan: bad argument");
  for (uint32_t j = 0; j < n_jobs; ++j)
    if (jobs[j].width == 0 || jobs[j].height == 0 ||
        static_cast<uint64_t>(jobs[j].width) * jobs[j].height > kPassItemsRecording)
      return tpt::fail(TPT_ERR_INVALID, "tpt_autotune_plan: bad image size");
  auto largest_pass = [&](const std::vector<tpt_job> &js, uint64_t pass_items) {
    uint64_t m = 0;
    for (const pass_plan &pp : plan_passes(js.data(), static_cast<uint32_t>(js.size()), pass_items))
      m = std::max<uint64_t>(m, pp.n_items);
    return m;
  };
  const std::vector<tpt_job> real(jobs, jobs + n_jobs);
  const bool recording = base && base->guiding;
  const uint64_t real_max = largest_pass(real, recording ? kPassItemsRecording : kPassItems);
  uint32_t n = 0;
  for (uint32_t c = 1u << 24; c <= (1u << 29); c <<= 1) {
    const uint32_t eff = static_cast<uint32_t>(std::min<uint64_t>(c, real_max));
    bool dup = false;
    for (uint32_t k = 0; k < n; ++k) dup = dup || caps[k] == eff;
    if (!dup) caps[n++] = eff;
  }
  *n_caps = n;
  uint32_t max_spp = 0;
  for (uint32_t j = 0; j < n_jobs; ++j) max_spp = std::max(max_spp, jobs[j].spp);
  const uint64_t want = 2ull * caps[n - 1];
  uint32_t k = calib_spp ? calib_spp : 2;
  std::vector<tpt_job> cj(real);
  for (;;) {
    for (uint32_t j = 0; j < n_jobs; ++j) cj[j].spp = std::min(jobs[j].spp, k);
    if (k >= max_spp || largest_pass(cj, kPassItems) >= want) break;
    k = k > 0x7FFFFFFFu ? max_spp : 2 * k;
  }
  *calib_spp_out = std::min(k, max_spp);
  return TPT_OK;
}

int tpt_autotune(tpt_scene *s, const tpt_job *jobs, uint32_t n_jobs, tpt_guide *guide, uint32_t calib_spp,
                 tpt_launch *best, float *best_ms) {
  if (!s || !jobs || n_jobs == 0 || !best) return tpt::fail(TPT_ERR_INVALID, "tpt_autotune: bad argument");
  uint32_t caps[TPT_AUTOTUNE_MAX_CAPS], n_caps = 0, cs_ = 0;
  int rc = tpt_autotune_plan(jobs, n_jobs, best, calib_spp, caps, &n_caps, &cs_);
  if (rc != TPT_OK) return rc;
  {
    // Candidates are given to the trials explicitly, past the render's own
    // memory check: a capacity whose render state for the real render would
    // not fit in 70% of the device memory free now (with what this scene's
    // state already holds) is not timed.  The smallest is always kept.
    uint64_t real_max = 0;
    for (const pass_plan &pp : plan_passes(jobs, n_jobs, best->guiding ? kPassItemsRecording : kPassItems))
      real_max = std::max<uint64_t>(real_max, pp.n_items);
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    const size_t budget = (free_b + s->state_bytes) / 10 * 7;
    const uint32_t g1 = static_cast<uint32_t>(s->n_cu) * 64u;                  // the largest grid tried
    const uint32_t g2 = guide ? static_cast<uint32_t>(s->n_cu) * 3u : 0u;  // k_guide_shade's largest grid
    uint32_t kept = 0;
    for (uint32_t k = 0; k < n_caps; ++k)
      if (kept == 0 || state_bytes_for(caps[k], real_max, g1, g2) <= budget) caps[kept++] = caps[k];
    n_caps = kept;
  }
  // calibration jobs: at most cs_ samples each, into scratch sums
  std::vector<tpt_job> cj(jobs, jobs + n_jobs);
  std::vector<float *> scratch(n_jobs, nullptr);
  auto release = [&]() {
    for (float *p : scratch) (void)hipFree(p);
  };
  for (uint32_t j = 0; j < n_jobs; ++j) {
    if (hipMalloc(&scratch[j], 3ull * jobs[j].width * jobs[j].height * sizeof(float)) != hipSuccess) {
      release();
      return tpt::fail(TPT_ERR_NOMEM, "tpt_autotune: scratch allocation failed");
    }
    cj[j].spp = std::min(jobs[j].spp, cs_);
    cj[j].d_sum = scratch[j];
    cj[j].d_sumsq = nullptr;
  }
  tpt_launch base = *best;
  base.flags = 0;
  base.guiding = 0;  // calibration renders do not record training rows
  float best_t = 3.0e38f;
  tpt_launch pick = base;
  auto trial = [&](uint32_t cap, uint32_t bpc, uint32_t graph) -> int {
    tpt_launch L = base;
    L.queue_capacity = cap;
    L.blocks_per_cu = bpc;
    L.graph_iters = graph;
    float t = 0.0f;
    for (int r = 0; r < 2; ++r) {  // the first run warms up (state growth, graph capture)
      const auto t0 = std::chrono::steady_clock::now();
      const int e = render_core(s, cj.data(), n_jobs, L, guide, nullptr, nullptr);
      if (e != TPT_OK) return e;
      t = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (t < best_t) {
      best_t = t;
      pick = L;
    }
    return TPT_OK;
  };
  for (uint32_t k = 0; k < n_caps && rc == TPT_OK; ++k) rc = trial(caps[k], 64, 0);
  const uint32_t cap_pick = pick.queue_capacity;
  for (uint32_t b : {16u, 32u})
    if (rc == TPT_OK) rc = trial(cap_pick, b, 0);
  if (rc == TPT_OK) rc = trial(cap_pick, pick.blocks_per_cu, 8);
  release();
  if (rc != TPT_OK) return rc;
  pick.flags = best->flags;
  pick.guiding = best->guiding;
  pick.guide_vertices = best->guide_vertices;
  *best = pick;
  if (best_ms) *best_ms = best_t;
  return TPT_OK;
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

