// Wavefront renderer kernels (replace color() + the sample loop of the
// reference: src/utils.cc:58-94, main.cpp:115-142).
//
// Work: a render call covers one or more jobs (an image of one scene of the
// scene pool: camera, resolution, samples, seed).  Its work items are the
// jobs' (sample, pixel) pairs, job after job, sample-major within a job.  A
// pass covers a run of whole sample layers of at most 2^31 items (2^26 while
// training rows are recorded); the pass table maps pass item w to its job,
// pixel and sample.  A fixed number of path records is in flight; a finished
// path writes its de_nan'ed sample to res[w] and claims the next item from a
// shared cursor, so the GPU stays full until the pass runs out of work
// whatever the mix of jobs.  At the end of a
// pass k_accumulate adds res to each job's per-pixel sums in sample order, so
// the result does not depend on scheduling or on how jobs share a pass.
//
// Queues hold path records by value, as planes of float4 (one 16-byte plane
// per field group), so every queue access is a coalesced dwordx4:
//   ray queue (ping-pong), 3 planes: (o, item) (d, depth | pass entry)
//                                    (throughput, diffuse count | NaN bits)
//   material queues (diffuse / other / terminal), the 3 ray planes plus
//                                    (t, slot, entry, material)
// Appends use a 64-lane ballot + popcount prefix and one atomic per wave, on
// one of 64 sub-queue counters (see NSUB).
//
// MULTI (template flag): the launch spans several jobs, so each lane looks up
// its own job and scene; without it every lane belongs to job 0 and those
// lookups are wave-uniform.
#ifndef TPT_CSRC_RENDER_H
#define TPT_CSRC_RENDER_H

#include "guide.h"
#include "shade.h"

#pragma clang fp contract(off)

namespace tptd {

// Material queues: lambertian hits (the guided ones in guided mode), every other
// scattering hit, and hits that end the path (emitters, depth limit).
enum { Q_DIFFUSE = 0, Q_OTHER = 1, Q_TERMINAL = 2 };
// Per-block statistics: the first ST_BASIC counters in every launch, the
// walk's operations by kind (op_counts) in the counting variant of k_trace.
enum {
  ST_SEGMENTS = 0, ST_LIGHT = 1, ST_NAN = 2, ST_PATHS = 3, ST_NODES = 4, ST_TAIL = 5, ST_BASIC = 8,
  ST_BOX = 8, ST_SPHERE = 9, ST_RECT = 10, ST_XFORM = 11, ST_COUNT = 12
};

// Every queue is split into NSUB regions, each with its own counter 128 bytes
// from the next: block b appends to region b % NSUB.  Wave atomics on one
// address serialise (about 11 ns each on MI355X); spread over 64 counters they
// cost about 0.3 ns, so appends stop limiting the bounce loop.  Consumers see
// the concatenation of the regions through a prefix sum kept in LDS.
constexpr uint32_t NSUB = 64;
constexpr uint32_t CSTRIDE = 32;  // words between counters
// counter word offsets
__host__ __device__ constexpr uint32_t rc_at(uint32_t parity, uint32_t k) { return (parity * NSUB + k) * CSTRIDE; }
__host__ __device__ constexpr uint32_t mc_at(uint32_t parity, uint32_t c, uint32_t k) {
  return (2 * NSUB + (parity * 3 + c) * NSUB + k) * CSTRIDE;
}
__host__ __device__ constexpr uint32_t cur_at(uint32_t k) { return (8 * NSUB + k) * CSTRIDE; }
// Nonzero from k_init until the pass's first k_shade: the first k_trace of a
// pass generates the camera rays of queue 0 itself (k_init writes none).
__host__ __device__ constexpr uint32_t first_at() { return 9 * NSUB * CSTRIDE; }
constexpr uint32_t COUNTER_WORDS = 10 * NSUB * CSTRIDE;

struct render_state {
  float4 *rq[2][3];    // ray queues by parity, 3 planes (path record); NSUB regions of `region` entries
  float4 *mq[3][4];    // material queues: the path record + the hit, same regions
  uint32_t region;     // entries per region
  uint32_t *counters;  // COUNTER_WORDS
  float *res;          // 3 floats per item of the pass
  unsigned long long *stats;  // ST_COUNT per block
};

// One image of one scene (device copy of a tpt_job).
struct job_dev {
  tpt_camera cam;
  uint32_t width, height, n_pix, max_depth;
  uint64_t seed;
  float background;
  uint32_t scene;
  float *sum, *sumsq;
};

// A scene of the pool: its segments and lights in the pooled arrays, and its
// box (position encoding of the guide).
struct scene_dev {
  uint32_t seg_begin, n_seg, light_begin, n_lights, light_mode, pad[3];
  float bounds[6], walk_ulp, pad2;  // walk_ulp: scene_view::walk_ulp of this scene
};

// Pass table: items [item_begin, next entry's item_begin) of the pass are
// samples [s_begin, s_begin + n_layers) of job `job`, sample-major;
// acc_begin is the prefix of the jobs' pixel counts (k_accumulate's index).
struct pass_ent {
  uint32_t item_begin, job, s_begin, n_layers, acc_begin, pad[3];
};
struct pass_hdr {
  uint32_t n_items;  // items of the pass
  uint32_t n_first;  // items [0, n_first) seed the first ray queue
  uint32_t n_ent;    // pass table entries
  uint32_t n_acc;    // sum of the entries' pixel counts
};

// Kernel parameters: constant over a render call (per-pass values live in the
// device-side pass header and table, so a captured bounce loop can be
// replayed in every pass).
struct render_params {
  const job_dev *jobs;
  const scene_dev *scenes;
  const pass_hdr *hdr;
  const pass_ent *pass;
  // guided mode
  const uint32_t *g_tiles;  // bf16 inference weight tiles (16x16x32 layout)
  const float *g_params;    // fp32 parameters (biases, mix logit)
  uint32_t guide_verts;     // diffuse vertices per path that use the guide (~0u: every one, plain mode)
  uint32_t rec_stride;      // record training rows from items with item % rec_stride == 0 (0: off)
  float4 *rec_side;         // per recorded path: 2 vertices x 5 float4
  float4 *rec_rows;         // training rows: 2 per recorded path, 4 float4 each
  uint32_t motion;          // some scene has moving spheres: rays need their time
  uint32_t media;           // some scene has constant_medium: queries need their media draws
  uint32_t media_scatter;   // TPT_RENDER_MEDIA_SCATTER
};

constexpr uint32_t REC_VERTS = 2;

struct path {
  v3 o, d, thr, rad;
  uint32_t item, depth;
  uint32_t ent;    // pass table entry of the item (its job)
  uint32_t ndiff;  // diffuse (lambertian) vertices so far
};

// Where a pass item comes from.
struct item_ctx {
  uint32_t ent, job, px, smp;
};

template <bool MULTI>
__device__ __forceinline__ uint32_t find_entry(const render_params &P, uint32_t n_ent, uint32_t item) {
  if (!MULTI) return 0;
  uint32_t lo = 0, hi = n_ent;  // invariant: pass[lo].item_begin <= item < pass[hi].item_begin
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (P.pass[mid].item_begin <= item) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ item_ctx item_in_entry(const render_params &P, uint32_t ent, uint32_t item) {
  const pass_ent e = P.pass[ent];
  const uint32_t n_pix = P.jobs[e.job].n_pix;
  const uint32_t local = item - e.item_begin;
  return item_ctx{ent, e.job, local % n_pix, e.s_begin + local / n_pix};
}

// The scene view of a job: the pooled arrays with that scene's segment and
// light ranges (without MULTI the caller's view already is job 0's scene).
template <bool MULTI>
__device__ __forceinline__ scene_view view_of(const scene_view &S, const render_params &P, uint32_t job) {
  if (!MULTI) return S;
  const scene_dev &sc = P.scenes[P.jobs[job].scene];
  scene_view v = S;
  v.segments = S.segments + 2 * sc.seg_begin;
  v.seg_xf = S.seg_xf ? S.seg_xf + sc.seg_begin : nullptr;
  v.n_segments = sc.n_seg;
  v.lights = S.lights + sc.light_begin;
  v.n_lights = sc.n_lights;
  v.light_mode = sc.light_mode;
  v.walk_ulp = sc.walk_ulp;
  return v;
}

__device__ __forceinline__ uint32_t lane_prefix(unsigned long long mask) {
  return __popcll(mask & ((1ull << __lane_id()) - 1ull));
}

// One atomic per wave; every lane of the wave must call it.
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool pred) {
  const unsigned long long mask = __ballot(pred);
  const uint32_t total = __popcll(mask);
  uint32_t base = 0;
  if (total) {
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;  // wave-uniform
    if (static_cast<int>(__lane_id()) == leader) base = atomicAdd(counter, total);
    base = __builtin_amdgcn_readlane(base, leader);
  }
  return base + lane_prefix(mask);
}

// Items [n_first, n_items) are split into NSUB contiguous shards, one cursor each.
__device__ __forceinline__ uint32_t shard_lo(const pass_hdr &H, uint32_t k) {
  return H.n_first + static_cast<uint32_t>((static_cast<uint64_t>(H.n_items - H.n_first) * k) / NSUB);
}

// Exclusive prefix over n (<= 4 * 64) region counts into pre[0..n]; one wave
// of the block computes it.  Ends with a block barrier.
__device__ __forceinline__ void region_prefix(const uint32_t *counters, const uint32_t *offsets, uint32_t n,
                                              uint32_t *pre) {
  if (threadIdx.x < 64) {
    uint32_t v[4] = {0, 0, 0, 0};
    uint32_t local = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t idx = threadIdx.x * 4 + j;
      v[j] = idx < n ? counters[offsets[idx]] : 0u;
      local += v[j];
    }
    uint32_t incl = local;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t o = __shfl_up(incl, off);
      if (__lane_id() >= static_cast<uint32_t>(off)) incl += o;
    }
    uint32_t run = incl - local;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t idx = threadIdx.x * 4 + j;
      if (idx <= n) pre[idx] = run;
      run += v[j];
    }
  }
  __syncthreads();
}

// Region r with pre[r] <= q < pre[r + 1] (pre has n + 1 entries, q < pre[n]).
__device__ __forceinline__ uint32_t find_region(const uint32_t *pre, uint32_t n, uint32_t q) {
  uint32_t lo = 0, hi = n;  // invariant: pre[lo] <= q < pre[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pre[mid] <= q) lo = mid; else hi = mid;
  }
  return lo;
}

// Lanes with `want` receive the next unclaimed work item (or >= n_items when
// the pass has none left).  One atomic per wave per shard tried; a wave starts
// at its own shard and moves on when that one runs dry.  All lanes call.
__device__ __forceinline__ uint32_t claim_item(const pass_hdr &H, uint32_t *counters, bool want) {
  unsigned long long need = __ballot(want);
  uint32_t item = 0xFFFFFFFFu;
  uint32_t k = (blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64) % NSUB;
  for (uint32_t attempt = 0; attempt < NSUB && need; ++attempt, k = (k + 1) % NSUB) {
    const uint32_t hi = shard_lo(H, k + 1);
    uint32_t *cur = counters + cur_at(k);
    if (__builtin_amdgcn_readfirstlane(__hip_atomic_load(cur, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >= hi)
      continue;
    const uint32_t cnt = __popcll(need);
    const int leader = __ffsll(static_cast<long long>(need)) - 1;  // wave-uniform
    uint32_t base = 0;
    if (static_cast<int>(__lane_id()) == leader) base = atomicAdd(cur, cnt);
    base = __builtin_amdgcn_readlane(base, leader);
    const uint32_t avail = base < hi ? (hi - base < cnt ? hi - base : cnt) : 0u;
    const bool mine = (need >> __lane_id()) & 1ull;
    const uint32_t rank = __popcll(need & ((1ull << __lane_id()) - 1ull));
    if (mine && rank < avail) item = base + rank;
    need &= ~__ballot(mine && rank < avail);
  }
  return item;
}

// Adds per-thread counts into this block's stats row (no global atomics).
__device__ __forceinline__ void block_stats(unsigned long long *stats, const uint32_t *v, int n) {
  __shared__ unsigned long long acc[ST_COUNT];
  if (threadIdx.x < ST_COUNT) acc[threadIdx.x] = 0;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    uint32_t x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if (__lane_id() == 0 && x) atomicAdd(&acc[k], static_cast<unsigned long long>(x));
  }
  __syncthreads();
  if (threadIdx.x < static_cast<uint32_t>(n)) stats[blockIdx.x * ST_COUNT + threadIdx.x] += acc[threadIdx.x];
}

// Path record, 48 bytes in 3 planes (depth and pass entry share a word: both
// are below 2^16).  A live path's radiance is 0 or NaN per
// channel: light is only added by the event that ends the path (an emitter or
// a miss), and the reference's `emitted + ...` form turns a channel NaN when a
// non-finite throughput meets a zero term.  So only its NaN bits are stored
// (bits 16..18 next to the diffuse-vertex count).
// Queue entries are read once.  Non-temporal loads of them leave L2 to the
// queues being filled, whose lines arrive in parts (each wave's run starts
// where the previous one ended) and are written to HBM twice when evicted
// half-filled.
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_once(const float4 *p) {
  const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

// Ray records are stored non-temporally: the next launch reads them long
// after L2 could hold them, and streaming them past L2 cut k_shade's HBM writes
// from 61 to 49 B per segment (7% faster).  The material-queue entries of
// k_trace keep plain stores (non-temporal ones wrote 6 B more per segment and
// saved no time).
__device__ __forceinline__ void st_once(float4 *p, float4 v) {
  __builtin_nontemporal_store(f4v{v.x, v.y, v.z, v.w}, reinterpret_cast<f4v *>(p));
}

// The record comes through a plain pointer.  Through a reference (known
// dereferenceable) the function's own simplification, before it is inlined,
// merges each plane's three field reads into one 16-byte load that straddles
// two fields of the caller's record; the caller's record then cannot be split
// into registers, and the compiler keeps it in LDS (10 KB per block) with an
// LDS round trip per field access.
__device__ __forceinline__ void store_ray(float4 *const *q, uint32_t at, const path *p) {
  const uint32_t nan_bits = (p->rad.x != p->rad.x ? 1u : 0u) | (p->rad.y != p->rad.y ? 2u : 0u) |
                            (p->rad.z != p->rad.z ? 4u : 0u);
  st_once(&q[0][at], make_float4(p->o.x, p->o.y, p->o.z, __uint_as_float(p->item)));
  st_once(&q[1][at], make_float4(p->d.x, p->d.y, p->d.z, __uint_as_float(p->depth | (p->ent << 16))));
  st_once(&q[2][at], make_float4(p->thr.x, p->thr.y, p->thr.z, __uint_as_float(p->ndiff | (nan_bits << 16))));
}

// ONCE: the entry's last read (non-temporal); k_guide_shade reads its
// entries a first time before the network passes.
template <bool ONCE = true>
__device__ __forceinline__ path load_ray(float4 *const *q, uint32_t at) {
  const float4 a = ONCE ? ld_once(&q[0][at]) : q[0][at], b = ONCE ? ld_once(&q[1][at]) : q[1][at],
               c = ONCE ? ld_once(&q[2][at]) : q[2][at];
  path p;
  p.o = v3{a.x, a.y, a.z};
  p.item = __float_as_uint(a.w);
  p.d = v3{b.x, b.y, b.z};
  p.depth = __float_as_uint(b.w) & 0xFFFFu;
  p.ent = __float_as_uint(b.w) >> 16;
  p.thr = v3{c.x, c.y, c.z};
  const uint32_t w = __float_as_uint(c.w);
  p.ndiff = w & 0xFFFFu;
  const float qnan = __uint_as_float(0x7FC00000u);
  p.rad = v3{(w & 0x10000u) ? qnan : 0.0f, (w & 0x20000u) ? qnan : 0.0f, (w & 0x40000u) ? qnan : 0.0f};
  return p;
}

__device__ __forceinline__ v3 cam_vec(const float *a) { return v3{a[0], a[1], a[2]}; }

// Camera path for work item `item` of job ctx.job: jitter (main.cpp:121-122,
// in double as there), then camera_with_blur::get_ray (src/camera.cc:23-31).
// Bounce 0 of the (pixel, sample) stream.
__device__ __forceinline__ path raygen(const render_params &P, uint32_t item, const item_ctx &c) {
  const job_dev &J = P.jobs[c.job];
  stream R;
  R.init(c.px, c.smp, 0, J.seed);
  const uint32_t i = c.px % J.width, j = c.px / J.width;
  const float u = static_cast<float>((static_cast<double>(static_cast<float>(i)) + static_cast<double>(R.next())) /
                                     static_cast<double>(static_cast<float>(J.width)));
  const float v = static_cast<float>((static_cast<double>(static_cast<float>(j)) + static_cast<double>(R.next())) /
                                     static_cast<double>(static_cast<float>(J.height)));
  const tpt_camera &C = J.cam;
  const v3 disk = in_unit_disk(R);
  const v3 rd = disk * C.lens_radius;
  const v3 off = cam_vec(C.u) * rd.x + cam_vec(C.v) * rd.y;
  (void)R.next();  // the ray time draw; ray_time() replays the stream up to it
  const v3 org = cam_vec(C.origin);
  path p;
  p.o = org + off;
  p.d = cam_vec(C.lower_left) + u * cam_vec(C.horizontal) + v * cam_vec(C.vertical) - org - off;
  p.thr = v3{1.0f, 1.0f, 1.0f};
  p.rad = v3{0.0f, 0.0f, 0.0f};
  p.item = item;
  p.depth = 0;
  p.ent = c.ent;
  p.ndiff = 0;
  return p;
}

// The time of the camera ray of item c (camera_with_blur::get_ray,
// src/camera.cc:26: time0 + drand * (time1 - time0), in double as there).
// Every ray of a path keeps it (ray(p, dir, r.time()) in color() and the
// materials), so it is recomputed from the item's bounce-0 stream instead of
// being stored in the path record; only scenes with moving spheres need it.
__device__ __forceinline__ float ray_time(const render_params &P, const item_ctx &c) {
  const job_dev &J = P.jobs[c.job];
  stream R;
  R.init(c.px, c.smp, 0, J.seed);
  (void)R.next();
  (void)R.next();
  (void)in_unit_disk(R);
  const float xi = R.next();
  return static_cast<float>(static_cast<double>(J.cam.time0) +
                            static_cast<double>(xi) * static_cast<double>(J.cam.time1 - J.cam.time0));
}

// Closest-hit context of a ray traced at `depth` on item c: its time, and its
// media draw from the stream of the bounce that shades the hit (counter
// (pixel, sample, depth + 1, 2^20 + leaf), key = the job's seed).
__device__ __forceinline__ ray_ctx make_ray_ctx(const render_params &P, const item_ctx &c, uint32_t depth) {
  ray_ctx rc{0.0f, 0u, 0u, 0u, 0u, 0u};
  if (P.motion) rc.time = ray_time(P, c);
  if (P.media) {
    const uint64_t seed = P.jobs[c.job].seed;
    rc = ray_ctx{rc.time, c.px, c.smp, depth + 1, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
  }
  return rc;
}

template <bool MULTI>
__device__ __forceinline__ path raygen_item(const render_params &P, const pass_hdr &H, uint32_t item) {
  return raygen(P, item, item_in_entry(P, find_entry<MULTI>(P, H.n_ent, item), item));
}

__global__ void k_zero_stats(unsigned long long *stats, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) stats[i] = 0;
}

// Start of a pass: resets every counter.  Ray queue 0 holds the pass's first
// n_first items, entry q being item q; its region counts are set as if runs
// of 64 items had been appended region by region, and the first-trace flag
// tells the pass's first k_trace to generate those camera rays itself instead
// of reading them (the records of a 2^29-item fill would cost a 26 GB write
// and read).
__global__ void __launch_bounds__(256) k_init(render_params P, render_state Q) {
  const pass_hdr H = *P.hdr;
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t w = gid; w < COUNTER_WORDS / CSTRIDE; w += stride) {
    const uint32_t set = w / NSUB, k = w % NSUB;
    uint32_t v = 0;
    if (set == 0) {
      // ray queue 0 region counts: the full runs of region k, and the last,
      // partial run if it falls there
      const uint32_t full = H.n_first / 64, part = H.n_first % 64;
      v = 64 * (full / NSUB + (k < full % NSUB ? 1u : 0u)) + (part && full % NSUB == k ? part : 0u);
    }
    if (set == 8) v = shard_lo(H, k);  // shard cursors
    if (set == 9) v = k == 0 ? 1u : 0u;  // first-trace flag
    Q.counters[w * CSTRIDE] = v;
  }
}

// SWEEP: wave-uniform traversal for small scenes (see segment_sweep).  SPH:
// every leaf of the scene (or pool) is a sphere; without the rect and list
// code the kernel fits 64 registers, 8 waves per SIMD.  MEDIA: the scenes hold
// constant_media (their code costs registers, so only this variant has it).  The slot/entry arrays come
// in as separate restrict arguments so that wave-uniform loads from them can
// use the scalar data path.  COUNT: also counts the walk's operations by kind
// (TPT_RENDER_COUNT_OPS; the other variants are built without that code).
// The sweep variant parks its path record in LDS during the walk (see below)
// and targets 6 waves per SIMD (80 VGPRs, 4 spilled): 4% faster on the Cornell
// frame than 5 waves with the record in registers, 2% faster than 7 waves.
constexpr uint32_t TRACE_THREADS = 256;  // k_trace's block size (its record stash holds one record per thread)
template <bool SWEEP, bool MULTI, bool SPH = false, bool MEDIA = false, bool COUNT = false>
__global__ void __launch_bounds__(TRACE_THREADS) __attribute__((amdgpu_waves_per_eu(SPH ? 8 : (MEDIA ? 1 : (SWEEP ? 6 : 5))))) k_trace(scene_view S, const tpt_slot *__restrict__ slots,
                                               const tpt_entry *__restrict__ entries, render_params P,
                                               render_state Q, uint32_t parity) {
  static_assert(!(SWEEP && MULTI), "the wave-uniform sweep needs one scene");
  static_assert(!(SWEEP && SPH), "the sweep takes every leaf kind");
  static_assert(!(MEDIA && (SWEEP || SPH)), "media go through the general walk");
  S.slots = slots;
  S.entries = entries;
  __shared__ uint32_t offs[NSUB];
  __shared__ uint32_t pre[NSUB + 1];
  if (threadIdx.x < NSUB) offs[threadIdx.x] = rc_at(parity, threadIdx.x);
  __syncthreads();
  region_prefix(Q.counters, offs, NSUB, pre);
  const uint32_t count = pre[NSUB];
  // the pass's first launch: entry q of queue 0 is the camera ray of item q
  const bool first = __builtin_amdgcn_readfirstlane(Q.counters[first_at()]) != 0u;
  // the next ray queue is filled by this iteration's k_shade
  if (blockIdx.x == 0 && threadIdx.x < NSUB) Q.counters[rc_at(parity ^ 1u, threadIdx.x)] = 0;
  float4 *const *in = Q.rq[parity];
  const uint32_t home = blockIdx.x % NSUB;
  uint32_t st[ST_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  const uint32_t stride = gridDim.x * blockDim.x;
  // The record read before the walk is stored here and read back after it:
  // only the ray stays in registers across the walk.  The compiler barrier
  // before the read-back keeps the compiler from forwarding the stored values
  // in registers.  Measured: Cornell +4% with the sweep at 6 waves;
  // sphere-only walk at 8 waves, 19 instead of 23 spilled VGPRs, batch64 +3%;
  // the general walk spills 2-11 instead of 6-13 VGPRs (no change in time on
  // box_list).
  __shared__ float4 stash[3 * TRACE_THREADS];
  for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride) {
    const uint32_t qi = base + threadIdx.x;
    const bool valid = qi < count;
    uint32_t cls = 3;
    // Without SWEEP only the ray and the (depth, entry) word stay live across
    // the traversal (fewer registers, more waves); the third plane is read
    // after it.  The sweep's waves are few and short of latency cover, so it
    // reads the whole record first (and parks it in LDS during the walk).
    uint32_t src = 0;
    float4 a{}, b{}, c0{};
    hit_t h{0.0f, kNone, kNone, 0};
    if (valid && first) {
      const path p = raygen_item<MULTI>(P, *P.hdr, qi);
      a = make_float4(p.o.x, p.o.y, p.o.z, __uint_as_float(p.item));
      b = make_float4(p.d.x, p.d.y, p.d.z, __uint_as_float(p.depth | (p.ent << 16)));
      c0 = make_float4(p.thr.x, p.thr.y, p.thr.z, __uint_as_float(p.ndiff));  // no NaN bits: radiance 0
    } else if (valid) {
      const uint32_t r = find_region(pre, NSUB, qi);
      src = r * Q.region + (qi - pre[r]);
      a = ld_once(&in[0][src]);
      b = ld_once(&in[1][src]);
      if (SWEEP) c0 = ld_once(&in[2][src]);
    }
    {
      stash[threadIdx.x] = a;
      stash[TRACE_THREADS + threadIdx.x] = b;
      stash[2 * TRACE_THREADS + threadIdx.x] = c0;
    }
    uint32_t depth = __float_as_uint(b.w) & 0xFFFFu;
    const uint32_t ent = MULTI ? __float_as_uint(b.w) >> 16 : 0u;
    const uint32_t job = MULTI ? P.pass[ent].job : 0u;
    const scene_view Sj = view_of<MULTI>(S, P, job);
    // (scenes with moving spheres or media use neither the sweep nor the
    // sphere-only walk)
    const ray_ctx rc = !SWEEP && !SPH && (P.motion | P.media) && valid
                           ? make_ray_ctx(P, item_in_entry(P, ent, __float_as_uint(a.w)), depth)
                           : ray_ctx{0.0f, 0u, 0u, 0u, 0u, 0u};
    op_counts oc;
    bool got;
    if (SPH) {
      // A wave of camera rays of one job walks nearly the same slots (about
      // the longest lane's walk in all): it sweeps the tree together, with
      // scalar slot loads and no divergence on the slot kind.
      const unsigned long long vm = __ballot(valid);
      const int first = vm ? __ffsll(static_cast<long long>(vm)) - 1 : 0;
      const uint32_t job_u = __builtin_amdgcn_readlane(job, first);
      const bool same = !valid || (depth == 0 && job == job_u);
      if (vm != 0ull && __ballot(!same) == 0ull)
        got = trace_world<true, SPH, false>(view_of<MULTI>(S, P, job_u), v3{a.x, a.y, a.z}, v3{b.x, b.y, b.z}, rc,
                                            0.001f, kFltMax, &h, &st[ST_NODES], valid, COUNT ? &oc : nullptr);
      else
        got = trace_world<false, SPH, MEDIA>(Sj, v3{a.x, a.y, a.z}, v3{b.x, b.y, b.z}, rc, 0.001f, kFltMax, &h,
                                             &st[ST_NODES], valid, COUNT ? &oc : nullptr);
    } else {
      got = trace_world<SWEEP, SPH, MEDIA>(Sj, v3{a.x, a.y, a.z}, v3{b.x, b.y, b.z}, rc, 0.001f, kFltMax, &h,
                                           &st[ST_NODES], valid, COUNT ? &oc : nullptr);
    }
    {
      asm volatile("" ::: "memory");
      a = stash[threadIdx.x];
      b = stash[TRACE_THREADS + threadIdx.x];
      c0 = stash[2 * TRACE_THREADS + threadIdx.x];
      depth = __float_as_uint(b.w) & 0xFFFFu;
    }
    if (COUNT) {
      st[ST_BOX] += oc.box;
      st[ST_SPHERE] += oc.sphere;
      st[ST_RECT] += oc.rect;
      st[ST_XFORM] += oc.xform;
    }
    const float4 c = SWEEP || !valid || first ? c0 : ld_once(&in[2][src]);
    if (valid) {
      ++st[ST_SEGMENTS];
      if (got) {
        h.mat &= TPT_PAYLOAD_MASK;
        const uint32_t kind = S.materials[h.mat].kind;
        if (kind == TPT_MAT_DIFFUSE_LIGHT || depth >= P.jobs[job].max_depth) {
          cls = Q_TERMINAL;
        } else {
          // In guided mode only the first guide_verts diffuse vertices of a
          // path go to the diffuse queue (where the network runs); later
          // ones are shaded with the plain sampler from the other queue.
          const uint32_t ndiff = __float_as_uint(c.w) & 0xFFFFu;
          cls = kind == TPT_MAT_LAMBERTIAN && ndiff < P.guide_verts ? Q_DIFFUSE : Q_OTHER;
        }
      } else {
        h.slot = kNone;
        cls = Q_TERMINAL;
      }
    }
#pragma unroll
    for (uint32_t k = 0; k < 3; ++k) {
      const uint32_t at = home * Q.region + wave_append(Q.counters + mc_at(parity, k, home), cls == k);
      if (cls == k) {
        Q.mq[k][0][at] = a;
        Q.mq[k][1][at] = b;
        Q.mq[k][2][at] = c;
        Q.mq[k][3][at] = make_float4(h.t, __uint_as_float(h.slot), __uint_as_float(h.entry), __uint_as_float(h.mat));
      }
    }
  }
  block_stats(Q.stats, st, COUNT ? ST_COUNT : ST_BASIC);
}

__device__ __forceinline__ float luminance(v3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// Training rows of a finished recorded path: c_k = luminance((R - R_k) / T'_k)
// with T'_k the throughput times albedo at vertex k (so c is the sample's
// contribution over its proposal density, without albedo).
__device__ __forceinline__ void emit_rows(const render_params &P, const path &p) {
  const uint32_t slot = p.item / P.rec_stride;
  for (uint32_t k = 0; k < REC_VERTS; ++k) {
    float4 *row = P.rec_rows + (static_cast<size_t>(slot) * REC_VERTS + k) * 4;
    if (k < min(p.ndiff, min(REC_VERTS, P.guide_verts))) {
      const float4 *sd = P.rec_side + (static_cast<size_t>(slot) * REC_VERTS + k) * 5;
      const float4 a = sd[0], b = sd[1], c = sd[2], d = sd[3], e = sd[4];
      // side: (p̂, n.x) (n.y, n.z, wo.x, wo.y) (wo.z, ω) (q, R) (T', -)
      const float Rk[3] = {d.y, d.z, d.w};
      const float Tk[3] = {e.x, e.y, e.z};
      const float Rf[3] = {p.rad.x, p.rad.y, p.rad.z};
      float cc[3];
      for (int ch = 0; ch < 3; ++ch) cc[ch] = Tk[ch] > 0.0f ? (Rf[ch] - Rk[ch]) / Tk[ch] : 0.0f;
      float cval = 0.2126f * cc[0] + 0.7152f * cc[1] + 0.0722f * cc[2];
      if (!(cval > 0.0f) || !(cval < 1e30f)) cval = 0.0f;
      row[0] = make_float4(a.x, a.y, a.z, a.w);
      row[1] = make_float4(b.x, b.y, b.z, b.w);
      row[2] = make_float4(c.x, c.y, c.z, c.w);
      row[3] = make_float4(cval, d.x, 0.0f, 0.0f);
    } else {
      row[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      row[0] = row[1] = row[2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
  }
}

// A finished path: de_nan (headers/utils.h:100-109), then the sample goes to
// res; in guided mode a recorded path also writes its training rows.
template <bool GUIDED>
__device__ __forceinline__ void finish_sample(const render_params &P, const render_state &Q, const path &p,
                                              uint32_t *st) {
  float c[3] = {p.rad.x, p.rad.y, p.rad.z};
  bool any_nan = false;
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if (c[k] != c[k]) {
      c[k] = 0.0f;
      any_nan = true;
    }
  st[ST_NAN] += any_nan ? 1u : 0u;
  ++st[ST_PATHS];
  float *r = Q.res + 3ull * p.item;
  r[0] = c[0], r[1] = c[1], r[2] = c[2];
  if (GUIDED && P.rec_stride && (p.item % P.rec_stride) == 0) emit_rows(P, p);
}

// Finished lanes claim the next work item; every continuing path goes to the
// next ray queue.  All lanes call.  (The record comes through a plain pointer,
// for the reason given at store_ray.)
template <bool MULTI>
__device__ __forceinline__ void claim_and_append(const render_params &P, const pass_hdr &H, const render_state &Q,
                                                 uint32_t *next_count, float4 *const *out, uint32_t home, path *p,
                                                 bool cont, bool finished) {
  const uint32_t item = claim_item(H, Q.counters, finished);
  if (finished && item < H.n_items) {
    *p = raygen_item<MULTI>(P, H, item);
    cont = true;
  }
  const uint32_t at = home * Q.region + wave_append(next_count, cont);
  if (cont) store_ray(out, at, p);
}

// The same with the record by reference, for k_shade.  Through the reference
// the record's three plane reads are merged into 16-byte loads (see
// store_ray), which keeps the first 40 bytes of k_shade's record (origin,
// direction, throughput, one radiance channel) in LDS instead of registers.
// The single-scene SIMPLE variant then fits 64 VGPRs (8 waves per SIMD)
// without spilling; with the record in registers it spills 6 VGPRs to scratch
// and measured 1.5-2.6% slower on the Cornell frame (2-3% slower at a 7-wave
// target, 1.5-3% with the record in an explicit LDS array).
template <bool MULTI>
__device__ __forceinline__ void claim_and_append_ref(const render_params &P, const pass_hdr &H, const render_state &Q,
                                                     uint32_t *next_count, float4 *const *out, uint32_t home, path &p,
                                                     bool cont, bool finished) {
  claim_and_append<MULTI>(P, H, Q, next_count, out, home, &p, cont, finished);
}

// One queued hit (or miss) of class cls with the plain sampler: the
// reference's color() step for it.  Updates p (radiance, throughput, next
// ray); true when the path goes on.  Shared by k_shade and k_tail.
template <bool SIMPLE>
__device__ __forceinline__ bool shade_hit(const scene_view &S, const scene_view &Sj, const render_params &P,
                                          const job_dev &J, const item_ctx &c, path &p, const hit_t &h, uint32_t cls,
                                          uint32_t *st) {
  bool cont = false;
  if (h.slot == kNone) {
    // miss: black (src/utils.cc:85-86) or the sky extension
    const float bgs = J.background;
    const v3 bg = bgs > 0.0f ? sky(p.d, bgs) : v3{0.0f, 0.0f, 0.0f};
    p.rad = p.rad + p.thr * bg;
  } else {
    v3 pt, n;
    float u, v;
    // (the job's view: a wrapped segment is looked up among its scene's segments only)
    hit_record(Sj, h, p.o, p.d,
               !SIMPLE && (P.motion | P.media) ? make_ray_ctx(P, c, p.depth) : ray_ctx{0.0f, 0u, 0u, 0u, 0u, 0u},
               &pt, &n, &u, &v);
    const tpt_material M = S.materials[h.mat];
    stream R;
    R.init(c.px, c.smp, p.depth + 1, J.seed);
    v3 nd{0.0f, 0.0f, 0.0f};
    if (cls == Q_TERMINAL) {
      // diffuse_light::emitted is one-sided (src/material.cc:79-86);
      // every other material emits nothing.
      v3 e{0.0f, 0.0f, 0.0f};
      if (M.kind == TPT_MAT_DIFFUSE_LIGHT && dot(n, p.d) < 0.0f) e = tex_value<SIMPLE>(S, M.tex, u, v, pt);
      p.rad = p.rad + p.thr * e;
    } else if (M.kind == TPT_MAT_METAL) {
      // metal::scatter (src/material.cc:88-98); specular: no emitted term.
      const v3 refl = reflect(unit(p.d), n);
      nd = refl + M.p[3] * in_unit_sphere(R);
      if (dot(nd, n) > 0.0f) {
        p.thr = p.thr * v3{M.p[0], M.p[1], M.p[2]};
        cont = true;
      } else {
        p.rad = p.rad + p.thr * v3{0.0f, 0.0f, 0.0f};
      }
    } else if (M.kind == TPT_MAT_ISOTROPIC) {
      // constant_medium's phase function.  The reference's
      // isotropic::scatter does not override material::scatter
      // (headers/material.h:74-80), so color() returns emitted (0): the
      // path ends black.  TPT_RENDER_MEDIA_SCATTER scatters as the book
      // intends (src/material.cc:72-77): random_in_unit_sphere, albedo.
      if (!SIMPLE && P.media_scatter) {
        nd = in_unit_sphere(R);
        p.thr = p.thr * tex_value<SIMPLE>(S, M.tex, u, v, pt);
        cont = true;
      } else {
        p.rad = p.rad + p.thr * v3{0.0f, 0.0f, 0.0f};
      }
    } else if (M.kind == TPT_MAT_DIELECTRIC) {
      // dielectric::scatter (src/material.cc:19-70).
      const float ior = M.p[0];
      const v3 reflected = reflect(p.d, n);
      v3 outward;
      float ni, cosine;
      if (dot(p.d, n) > 0.0f) {
        outward = -n;
        ni = ior;
        cosine = dot(p.d, n) / len(p.d);
        cosine = sqrtf(1.0f - ior * ior * (1.0f - cosine * cosine));
      } else {
        outward = n;
        ni = 1.0f / ior;
        cosine = -dot(p.d, n) / len(p.d);
      }
      v3 refracted{0.0f, 0.0f, 0.0f};
      const float prob = refract(p.d, outward, ni, &refracted) ? schlick(cosine, ior) : 1.0f;
      nd = R.next() < prob ? reflected : refracted;
      cont = true;
    } else {
      // lambertian + mixture_pdf(hitable_pdf(lights), cosine_pdf)
      // (src/material.cc:3-17, src/utils.cc:73-81); cosine_pdf::value
      // is cos θ without 1/π, as in the reference.
      const v3 atten = tex_value<SIMPLE>(S, M.tex, u, v, pt);
      const onb uvw(n);
      float pdf;
      if (Sj.n_lights > 0) {
        if (R.next() < 0.5f) {
          nd = light_random<SIMPLE>(Sj, pt, R);
        } else {
          nd = uvw.local(on_hemisphere(R));
        }
        const float pl = light_pdf<SIMPLE>(Sj, pt, nd);
        ++st[ST_LIGHT];
        float pc = dot(unit(nd), uvw.w);
        pc = pc > 0.0f ? pc : 0.0f;
        pdf = 0.5f * pl + 0.5f * pc;
      } else {
        nd = uvw.local(on_hemisphere(R));
        const float pc = dot(unit(nd), uvw.w);
        pdf = pc > 0.0f ? pc : 0.0f;
      }
      float spdf = dot(n, unit(nd));
      spdf = spdf < 0.0f ? 0.0f : spdf / kPi;
      const v3 w = (atten * spdf) / pdf;
      p.rad = p.rad + p.thr * v3{0.0f, 0.0f, 0.0f};  // carries a NaN throughput into the sample
      p.thr = p.thr * w;
      cont = true;
      p.ndiff += p.ndiff < 0xFFFFu ? 1u : 0u;  // 16 bits in the record
    }
    if (cont) {
      p.o = pt;
      p.d = nd;
      p.depth += 1;
    }
  }
  return cont;
}

// Shading of the material queues.  Plain mode: all three classes.  Guided
// mode: the other and terminal classes; the diffuse class (the vertices that
// use the guide) is shaded by k_guide_shade.  SIMPLE: constant textures, at
// most one rect light (see shade.h) and no moving spheres.  The occupancy targets below make the
// register allocator fit 96 registers (5 waves per SIMD instead of 4) without
// spilling, or, for the single-scene SIMPLE variant, 64 (8 waves per SIMD;
// 40 bytes of the record are held in LDS, see claim_and_append_ref).
template <bool GUIDED, bool MULTI, bool SIMPLE = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SIMPLE ? (MULTI ? 5 : 8) : (GUIDED ? 1 : 5)))) k_shade(scene_view S, render_params P, render_state Q, uint32_t parity) {
  const pass_hdr H = *P.hdr;
  // material regions in class-major order: region r = (class - C0) * NSUB + k
  constexpr uint32_t C0 = GUIDED ? 1u : 0u, NR = (3u - C0) * NSUB;
  __shared__ uint32_t offs[NR];
  __shared__ uint32_t pre[NR + 1];
  if (threadIdx.x < NR) offs[threadIdx.x] = mc_at(parity, C0 + threadIdx.x / NSUB, threadIdx.x % NSUB);
  __syncthreads();
  region_prefix(Q.counters, offs, NR, pre);
  const uint32_t total = pre[NR];
  if (blockIdx.x == 0 && threadIdx.x < 3 * NSUB)
    Q.counters[mc_at(parity ^ 1u, threadIdx.x / NSUB, threadIdx.x % NSUB)] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) Q.counters[first_at()] = 0;  // the pass's first trace has run
  const uint32_t home = blockIdx.x % NSUB;
  uint32_t *next_count = Q.counters + rc_at(parity ^ 1u, home);
  float4 *const *out = Q.rq[parity ^ 1u];
  uint32_t st[ST_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += stride) {
    const uint32_t qi = base + threadIdx.x;
    const bool valid = qi < total;
    bool cont = false, finished = false;
    path p{};
    if (valid) {
      const uint32_t r = find_region(pre, NR, qi);
      const uint32_t cls = C0 + r / NSUB;
      const uint32_t off = (r % NSUB) * Q.region + (qi - pre[r]);
      p = load_ray(Q.mq[cls], off);
      const float4 hw = ld_once(&Q.mq[cls][3][off]);
      const hit_t h{hw.x, __float_as_uint(hw.y), __float_as_uint(hw.z), __float_as_uint(hw.w)};
      const item_ctx c = item_in_entry(P, MULTI ? p.ent : 0u, p.item);
      const scene_view Sj = view_of<MULTI>(S, P, c.job);
      cont = shade_hit<SIMPLE>(S, Sj, P, P.jobs[c.job], c, p, h, cls, st);
      if (!cont) {
        finish_sample<GUIDED>(P, Q, p, st);
        finished = true;
      }
    }
    claim_and_append_ref<MULTI>(P, H, Q, next_count, out, home, p, cont, finished);
  }
  block_stats(Q.stats, st, ST_BASIC);
}

// Guided mode, diffuse queue: network inference fused with the guided
// shading.  Each lane gathers one entry (path, hit record, encoding inputs);
// the wave then runs four MFMA passes of 16 entries, lane l taking the inputs
// of entry 16j + l%16 by shuffle.  A pass leaves outputs 16mt + 4(l/16) .. +3
// of that entry in lane l's D tiles; they go to the wave's rows of an LDS
// table (bf16) from which every lane reads its own entry's 40 outputs.
//
// One-sample MIS of the guide g with the reference sampler q (the
// reference's mixture_pdf(hitable_pdf, cosine_pdf), src/material.cc:3-17,
// src/utils.cc:73-81): directions come from m = a g + (1 - a) q and are
// weighted atten (cos/π) q / (p_ref m), p_ref being the reference's mixture
// value (cosine_pdf::value is cos θ without 1/π), so the expectation is the
// reference estimator's for any a and g.
//
// SIMPLE: as for k_shade (constant textures, at most one rect light, no
// moving spheres).
// Block size of k_guide_shade: its LDS holds the inference tiles (30 KB) once
// per block plus 80 B of outputs per thread.  SIMPLE fits 128 VGPRs, so
// 512-thread blocks (73 KB, two per CU) give 4 waves per SIMD; the general
// variant needs more registers and runs 3 blocks of 256 per CU.
__host__ __device__ constexpr uint32_t guide_shade_threads(bool simple) { return simple ? 512u : 256u; }
__host__ __device__ constexpr uint32_t guide_shade_blocks_per_cu(bool simple) { return simple ? 2u : 3u; }

template <bool MULTI, bool SIMPLE = false>
__global__ void __launch_bounds__(guide_shade_threads(SIMPLE)) __attribute__((amdgpu_waves_per_eu(SIMPLE ? 4 : 3))) k_guide_shade(scene_view S, render_params P, render_state Q, uint32_t parity) {
  const pass_hdr H = *P.hdr;
  __shared__ uint32_t wt[GX_TILES * GX_TILE_WORDS];
  __shared__ __align__(16) float bias[G_BIAS];
  __shared__ uint32_t offs[NSUB];
  __shared__ uint32_t pre[NSUB + 1];
  __shared__ uint2 outs[guide_shade_threads(SIMPLE) * 10];  // entry e: outputs 4k..4k+3 as bf16 in outs[10 e + k]
  guide_stage_weights(wt, P.g_tiles, GX_TILES * GX_TILE_WORDS);
  for (int i = threadIdx.x; i < G_BIAS; i += blockDim.x) bias[i] = P.g_params[G_B1 + i];
  if (threadIdx.x < NSUB) offs[threadIdx.x] = mc_at(parity, Q_DIFFUSE, threadIdx.x);
  __syncthreads();
  region_prefix(Q.counters, offs, NSUB, pre);
  const uint32_t total = pre[NSUB];
  const float alpha_net = 1.0f / (1.0f + __expf(-P.g_params[G_NPARAM_PAD]));
  const int l = __lane_id(), fg = l >> 4, sl = l & 15;
  const uint32_t wave_base = (threadIdx.x / 64) * 64;
  const uint32_t home = blockIdx.x % NSUB;
  uint32_t *next_count = Q.counters + rc_at(parity ^ 1u, home);
  float4 *const *out = Q.rq[parity ^ 1u];
  uint32_t st[ST_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
    const uint32_t qi = base + threadIdx.x;
    const bool valid = qi < total;
    // Phase 1: the entry's hit record and encoding inputs.  The hit point,
    // (u, v) and material stay live across the network passes (the normal is
    // in9[3..5]); phase 2 reloads only the path record.
    uint32_t off = 0, hmat = 0;
    float in9[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 1.0f};
    v3 pt{0.0f, 0.0f, 0.0f};
    float u = 0.0f, v = 0.0f;
    if (valid) {
      const uint32_t r = find_region(pre, NSUB, qi);
      off = r * Q.region + (qi - pre[r]);
      const path p = load_ray<false>(Q.mq[Q_DIFFUSE], off);
      const float4 hw = ld_once(&Q.mq[Q_DIFFUSE][3][off]);
      const hit_t h{hw.x, __float_as_uint(hw.y), __float_as_uint(hw.z), __float_as_uint(hw.w)};
      hmat = h.mat;
      v3 n;
      // (the diffuse queue holds lambertian hits only: no media; the job's
      // view: a wrapped segment is looked up among its scene's segments only)
      const ray_ctx rc{!SIMPLE && P.motion ? ray_time(P, item_in_entry(P, MULTI ? p.ent : 0u, p.item)) : 0.0f, 0u, 0u, 0u, 0u, 0u};
      hit_record(view_of<MULTI>(S, P, MULTI ? P.pass[p.ent].job : 0u), h, p.o, p.d, rc, &pt, &n, &u, &v);
      const v3 wo = -unit(p.d);
      const float pv[3] = {pt.x, pt.y, pt.z};
      const float *bounds = P.scenes[P.jobs[MULTI ? P.pass[p.ent].job : 0u].scene].bounds;
      for (int k = 0; k < 3; ++k) in9[k] = (pv[k] - bounds[k]) / fmaxf(bounds[3 + k] - bounds[k], 1e-20f);
      in9[3] = n.x, in9[4] = n.y, in9[5] = n.z;
      in9[6] = wo.x, in9[7] = wo.y, in9[8] = wo.z;
      if (P.rec_stride && (p.item % P.rec_stride) == 0 && p.ndiff < REC_VERTS) {
        // training-row side data: encoding inputs now, the rest after sampling
        float4 *sd = P.rec_side + (static_cast<size_t>(p.item / P.rec_stride) * REC_VERTS + p.ndiff) * 5;
        sd[0] = make_float4(in9[0], in9[1], in9[2], in9[3]);
        sd[1] = make_float4(in9[4], in9[5], in9[6], in9[7]);
      }
    }
#pragma unroll 1
    for (uint32_t j = 0; j < 4; ++j) {
      if (base + wave_base + 16 * j >= total) break;  // wave-uniform
      const int src = static_cast<int>(16 * j) + sl;
      float x9[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) x9[k] = __shfl(in9[k], src);
      bf16x8 xb[2];
      guide_xfeatures(fg, fg == 0 ? x9[0] : (fg == 1 ? x9[1] : x9[2]), x9 + 3, xb);
      f32x4 o[3];
      guide_forward_x32(xb, wt, bias, o);
      uint2 *row = outs + (wave_base + 16 * j + sl) * 10;
#pragma unroll
      for (int mt = 0; mt < 3; ++mt) {
        const int k = 4 * mt + fg;
        if (k < 10) {
          const bf16x4 b = pack4(o[mt][0], o[mt][1], o[mt][2], o[mt][3]);
          row[k] = make_uint2(static_cast<uint32_t>(static_cast<uint16_t>(b[0])) |
                                  (static_cast<uint32_t>(static_cast<uint16_t>(b[1])) << 16),
                              static_cast<uint32_t>(static_cast<uint16_t>(b[2])) |
                                  (static_cast<uint32_t>(static_cast<uint16_t>(b[3])) << 16));
        }
      }
    }
    // Each wave reads only the table rows its own MFMA passes wrote, so a
    // wave-level barrier suffices (a wave's LDS operations complete in
    // order) and the block's waves need not wait for each other.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // Phase 2: guided shading of this lane's entry.
    bool cont = false, finished = false;
    path p{};
    if (valid) {
      p = load_ray(Q.mq[Q_DIFFUSE], off);
      const item_ctx c = item_in_entry(P, MULTI ? p.ent : 0u, p.item);
      const scene_view Sj = view_of<MULTI>(S, P, c.job);
      const v3 n{in9[3], in9[4], in9[5]};
      vmf_mix mix;
      const bool finite = decode_mix_bf16(reinterpret_cast<const uint16_t *>(outs + threadIdx.x * 10), &mix);
      mix_norms(&mix);
      // a non-finite network output falls back to the reference sampler at this vertex
      const float alpha = finite ? alpha_net : 0.0f;
      const tpt_material M = S.materials[hmat];
      const bool lights = Sj.n_lights > 0;
      stream R;
      R.init(c.px, c.smp, p.depth + 1, P.jobs[c.job].seed);
      const v3 atten = tex_value<SIMPLE>(S, M.tex, u, v, pt);
      const onb uvw(n);
      v3 nd;
      if (R.next() < alpha) {
        const float u0 = R.next(), u1 = R.next(), u2 = R.next();
        nd = mix_sample(mix, u0, u1, u2);
      } else if (lights && R.next() < 0.5f) {
        nd = light_random<SIMPLE>(Sj, pt, R);
      } else {
        nd = uvw.local(on_hemisphere(R));
      }
      const v3 ud = unit(nd);
      // the mixture density first, so the mixture is dead during light_pdf
      const float g = alpha > 0.0f ? mix_pdf(mix, ud) : 0.0f;
      const float pl = lights ? light_pdf<SIMPLE>(Sj, pt, nd) : 0.0f;
      st[ST_LIGHT] += lights ? 1u : 0u;
      float pc = dot(ud, uvw.w);
      pc = pc > 0.0f ? pc : 0.0f;
      const float p_ref = lights ? 0.5f * pl + 0.5f * pc : pc;
      const float q = lights ? 0.5f * pl + 0.5f * (pc / kPi) : pc / kPi;
      const float m = alpha > 0.0f ? alpha * g + (1.0f - alpha) * q : q;
      float spdf = dot(n, ud);
      spdf = spdf < 0.0f ? 0.0f : spdf / kPi;
      const v3 w = (atten * (spdf * q)) / (p_ref * m);
      if (P.rec_stride && (p.item % P.rec_stride) == 0 && p.ndiff < REC_VERTS) {
        float4 *sd = P.rec_side + (static_cast<size_t>(p.item / P.rec_stride) * REC_VERTS + p.ndiff) * 5;
        const v3 tp = p.thr * atten;
        sd[2] = make_float4(-unit(p.d).z, ud.x, ud.y, ud.z);
        sd[3] = make_float4(q, p.rad.x, p.rad.y, p.rad.z);
        sd[4] = make_float4(tp.x, tp.y, tp.z, 0.0f);
      }
      p.thr = p.thr * w;
      p.ndiff += p.ndiff < 0xFFFFu ? 1u : 0u;  // 16 bits in the record
      // Below the surface and off the light the integrand is zero (the
      // reference sampler never goes there): the path ends with what it has.
      cont = spdf > 0.0f && q > 0.0f;
      if (cont) {
        p.o = pt;
        p.d = nd;
        p.depth += 1;
      } else {
        finish_sample<true>(P, Q, p, st);
        finished = true;
      }
    }
    claim_and_append<MULTI>(P, H, Q, next_count, out, home, &p, cont, finished);
  }
  block_stats(Q.stats, st, ST_BASIC);
}

// Fused tail of a pass.  Once every item of the pass is claimed and few paths
// are left, a bounce costs a trace + shade launch pair whose time is mostly
// the latency of one wave's work.  k_tail instead takes the live records of
// ray queue `parity` and carries each path through all its remaining bounces
// in registers: the trace of k_trace (the same walk), then shade_hit, the
// shading of k_shade, until the path ends.  Same functions and Philox streams,
// so every sample equals the bounce loop's bit for bit.  SIMPLE scenes only
// (constant textures, at most one rect light, no motion or media).
//
// Guided mode: a path that reaches a guided diffuse vertex needs the network
// (k_guide_shade, a wave-wide MFMA pass); it is handed back instead, its
// record and hit appended to diffuse material queue `parity` exactly as
// k_trace appends them, for the caller's next shade launch.  Ray queue
// parity ^ 1 (where that launch appends) is cleared here, as k_trace does.
template <bool SWEEP, bool MULTI, bool SPH, bool GUIDED>
__global__ void __launch_bounds__(256) k_tail(scene_view S, const tpt_slot *__restrict__ slots,
                                              const tpt_entry *__restrict__ entries, render_params P,
                                              render_state Q, uint32_t parity) {
  static_assert(!(SWEEP && MULTI), "the wave-uniform sweep needs one scene");
  static_assert(!(SWEEP && SPH), "the sweep takes every leaf kind");
  S.slots = slots;
  S.entries = entries;
  __shared__ uint32_t offs[NSUB];
  __shared__ uint32_t pre[NSUB + 1];
  if (threadIdx.x < NSUB) offs[threadIdx.x] = rc_at(parity, threadIdx.x);
  __syncthreads();
  region_prefix(Q.counters, offs, NSUB, pre);
  const uint32_t count = pre[NSUB];
  if (blockIdx.x == 0 && threadIdx.x < NSUB) Q.counters[rc_at(parity ^ 1u, threadIdx.x)] = 0;
  float4 *const *in = Q.rq[parity];
  const uint32_t home = blockIdx.x % NSUB;
  uint32_t st[ST_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride) {
    const uint32_t qi = base + threadIdx.x;
    bool live = qi < count;
    path p{};
    if (live) {
      const uint32_t r = find_region(pre, NSUB, qi);
      p = load_ray(in, r * Q.region + (qi - pre[r]));
    }
    // wave-uniform: the sweep needs every lane of the wave
    while (__ballot(live) != 0ull) {
      const item_ctx c = live ? item_in_entry(P, MULTI ? p.ent : 0u, p.item) : item_ctx{0u, 0u, 0u, 0u};
      const scene_view Sj = view_of<MULTI>(S, P, c.job);
      hit_t h{0.0f, kNone, kNone, 0};
      const bool got = trace_world<SWEEP, SPH, false>(Sj, p.o, p.d, ray_ctx{0.0f, 0u, 0u, 0u, 0u, 0u}, 0.001f,
                                                      kFltMax, &h, &st[ST_NODES], live);
      uint32_t cls = Q_TERMINAL;
      if (live) {
        ++st[ST_SEGMENTS];
        ++st[ST_TAIL];
        if (got) {
          h.mat &= TPT_PAYLOAD_MASK;
          const uint32_t kind = S.materials[h.mat].kind;
          if (kind != TPT_MAT_DIFFUSE_LIGHT && p.depth < P.jobs[c.job].max_depth)
            cls = kind == TPT_MAT_LAMBERTIAN && p.ndiff < P.guide_verts ? Q_DIFFUSE : Q_OTHER;
        } else {
          h.slot = kNone;
        }
      }
      if (GUIDED) {
        const bool back = live && cls == Q_DIFFUSE;
        const uint32_t at = home * Q.region + wave_append(Q.counters + mc_at(parity, Q_DIFFUSE, home), back);
        if (back) {
          store_ray(Q.mq[Q_DIFFUSE], at, &p);
          Q.mq[Q_DIFFUSE][3][at] =
              make_float4(h.t, __uint_as_float(h.slot), __uint_as_float(h.entry), __uint_as_float(h.mat));
          live = false;
        }
      }
      if (live && !shade_hit<true>(S, Sj, P, P.jobs[c.job], c, p, h, cls, st)) {
        finish_sample<GUIDED>(P, Q, p, st);
        live = false;
      }
    }
  }
  block_stats(Q.stats, st, ST_BASIC);
}

// Each job's sum/sumsq += its samples of the pass, in sample order per pixel.
template <bool MULTI>
__global__ void __launch_bounds__(256) k_accumulate(render_params P, const float *__restrict__ res) {
  const pass_hdr H = *P.hdr;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < H.n_acc; t += gridDim.x * blockDim.x) {
    uint32_t e = 0;
    if (MULTI) {
      uint32_t lo = 0, hi = H.n_ent;  // invariant: pass[lo].acc_begin <= t < pass[hi].acc_begin
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (P.pass[mid].acc_begin <= t) lo = mid; else hi = mid;
      }
      e = lo;
    }
    const pass_ent E = P.pass[e];
    const job_dev &J = P.jobs[E.job];
    const uint32_t px = t - E.acc_begin;
    float *sum = J.sum + 3ull * px;
    float a0 = sum[0], a1 = sum[1], a2 = sum[2];
    float q0 = 0.0f, q1 = 0.0f, q2 = 0.0f;
    float *sq = J.sumsq ? J.sumsq + 3ull * px : nullptr;
    if (sq) q0 = sq[0], q1 = sq[1], q2 = sq[2];
    for (uint32_t s = 0; s < E.n_layers; ++s) {
      const float *r = res + 3ull * (E.item_begin + static_cast<uint64_t>(s) * J.n_pix + px);
      const float c0 = r[0], c1 = r[1], c2 = r[2];
      a0 += c0, a1 += c1, a2 += c2;
      q0 += c0 * c0, q1 += c1 * c1, q2 += c2 * c2;
    }
    sum[0] = a0, sum[1] = a1, sum[2] = a2;
    if (sq) sq[0] = q0, sq[1] = q1, sq[2] = q2;
  }
}

}  // namespace tptd

#endif  // TPT_CSRC_RENDER_H
