/* tpt.h — C ABI of the MI355X wavefront path tracer (libtpt.so).
 *
 * The reference (BlurryLight/tiny-path-tracer) has no FFI: its render path is
 * two C++ entry points, and each C-ABI function below replaces one of them or
 * the driver loop around them:
 *
 *   hitable::hit(r, t_min, t_max, rec)          headers/hitable.h:32-33
 *     -> tpt_trace_closest / tpt_trace_closest_record   (batched, device)
 *   color(r, world, light_shape, depth, max_depth)  headers/utils.h:26-27,
 *   with the per-row sample loop of main.cpp:109-156
 *     -> tpt_render                                 (one call per frame/pass)
 *   the object graph built by the scene API (headers/ of the reference)
 *     -> tpt_flat_scene (host arrays, produced by the C++ flattener
 *        tpt::flatten in host/flatten.h) -> tpt_scene_create (device upload)
 *
 * Conventions: plain pointers and sizes; every function returns an int status
 * (TPT_OK = 0, negative on error) and tpt_last_error() returns the message of
 * the calling thread's last failure.  No C++ exception crosses this boundary.
 * Pointers named d_* are device (HBM) pointers; `stream` is a hipStream_t
 * (NULL = the default stream).  A scene handle may be used from one stream at
 * a time; distinct handles are independent.
 */
#ifndef TPT_H
#define TPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TPT_ABI_VERSION 2

enum tpt_status {
  TPT_OK = 0,
  TPT_ERR_INVALID = -1,     /* bad argument */
  TPT_ERR_HIP = -2,         /* a HIP runtime call failed */
  TPT_ERR_UNSUPPORTED = -3, /* hitable/material/texture type not on the device path */
  TPT_ERR_NOMEM = -4,
  TPT_ERR_NODEVICE = -5     /* no gfx950 device */
};

/* ---------------------------------------------------------------------------
 * Flattened scene (host arrays).  Layout is documented per record; all
 * records are 16-byte aligned so one lane fetches one record with dwordx4
 * loads.
 *
 * World = a sequence of segments evaluated like the reference's
 * hitable_list::hit (src/hitable_list.cc:38-51): each segment is intersected
 * with t_max = the closest hit so far.  A segment is a pre-order range
 * [start, end) of slots holding one bvh_node tree (or a single leaf).  Within
 * a segment the device walks the slots in pre-order with skip links, which
 * visits exactly the nodes bvh_node::hit visits (src/hitable.cc:63-90: both
 * children, caller's t_max, no shrinking) in the same left-to-right order.
 * ------------------------------------------------------------------------- */

/* slot.b >> 28 */
#define TPT_SLOT_BOX 0u
#define TPT_SLOT_SPHERE 1u
#define TPT_SLOT_RECT 2u
#define TPT_SLOT_LIST 3u
#define TPT_SLOT_MSPHERE 4u
#define TPT_SLOT_MEDIUM 5u
#define TPT_PAYLOAD_MASK 0x0FFFFFFFu
#define TPT_NO_RIGHT 0x0FFFFFFFu

/* Rect axis (normal axis): 0 = yz_rect (x = k), 1 = xz_rect (y = k), 2 = xy_rect (z = k). */

/* 32-byte pre-order slot.
 *  BOX:    f[0..2] box min, f[3..5] box max; a = skip (first slot after the
 *          subtree); b payload = index of the right child slot, or TPT_NO_RIGHT
 *          for an n==1 node (left == right in the reference).
 *  SPHERE: f[0..2] centre, f[3] radius; a = material; b payload = leaf id.
 *  RECT:   f[0] k, f[1] a0, f[2] a1, f[3] b0, f[4] b1 (in-plane bounds, in
 *          the order of the reference constructor); a = material | axis<<28 |
 *          flip<<30; b payload = leaf id.
 *  MSPHERE: a moving_sphere (headers/sphere.h:23-40): f[0] = index into
 *          motions (as uint32 bits); a = material; b payload = leaf id.
 *  MEDIUM: a constant_medium (headers/hitable.h:55-66): f[0] = first entry of
 *          its boundary's primitives, f[1] = their count (both as uint32 bits),
 *          f[2] = density; a = material (its isotropic phase function);
 *          b payload = leaf id.  The boundary entries are evaluated like a LIST
 *          (hitable_list order, shrinking t_max) and hold no MEDIUM.
 *  LIST:   a = first entry; f[0] = leaf id of the first entry (as uint32 bits),
 *          f[1] = entry count (as uint32 bits).  A leaf made of wrappers
 *          (translate / rotate_y / flip_normal), box and hitable_list around
 *          primitives; evaluated with the list's shrinking t_max.
 * Leaf ids number primitive occurrences in depth-first order of the object
 * graph (bvh left before right, lists in order, an n==1 node's duplicate child
 * counted once); tpt_trace_closest reports them. */
typedef struct tpt_slot {
  float f[6];
  uint32_t a;
  uint32_t b;
} tpt_slot;

/* 32-byte entry of a LIST leaf: one primitive seen through a wrapper chain.
 *  f[0..4] primitive parameters as in tpt_slot (sphere: centre, radius;
 *  rect: k, a0, a1, b0, b1; moving sphere: f[0] = motion index bits;
 *  medium: f[0..2] as the MEDIUM slot);
 *  xform = index into xforms or 0xFFFFFFFF (always none for a medium);
 *  a = material | axis<<28 | flip<<30 (flip = parity of flip_normal wrappers);
 *  b = TPT_SLOT_SPHERE, TPT_SLOT_RECT, TPT_SLOT_MSPHERE or TPT_SLOT_MEDIUM. */
typedef struct tpt_entry {
  float f[5];
  uint32_t xform;
  uint32_t a;
  uint32_t b;
} tpt_entry;

/* Wrapper chain, outermost first, at most 4 operations.  kinds: 2 bits per op
 * (0 = translate by p[i][0..2], 1 = rotate_y with p[i][0] = sin, p[i][1] = cos,
 * values from rotate_y's constructor). */
#define TPT_XFORM_TRANSLATE 0u
#define TPT_XFORM_ROTATE_Y 1u
#define TPT_XFORM_MAX_OPS 4
typedef struct tpt_xform {
  uint32_t nops;
  uint32_t kinds;
  float p[TPT_XFORM_MAX_OPS][3];
  uint32_t pad[2];
} tpt_xform;

#define TPT_MAT_LAMBERTIAN 0u
#define TPT_MAT_METAL 1u
#define TPT_MAT_DIELECTRIC 2u
#define TPT_MAT_DIFFUSE_LIGHT 3u
#define TPT_MAT_ISOTROPIC 4u
/* lambertian: tex = albedo texture.  metal: p[0..2] albedo, p[3] fuzz.
 * dielectric: p[0] ior.  diffuse_light: tex = emission texture.
 * isotropic (constant_medium's phase function): tex = albedo texture. */
typedef struct tpt_material {
  uint32_t kind;
  uint32_t tex;
  float p[6];
} tpt_material;

#define TPT_TEX_CONSTANT 0u
#define TPT_TEX_CHECKER 1u
#define TPT_TEX_PERLIN 2u
#define TPT_TEX_IMAGE 3u
/* constant: p[0..2] colour.  checker: a = odd, b = even.  perlin: p[0] scale,
 * a = octaves, b = index of its Perlin tables (0 in a single scene).  image:
 * a = byte offset into texels, b = width, c = height. */
typedef struct tpt_texture {
  uint32_t kind;
  uint32_t a, b, c;
  float p[4];
} tpt_texture;

/* Perlin tables (shared by every Perlin texture of a scene, as the static
 * tables of headers/perlin_noise.h:13-16 are). */
typedef struct tpt_perlin {
  float grad[256][4];
  int32_t perm[3][256];
} tpt_perlin;

#define TPT_LIGHT_NONE 0u    /* hitable without pdf_value/random overrides: pdf 0, direction (1,0,0) */
#define TPT_LIGHT_XZ_RECT 1u /* p[0..4] = x0, x1, z0, z1, k */
#define TPT_LIGHT_SPHERE 2u  /* p[0..3] = centre, radius */
typedef struct tpt_light {
  uint32_t kind;
  float weight; /* product of the 1/list_size factors above it */
  float p[6];
} tpt_light;

/* moving_sphere parameters (headers/sphere.h:23-40, src/sphere.cc:3-11):
 * centre(time) = center0 + (time - time0) / (time1 - time0) * (center1 - center0)
 * in float, as moving_sphere::center computes it. */
typedef struct tpt_motion {
  float center0[3];
  float radius;
  float center1[3];
  float time0;
  float time1;
  float pad[3];
} tpt_motion;

typedef struct tpt_flat_scene {
  const tpt_slot *slots;
  uint32_t n_slots;
  const uint32_t *segments; /* n_segments pairs (start, end) */
  uint32_t n_segments;
  const tpt_entry *entries;
  uint32_t n_entries;
  const tpt_xform *xforms;
  uint32_t n_xforms;
  const tpt_material *materials;
  uint32_t n_materials;
  const tpt_texture *textures;
  uint32_t n_textures;
  const uint8_t *texels;
  uint64_t n_texels;
  const tpt_perlin *perlin; /* NULL when no Perlin texture */
  const tpt_light *lights;  /* light list flattened; n_lights == 0: no light sampling */
  uint32_t n_lights;
  uint32_t light_mode; /* 0: the light root is one shape (no pick draw); 1: a hitable_list of
                          shapes, picked with int(u*n) as hitable_list::random does; 2: nested
                          lists, picked by cumulative weight */
  uint32_t n_leaves; /* number of leaf ids */
  uint32_t max_depth; /* deepest bvh_node nesting */
  float bounds[6];    /* world bounding_box (min xyz, max xyz); normalises guiding inputs */
  const tpt_motion *motions; /* moving spheres (MSPHERE slots and entries index it) */
  uint32_t n_motions;
  /* One word per segment, or NULL: a segment that is a bvh_node seen through
   * translate / rotate_y / flip_normal wrappers (headers/rect_box.h:46-113)
   * holds its xform index | flip<<31; TPT_SEGMENT_PLAIN for the others. */
  const uint32_t *segment_xforms;
} tpt_flat_scene;

#define TPT_SEGMENT_PLAIN 0x7FFFFFFFu
#define TPT_SEGMENT_FLIP 0x80000000u

/* Camera fields as computed by camera_with_blur's constructor
 * (src/camera.cc:2-21); ray generation follows get_ray (src/camera.cc:23-31). */
typedef struct tpt_camera {
  float origin[3];
  float lower_left[3];
  float horizontal[3];
  float vertical[3];
  float u[3], v[3], w[3];
  float lens_radius;
  float time0, time1;
} tpt_camera;

/* ---------------------------------------------------------- scene library ---
 * Counterpart of the reference's scene factories (random_scene(),
 * cornell_box(), ... headers/utils.h:29-36) plus the benchmark scenes.  Builds
 * the named scene through the C++ scene API on a fresh thread whose drand_r
 * generator is seeded with `seed` (5489 = the reference's default state),
 * flattens it and keeps the host arrays.  Names (tpt_scene_names()):
 *   random_spheres          one-weekend spheres, 400x400, 16 spp, depth 50
 *   random_spheres_textured the same geometry with checker/Perlin/image
 *                           textures, 1600x1600, 300 spp, depth 50
 *   cornell555              the 555-unit Cornell box, 1200x1200, 2048 spp
 *   random_spheres_n<N>     random spheres with about N small spheres
 *   ref_cornell_box, ref_sphere_cornell_box, ref_two_perlin_spheres,
 *   ref_light_spheres, ref_two_checker_spheres   the reference's scenes
 *   box_list                instanced boxes in a plain hitable_list */
typedef struct tpt_host_scene tpt_host_scene;
typedef struct tpt_scene_info {
  uint32_t width, height, spp, max_depth; /* the scene's default render settings */
  float background;                       /* sky scale used when the scene has no light */
  uint32_t n_leaves;
  float lookfrom[3], lookat[3];           /* its camera: vup (0,1,0), vfov degrees, aperture, focus */
  float vfov, aperture, focus;
  float time0, time1;                     /* its shutter */
} tpt_scene_info;

const char *tpt_scene_names(void);
int tpt_host_scene_build(const char *name, uint32_t seed, tpt_host_scene **out);
int tpt_host_scene_flat(const tpt_host_scene *hs, tpt_flat_scene *flat);
int tpt_host_scene_camera(const tpt_host_scene *hs, tpt_camera *camera);
int tpt_host_scene_info(const tpt_host_scene *hs, tpt_scene_info *info);
int tpt_host_scene_destroy(tpt_host_scene *hs);

/* load_image_texture (src/utils.cc:236-240, stb_image there): decodes a
 * baseline JPEG or an 8-bit PPM (P6/P3) into *out (width*height*3 bytes of
 * RGB, rows top first; free with tpt_image_free). */
int tpt_image_load(const char *path, uint8_t **out, int *width, int *height);
void tpt_image_free(uint8_t *pixels);

/* The fields of camera_with_blur(lookfrom, lookat, vup, vfov, aspect,
 * aperture, focus_dist, t0, t1) as its constructor computes them
 * (src/camera.cc:2-21): for a driver that sets fov, aspect, aperture and the
 * shutter from its configuration (main.cpp:87-91). */
int tpt_camera_make(const float lookfrom[3], const float lookat[3], const float vup[3], float vfov, float aspect,
                    float aperture, float focus_dist, float t0, float t1, tpt_camera *out);

/* ------------------------------------------------------------- lifecycle --- */
typedef struct tpt_scene tpt_scene;
typedef struct tpt_guide tpt_guide; /* learned guiding proposal, below */

int tpt_abi_version(void);
const char *tpt_last_error(void);
/* Number of visible gfx950 devices (0 when none); never fails. */
int tpt_device_count(void);

/* Upload a flattened scene to the current device.  The flat arrays are copied;
 * the caller may free them afterwards. */
int tpt_scene_create(const tpt_flat_scene *flat, tpt_scene **out);
int tpt_scene_destroy(tpt_scene *scene);

/* ------------------------------------------------------------ closest hit ---
 * Replaces hitable::hit (headers/hitable.h:32-33) for a batch of n rays in SoA
 * layout: d_rays = 6 arrays of n floats back to back (ox, oy, oz, dx, dy, dz).
 * Outputs per ray: d_leaf = leaf id of the closest hit or -1 on a miss, d_t =
 * hit distance (undefined on a miss).  Results equal the reference's
 * recursive hit() bit for bit (same IEEE single operations, FP64 sphere roots,
 * the same tie and NaN rules). */
int tpt_trace_closest(tpt_scene *scene, const float *d_rays, uint64_t n, float t_min, float t_max,
                      int32_t *d_leaf, float *d_t, void *stream);

/* As tpt_trace_closest, and also writes the whole hit_record: d_rec = 8 arrays
 * of n floats (u, v, px, py, pz, nx, ny, nz) and d_mat = material index. */
int tpt_trace_closest_record(tpt_scene *scene, const float *d_rays, uint64_t n, float t_min,
                             float t_max, int32_t *d_leaf, float *d_t, float *d_rec,
                             int32_t *d_mat, void *stream);

/* Rays with a time (ray::time(), headers/ray.h): d_time = n floats, the time
 * each ray's moving spheres are placed at (NULL: time 0, the default of the
 * reference's ray constructor).  d_rec and d_mat may be NULL.
 *
 * constant_medium::hit draws random numbers (src/hitable.cc:98-128): in every
 * closest-hit entry point, ray i's medium with leaf id L draws word k of
 * Philox4x32-10(counter = (i mod 2^32, i / 2^32, 0, 2^20 + L), key = 0): word 0
 * for the free path, words 1-3 for the arbitrary normal of a hit (u = v = 0). */
int tpt_trace_closest_timed(tpt_scene *scene, const float *d_rays, const float *d_time, uint64_t n,
                            float t_min, float t_max, int32_t *d_leaf, float *d_t, float *d_rec,
                            int32_t *d_mat, void *stream);

/* ----------------------------------------------------------------- render ---
 * Replaces color() (src/utils.cc:58-94) driven by the sample loop of
 * main.cpp:115-134: renders spp samples per pixel, sample indices
 * [sample_offset, sample_offset + spp), and ADDS per-pixel sums to the
 * caller's device buffers: d_sum (3·width·height floats, RGB interleaved,
 * row 0 = bottom row j = 0 as in the reference) and d_sumsq (same shape,
 * sum of squares; may be NULL).  Each sample is de_nan'ed per channel
 * (headers/utils.h:100-109) before it is added. */
typedef struct tpt_render_desc {
  uint32_t width, height;
  uint32_t spp;
  uint32_t max_depth;     /* color()'s depth < max_depth gate */
  uint64_t seed;          /* Philox key */
  uint32_t sample_offset; /* first sample index of this pass */
  float background;       /* 0: miss returns black (reference); > 0: sky gradient times this */
  uint32_t guiding;       /* tpt_render: must be 0; tpt_render_guided: training-record stride (0 = none) */
  uint32_t queue_capacity; /* paths in flight, at most 2^29; 0 = automatic (2^29 for a single scene traced by
                              the wave-uniform sweep: at most 64 slots, no motion, media or wrapped trees;
                              otherwise half the largest pass rounded down to a power of two between 2^27 and
                              2^29; at most the pass, while the render state fits in 70% of the device memory
                              free at the call) */
  uint32_t flags;          /* TPT_RENDER_* flags or 0 */
  uint32_t guide_vertices; /* tpt_render_guided: diffuse vertices per path sampled with the guide,
                              the rest with the reference sampler (0 = 2); ignored by tpt_render */
} tpt_render_desc;

/* flags: time every trace / shade launch with HIP events (fills the *_ms and
 * *_launches fields of tpt_render_stats). */
#define TPT_RENDER_PROFILE 1u
/* flags: constant_medium scatters as the book intends (isotropic: a direction
 * from random_in_unit_sphere, throughput times the albedo).  Without it a
 * medium behaves as the reference's does: isotropic::scatter does not
 * override material::scatter (headers/material.h:74-80), so a path that the
 * medium stops ends there, black. */
#define TPT_RENDER_MEDIA_SCATTER 2u
/* flags: drain every pass with trace / shade launches only.  By default, once
 * a pass has handed out all its work and few paths are left, one fused launch
 * (k_tail) carries those paths through their remaining bounces (scenes with
 * constant textures, at most one rect light, no motion or media, on the
 * small-scene or sphere-only walk); every sample is the same bit for bit. */
#define TPT_RENDER_NO_TAIL 4u
/* flags: count the closest-hit walk's operations by kind (box_tests,
 * sphere_tests, rect_tests, xforms of tpt_render_stats), for the counted VALU
 * work of a render.  Runs a counting variant of the trace kernel and no fused
 * tail; the sums are the same bit for bit. */
#define TPT_RENDER_COUNT_OPS 8u

typedef struct tpt_render_stats {
  uint64_t paths;          /* camera paths completed */
  uint64_t segments;       /* closest-hit queries (ray segments) */
  uint64_t light_queries;  /* light-pdf intersections */
  uint64_t nan_samples;    /* samples with a NaN channel zeroed by de_nan */
  double kernel_ms;        /* GPU time of the whole pass (events around it) */
  uint64_t nodes_visited;  /* pre-order slots visited by closest-hit queries */
  double trace_ms, shade_ms;          /* with TPT_RENDER_PROFILE */
  uint64_t trace_launches, shade_launches;
  uint64_t tail_segments;  /* segments traced by the fused tail launches (included in segments) */
  double tail_ms;          /* with TPT_RENDER_PROFILE: time of the fused tail launches */
  uint64_t tail_launches;
  /* with TPT_RENDER_COUNT_OPS (else 0): slab tests, sphere tests, rect tests
   * and wrapper chains (translate / rotate_y) applied by the closest-hit walks */
  uint64_t box_tests, sphere_tests, rect_tests, xforms;
} tpt_render_stats;

int tpt_render(tpt_scene *scene, const tpt_camera *camera, const tpt_render_desc *desc, float *d_sum,
               float *d_sumsq, tpt_render_stats *stats, void *stream);

/* tpt_render without allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code running under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * arguments, the same sums bit for bit and the same stats as tpt_render
 * (kernel_ms is host wall time), but it calls no hipMalloc / hipFree /
 * hipHostMalloc / hipEventCreate: the render state (queues, counters,
 * per-sample results, job and pass tables, statistics) lives in one statically
 * allocated device array of TPT_NOALLOC_SCRATCH_BYTES, which each GPU that the
 * library runs on holds until the process ends.  Limits: at most
 * TPT_NOALLOC_PATHS path records in flight (desc->queue_capacity 0 = that
 * many), passes of at most TPT_NOALLOC_PASS_ITEMS samples, so an image of at
 * most that many pixels; larger values are refused with TPT_ERR_INVALID, as
 * are TPT_RENDER_PROFILE and a nonzero desc->guiding.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams or
 * threads) race and must be serialised by the caller.  The scene's device must
 * be current. */
#define TPT_NOALLOC_SCRATCH_BYTES (144u << 20)
#define TPT_NOALLOC_PATHS (1u << 17)
#define TPT_NOALLOC_PASS_ITEMS (3u << 19)
int tpt_render_noalloc(tpt_scene *scene, const tpt_camera *camera, const tpt_render_desc *desc, float *d_sum,
                       float *d_sumsq, tpt_render_stats *stats, void *stream);

/* ----------------------------------------------------- batched rendering ---
 * No counterpart in the reference, which renders one scene per run: many
 * renders (scenes, resolutions, sample counts) in one launch sequence.
 *
 * tpt_scene_create_pool packs several flattened scenes into one handle: their
 * BVH slots, primitives, transforms, materials, textures and lights go into
 * one pool with indices rebased.  tpt_trace_closest and tpt_render use scene
 * 0 of a pool. */
int tpt_scene_create_pool(const tpt_flat_scene *scenes, uint32_t n_scenes, tpt_scene **out);

/* One render of a pooled scene (the fields of tpt_render_desc, its camera and
 * its sums).  Samples [sample_offset, sample_offset + spp) of every pixel are
 * added to d_sum / d_sumsq (3 * width * height floats each; d_sumsq may be
 * NULL). */
typedef struct tpt_job {
  uint32_t scene;
  uint32_t width, height, spp, max_depth;
  uint32_t sample_offset;
  uint64_t seed;
  float background;
  uint32_t reserved;
  tpt_camera camera;
  float *d_sum, *d_sumsq;
} tpt_job;

/* Launch settings of one tpt_render_jobs call. */
typedef struct tpt_launch {
  uint32_t queue_capacity;  /* path records in flight, at most 2^29; 0 = automatic, as for tpt_render */
  uint32_t blocks_per_cu;   /* trace/shade grid = CUs x this (256-thread blocks); 0 = 64 for passes of
                               at least 2^25 items, else 32 */
  uint32_t graph_iters;     /* bounce iterations per replay of a captured hipGraph (rounded up to
                               even); 0 = one launch at a time */
  uint32_t flags;           /* TPT_RENDER_* flags or 0 (profiling launches one at a time) */
  uint32_t guiding;         /* with a guide: training-record stride (0 = none) */
  uint32_t guide_vertices;  /* with a guide: guided diffuse vertices per path (0 = 2) */
} tpt_launch;

/* Renders every job through one shared work cursor (jobs of different scenes
 * share every launch).  Each job's result equals tpt_render (guide NULL) or
 * tpt_render_guided of that job alone, bit for bit. */
int tpt_render_jobs(tpt_scene *scene, const tpt_job *jobs, uint32_t n_jobs, const tpt_launch *launch,
                    tpt_guide *guide, tpt_render_stats *stats, void *stream);

/* Autotuner: times calibration renders of the jobs (into scratch sums) in
 * three stages: the queue capacities of tpt_autotune_plan (those whose render
 * state fits in 70% of the device memory free at the call) at 64 blocks per CU
 * without graph replay, then grid sizes of 16 and 32 blocks per CU at the best
 * capacity, then graph replay of 8 iterations at the best of those.  Writes
 * the fastest settings into *best (its flags, guiding and guide_vertices are
 * kept).  *best_ms: that calibration's wall time. */
int tpt_autotune(tpt_scene *scene, const tpt_job *jobs, uint32_t n_jobs, tpt_guide *guide, uint32_t calib_spp,
                 tpt_launch *best, float *best_ms);

/* The calibration plan of tpt_autotune (host only; no device needed).  A
 * render clips its queue capacity to its largest pass, so the candidates
 * 2^24..2^29 are clipped to the largest pass of the real render and equal
 * ones are dropped: caps[0..*n_caps) are distinct (at most
 * TPT_AUTOTUNE_MAX_CAPS).  The calibration jobs take
 * min(spp, *calib_spp_out) samples each; that cap starts at calib_spp (0 = 2)
 * and doubles until a calibration pass holds at least twice the largest
 * candidate, or the calibration is the whole render.  base: the launch whose
 * guiding stride sets the real render's pass length (NULL: none). */
#define TPT_AUTOTUNE_MAX_CAPS 6
int tpt_autotune_plan(const tpt_job *jobs, uint32_t n_jobs, const tpt_launch *base, uint32_t calib_spp,
                      uint32_t caps[TPT_AUTOTUNE_MAX_CAPS], uint32_t *n_caps, uint32_t *calib_spp_out);

/* ------------------------------------------------------- learned guiding ---
 * No counterpart in the reference: an online-trained proposal, mixed into the
 * reference's mixture_pdf(hitable_pdf, cosine_pdf) by one-sample MIS so that
 * the estimator keeps the reference's expectation.  A 42 -> 3x64 ReLU -> 40
 * MLP (frequency-encoded hit point, normal, outgoing direction) outputs an
 * 8-lobe von Mises-Fisher mixture g.  Directions are drawn from
 * m = a*g + (1-a)*q, q = 0.5*p_light + 0.5*cos/pi (the reference sampler's true
 * density), and weighted atten*(cos/pi)*q / (p_ref * m), p_ref being the
 * reference's mixture value 0.5*p_light + 0.5*cos; a = 0 is the reference
 * estimator exactly.  Forward, backward and Adam run as BF16 MFMA kernels with
 * FP32 accumulation.
 *
 * Parameters: TPT_GUIDE_NPARAM floats (fp32, padded layout: W1 64x48, W2 64x64,
 * W3 64x64, W4 48x64, b1 64, b2 64, b3 64, b4 48; padding entries stay zero)
 * followed by the mix-weight logit (a = sigmoid).
 * Training row (TPT_GUIDE_ROW floats): p_hat[3] (hit point mapped to [0,1]^3
 * by the scene bounds), n[3], wo[3], omega[3] (unit sampled direction),
 * c (luminance of the sample's contribution over its proposal density, without
 * albedo; rows with c <= 0 are ignored), q (reference sampler density at omega),
 * 2 unused. */
#define TPT_GUIDE_NPARAM 14576
#define TPT_GUIDE_ROW 16

int tpt_guide_create(uint64_t seed, tpt_guide **out);
int tpt_guide_destroy(tpt_guide *guide);
/* host_params: TPT_GUIDE_NPARAM + 1 floats (the last is the mix-weight logit). */
int tpt_guide_get_params(tpt_guide *guide, float *host_params);
int tpt_guide_set_params(tpt_guide *guide, const float *host_params);
/* Raw network outputs for n inputs: d_in n x 9 (p_hat, n, wo), d_out n x 40. */
int tpt_guide_infer(tpt_guide *guide, const float *d_in, uint64_t n, float *d_out, void *stream);
/* Gradient of the training loss -(1/n) sum_j c_j log g(omega_j) (and of the
 * mix-weight loss -(1/n) sum c_j log m(omega_j)) over d_rows (n x
 * TPT_GUIDE_ROW) into d_grad (TPT_GUIDE_NPARAM + 1 floats); *loss receives the
 * first loss.  No update. */
int tpt_guide_gradient(tpt_guide *guide, const float *d_rows, uint64_t n, float *d_grad, float *loss, void *stream);
/* iters Adam steps (beta 0.9/0.999, eps 1e-8) at learning rate lr, each on the
 * next n_batch rows of d_rows (n rows, used cyclically).  d_rows == NULL: the
 * rows recorded by the last tpt_render_guided. */
int tpt_guide_train(tpt_guide *guide, const float *d_rows, uint64_t n, uint64_t n_batch, uint32_t iters, float lr,
                    float *loss, void *stream);
/* tpt_render with the guided estimator.  desc->guiding = N > 0 also records
 * training rows from every N-th path (2 diffuse vertices each) into the
 * guide's batch, returned by tpt_guide_batch. */
int tpt_render_guided(tpt_scene *scene, tpt_guide *guide, const tpt_camera *camera, const tpt_render_desc *desc,
                      float *d_sum, float *d_sumsq, tpt_render_stats *stats, void *stream);
int tpt_guide_batch(tpt_guide *guide, const float **d_rows, uint64_t *n_rows);

#ifdef __cplusplus
}
#endif

#endif /* TPT_H */
