/*
 * pt_render.h -- C ABI of the MI355X path-tracing core.
 *
 * This is the drop-in boundary for the reference renderer's GPU render path:
 * every entry point replaces one member of class cuda_renderer or one free
 * function declared in src/cuda/cuda_renderer.h of the reference
 * (AdmiralHeathCL/cuda-path-tracer).  Arguments are plain pointers, sizes and
 * POD records; nothing here names a C++ or torch type.
 *
 * Record layouts are byte-identical to the reference's transfer records
 * (src/flat/flat_data.h:14-40, src/renderer/transport.h:12, src/core/ray.h),
 * so a caller can hand its flat_scene vectors over without conversion.
 *
 * Errors: every function returns a pt_status.  PT_EINVAL corresponds to the
 * reference throwing std::invalid_argument; every other non-zero status
 * corresponds to std::runtime_error.  pt_last_error() returns the message of
 * the last failure on the calling thread.
 *
 * Threading: a renderer is used from one host thread at a time; every call
 * returns after its GPU work has finished (the reference synchronises in
 * reset and add_samples).
 */
#ifndef PT_RENDER_H
#define PT_RENDER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- transfer records (src/flat/flat_data.h:14-40) ---- */
typedef struct pt_vec3 { float x, y, z; } pt_vec3;                                       /* 12 B */
typedef struct pt_ray { pt_vec3 origin, direction; float time; } pt_ray;                  /* 28 B, src/core/ray.h:30-32 */
typedef struct pt_flat_triangle {                                                         /* 116 B */
    pt_vec3 v0, v1, v2, n0, n1, n2, uv0, uv1, uv2;
    int32_t material_id, smooth;
} pt_flat_triangle;
typedef struct pt_flat_sphere { pt_vec3 center; float radius; int32_t material_id; } pt_flat_sphere; /* 20 B */
typedef struct pt_flat_node {                                                             /* 40 B */
    pt_vec3 minimum, maximum;
    int32_t type, left, right, primitive;
} pt_flat_node;
typedef struct pt_flat_material { pt_vec3 albedo; float fuzz, ref_idx; int32_t type, texture_id; } pt_flat_material; /* 28 B */
typedef struct pt_flat_texture {                                                          /* 60 B */
    pt_vec3 color, scale, offset;
    int32_t type, image_id, width, height, byte_offset, clamp;
} pt_flat_texture;
typedef struct pt_flat_camera {                                                           /* 84 B */
    pt_vec3 origin, lower_left_corner, horizontal, vertical, u, v;
    float lens_radius, time0, time1;
} pt_flat_camera;
typedef struct pt_flat_hit {                                                              /* 48 B */
    float t, u, v;
    pt_vec3 p, normal;
    int32_t material_id, primitive_type, primitive_id;
} pt_flat_hit;
typedef struct pt_trace_result { pt_flat_hit rec; int32_t hit, error; } pt_trace_result; /* 56 B, transport.h:12 */

enum pt_primitive_type { PT_BRANCH = 0, PT_TRIANGLE = 1, PT_SPHERE = 2 };               /* flat_data.h:10 */
enum pt_material_type { PT_DIFFUSE = 0, PT_METAL = 1, PT_GLASS = 2, PT_LIGHT = 3 };      /* flat_data.h:11 */
enum pt_texture_type { PT_CONSTANT = 0, PT_IMAGE = 1, PT_MAP = 2 };                      /* flat_data.h:12 */
enum pt_render_debug {                                                                   /* transport.h:10 */
    PT_DEBUG_SHADED = 0, PT_DEBUG_NORMALS = 1, PT_DEBUG_TRIANGLE = 2,
    PT_DEBUG_MATERIAL = 3, PT_DEBUG_BVH = 4, PT_DEBUG_INTERSECTIONS = 5
};

/* Host view of one flat_scene (flat_data.h:47-56).  Arrays are read during
 * the call only; the library deep-copies them to the device. */
typedef struct pt_scene_desc {
    const pt_flat_triangle *triangles; int64_t triangle_count;
    const pt_flat_sphere *spheres; int64_t sphere_count;
    const pt_flat_node *nodes; int64_t node_count;
    const pt_flat_material *materials; int64_t material_count;
    const pt_flat_texture *textures; int64_t texture_count;
    const uint8_t *pixels; int64_t pixel_bytes;
    int32_t root, max_depth;
} pt_scene_desc;

/* Constructor arguments of cuda_renderer (cuda_renderer.h:12-13) plus the
 * path-length limit, which the reference fixes at 50 (transport.h:174,193). */
typedef struct pt_settings {
    int32_t width, height;
    uint64_t seed;
    int32_t brute;          /* bool brute = false */
    int32_t debug;          /* render_debug debug = DEBUG_SHADED */
    int32_t direct_light;   /* bool direct_light = false */
    int32_t max_path_depth; /* 0 selects the reference's 50 */
} pt_settings;

typedef enum pt_status {
    PT_OK = 0,
    PT_EINVAL = 1,    /* std::invalid_argument in the reference */
    PT_ESCENE = 2,    /* invalid flat scene data (flat_validate.h:11, runtime_error) */
    PT_EDEVICE = 3,   /* no usable device, or a HIP runtime error */
    PT_EOVERFLOW = 4, /* traversal stack overflow (cuda_renderer.cu:204) */
    PT_ENOMEM = 5
} pt_status;

typedef struct pt_renderer pt_renderer;

/* Message of the last failing call on this thread ("" if none). */
const char *pt_last_error(void);

/* cuda_device_available (cuda_renderer.h:8, .cu:128-140): PT_OK and the
 * device name, or PT_EDEVICE and the reason, written NUL-terminated into
 * description[0..capacity). */
pt_status pt_device_available(char *description, size_t capacity);
/* Page-locked host memory for read-backs (no counterpart in the reference):
 * device-to-host copies into it run at the link's rate instead of through a
 * staging copy.  Free with pt_host_free. */
pt_status pt_host_alloc(size_t bytes, void **out);
pt_status pt_host_free(void *p);

/* Linear BVH built on the device (no counterpart in the reference, whose
 * hierarchy is built on the host, src/bvh/bvh.h:132-171; for callers that
 * rebuild large or changing scenes): the flat hierarchy of the given
 * triangles and spheres (primitive p < triangle_count is triangle p, the
 * rest spheres), one primitive per leaf with the reference's leaf boxes
 * (triangle.h:96-109, sphere.h:69-75) and branch boxes the union of their
 * children (aabb.h:40-56).  Karras 2012: 63-bit Morton codes of the box
 * centroids, radix sort, independent splits, bottom-up boxes.  Writes
 * 2n - 1 nodes (n = triangle_count + sphere_count) into nodes (host memory,
 * node_capacity entries), *root = 0 and *max_depth; primitives keep their
 * indices.  n == 0 gives *root = -1, *max_depth = 0.  PT_EINVAL for bad
 * arrays or a short node array, PT_ESCENE for a non-finite vertex, a bad
 * sphere, or a tree deeper than the 64-level traversal stack (then use
 * pts_build).  *milliseconds (optional) is the device time of the build.
 * The output is identical for identical input. */
pt_status pt_build_bvh(const pt_flat_triangle *triangles, int64_t triangle_count, const pt_flat_sphere *spheres,
                       int64_t sphere_count, pt_flat_node *nodes, int64_t node_capacity, int32_t *root,
                       int32_t *max_depth, float *milliseconds);

/* cuda_renderer::cuda_renderer (cuda_renderer.h:12-13, .cu:177-183):
 * validates scene, camera, mode and size, uploads, then reset(camera). */
pt_status pt_renderer_create(const pt_scene_desc *scene, const pt_flat_camera *camera,
                             const pt_settings *settings, pt_renderer **out);
/* ~cuda_renderer */
pt_status pt_renderer_destroy(pt_renderer *renderer);
/* cuda_renderer::reset (cuda_renderer.h:17, .cu:186-194) */
pt_status pt_renderer_reset(pt_renderer *renderer, const pt_flat_camera *camera);
/* cuda_renderer::add_samples (cuda_renderer.h:18, .cu:196-213) */
pt_status pt_renderer_add_samples(pt_renderer *renderer, int32_t count);
/* add_samples without any allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory while they render -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid it.  The
 * results (sums, PCG states, ray count, image) equal pt_renderer_add_samples'.
 * The call's scratch (launch control block, traversal stack entries beyond
 * LDS) is a fixed 64 MiB device array per GPU, held from the library's load
 * until the process ends and shared by every such call on that GPU: calls
 * that overlap on one GPU race.  The launch keeps as many resident path-lane
 * blocks as the array's stack region holds (at least 181 at the deepest
 * hierarchy validation allows).  Shaded mode needs the 4-wide traversal:
 * brute-force and exact-traversal renderers get PT_EINVAL; diagnostic views
 * are supported.  Synchronous, like add_samples. */
pt_status pt_renderer_add_samples_noalloc(pt_renderer *renderer, int32_t count);
/* cuda_renderer::set_block_shape (cuda_renderer.h:19, .cu:215-228): one of
 * 8x8, 16x8, 16x16, 32x8; otherwise PT_EINVAL. */
pt_status pt_renderer_set_block_shape(pt_renderer *renderer, int32_t width, int32_t height);
/* cuda_renderer::image (cuda_renderer.h:20, .cu:230-234): width*height*3
 * floats of sum /= float(samples), row 0 at the bottom.  The scaling runs on
 * the device (the same float products); the copy to rgb is fastest when rgb
 * is page-locked (pt_host_alloc). */
pt_status pt_renderer_read_image(const pt_renderer *renderer, float *rgb, size_t float_count);
/* cuda_renderer::rgba_image (cuda_renderer.h:21, .cu:241-244) */
pt_status pt_renderer_read_rgba(const pt_renderer *renderer, uint8_t *rgba, size_t bytes);
/* cuda_renderer::write_rgba (cuda_renderer.h:22, .cu:235-240): packs into a
 * caller-owned device buffer (e.g. a mapped GL buffer); returns after the
 * pack kernel has finished. */
pt_status pt_renderer_write_rgba_device(const pt_renderer *renderer, void *device_rgba, size_t bytes);
/* sample_count / kernel_milliseconds / rays_traced (cuda_renderer.h:23-25) */
pt_status pt_renderer_stats(const pt_renderer *renderer, int32_t *samples, float *kernel_ms, uint64_t *rays);

/* cuda_query_rays (cuda_renderer.h:32, .cu:248-260): closest hit per ray on
 * [0.001, FLT_MAX).  count may be 0. */
pt_status pt_query_rays(const pt_scene_desc *scene, const pt_ray *rays, size_t count, int32_t brute,
                        pt_trace_result *results);

/* ---- extensions without a counterpart in cuda_renderer.h ---- */

/* Traversal counters of pt_query_rays (trace_stats, transport.h:11): branch
 * nodes and leaves popped per ray, in the reference traversal order. */
pt_status pt_query_rays_stats(const pt_scene_desc *scene, const pt_ray *rays, size_t count, int32_t brute,
                              pt_trace_result *results, int32_t *boxes, int32_t *primitives);

/* Closest hit of each ray as the shaded path finds it (no counterpart in the
 * reference; for verification): the 4-wide nearest-first traversal with its
 * reference-order retrace.  Results equal pt_query_rays' bit for bit;
 * retraced[i] (optional) is 1 where the retrace ran.  count may be 0. */
pt_status pt_query_rays_wide(const pt_scene_desc *scene, const pt_ray *rays, size_t count, pt_trace_result *results,
                             int32_t *retraced);

/* Raw accumulation state, the complete resumable state of a renderer
 * (cuda_renderer.cu:145-155): sum (3 floats per pixel) and PCG state. */
pt_status pt_renderer_read_state(const pt_renderer *renderer, float *sum, uint64_t *rng_state, size_t pixels);
pt_status pt_renderer_write_state(pt_renderer *renderer, const float *sum, const uint64_t *rng_state,
                                  size_t pixels, int32_t samples, uint64_t rays);

/* Benchmark/diagnostic counters of the last add_samples: bounce iterations
 * run and HIP-event milliseconds of the closest-hit/any-hit trace kernel. */
typedef struct pt_launch_stats {
    int32_t iterations;
    int32_t trace_launches;
    float trace_ms;
    float shade_ms;
    uint64_t extend_rays;
    uint64_t shadow_rays;
    float regen_ms;          /* end-of-sample accumulation and camera-ray regeneration */
    uint64_t box_tests;      /* filled when pt_tuning.count_tests is set */
    uint64_t triangle_tests;
    uint64_t sphere_tests;
    uint64_t drained_paths;  /* paths finished on path lanes (or the drain kernel) */
    float drain_ms;
    int32_t split_launches;  /* path-lane launches run in two passes (pt_tuning.path_first_samples) */
    uint64_t retraced_queries; /* closest-hit queries of the 4-wide traversal traced again in the reference's
                                  order (order-sensitive cases); counted by the path-lane and persistent-lane
                                  trace kernels */
    /* counting pass (count_tests) on the 4-wide persistent-lane trace kernel: box_tests then counts wide
       nodes visited and triangle_tests/sphere_tests leaf records tested; these count the closest-hit queries
       that hit and the binary branch and leaf pops of the retraces */
    uint64_t hit_queries;
    uint64_t retrace_box_tests;
    uint64_t retrace_primitive_tests;
} pt_launch_stats;
pt_status pt_renderer_last_launch_stats(const pt_renderer *renderer, pt_launch_stats *stats);

/* Queue sizes of each bounce iteration of the last shaded add_samples:
 * extension[i] closest-hit queries and shadow[i] any-hit queries traced by
 * iteration i, the pixel partitions' iterations one partition after the
 * other.  Writes min(capacity, iterations) entries; *count receives the
 * number of iterations of all partitions. */
pt_status pt_renderer_iteration_log(const pt_renderer *renderer, uint32_t *extension, uint32_t *shadow,
                                    size_t capacity, int32_t *count);

/* First-hit feature buffers for the denoiser (no counterpart in the
 * reference): per pixel, the albedo (diffuse texture value, metal albedo,
 * 1 for glass, emission clamped to 1 for lights) and shading normal of each
 * camera ray's first hit, averaged over the samples like the image (0 where
 * the ray escapes).  Shaded mode only; enable right after create or reset
 * (no samples yet).  Random-number use and radiance are unchanged. */
pt_status pt_renderer_enable_aovs(pt_renderer *renderer, int32_t enable);
/* count = 3 * width * height floats per buffer; either pointer may be NULL. */
pt_status pt_renderer_read_aovs(const pt_renderer *renderer, float *albedo, float *normal, size_t count);

/* Per-renderer tuning of the wavefront loop (0 keeps the current value).
 * Defaults: 128, 256, 16, 8, 0, 0, 0, 1073741824 (every path on path lanes), 48, 1, 16, 32, 12, 4, 4, 8. */
typedef struct pt_tuning {
    int32_t trace_block;     /* threads per block of the trace kernel, multiple of 64 */
    int32_t shade_block;     /* threads per block of the shade kernel, multiple of 64 */
    int32_t blocks_per_cu;   /* persistent trace blocks per compute unit */
    int32_t check_interval;  /* bounce iterations between host checks of the live count */
    int32_t time_kernels;    /* 1: record per-kernel HIP events (trace_ms, shade_ms) */
    int32_t count_tests;     /* 1: count box/triangle/sphere tests (slower; for roofline bytes) */
    int32_t exact_traversal; /* 1: shaded path on the reference-order binary traversal instead of the
                                4-wide nearest-first one (both return the same closest hit) */
    int32_t drain_threshold; /* live paths at or below which the wavefront bounce loop hands every
                                remaining path to path lanes (the default, 2^30, does so from the
                                start; -1: never) */
    int32_t lane_refill;     /* wide traversal with persistent lanes: idle lanes of a wave that
                                trigger a claim of new queries (1..64; -1: one batch per wave) */
    int32_t partitions;      /* pixel partitions whose bounce loops run concurrently on their own
                                streams (1..8) */
    int32_t handoff_steps;   /* persistent lanes: steps a wave keeps tracing once the queues are
                                empty before it saves the queries still running and ends; the next
                                bounce iteration resumes them first and shades their paths then
                                (1..65536; -1: never) */
    int32_t path_shade_batch;  /* path lanes: lanes of a wave that wait to shade
                                  before the wave shades them together (1..64) */
    int32_t path_blocks_per_cu; /* path lanes: resident 256-thread blocks per compute unit, each lane
                                  taking pixel after pixel (-1: one lane per remaining pixel) */
    int32_t path_first_samples; /* path lanes over every pixel of a call with more samples than this:
                                  a first pass of this many samples per pixel, then the rest of the
                                  samples with the pixels ordered by the cost of their first ones,
                                  most expensive first (1..65536; -1: one pass) */
    int32_t path_sort_group;    /* that ordering moves groups of this many consecutive pixels by their
                                  summed cost, each group keeping its pixels' order (1..4096) */
    int32_t path_trace_steps;   /* path lanes: traversal steps a tracing lane takes per pass of its
                                  state machine before the wave checks for shading and new pixels
                                  (1..64) */
} pt_tuning;
pt_status pt_renderer_set_tuning(pt_renderer *renderer, const pt_tuning *tuning);

/* ---------------- batched rendering (no counterpart in the reference) ----------------
 * Many views of many scenes rendered together: the scenes' arrays are
 * merged on the device, each scene keeping its own hierarchy and lights, and
 * the pixels of all views share one schedule -- by default one persistent
 * path-lane launch, or (drain_threshold below the pixel count) the wavefront
 * bounce loop captured once into a HIP graph and replayed, its tail on path
 * lanes.  Shaded mode, 4-wide traversal.  Each view's image equals a single
 * renderer's image of it (same scene, camera, size and seed). */
typedef struct pt_batch pt_batch;
typedef struct pt_batch_view {
    uint64_t seed;           /* the view's render seed (pt_settings.seed of a single render) */
    int32_t scene;           /* index into the scene array */
    int32_t width, height;
    pt_flat_camera camera;
} pt_batch_view;             /* 104 B */
typedef struct pt_batch_trial { pt_tuning tuning; float mrays; } pt_batch_trial;

pt_status pt_batch_create(const pt_scene_desc *scenes, int32_t scene_count, const pt_batch_view *views,
                          int32_t view_count, int32_t direct_light, int32_t max_path_depth, pt_batch **out);
pt_status pt_batch_destroy(pt_batch *batch);
/* Zero every view's accumulation and reseed its PCG streams. */
pt_status pt_batch_reset(pt_batch *batch);
/* `count` more samples for every pixel of every view; synchronous. */
pt_status pt_batch_add_samples(pt_batch *batch, int32_t count);
/* Averaged linear RGB of one view (3 * width * height floats, row 0 at the bottom). */
pt_status pt_batch_read_image(const pt_batch *batch, int32_t view, float *rgb, size_t count);
pt_status pt_batch_stats(const pt_batch *batch, int32_t *samples, float *milliseconds, uint64_t *rays);
pt_status pt_batch_last_launch_stats(const pt_batch *batch, pt_launch_stats *stats);
/* Fields used: trace_block, shade_block, blocks_per_cu, check_interval (bounce
 * iterations per graph replay, rounded down to even), drain_threshold (live
 * paths at which path lanes take over; -1 never), lane_refill, handoff_steps,
 * path_shade_batch, path_blocks_per_cu, path_first_samples, path_sort_group, path_trace_steps. */
pt_status pt_batch_get_tuning(const pt_batch *batch, pt_tuning *tuning);
pt_status pt_batch_set_tuning(pt_batch *batch, const pt_tuning *tuning);
/* Autotuner: coordinate descent over workgroup sizes, persistent blocks per
 * CU, lane refill, the hand-off step budget, the drain (compaction) threshold and the graph length,
 * each candidate timed on probe_samples samples of the whole batch; `passes`
 * sweeps.  Adopts the best tuning, resets the accumulation, and reports the
 * best rate in Mrays/s and up to trial_capacity trials. */
pt_status pt_batch_autotune(pt_batch *batch, int32_t probe_samples, int32_t passes, pt_tuning *best, float *best_mrays,
                            pt_batch_trial *trials, int32_t trial_capacity, int32_t *trial_count);

#ifdef __cplusplus
}
#endif
#endif
