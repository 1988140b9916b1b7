// pt_device.h -- device-side scene view, path-state arrays and launch
// interface shared by the kernels (pt_kernels.hip) and the host runtime
// (pt_api.hip).  Internal header; the public boundary is include/pt_render.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "pt_render.h"

namespace pt {

// Light table record (src/renderer/scene_view.h:8).
struct light_rec { int32_t type, primitive; float cumulative_area; };
static_assert(sizeof(light_rec) == 12, "flat_light layout");

// Child reference used by the traversal layout: >= 0 is a branch node in the
// compact branch array; < 0 is a leaf, ~ref = primitive*2 + (sphere ? 1 : 0).
__host__ __device__ inline int32_t leaf_ref(int32_t primitive, bool sphere) { return ~(primitive * 2 + (sphere ? 1 : 0)); }

// Everything a kernel needs to read the uploaded scene.
struct scene_dev {
    // Traversal layout (built from the flat arrays at upload).
    const float4 *bnodes;      // 2 per branch: {lo.x,lo.y,lo.z,hi.x} {hi.y,hi.z,left_ref,right_ref}
    const float4 *tri_isect;   // 3 per triangle: {v0, 1e-7f*|e1 x e2|} {e1, |e1 x e2|} {e2, 0}
    const float4 *sph_isect;   // 1 per sphere: {center, radius}
    const float4 *shade_tri;   // 5 per triangle: {outward unit normal, material} {n0, smooth} {n1, uv0.x} {n2, uv0.y} {uv1.xy, uv2.xy}
    int32_t root_ref;          // -2^31 when the scene is empty
    // Original transfer arrays (shading, brute mode).
    const pt_flat_triangle *triangles;
    const pt_flat_sphere *spheres;
    const pt_flat_node *nodes;
    int32_t node_count;
    const pt_flat_material *materials;
    const pt_flat_texture *textures;
    const uint8_t *pixels;
    const light_rec *lights;
    int32_t light_count;
    float light_area;
    // cos/sin of the 2^24 angles 2*float(pi)*k*2^-24 as the host libm's
    // cosf/sinf return them: computed on the device in double and rounded,
    // except for the k flagged in trig_bits, whose host values are
    // trig_exceptions[trig_rank[k/64] + (set bits of word k/64 below k)].
    // log2_table: log2f(1+n) of the host libm for n < log2_count.
    const unsigned long long *trig_bits;
    const uint32_t *trig_rank;
    const float2 *trig_exceptions;
    const float *log2_table;
    int32_t log2_count;
    // x^5 of the host libm's pow where it differs from the correctly rounded
    // value (pt_libm.h): table-domain indices, ascending, and host values.
    const int32_t *pow5_keys;
    const double *pow5_values;
    int32_t pow5_count;
    int32_t stack_capacity;    // per-thread traversal stack entries (<= 64)
    // 4-wide layout for the shaded path (nearest-first traversal):
    //   wnodes: 8 float4 per node = lo.x[4] lo.y[4] lo.z[4] hi.x[4] hi.y[4] hi.z[4] ref[4] pad[4]
    //   leaves: 3 float4 per depth-first leaf position, the tri_isect record of
    //           its triangle with w of the third = primitive code; a sphere leaf
    //           is {center, radius} {0,0,0,-1} {0,0,0,code}
    // ref >= 0 is a wide node, wide_empty an unused slot, otherwise a leaf
    // cluster ~(first << 3 | (count - 1)) over leaf positions.
    const float4 *wnodes;
    const float4 *leaves;
    int32_t wroot_ref;         // empty_root when the scene is empty or wide traversal is off
    // Entries of the wide traversal's stack: 3 * (wide depth) + 1, and at
    // least stack_capacity, since a reference-order retrace (pt_traverse.h)
    // uses the same stack for its binary entries.
    int32_t wide_stack_capacity;
};

constexpr int32_t wide_empty = INT32_MIN + 1;
constexpr int wide_leaf_max = 2;   // leaves per leaf cluster (1, 2, 3, 4, 8 measured: 2 and 3 fastest)
constexpr int wide_lds_entries = 12;  // per-lane stack entries kept in LDS
constexpr int path_lane_block = 256;  // threads per path-lane block (the kernel's stack layout assumes it)

constexpr int32_t empty_root = INT32_MIN;
// Wide node records are addressed by 32-bit byte offsets (node * 128): a
// layout with more nodes is not built (the binary traversal runs instead).
constexpr size_t wide_node_limit = size_t(1) << 25;

// Per-pixel path state, structure of arrays in HBM.
// Traversal state of queries that a trace launch hands on to the next one
// (persistent lanes, see trace_lanes_kernel): slot k holds the query's
// pixel (~pixel for a shadow query), current node, stack depth and leaf rank
// of the closest hit so far (-1: none), that hit, and its stack entries.
struct saved_queries {
    int4 *head;         // pixel or ~pixel, cur, count, best_rank
    int4 *best;         // t, u, v (bit patterns), code
    int2 *stack;        // entry j of slot k at [k * stride + j]
};

struct paths_dev {
    float4 *ray_o;      // origin.xyz, time
    float4 *ray_d;      // direction.xyz; w = 1 while the extension query is handed on between launches
    float4 *thr;        // throughput.xyz, previous_pdf
    float4 *rad;        // radiance.xyz, unused
    int2 *ctr;          // depth, samples left in this add_samples call
    int4 *hit;          // t, u, v (bit patterns), code = primitive*2+sphere or -1
    float4 *sh_o;       // shadow origin.xyz, t_max
    float4 *sh_d;       // shadow direction.xyz; w = 1 while the shadow query is handed on
    float4 *sh_c;       // unblocked contribution.xyz, unused
    uint64_t *rng;      // PCG state; the increment is (pixel << 1) | 1
    float *sum;         // 3 floats per pixel (vec3 sum[], cuda_renderer.cu:145)
    int32_t *queue[2];  // extension queues (pixel ids; ~pixel: a path waiting for a handed-on query), ping-pong
    int32_t *shadow_queue;
    int32_t *finish_queue; // paths whose sample ended this bounce
    int2 *spill;        // wide traversal stack entries beyond wide_lds_entries
    float *aov;         // optional: 6 floats per pixel, first-hit albedo.xyz and normal.xyz summed over samples
    saved_queries saved[2]; // written by the trace launch of parity p ^ 1, resumed by the one of parity p
    int32_t save_capacity;  // slots per buffer
    int32_t save_stride;    // stack entries per slot (the wide stack capacity)
};

// Device control block of the bounce loop.
struct loop_ctl {
    uint32_t n_ext[2];
    uint32_t n_shadow;
    uint32_t n_finish;
    uint32_t work;           // persistent-fetch cursor of the trace kernel
    uint32_t n_saved[2];     // queries handed on into saved[p] (may exceed the capacity: then only that many)
    uint32_t n_waiting;      // ~pixel entries the last shade stage queued (no ray to trace)
    unsigned long long rays; // trace_world calls, shadow queries included
    unsigned long long ext_rays, shadow_rays;
    uint32_t iterations;
    int32_t failure;         // traversal stack overflow
    uint32_t *log;           // per-iteration {extension, shadow} counts, log_capacity pairs
    uint32_t log_capacity;
    // Test counts of a counting pass (trace_stats of every query, split by
    // primitive kind): box tests, triangle tests, sphere tests.
    unsigned long long box_tests, triangle_tests, sphere_tests;
    // Closest-hit queries of the 4-wide traversal traced again in the
    // reference's order (order-sensitive cases, see pt_traverse.h).
    unsigned long long retraces;
    // Counting pass of the 4-wide persistent-lane trace kernel: closest-hit
    // queries that hit, and the binary branch and leaf pops of the retraces
    // (box_tests/triangle_tests/sphere_tests then count wide nodes visited and
    // leaf records tested).
    unsigned long long hits, retrace_branches, retrace_leaves;
};

struct frame_params {
    pt_flat_camera camera;
    int32_t width, height;
    int32_t max_depth;      // path-length limit (50 in the reference)
    int32_t direct_light;
    int32_t brute;
    int32_t debug;
    int32_t exact_traversal; // shaded path on the reference-order binary traversal
};

// ---- launchers (pt_kernels.hip) ----
// Device cos/sin (double, rounded) of every angle index k in [0, 2^24).
void launch_trig_values(float2 *out, hipStream_t stream);
void launch_query(const scene_dev& s, const pt_ray *rays, pt_trace_result *out, int32_t *boxes, int32_t *prims,
                  int32_t count, bool brute, hipStream_t stream);
// Closest hit of each ray through the 4-wide traversal with its
// reference-order retrace (requires the wide layout); `spill` holds
// grid * block * wide_spill_entries entries.
void launch_wide_query(const scene_dev& s, const pt_ray *rays, pt_trace_result *out, int32_t *retraced, int2 *spill,
                       int32_t count, int block, hipStream_t stream);
// Batched rendering: many views, each of one scene of a merged scene set
// (nodes, primitives, materials and lights of all scenes concatenated; each
// scene keeps its own hierarchy root and light range).  Pixels of all views
// form one index space; each path's view gives its camera, root, lights
// and its pixel index within the view (for the PCG stream).
struct view_rec {
    pt_flat_camera camera;
    int32_t width, height;
    int64_t pixel_offset;      // first pixel of the view in the batch
    uint64_t seed;
    int32_t wroot_ref;         // the view's scene's hierarchy root
    int32_t broot_ref;         // the same scene's root in the binary layout (reference-order retraces)
    int32_t light_offset, light_count;
    float light_area;
};
struct batch_dev {
    const view_rec *views;
    const uint16_t *pixel_view;   // view of every batch pixel
};
void launch_init_pixels(float *sum, uint64_t *rng, int64_t count, uint64_t seed, loop_ctl *ctl, hipStream_t stream);
void launch_debug(const scene_dev& s, const frame_params& f, float *sum, uint64_t *rng, int32_t samples,
                  int32_t *failure, dim3 block, hipStream_t stream);
// First camera rays of one pixel partition (chunks c with c % parts == part).
void launch_start(const scene_dev& s, const frame_params& f, const paths_dev& p, loop_ctl *ctl, int32_t samples,
                  int part, int parts, int chunk, int64_t count, hipStream_t stream, const batch_dev *batch = nullptr);
void launch_clear_ctl(loop_ctl *ctl, hipStream_t stream);
void launch_trace(const scene_dev& s, const frame_params& f, const paths_dev& p, loop_ctl *ctl, int parity,
                  int grid, int block, bool count_tests, hipStream_t stream);
// Wide-layout trace stage with persistent lanes (refill: idle lanes that
// trigger a new claim, 1..64; hand_steps: steps a wave traces once the
// queues are empty before handing the queries still running on to the next
// launch, -1 never).
void launch_trace_lanes(const scene_dev& s, const frame_params& f, const paths_dev& p, loop_ctl *ctl, int parity,
                        int grid, int block, int refill, int hand_steps, bool count_tests, hipStream_t stream,
                        const batch_dev *batch = nullptr);
void launch_bookkeep(loop_ctl *ctl, int parity, hipStream_t stream);
void launch_shade(const scene_dev& s, const frame_params& f, const paths_dev& p, loop_ctl *ctl, int parity,
                  int grid, int block, hipStream_t stream, const batch_dev *batch = nullptr);
void launch_regen(const frame_params& f, const paths_dev& p, loop_ctl *ctl, int parity, int grid, int block,
                  hipStream_t stream, const batch_dev *batch = nullptr);
// Finishes the pending shadow queries (and the queries handed on by the last
// trace launch), then every queued path to its end: on path lanes with the
// 4-wide traversal, otherwise one thread per path (drain_kernel).
// `shadows`/`saved`/`paths` are the host's copies of the queue sizes.
// Path lanes: lanes of a wave waiting to shade before the wave shades them,
// the resident grid (blocks 0: one lane per path) and the traversal steps a
// tracing lane takes per pass of its state machine.
struct path_lanes { int shade_batch; int blocks; int trace_steps; };
void launch_drain(const scene_dev& s, const frame_params& f, const paths_dev& p, loop_ctl *ctl, int parity,
                  uint32_t shadows, uint32_t saved, uint32_t paths, int block, path_lanes lanes, hipStream_t stream,
                  const batch_dev *batch = nullptr);
// A path-lane launch over every pixel from a sample boundary may be split in
// two passes: PATHS_FIRST runs at most `first_samples` samples of each pixel
// and writes the pixel's cost (loop passes its lane spent on it; 0 when the
// pixel is done) at its queue position; PATHS_REST starts each queued pixel
// at its next sample.  PATHS_FRESH (single renderers) takes pixel indices
// 0 .. n-1 without a queue and runs `first_samples` samples of each from its
// stored PCG state and sum, with no other per-pixel state.  PATHS_WHOLE takes
// queued paths where they stand and
// runs them to the end.
enum path_mode : int { PATHS_WHOLE = 0, PATHS_FIRST = 1, PATHS_REST = 2, PATHS_FRESH = 3 };
struct path_pass {
    path_mode mode = PATHS_WHOLE;
    int first_samples = 0;
    uint32_t *costs = nullptr;
};
// Between the passes: groups of `group` consecutive queue positions are
// ordered by their summed cost, most expensive first, each group keeping its
// pixels' order.  launch_group_keys writes an 8-bit key and the index of
// every group (a last, partial group gets key 0, so that it stays last);
// after the (key, index) pairs are sorted in descending key order,
// launch_group_gather writes the reordered queue.
void launch_group_keys(const uint32_t *costs, uint32_t n, int group, uint8_t *keys, uint32_t *ids, hipStream_t stream);
void launch_group_gather(const int32_t *queue, const uint32_t *ids, uint32_t n, int group, int32_t *out,
                         hipStream_t stream);
// Lanes of the path-lane kernel resident per compute unit: 4 waves of 64 on
// each of the 4 SIMDs (the kernel is compiled for 4 waves per SIMD).
constexpr int path_lanes_per_cu = 1024;
// Path lanes over queue[0 .. ctl->n_ext[parity]) (`paths`: the host's copy of
// that count, for the grid).  Requires the 4-wide traversal (wide_traversal).
void launch_path_lanes(const scene_dev& s, const frame_params& f, const paths_dev& p, loop_ctl *ctl, int parity,
                       const int32_t *queue, uint32_t paths, int block, path_lanes lanes, path_pass pass,
                       hipStream_t stream, const batch_dev *batch = nullptr);
void launch_zero_work(loop_ctl *ctl, hipStream_t stream);
// Zero a control block and set n_ext[0] = n (the item count of a launch).
void launch_fresh_ctl(loop_ctl *ctl, uint32_t n, hipStream_t stream);
void launch_pack(const float *sum, uint8_t *rgba, int64_t count, int32_t samples, hipStream_t stream);
void launch_scale(const float *sum, float *out, int64_t count, float k, hipStream_t stream);
void launch_add_rays(loop_ctl *ctl, unsigned long long rays, hipStream_t stream);

// Dynamic LDS bytes the traversal kernels need per block.
inline size_t stack_bytes(int capacity, int threads) { return size_t(capacity) * size_t(threads) * sizeof(int32_t); }
// The wide traversal keeps (ref, entry distance) pairs, wide_lds_entries of them in LDS.
inline size_t wide_stack_bytes(int capacity, int threads) {
    return 2 * stack_bytes(std::min(capacity, wide_lds_entries), threads);
}
// Global overflow entries per lane of the wide traversal.
inline int wide_spill_entries(int capacity) { return std::max(0, capacity - wide_lds_entries); }
// Whether the shaded path runs on the 4-wide traversal; LDS is sized for it
// only then, so the wide kernels (path lanes) must not run otherwise.
inline bool wide_traversal(const scene_dev& s, const frame_params& f) {
    return !f.brute && s.wroot_ref != empty_root && !f.exact_traversal;
}
inline size_t trace_lds_bytes(const scene_dev& s, const frame_params& f, int threads) {
    size_t exact = stack_bytes(s.stack_capacity, threads);
    return wide_traversal(s, f) ? std::max(exact, wide_stack_bytes(s.wide_stack_capacity, threads)) : exact;
}

void launch_batch_init(const batch_dev& b, float *sum, uint64_t *rng, int64_t count, loop_ctl *ctl, hipStream_t stream);

// Device view of a renderer's accumulation for the denoiser (same library).
struct frame_view {
    const float *sum;   // 3 floats per pixel
    const float *aov;   // 6 floats per pixel or nullptr
    int32_t width, height, samples;
    hipStream_t stream;
};
bool renderer_frame(const pt_renderer *renderer, frame_view *out);

// Linear BVH of nt triangles and ns spheres (nt + ns >= 1, validated by the
// caller) built on the current device: 2 (nt + ns) - 1 flat nodes into
// nodes_out (host memory), root 0; the device time of the build in
// *milliseconds (may be null).  pt_lbvh.hip.
void build_lbvh(const pt_flat_triangle *triangles, int64_t nt, const pt_flat_sphere *spheres, int64_t ns,
                pt_flat_node *nodes_out, float *milliseconds);

} // namespace pt
