/*
 * pt_oracle.h -- CPU oracle for the flattened-BVH render path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * renderer's host transport code; it is the checker the HIP library is
 * compared against and the "port" CPU baseline.  Only tests, the smoke entry
 * and the benchmark's CPU-baseline leg may load it.  The product never links
 * or calls it.
 *
 * Every function cites the reference file:line it restates.  Record layouts
 * are byte-identical to src/flat/flat_data.h:14-40 (sizes asserted in the .c).
 */
#ifndef PT_ORACLE_H
#define PT_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y, z; } po_vec3;
typedef struct { po_vec3 origin, direction; float time; } po_ray;

typedef struct {
    po_vec3 v0, v1, v2, n0, n1, n2, uv0, uv1, uv2;
    int32_t material_id, smooth;
} po_triangle;
typedef struct { po_vec3 center; float radius; int32_t material_id; } po_sphere;
typedef struct { po_vec3 minimum, maximum; int32_t type, left, right, primitive; } po_node;
typedef struct { po_vec3 albedo; float fuzz, ref_idx; int32_t type, texture_id; } po_material;
typedef struct {
    po_vec3 color, scale, offset;
    int32_t type, image_id, width, height, byte_offset, clamp;
} po_texture;
typedef struct {
    po_vec3 origin, lower_left_corner, horizontal, vertical, u, v;
    float lens_radius, time0, time1;
} po_camera;
typedef struct {
    float t, u, v;
    po_vec3 p, normal;
    int32_t material_id, primitive_type, primitive_id;
} po_hit;
typedef struct { po_hit rec; int32_t hit, error; } po_result;
typedef struct { uint64_t state, increment; } po_rng;
typedef struct { int32_t type, primitive; float cumulative_area; } po_light;

enum { PO_BRANCH = 0, PO_TRIANGLE = 1, PO_SPHERE = 2 };
enum { PO_DIFFUSE = 0, PO_METAL = 1, PO_GLASS = 2, PO_LIGHT = 3 };
enum { PO_CONSTANT = 0, PO_IMAGE = 1, PO_MAP = 2 };
enum { PO_SHADED = 0, PO_NORMALS, PO_TRIANGLE_ID, PO_MATERIAL_ID, PO_BVH, PO_INTERSECTIONS };

/* Host arrays of one flattened scene (src/flat/flat_data.h:47-56). */
typedef struct {
    const po_triangle *triangles; int64_t triangle_count;
    const po_sphere *spheres; int64_t sphere_count;
    const po_node *nodes; int64_t node_count;
    const po_material *materials; int64_t material_count;
    const po_texture *textures; int64_t texture_count;
    const uint8_t *pixels; int64_t pixel_bytes;
    int32_t root, max_depth;
} po_scene;

/* ---- building blocks, exported for unit tests ---- */
uint32_t po_rng_next(po_rng *r);
void po_rng_seed(po_rng *r, uint64_t value, uint64_t pixel);
float po_rng_uniform(po_rng *r);
po_vec3 po_sample_ball(po_rng *r);
po_vec3 po_sample_disk(po_rng *r);
float po_mis_weight(float a, float b);
po_vec3 po_cosine_direction(po_vec3 normal, po_rng *r);
po_vec3 po_id_color(int32_t id);
int po_trace_box(const po_node *n, const po_ray *r, float low, float high);
int po_triangle_hit(const po_triangle *t, const po_ray *r, float tmin, float tmax, po_hit *out);
int po_sphere_hit(const po_sphere *s, const po_ray *r, float tmin, float tmax, po_hit *out);
void po_sphere_uv(po_vec3 p, float *u, float *v);
float po_light_direction_pdf(float area, po_vec3 from, po_vec3 point, po_vec3 normal);
int po_pack_channel(float linear);

/* Light table: returns count written (<= cap) or -1 on an invalid area. */
int64_t po_scene_lights(const po_scene *s, po_light *out, int64_t cap);
/* 0 = valid, -1 = invalid (src/flat/flat_validate.h:10-83). */
int po_validate_scene(const po_scene *s);
int po_validate_camera(const po_camera *c);

/* Closest-hit queries; boxes/primitives may be NULL (trace_stats). */
void po_query_rays(const po_scene *s, const po_ray *rays, int64_t n, int brute,
                   po_result *out, int32_t *boxes, int32_t *primitives);
/* One trace_color call (src/renderer/transport.h:169-230). */
po_vec3 po_trace_color(const po_scene *s, const po_light *lights, int64_t light_count,
                       po_ray r, po_rng *rng, int brute, int debug, int direct_light,
                       int max_depth, int *error, uint64_t *rays);

/* Progressive renderer (tests/support/portable_renderer.h:11-57). */
typedef struct po_renderer po_renderer;
po_renderer *po_renderer_create(const po_scene *s, const po_camera *c, int width, int height,
                                uint64_t seed, int brute, int debug, int direct_light, int max_depth);
void po_renderer_destroy(po_renderer *r);
int po_renderer_reset(po_renderer *r, const po_camera *c);
/* Returns 0, -1 for an invalid count, -2 for a traversal stack overflow. */
int po_renderer_add_samples(po_renderer *r, int count, int threads);
/* Row range variant used to time a bounded slice: renders rows [y0,y1) only. */
int po_renderer_add_samples_rows(po_renderer *r, int count, int threads, int y0, int y1);
void po_renderer_image(const po_renderer *r, float *rgb);
void po_renderer_sum(const po_renderer *r, float *rgb);
void po_renderer_rng(const po_renderer *r, uint64_t *state);
int po_renderer_sample_count(const po_renderer *r);
uint64_t po_renderer_rays(const po_renderer *r);
void po_renderer_rgba(const po_renderer *r, uint8_t *rgba);

#ifdef __cplusplus
}
#endif
#endif
