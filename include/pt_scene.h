/*
 * pt_scene.h -- host-side scene assembly for the path-tracing core: a BVH
 * builder that emits the reference's flat transfer format, the camera basis,
 * and seeded synthetic scenes used by the tests and the benchmark.
 *
 * The builder replaces the reference's host BVH construction and export
 * (src/bvh/bvh.h:132-171 bvh_node, src/flat/flat_export.h:121-159
 * flat_exporter::object): nodes are emitted depth-first, left subtree before
 * right, one primitive per leaf, triangles and spheres renumbered in leaf
 * order, leaf bounds as the reference computes them (padded triangle boxes,
 * src/geometry/triangle.h:96-109; sphere boxes, src/geometry/sphere.h:69-75)
 * and branch bounds as the union of the children (src/bvh/aabb.h:40-56).
 * Unlike the reference's random-axis median split it uses binned SAH, with a
 * balanced median split wherever the 64-level depth limit could be reached.
 */
#ifndef PT_SCENE_H
#define PT_SCENE_H

#include "pt_render.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pts_scene pts_scene;

enum pts_split { PTS_SPLIT_SAH = 0, PTS_SPLIT_MEDIAN = 1, PTS_SPLIT_REFERENCE = 2 };

/* Build a flat scene (nodes + leaf-ordered primitives) over the given
 * primitives.  Materials, textures and pixels are copied unchanged.
 * Returns NULL and sets pt_scene_last_error() on invalid input. */
pts_scene *pts_build(const pt_flat_triangle *triangles, int64_t triangle_count,
                     const pt_flat_sphere *spheres, int64_t sphere_count,
                     const pt_flat_material *materials, int64_t material_count,
                     const pt_flat_texture *textures, int64_t texture_count,
                     const uint8_t *pixels, int64_t pixel_bytes, int32_t split);

/* pts_build with a seed for PTS_SPLIT_REFERENCE: the reference's own
 * hierarchy (src/bvh/bvh.h:132-171 bvh_node over [triangles..., spheres...]:
 * an axis drawn by drand48 after srand48(seed), objects stably sorted on
 * their boxes' minimum along it, halves at n/2, a lone object aliased as both
 * children), exported like src/flat/flat_export.h:85-123 (nodes depth first,
 * materials and textures renumbered in first-encounter order, images before
 * the maps that read them).  Other split methods ignore the seed. */
pts_scene *pts_build_seeded(const pt_flat_triangle *triangles, int64_t triangle_count,
                            const pt_flat_sphere *spheres, int64_t sphere_count,
                            const pt_flat_material *materials, int64_t material_count,
                            const pt_flat_texture *textures, int64_t texture_count,
                            const uint8_t *pixels, int64_t pixel_bytes, int32_t split, int64_t seed);

/* OBJ/MTL import (src/scene/obj_loader.h:120-237) staged as the reference's
 * obj_scene (src/scene/obj_scene.h:10-31): the mesh after the transform
 * (src/transforms/mesh_transform.h: scale, X/Y/Z rotation in degrees, then
 * translation), a floor sphere and a light sphere.  PTS_SPLIT_REFERENCE with
 * seed s gives the arrays the reference's srand48(s); obj_scene;
 * flatten_scene gives.  Textures: PNG and binary PNM. */
typedef struct pts_obj_options {
    float scale[3], rotate[3], translate[3];
    float aspect;       /* width / height, for the camera distance */
    int32_t split;      /* enum pts_split */
    int64_t seed;       /* srand48 seed of PTS_SPLIT_REFERENCE */
} pts_obj_options;
typedef struct pts_obj_info {
    pt_vec3 lookfrom, lookat;  /* camera placement of obj_scene (up (0,1,0), vfov 40) */
    float focus_dist;
    int64_t source_vertices, shape_count, mesh_triangles;
} pts_obj_info;
pts_scene *pts_load_obj(const char *filename, const pts_obj_options *options, pts_obj_info *info);
/* Loader warnings of a scene from pts_load_obj ("" otherwise). */
const char *pts_warnings(const pts_scene *scene);

/* Decode a PNG, binary PNM or JPEG (baseline or progressive) file to 8-bit
 * RGB rows, top row first, as the reference's stbi_load(..., STBI_rgb)
 * returns them (src/textures/texture.h:65-72); free with pts_free_image.
 * Returns 0, or -1 with pt_scene_last_error() set. */
int32_t pts_load_image(const char *path, int32_t *width, int32_t *height, uint8_t **rgb);
void pts_free_image(uint8_t *rgb);

/* Write a linear RGB image (row 0 at the bottom, as pt_renderer_read_image
 * returns it) as the reference application's ASCII PPM (src/app/main.cpp:40-46,
 * channels through src/display/color.h:6-11).  path "-" writes to stdout.
 * Returns 0, or -1 with pt_scene_last_error() set. */
int32_t pts_write_ppm(const char *path, const float *rgb, int32_t width, int32_t height);

/* Cornell-type room: 5 walls (10 triangles), a 2-triangle ceiling light of
 * emission 15 and three seeded spheres (Lambertian, metal, glass). */
pts_scene *pts_cornell(uint64_t seed, int32_t width, int32_t height, pt_flat_camera *camera);

/* Geodesic icosphere of 20*frequency^2 triangles, radially displaced by
 * 3-octave seeded value noise (amplitude 0.05 of the radius), smooth
 * per-vertex normals, per-cluster materials (60% diffuse, 25% metal, 15%
 * glass), a 2-triangle floor and 4 rectangular lights (emission U(5,20)). */
pts_scene *pts_icosphere(int32_t frequency, uint64_t seed, int32_t width, int32_t height, pt_flat_camera *camera);

/* Seeded soup of Gaussian-clustered triangles with a floor and 2 lights. */
pts_scene *pts_soup(int64_t triangle_count, uint64_t seed, int32_t width, int32_t height, pt_flat_camera *camera);

/* Array sizes: triangles, spheres, nodes, materials, textures, pixel bytes;
 * root_depth = {root, max_depth}. */
void pts_counts(const pts_scene *scene, int64_t counts[6], int32_t root_depth[2]);
/* Copy the arrays out (each pointer may be NULL to skip that array). */
void pts_copy(const pts_scene *scene, pt_flat_triangle *triangles, pt_flat_sphere *spheres,
              pt_flat_node *nodes, pt_flat_material *materials, pt_flat_texture *textures, uint8_t *pixels);
/* View of the scene's arrays, valid while the scene lives. */
void pts_desc(const pts_scene *scene, pt_scene_desc *desc);
void pts_free(pts_scene *scene);

/* Camera basis of the reference's camera constructor (src/core/camera.h:16-43). */
void pts_camera(const float lookfrom[3], const float lookat[3], const float vup[3], float vfov, float aspect,
                float aperture, float focus_dist, float time0, float time1, pt_flat_camera *out);

const char *pt_scene_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
