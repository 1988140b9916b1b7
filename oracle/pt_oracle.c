/*
 * pt_oracle.c -- CPU oracle (restatement) of the reference's render path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the "port"
 * CPU baseline.  Never linked into the product.
 *
 * Built with gcc -O3 -ffp-contract=off and no -march, mirroring the reference
 * Release build, which compiles without FMA contraction (CMakeLists.txt:19-20).
 * Every float expression keeps the reference's association order and its
 * float/double promotions; the comment on each function names the lines.
 * Transcendentals are glibc's, as in the reference's host build.
 *
 * Parity pins (see tests/test_oracle_*.py): PCG known-answer stream, analytic
 * primitive cases, the pruning/tie scene, MIS quadrature, RGBA bytes, and
 * bit-equality with the compiled reference (oracle/_ref) on golden scenes.
 */
#define _DEFAULT_SOURCE 1
#include "pt_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

_Static_assert(sizeof(po_vec3) == 12, "vec3 layout");
_Static_assert(sizeof(po_ray) == 28, "ray layout");
_Static_assert(sizeof(po_triangle) == 116, "flat_triangle layout");
_Static_assert(sizeof(po_sphere) == 20, "flat_sphere layout");
_Static_assert(sizeof(po_node) == 40, "flat_node layout");
_Static_assert(sizeof(po_material) == 28, "flat_material layout");
_Static_assert(sizeof(po_texture) == 60, "flat_texture layout");
_Static_assert(sizeof(po_camera) == 84, "flat_camera layout");
_Static_assert(sizeof(po_hit) == 48, "flat_hit layout");
_Static_assert(sizeof(po_result) == 56, "trace_result layout");
_Static_assert(sizeof(po_light) == 12, "flat_light layout");

#define STACK_CAPACITY 64 /* src/flat/flat_data.h:52 */
static const float PI_F = (float)M_PI;

/* ---------------- vec3 (src/core/vec3.h:49-162) ---------------- */
static inline po_vec3 v3(float x, float y, float z) { po_vec3 r = {x, y, z}; return r; }
static inline po_vec3 add(po_vec3 a, po_vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline po_vec3 sub(po_vec3 a, po_vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline po_vec3 mul(po_vec3 a, po_vec3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
/* operator*(float,vec3) and operator*(vec3,float) both compute t*e[i] (vec3.h:93-103). */
static inline po_vec3 scale(float t, po_vec3 v) { return v3(t * v.x, t * v.y, t * v.z); }
/* operator/(vec3,float): per-component float division (vec3.h:97-99). */
static inline po_vec3 divs(po_vec3 v, float t) { return v3(v.x / t, v.y / t, v.z / t); }
static inline po_vec3 neg(po_vec3 v) { return v3(-v.x, -v.y, -v.z); }
static inline float dot(po_vec3 a, po_vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* cross keeps the negated middle difference (vec3.h:109-115). */
static inline po_vec3 cross(po_vec3 a, po_vec3 b) {
    return v3(a.y * b.z - a.z * b.y, -(a.x * b.z - a.z * b.x), a.x * b.y - a.y * b.x);
}
static inline float length(po_vec3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
static inline float squared_length(po_vec3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
/* operator/=(float): multiply by the double reciprocal rounded to float (vec3.h:152-158). */
static inline po_vec3 div_assign(po_vec3 v, float t) {
    float k = (float)(1.0 / t);
    return v3(v.x * k, v.y * k, v.z * k);
}
static inline po_vec3 unit_vector(po_vec3 v) { return divs(v, length(v)); }
static inline float comp(po_vec3 v, int axis) { return axis == 0 ? v.x : axis == 1 ? v.y : v.z; }
/* ray::point_at_parameter = A + t*B (src/core/ray.h:26-28). */
static inline po_vec3 point_at(const po_ray *r, float t) { return add(r->origin, scale(t, r->direction)); }
static inline int finite3(po_vec3 v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }

/* ---------------- PCG stream (src/renderer/random.h:8-23) ---------------- */
uint32_t po_rng_next(po_rng *r) {
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->increment;
    uint32_t shifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rotation = (uint32_t)(old >> 59u);
    return (shifted >> rotation) | (shifted << ((-rotation) & 31u));
}
void po_rng_seed(po_rng *r, uint64_t value, uint64_t pixel) {
    r->state = 0; r->increment = (pixel << 1u) | 1u;
    po_rng_next(r); r->state += value; po_rng_next(r);
}
float po_rng_uniform(po_rng *r) { return (float)(po_rng_next(r) >> 8) * 0x1p-24f; }

/* random.h:25-33: three draws per try, accept squared length < 1. */
po_vec3 po_sample_ball(po_rng *r) {
    po_vec3 p;
    do {
        float x = po_rng_uniform(r), y = po_rng_uniform(r), z = po_rng_uniform(r);
        p = sub(scale(2.0f, v3(x, y, z)), v3(1, 1, 1));
    } while (squared_length(p) >= 1.0f);
    return p;
}
/* random.h:35-42 */
po_vec3 po_sample_disk(po_rng *r) {
    po_vec3 p;
    do {
        float x = po_rng_uniform(r), y = po_rng_uniform(r);
        p = sub(scale(2.0f, v3(x, y, 0)), v3(1, 1, 0));
    } while (squared_length(p) >= 1.0f);
    return p;
}

/* ---------------- light sampling (src/renderer/light_sampling.h) ---------------- */
/* light_sampling.h:10-16 */
float po_mis_weight(float a, float b) {
    float s = fmaxf(a, b);
    if (!(s > 0)) return 0;
    a /= s; b /= s;
    return a * a / (a * a + b * b);
}
/* light_sampling.h:18-25 */
po_vec3 po_cosine_direction(po_vec3 normal, po_rng *r) {
    float u = po_rng_uniform(r), v = po_rng_uniform(r);
    float radius = sqrtf(u), angle = 2 * PI_F * v;
    po_vec3 helper = fabsf(normal.x) > 0.9f ? v3(0, 1, 0) : v3(1, 0, 0);
    po_vec3 tangent = unit_vector(cross(helper, normal));
    po_vec3 bitangent = cross(normal, tangent);
    po_vec3 a = scale(radius * cosf(angle), tangent);
    po_vec3 b = scale(radius * sinf(angle), bitangent);
    return add(add(a, b), scale(sqrtf(1 - u), normal));
}

/* src/geometry/sphere_uv.h:7-13: float atan2/asin, then double arithmetic. */
void po_sphere_uv(po_vec3 p, float *u, float *v) {
    float phi = atan2f(p.z, p.x);
    float theta = asinf(p.y);
    *u = (float)(1 - (phi + M_PI) / (2 * M_PI));
    *v = (float)((theta + M_PI / 2) / M_PI);
}

typedef struct { po_vec3 point, normal; float u, v; int32_t material; } light_sample;

typedef struct {
    const po_scene *s;
    const po_light *lights;
    int64_t light_count;
    float light_area;
} view;

/* light_sampling.h:27-54 */
static light_sample sample_light(const view *sc, po_rng *r) {
    float choice = po_rng_uniform(r) * sc->light_area;
    int low = 0, high = (int)sc->light_count - 1;
    while (low < high) {
        int middle = low + (high - low) / 2;
        if (choice < sc->lights[middle].cumulative_area) high = middle;
        else low = middle + 1;
    }
    const po_light *light = &sc->lights[low];
    float u = po_rng_uniform(r), v = po_rng_uniform(r);
    light_sample res;
    if (light->type == PO_TRIANGLE) {
        const po_triangle *t = &sc->s->triangles[light->primitive];
        float root = sqrtf(u), a = 1 - root, b = root * (1 - v), c = root * v;
        res.point = add(add(scale(a, t->v0), scale(b, t->v1)), scale(c, t->v2));
        res.normal = unit_vector(cross(sub(t->v1, t->v0), sub(t->v2, t->v0)));
        po_vec3 uv = add(add(scale(a, t->uv0), scale(b, t->uv1)), scale(c, t->uv2));
        res.u = uv.x; res.v = uv.y; res.material = t->material_id;
    } else {
        const po_sphere *sp = &sc->s->spheres[light->primitive];
        float z = 1 - 2 * u, radius = sqrtf(fmaxf(0, 1 - z * z)), angle = 2 * PI_F * v;
        res.normal = v3(radius * cosf(angle), z, radius * sinf(angle));
        res.point = add(sp->center, scale(fabsf(sp->radius), res.normal));
        if (sp->radius < 0) res.normal = neg(res.normal);
        po_sphere_uv(res.normal, &res.u, &res.v);
        res.material = sp->material_id;
    }
    return res;
}

/* light_sampling.h:56-64 */
float po_light_direction_pdf(float area, po_vec3 from, po_vec3 point, po_vec3 normal) {
    po_vec3 delta = sub(point, from);
    float d2 = squared_length(delta);
    if (!(area > 0) || !(d2 > 0)) return 0;
    float cosine = fabsf(dot(normal, delta)) / sqrtf(d2);
    if (!(cosine > 0)) return 0;
    float pdf = d2 / (cosine * area);
    return isfinite(pdf) ? pdf : 0;
}

/* light_sampling.h:66-73 */
static float hit_light_pdf(const view *sc, po_vec3 from, const po_hit *hit) {
    po_vec3 normal = hit->normal;
    if (hit->primitive_type == PO_TRIANGLE) {
        const po_triangle *t = &sc->s->triangles[hit->primitive_id];
        normal = unit_vector(cross(sub(t->v1, t->v0), sub(t->v2, t->v0)));
    }
    return po_light_direction_pdf(sc->light_area, from, hit->p, normal);
}

/* ---------------- primitive tests (src/flat/flat_intersect.h) ---------------- */
/* flat_intersect.h:8-43: Moller-Trumbore without culling. */
int po_triangle_hit(const po_triangle *face, const po_ray *r, float t_min, float t_max, po_hit *rec) {
    po_vec3 edge1 = sub(face->v1, face->v0);
    po_vec3 edge2 = sub(face->v2, face->v0);
    po_vec3 outward = cross(edge1, edge2);
    float normal_length = length(outward);
    float direction_length = length(r->direction);
    if (!(normal_length > 0) || !(direction_length > 0)) return 0;
    po_vec3 pvec = cross(r->direction, edge2);
    float det = dot(edge1, pvec);
    if (fabsf(det) <= 0.0000001f * normal_length * direction_length) return 0;
    float inv_det = 1.0f / det;
    po_vec3 tvec = sub(r->origin, face->v0);
    float u = dot(tvec, pvec) * inv_det;
    if (!(u >= 0 && u <= 1)) return 0;
    po_vec3 qvec = cross(tvec, edge1);
    float v = dot(r->direction, qvec) * inv_det;
    if (!(v >= 0 && u + v <= 1)) return 0;
    float t = dot(edge2, qvec) * inv_det;
    if (!(t > t_min && t < t_max)) return 0;
    outward = div_assign(outward, normal_length);
    po_vec3 normal = outward;
    if (face->smooth) {
        float w = 1 - u - v;
        po_vec3 interpolated = add(add(scale(w, face->n0), scale(u, face->n1)), scale(v, face->n2));
        float ls = squared_length(interpolated);
        if (ls > 0 && isfinite(ls)) {
            normal = unit_vector(interpolated);
            if (dot(normal, outward) < 0) normal = neg(normal);
            if (dot(normal, outward) == 0) normal = outward;
        }
    }
    rec->t = t; rec->p = point_at(r, t); rec->normal = normal;
    float w = 1 - u - v;
    po_vec3 uv = add(add(scale(w, face->uv0), scale(u, face->uv1)), scale(v, face->uv2));
    rec->u = uv.x; rec->v = uv.y; rec->material_id = face->material_id;
    return 1;
}

/* flat_intersect.h:49-67 */
int po_sphere_hit(const po_sphere *sp, const po_ray *r, float t_min, float t_max, po_hit *rec) {
    po_vec3 oc = sub(r->origin, sp->center);
    float a = dot(r->direction, r->direction);
    float b = dot(oc, r->direction);
    float c = dot(oc, oc) - sp->radius * sp->radius;
    float disc = b * b - a * c;
    if (disc > 0) {
        float t = (-b - sqrtf(disc)) / a;
        if (!(t < t_max && t > t_min)) t = (-b + sqrtf(disc)) / a;
        if (t < t_max && t > t_min) {
            rec->t = t; rec->p = point_at(r, t);
            rec->normal = divs(sub(rec->p, sp->center), sp->radius);
            po_sphere_uv(rec->normal, &rec->u, &rec->v);
            rec->material_id = sp->material_id;
            return 1;
        }
    }
    return 0;
}

/* ---------------- traversal (src/renderer/transport.h:14-85) ---------------- */
/* transport.h:14-31 */
int po_trace_box(const po_node *n, const po_ray *r, float low, float high) {
    for (int axis = 0; axis < 3; axis++) {
        float d = comp(r->direction, axis);
        float o = comp(r->origin, axis);
        if (d == 0) {
            if (o < comp(n->minimum, axis) || o > comp(n->maximum, axis)) return 0;
            continue;
        }
        float inverse = 1.0f / d;
        float a = (comp(n->minimum, axis) - o) * inverse;
        float b = (comp(n->maximum, axis) - o) * inverse;
        if (inverse < 0) { float swap = a; a = b; b = swap; }
        low = a > low ? a : low;
        high = b < high ? b : high;
        if (high <= low) return 0;
    }
    return high > low;
}

/* transport.h:33-41 */
static int trace_leaf(const po_scene *s, const po_node *n, const po_ray *r, float low, float high, po_hit *hit) {
    int found;
    if (n->type == PO_TRIANGLE) found = po_triangle_hit(&s->triangles[n->primitive], r, low, high, hit);
    else found = po_sphere_hit(&s->spheres[n->primitive], r, low, high, hit);
    if (found) { hit->primitive_type = n->type; hit->primitive_id = n->primitive; }
    return found;
}

/* transport.h:43-85 */
static int trace_world(const po_scene *s, const po_ray *r, float low, float high, po_hit *hit,
                       int brute, int *error, int32_t *boxes, int32_t *primitives) {
    if (s->root < 0) return 0;
    int found = 0;
    float closest = high;
    if (brute) {
        for (int64_t i = 0; i < s->node_count; i++) {
            const po_node *n = &s->nodes[i];
            if (n->type == PO_BRANCH) continue;
            (*primitives)++;
            po_hit candidate;
            if (trace_leaf(s, n, r, low, high, &candidate) && (!found || candidate.t <= closest)) {
                *hit = candidate; closest = candidate.t; found = 1;
            }
        }
        return found;
    }
    int stack[STACK_CAPACITY], count = 0;
    stack[count++] = s->root;
    while (count) {
        const po_node *n = &s->nodes[stack[--count]];
        if (n->type == PO_BRANCH) {
            (*boxes)++;
            float limit = found ? nextafterf(closest, high) : high;
            if (!po_trace_box(n, r, low, limit)) continue;
            int needed = n->left == n->right ? 1 : 2;
            if (count + needed > STACK_CAPACITY) { *error = 1; return 0; }
            if (n->left != n->right) stack[count++] = n->right;
            stack[count++] = n->left;
        } else {
            (*primitives)++;
            po_hit candidate;
            float limit = found ? nextafterf(closest, high) : high;
            if (trace_leaf(s, n, r, low, limit, &candidate) && (!found || candidate.t <= closest)) {
                *hit = candidate; closest = candidate.t; found = 1;
            }
        }
    }
    return found;
}

void po_query_rays(const po_scene *s, const po_ray *rays, int64_t n, int brute,
                   po_result *out, int32_t *boxes, int32_t *primitives) {
    /* src/cuda/cuda_renderer.cu:111-120: trace_result{} then trace_world(0.001, FLT_MAX). */
    for (int64_t i = 0; i < n; i++) {
        po_result res;
        memset(&res, 0, sizeof res);
        res.rec.material_id = res.rec.primitive_type = res.rec.primitive_id = -1;
        int error = 0; int32_t b = 0, p = 0;
        res.hit = trace_world(s, &rays[i], 0.001f, FLT_MAX, &res.rec, brute, &error, &b, &p);
        res.error = error;
        out[i] = res;
        if (boxes) boxes[i] = b;
        if (primitives) primitives[i] = p;
    }
}

/* ---------------- textures and scattering (transport.h:87-155) ---------------- */
/* transport.h:87-101: nearest texel, V flipped, double offsets and byte scale. */
static po_vec3 trace_image(const po_scene *s, const po_texture *tex, float u, float v) {
    u = fminf(1.0f, fmaxf(0.0f, u));
    v = fminf(1.0f, fmaxf(0.0f, v));
    int i = (int)(u * (float)tex->width);
    int j = (int)((double)((1 - v) * (float)tex->height) - 0.001);
    if (i < 0) i = 0;
    if (j < 0) j = 0;
    if (i > tex->width - 1) i = tex->width - 1;
    if (j > tex->height - 1) j = tex->height - 1;
    size_t offset = (size_t)tex->byte_offset + 3 * (size_t)i + 3 * (size_t)tex->width * (size_t)j;
    float red = (float)((int)s->pixels[offset] / 255.0);
    float green = (float)((int)s->pixels[offset + 1] / 255.0);
    float blue = (float)((int)s->pixels[offset + 2] / 255.0);
    return v3(red, green, blue);
}
/* transport.h:103-111 */
static po_vec3 trace_texture(const po_scene *s, int32_t id, float u, float v) {
    const po_texture *tex = &s->textures[id];
    if (tex->type == PO_CONSTANT) return tex->color;
    if (tex->type == PO_IMAGE) return trace_image(s, tex, u, v);
    u = u * tex->scale.x + tex->offset.x;
    v = v * tex->scale.y + tex->offset.y;
    if (!tex->clamp) { u -= floorf(u); v -= floorf(v); }
    return mul(tex->color, trace_image(s, &s->textures[tex->image_id], u, v));
}
/* transport.h:113 */
static po_vec3 reflect(po_vec3 v, po_vec3 n) { return sub(v, scale(2 * dot(v, n), n)); }

/* transport.h:115-155 */
static int trace_scatter(const po_scene *s, const po_material *mat, const po_ray *in, const po_hit *rec,
                         po_rng *rng, po_vec3 *attenuation, po_ray *scattered) {
    if (mat->type == PO_DIFFUSE) {
        po_vec3 target = add(add(rec->p, rec->normal), po_sample_ball(rng));
        scattered->origin = rec->p; scattered->direction = sub(target, rec->p); scattered->time = in->time;
        *attenuation = trace_texture(s, mat->texture_id, rec->u, rec->v);
        return 1;
    }
    if (mat->type == PO_METAL) {
        po_vec3 reflected = reflect(unit_vector(in->direction), rec->normal);
        po_vec3 fuzz = scale(mat->fuzz, po_sample_ball(rng));
        scattered->origin = rec->p; scattered->direction = add(reflected, fuzz); scattered->time = in->time;
        *attenuation = mat->albedo;
        return dot(scattered->direction, rec->normal) > 0;
    }
    if (mat->type == PO_GLASS) {
        po_vec3 outward, refracted = v3(0, 0, 0);
        po_vec3 reflected = reflect(in->direction, rec->normal);
        float ratio, cosine;
        if (dot(in->direction, rec->normal) > 0) {
            outward = neg(rec->normal); ratio = mat->ref_idx;
            cosine = mat->ref_idx * dot(in->direction, rec->normal) / length(in->direction);
        } else {
            outward = rec->normal; ratio = 1.0f / mat->ref_idx;
            cosine = -dot(in->direction, rec->normal) / length(in->direction);
        }
        po_vec3 uv = unit_vector(in->direction);
        float dt = dot(uv, outward);
        float discriminant = 1.0f - ratio * ratio * (1 - dt * dt);
        float probability = 1;
        if (discriminant > 0) {
            refracted = sub(scale(ratio, sub(uv, scale(dt, outward))), scale(sqrtf(discriminant), outward));
            float r0 = (1 - mat->ref_idx) / (1 + mat->ref_idx);
            r0 *= r0;
            /* Host pow(float,int) promotes to double (transport.h:147). */
            probability = (float)((double)r0 + (double)(1 - r0) * pow((double)(1 - cosine), 5.0));
        }
        scattered->origin = rec->p; scattered->time = in->time;
        if (po_rng_uniform(rng) < probability) scattered->direction = reflected;
        else scattered->direction = refracted;
        *attenuation = v3(1, 1, 1);
        return 1;
    }
    return 0;
}

/* transport.h:157-162 */
static po_ray trace_camera(const po_camera *c, float u, float v, po_rng *rng) {
    po_vec3 disk = scale(c->lens_radius, po_sample_disk(rng));
    po_vec3 offset = add(scale(disk.x, c->u), scale(disk.y, c->v));
    float time = c->time0 + po_rng_uniform(rng) * (c->time1 - c->time0);
    po_ray r;
    r.origin = add(c->origin, offset);
    r.direction = sub(sub(add(add(c->lower_left_corner, scale(u, c->horizontal)), scale(v, c->vertical)), c->origin), offset);
    r.time = time;
    return r;
}

/* transport.h:164-167 */
po_vec3 po_id_color(int32_t id) {
    uint32_t bits = ((uint32_t)id + 1u) * 2654435761u;
    return v3((float)(bits & 255) / 255.0f, (float)((bits >> 8) & 255) / 255.0f, (float)((bits >> 16) & 255) / 255.0f);
}

/* transport.h:169-230, with the path-length limit 50 lifted to max_depth. */
static po_vec3 trace_color(const view *sc, po_ray current, po_rng *rng, int brute, int debug,
                           int *error, int direct_light, int max_depth, uint64_t *rays) {
    const po_scene *s = sc->s;
    po_vec3 throughput = v3(1, 1, 1), radiance = v3(0, 0, 0);
    float previous_pdf = 0;
    for (int depth = 0; depth <= max_depth; depth++) {
        po_hit rec;
        int32_t boxes = 0, prims = 0;
        if (rays) ++*rays;
        int hit = trace_world(s, &current, 0.001f, FLT_MAX, &rec, brute, error, &boxes, &prims);
        if (debug == PO_BVH || debug == PO_INTERSECTIONS) {
            float cost = log2f(1.0f + (float)(debug == PO_BVH ? boxes : prims)) / 10.0f;
            cost = fminf(1, cost);
            return v3(cost, cost * cost, 1 - cost);
        }
        if (!hit || *error) return radiance;
        if (debug == PO_NORMALS) return scale(0.5f, add(rec.normal, v3(1, 1, 1)));
        if (debug == PO_TRIANGLE_ID)
            return rec.primitive_type == PO_TRIANGLE ? po_id_color(rec.primitive_id) : v3((float)0.3, (float)0.3, (float)0.3);
        if (debug == PO_MATERIAL_ID) return po_id_color(rec.material_id);
        const po_material *mat = &s->materials[rec.material_id];
        if (mat->type == PO_LIGHT) {
            float weight = previous_pdf > 0 ? po_mis_weight(previous_pdf, hit_light_pdf(sc, current.origin, &rec)) : 1;
            return add(radiance, mul(scale(weight, throughput), trace_texture(s, mat->texture_id, rec.u, rec.v)));
        }
        if (depth == max_depth) break;
        po_ray scattered; po_vec3 attenuation;
        if (direct_light && mat->type == PO_DIFFUSE) {
            attenuation = trace_texture(s, mat->texture_id, rec.u, rec.v);
            if (sc->light_count) {
                light_sample light = sample_light(sc, rng);
                po_vec3 delta = sub(light.point, rec.p);
                float distance = length(delta);
                if (distance > 0.002f && isfinite(distance)) {
                    po_vec3 direction = divs(delta, distance);
                    float cosine = fmaxf(0, dot(rec.normal, direction));
                    float pdf = po_light_direction_pdf(sc->light_area, rec.p, light.point, light.normal);
                    if (cosine > 0 && pdf > 0) {
                        po_hit blocker; int32_t sb = 0, sp = 0;
                        if (rays) ++*rays;
                        po_ray shadow = {rec.p, direction, current.time};
                        int blocked = trace_world(s, &shadow, 0.001f, distance - 0.001f, &blocker, brute, error, &sb, &sp);
                        if (!blocked && !*error) {
                            const po_material *emitter = &s->materials[light.material];
                            po_vec3 emission = trace_texture(s, emitter->texture_id, light.u, light.v);
                            float bsdf_pdf = cosine / PI_F;
                            float w = bsdf_pdf * po_mis_weight(pdf, bsdf_pdf) / pdf;
                            radiance = add(radiance, scale(w, mul(mul(throughput, attenuation), emission)));
                        }
                    }
                }
            }
            po_vec3 direction = po_cosine_direction(rec.normal, rng);
            previous_pdf = fmaxf(0, dot(rec.normal, direction)) / PI_F;
            throughput = mul(throughput, attenuation);
            current.origin = rec.p; current.direction = direction;
            continue;
        }
        previous_pdf = 0;
        if (!trace_scatter(s, mat, &current, &rec, rng, &attenuation, &scattered)) break;
        throughput = mul(throughput, attenuation);
        current = scattered;
    }
    return radiance;
}

static po_vec3 trace_pixel_sample(const view *sc, const po_camera *c, int x, int y, int width, int height,
                                  po_rng *rng, int brute, int debug, int *error, int direct_light,
                                  int max_depth, uint64_t *rays) {
    /* transport.h:232-239 */
    float u = ((float)x + po_rng_uniform(rng)) / (float)width;
    float v = ((float)y + po_rng_uniform(rng)) / (float)height;
    po_ray r = trace_camera(c, u, v, rng);
    return trace_color(sc, r, rng, brute, debug, error, direct_light, max_depth, rays);
}

/* ---------------- light table (src/renderer/scene_view.h:10-28) ---------------- */
int64_t po_scene_lights(const po_scene *s, po_light *out, int64_t cap) {
    double area = 0;
    int64_t n = 0;
    for (int64_t i = 0; i < s->triangle_count; i++) {
        const po_triangle *t = &s->triangles[i];
        if (s->materials[t->material_id].type != PO_LIGHT) continue;
        double size = 0.5 * length(cross(sub(t->v1, t->v0), sub(t->v2, t->v0)));
        if (!(size > 0) || !isfinite(size)) continue;
        area += size;
        if (!isfinite(area) || area > FLT_MAX) return -1;
        if (n < cap) { out[n].type = PO_TRIANGLE; out[n].primitive = (int32_t)i; out[n].cumulative_area = (float)area; }
        n++;
    }
    for (int64_t i = 0; i < s->sphere_count; i++) {
        const po_sphere *p = &s->spheres[i];
        if (s->materials[p->material_id].type != PO_LIGHT) continue;
        double size = 4 * M_PI * (double)p->radius * p->radius;
        if (!(size > 0) || !isfinite(size)) continue;
        area += size;
        if (!isfinite(area) || area > FLT_MAX) return -1;
        if (n < cap) { out[n].type = PO_SPHERE; out[n].primitive = (int32_t)i; out[n].cumulative_area = (float)area; }
        n++;
    }
    return n;
}

/* ---------------- validation (src/flat/flat_validate.h:10-83) ---------------- */
typedef struct {
    const po_scene *s;
    int *heights;
    char *active;
    int64_t visited;
    int bad;
} vstate;

static int valid_id(int32_t id, int64_t count) { return id >= 0 && (int64_t)id < count; }

static int visit(vstate *st, int32_t id, int depth) {
    const po_scene *s = st->s;
    if (!(valid_id(id, s->node_count) && depth <= STACK_CAPACITY && !st->active[id])) { st->bad = 1; return 0; }
    if (st->heights[id]) {
        if (!(depth + st->heights[id] - 1 <= STACK_CAPACITY)) st->bad = 1;
        return st->heights[id];
    }
    st->active[id] = 1;
    st->visited++;
    const po_node *n = &s->nodes[id];
    if (!(finite3(n->minimum) && finite3(n->maximum))) { st->bad = 1; return 0; }
    for (int a = 0; a < 3; a++) if (!(comp(n->minimum, a) <= comp(n->maximum, a))) { st->bad = 1; return 0; }
    int height = 1;
    if (n->type == PO_BRANCH) {
        int left = visit(st, n->left, depth + 1);
        if (st->bad) return 0;
        int right = n->left == n->right ? left : visit(st, n->right, depth + 1);
        if (st->bad) return 0;
        height += left > right ? left : right;
        int32_t children[2] = {n->left, n->right};
        for (int c = 0; c < 2; c++) {
            const po_node *ch = &s->nodes[children[c]];
            for (int a = 0; a < 3; a++)
                if (!(comp(n->minimum, a) <= comp(ch->minimum, a) && comp(n->maximum, a) >= comp(ch->maximum, a))) { st->bad = 1; return 0; }
        }
    } else if (n->type == PO_TRIANGLE) {
        if (!valid_id(n->primitive, s->triangle_count)) { st->bad = 1; return 0; }
        const po_triangle *t = &s->triangles[n->primitive];
        po_vec3 ps[3] = {t->v0, t->v1, t->v2};
        for (int k = 0; k < 3; k++)
            for (int a = 0; a < 3; a++)
                if (!(comp(ps[k], a) >= comp(n->minimum, a) && comp(ps[k], a) <= comp(n->maximum, a))) { st->bad = 1; return 0; }
    } else if (n->type == PO_SPHERE) {
        if (!valid_id(n->primitive, s->sphere_count)) { st->bad = 1; return 0; }
        const po_sphere *sp = &s->spheres[n->primitive];
        for (int a = 0; a < 3; a++)
            if (!(comp(sp->center, a) - sp->radius >= comp(n->minimum, a) && comp(sp->center, a) + sp->radius <= comp(n->maximum, a))) { st->bad = 1; return 0; }
    } else { st->bad = 1; return 0; }
    st->active[id] = 0;
    st->heights[id] = height;
    return height;
}

int po_validate_scene(const po_scene *s) {
#define REQUIRE(c) do { if (!(c)) return -1; } while (0)
    for (int64_t i = 0; i < s->texture_count; i++) {
        const po_texture *t = &s->textures[i];
        REQUIRE(finite3(t->color) && finite3(t->scale) && finite3(t->offset));
        REQUIRE(t->type >= PO_CONSTANT && t->type <= PO_MAP);
        if (t->type == PO_IMAGE) {
            REQUIRE(t->width > 0 && t->height > 0 && t->byte_offset >= 0);
            uint64_t end = (uint64_t)t->byte_offset + 3 * (uint64_t)t->width * (uint64_t)t->height;
            REQUIRE(end <= (uint64_t)s->pixel_bytes);
        } else if (t->type == PO_MAP) {
            REQUIRE(valid_id(t->image_id, s->texture_count));
            REQUIRE(s->textures[t->image_id].type == PO_IMAGE && (t->clamp == 0 || t->clamp == 1));
        }
    }
    for (int64_t i = 0; i < s->material_count; i++) {
        const po_material *m = &s->materials[i];
        REQUIRE(m->type >= PO_DIFFUSE && m->type <= PO_LIGHT && finite3(m->albedo));
        REQUIRE(isfinite(m->fuzz) && isfinite(m->ref_idx));
        if (m->type == PO_DIFFUSE || m->type == PO_LIGHT) REQUIRE(valid_id(m->texture_id, s->texture_count));
        if (m->type == PO_GLASS) REQUIRE(m->ref_idx > 0);
    }
    for (int64_t i = 0; i < s->triangle_count; i++) {
        const po_triangle *t = &s->triangles[i];
        REQUIRE(valid_id(t->material_id, s->material_count) && (t->smooth == 0 || t->smooth == 1));
        REQUIRE(finite3(t->v0) && finite3(t->v1) && finite3(t->v2) && finite3(t->n0) && finite3(t->n1) && finite3(t->n2));
        REQUIRE(finite3(t->uv0) && finite3(t->uv1) && finite3(t->uv2));
    }
    for (int64_t i = 0; i < s->sphere_count; i++) {
        const po_sphere *p = &s->spheres[i];
        REQUIRE(valid_id(p->material_id, s->material_count) && finite3(p->center) && isfinite(p->radius) && p->radius > 0);
    }
    if (s->root == -1) {
        REQUIRE(s->node_count == 0 && s->triangle_count == 0 && s->sphere_count == 0 && s->max_depth == 0);
        return 0;
    }
    REQUIRE(valid_id(s->root, s->node_count));
    vstate st;
    st.s = s; st.visited = 0; st.bad = 0;
    st.heights = (int *)calloc((size_t)s->node_count, sizeof(int));
    st.active = (char *)calloc((size_t)s->node_count, 1);
    int depth = visit(&st, s->root, 1);
    int ok = !st.bad && depth == s->max_depth && st.visited == s->node_count;
    free(st.heights); free(st.active);
    return ok ? 0 : -1;
#undef REQUIRE
}

/* src/renderer/render_validation.h:16-22 */
int po_validate_camera(const po_camera *c) {
    po_vec3 vs[6] = {c->origin, c->lower_left_corner, c->horizontal, c->vertical, c->u, c->v};
    for (int i = 0; i < 6; i++) if (!finite3(vs[i])) return -1;
    if (!isfinite(c->lens_radius) || c->lens_radius < 0 || !isfinite(c->time0) ||
        !isfinite(c->time1) || c->time1 < c->time0) return -1;
    return 0;
}

po_vec3 po_trace_color(const po_scene *s, const po_light *lights, int64_t light_count,
                       po_ray r, po_rng *rng, int brute, int debug, int direct_light,
                       int max_depth, int *error, uint64_t *rays) {
    view sc = {s, lights, light_count, light_count ? lights[light_count - 1].cumulative_area : 0};
    return trace_color(&sc, r, rng, brute, debug, error, direct_light, max_depth, rays);
}

/* src/display/color.h:6-11 and cuda_renderer.cu:136-141: 255.99 * sqrt in double. */
int po_pack_channel(float linear) {
    if (!(linear > 0)) return 0;
    if (linear >= 1) return 255;
    float gamma = sqrtf(linear);
    return (int)(255.99 * gamma);
}

/* ---------------- progressive renderer (tests/support/portable_renderer.h) ---------------- */
struct po_renderer {
    po_scene scene;
    po_camera camera;
    po_light *lights;
    int64_t light_count;
    int width, height, samples, brute, debug, direct_light, max_depth;
    uint64_t seed, rays;
    po_vec3 *sum;
    po_rng *rng;
};

po_renderer *po_renderer_create(const po_scene *s, const po_camera *c, int width, int height,
                                uint64_t seed, int brute, int debug, int direct_light, int max_depth) {
    if (width <= 0 || height <= 0 || (uint64_t)width * (uint64_t)height > (uint64_t)INT32_MAX) return NULL;
    if (debug < PO_SHADED || debug > PO_INTERSECTIONS) return NULL;
    if (po_validate_scene(s) != 0 || po_validate_camera(c) != 0) return NULL;
    po_renderer *r = (po_renderer *)calloc(1, sizeof *r);
    r->scene = *s;
    int64_t n = po_scene_lights(s, NULL, 0);
    if (n < 0) { free(r); return NULL; }
    r->lights = (po_light *)calloc((size_t)(n ? n : 1), sizeof(po_light));
    r->light_count = po_scene_lights(s, r->lights, n);
    r->width = width; r->height = height; r->seed = seed;
    r->brute = brute; r->debug = debug; r->direct_light = direct_light; r->max_depth = max_depth;
    size_t count = (size_t)width * (size_t)height;
    r->sum = (po_vec3 *)calloc(count, sizeof(po_vec3));
    r->rng = (po_rng *)calloc(count, sizeof(po_rng));
    po_renderer_reset(r, c);
    return r;
}

void po_renderer_destroy(po_renderer *r) {
    if (!r) return;
    free(r->lights); free(r->sum); free(r->rng); free(r);
}

int po_renderer_reset(po_renderer *r, const po_camera *c) {
    if (po_validate_camera(c) != 0) return -1;
    r->camera = *c; r->samples = 0; r->rays = 0;
    size_t count = (size_t)r->width * (size_t)r->height;
    for (size_t i = 0; i < count; i++) { r->sum[i] = v3(0, 0, 0); po_rng_seed(&r->rng[i], r->seed, i); }
    return 0;
}

typedef struct {
    po_renderer *r;
    int count, thread, threads, y0, y1, error;
    uint64_t rays;
} job;

static void *render_rows(void *arg) {
    job *j = (job *)arg;
    po_renderer *r = j->r;
    view sc = {&r->scene, r->lights, r->light_count, r->light_count ? r->lights[r->light_count - 1].cumulative_area : 0};
    /* Rows are interleaved across threads; each pixel has one RNG stream and
       one writer, so the image is identical for any thread count. */
    for (int y = j->y0 + j->thread; y < j->y1; y += j->threads) {
        for (int x = 0; x < r->width; x++) {
            size_t i = (size_t)y * (size_t)r->width + (size_t)x;
            int error = 0;
            po_vec3 acc = r->sum[i];
            for (int s = 0; s < j->count; s++)
                acc = add(acc, trace_pixel_sample(&sc, &r->camera, x, y, r->width, r->height, &r->rng[i],
                                                  r->brute, r->debug, &error, r->direct_light, r->max_depth, &j->rays));
            r->sum[i] = acc;
            if (error) j->error = 1;
        }
    }
    return NULL;
}

int po_renderer_add_samples_rows(po_renderer *r, int count, int threads, int y0, int y1) {
    if (count <= 0 || count > INT32_MAX - r->samples) return -1;
    if (threads < 1) threads = 1;
    if (y0 < 0) y0 = 0;
    if (y1 > r->height) y1 = r->height;
    job *jobs = (job *)calloc((size_t)threads, sizeof(job));
    pthread_t *ids = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    for (int t = 0; t < threads; t++) {
        jobs[t].r = r; jobs[t].count = count; jobs[t].thread = t; jobs[t].threads = threads;
        jobs[t].y0 = y0; jobs[t].y1 = y1;
        if (threads > 1) pthread_create(&ids[t], NULL, render_rows, &jobs[t]);
    }
    if (threads == 1) render_rows(&jobs[0]);
    int error = 0;
    for (int t = 0; t < threads; t++) {
        if (threads > 1) pthread_join(ids[t], NULL);
        r->rays += jobs[t].rays;
        error |= jobs[t].error;
    }
    free(jobs); free(ids);
    if (error) return -2;
    r->samples += count;
    return 0;
}

int po_renderer_add_samples(po_renderer *r, int count, int threads) {
    return po_renderer_add_samples_rows(r, count, threads, 0, r->height);
}

/* portable_renderer.h:37-41: divide via operator/=(float(samples)). */
void po_renderer_image(const po_renderer *r, float *rgb) {
    size_t count = (size_t)r->width * (size_t)r->height;
    for (size_t i = 0; i < count; i++) {
        po_vec3 p = r->sum[i];
        if (r->samples) p = div_assign(p, (float)r->samples);
        rgb[3 * i] = p.x; rgb[3 * i + 1] = p.y; rgb[3 * i + 2] = p.z;
    }
}
void po_renderer_sum(const po_renderer *r, float *rgb) {
    memcpy(rgb, r->sum, (size_t)r->width * (size_t)r->height * sizeof(po_vec3));
}
void po_renderer_rng(const po_renderer *r, uint64_t *state) {
    size_t count = (size_t)r->width * (size_t)r->height;
    for (size_t i = 0; i < count; i++) state[i] = r->rng[i].state;
}
int po_renderer_sample_count(const po_renderer *r) { return r->samples; }
uint64_t po_renderer_rays(const po_renderer *r) { return r->rays; }

/* cuda_renderer.cu:99-109: sum/float(samples) with true division, then the channel rule. */
void po_renderer_rgba(const po_renderer *r, uint8_t *rgba) {
    size_t count = (size_t)r->width * (size_t)r->height;
    for (size_t i = 0; i < count; i++) {
        po_vec3 c = r->samples ? divs(r->sum[i], (float)r->samples) : v3(0, 0, 0);
        float ch[3] = {c.x, c.y, c.z};
        for (int k = 0; k < 3; k++) {
            float value = ch[k];
            float gamma = value > 0 && value < 1 ? sqrtf(value) : 0;
            rgba[4 * i + k] = (uint8_t)(!(value > 0) ? 0 : value >= 1 ? 255 : (int)(255.99 * gamma));
        }
        rgba[4 * i + 3] = 255;
    }
}
