// ref_shim.cpp -- C entry points over the reference's own host render path.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/build_ref.sh against the
// reference headers where they lie (src/, tests/support/, third_party/) into
// oracle/_ref/libptref.so.  It is used to pin the C oracle bit-for-bit, to
// export golden fixtures, and optionally as the "reference" CPU baseline.
// Nothing here is part of the product.
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <thread>
#include <vector>

#include "flat/flat_export.h"
#include "renderer/render_validation.h"
#include "renderer/transport.h"
#include "display/display_pixels.h"
#include "support/flat_cpu.h"
#include "support/portable_renderer.h"

static_assert(sizeof(flat_triangle) == 116 && sizeof(flat_node) == 40 && sizeof(flat_sphere) == 20 &&
              sizeof(flat_material) == 28 && sizeof(flat_texture) == 60 && sizeof(flat_camera) == 84 &&
              sizeof(flat_hit) == 48 && sizeof(trace_result) == 56 && sizeof(ray) == 28 &&
              sizeof(pixel_rng) == 16 && sizeof(flat_light) == 12, "record layout");

extern "C" {

struct ref_scene_in {
    const flat_triangle *triangles; int64_t triangle_count;
    const flat_sphere *spheres; int64_t sphere_count;
    const flat_node *nodes; int64_t node_count;
    const flat_material *materials; int64_t material_count;
    const flat_texture *textures; int64_t texture_count;
    const unsigned char *pixels; int64_t pixel_bytes;
    int32_t root, max_depth;
};

}

static flat_scene to_scene(const ref_scene_in *in) {
    flat_scene s;
    s.triangles.assign(in->triangles, in->triangles + in->triangle_count);
    s.spheres.assign(in->spheres, in->spheres + in->sphere_count);
    s.nodes.assign(in->nodes, in->nodes + in->node_count);
    s.materials.assign(in->materials, in->materials + in->material_count);
    s.textures.assign(in->textures, in->textures + in->texture_count);
    s.pixels.assign(in->pixels, in->pixels + in->pixel_bytes);
    s.root = in->root; s.max_depth = in->max_depth;
    return s;
}

struct ref_scene_out {
    flat_scene scene;
    flat_camera camera;
};

extern "C" {

int ref_validate(const ref_scene_in *in) {
    try { validate_flat_scene(to_scene(in)); return 0; } catch(const std::runtime_error&) { return -1; }
}

int64_t ref_scene_lights(const ref_scene_in *in, flat_light *out, int64_t cap) {
    try {
        auto lights = scene_lights(to_scene(in));
        for(size_t i = 0; i < lights.size() && int64_t(i) < cap; i++) out[i] = lights[i];
        return int64_t(lights.size());
    } catch(const std::invalid_argument&) { return -1; }
}

void ref_query(const ref_scene_in *in, const ray *rays, int64_t n, int brute, trace_result *out,
               int32_t *boxes, int32_t *primitives) {
    flat_scene s = to_scene(in);
    scene_view view = host_view(s);
    for(int64_t i = 0; i < n; i++) {
        trace_result result{};
        trace_stats stats;
        int error = 0;
        result.hit = trace_world(view, rays[i], 0.001f, FLT_MAX, result.rec, brute != 0, error, stats);
        result.error = error;
        out[i] = result;
        if(boxes) boxes[i] = stats.boxes;
        if(primitives) primitives[i] = stats.primitives;
    }
}

// The older non-pruning traversal of tests/support/flat_cpu.h:14-46.
void ref_flat_hit_world(const ref_scene_in *in, const ray *rays, int64_t n, trace_result *out) {
    flat_scene s = to_scene(in);
    for(int64_t i = 0; i < n; i++) {
        trace_result result{};
        result.hit = flat_hit_world(s, rays[i], 0.001f, FLT_MAX, result.rec);
        out[i] = result;
    }
}

void ref_rng_stream(uint64_t seed, uint64_t pixel, int64_t n, uint32_t *out_next, float *out_uniform) {
    pixel_rng a, b; a.seed(seed, pixel); b.seed(seed, pixel);
    for(int64_t i = 0; i < n; i++) { out_next[i] = a.next(); out_uniform[i] = b.uniform(); }
}

void ref_samplers(uint64_t seed, uint64_t pixel, int64_t n, float *ball, float *disk, uint64_t *final_state) {
    pixel_rng r; r.seed(seed, pixel);
    for(int64_t i = 0; i < n; i++) {
        vec3 b = sample_ball(r); vec3 d = sample_disk(r);
        for(int k = 0; k < 3; k++) { ball[3*i+k] = b[k]; disk[3*i+k] = d[k]; }
    }
    *final_state = r.state;
}

// One trace_color from a given RNG state; returns radiance, updates state and ray count.
void ref_trace_color(const ref_scene_in *in, const ray *r, uint64_t *state, uint64_t increment, int brute,
                     int debug, int direct_light, float *out, uint64_t *rays, int *error) {
    flat_scene s = to_scene(in);
    auto lights = scene_lights(s);
    scene_view view = host_view(s, lights);
    pixel_rng rng{*state, increment};
    vec3 c = trace_color(*r, view, rng, brute != 0, render_debug(debug), *error, direct_light != 0, rays);
    out[0] = c.x(); out[1] = c.y(); out[2] = c.z();
    *state = rng.state;
}

// portable_renderer over a sequence of add_samples batches; image() out.
int ref_render(const ref_scene_in *in, const flat_camera *cam, int w, int h, uint64_t seed, int brute, int debug,
               int direct_light, const int *batches, int nbatches, float *image, uint64_t *rays) {
    try {
        portable_renderer r(to_scene(in), *cam, w, h, seed, brute != 0, render_debug(debug), direct_light != 0);
        for(int b = 0; b < nbatches; b++) r.add_samples(batches[b]);
        auto img = r.image();
        for(size_t i = 0; i < img.size(); i++) for(int k = 0; k < 3; k++) image[3*i+k] = img[i][k];
        *rays = r.rays_traced();
        return 0;
    } catch(const std::invalid_argument&) { return -1; } catch(const std::runtime_error&) { return -2; }
}

// Rows [y0,y1) of portable_renderer::add_samples, spread over threads with the
// reference's own trace_pixel_sample; used as the "reference" CPU baseline.
int ref_render_rows(const ref_scene_in *in, const flat_camera *cam, int w, int h, uint64_t seed, int direct_light,
                    int samples, int y0, int y1, int threads, float *sum_rows, uint64_t *rays_out) {
    flat_scene s = to_scene(in);
    auto lights = scene_lights(s);
    scene_view view = host_view(s, lights);
    std::vector<uint64_t> rays(size_t(threads), 0);
    std::vector<int> errors(size_t(threads), 0);
    auto work = [&](int t) {
        for(int y = y0 + t; y < y1; y += threads) for(int x = 0; x < w; x++) {
            size_t i = size_t(y)*w + x;
            pixel_rng rng; rng.seed(seed, i);
            vec3 acc(0,0,0);
            int error = 0;
            for(int k = 0; k < samples; k++)
                acc += trace_pixel_sample(view, *cam, x, y, w, h, rng, false, DEBUG_SHADED, error, direct_light != 0, &rays[t]);
            size_t o = size_t(y-y0)*w + x;
            for(int c = 0; c < 3; c++) sum_rows[3*o+c] = acc[c];
            if(error) errors[t] = 1;
        }
    };
    std::vector<std::thread> pool;
    for(int t = 1; t < threads; t++) pool.emplace_back(work, t);
    work(0);
    for(auto& th : pool) th.join();
    uint64_t total = 0; int error = 0;
    for(int t = 0; t < threads; t++) { total += rays[t]; error |= errors[t]; }
    *rays_out = total;
    return error ? -2 : 0;
}

void ref_display_pixels(const float *rgb, int64_t n, unsigned char *out) {
    std::vector<vec3> linear(static_cast<size_t>(n));
    for(int64_t i = 0; i < n; i++) linear[i] = vec3(rgb[3*i], rgb[3*i+1], rgb[3*i+2]);
    auto bytes = display_pixels(linear);
    std::memcpy(out, bytes.data(), bytes.size());
}

// Import an OBJ asset with the reference's loader, BVH and exporter (srand48-seeded).
void *ref_load_obj(const char *path, float aspect, long srand_seed, const float *scale3, const float *rotate3,
                   const float *translate3, float vfov, float aperture, float t0, float t1) {
    try {
        srand48(srand_seed);
        obj_scene source(path, aspect, mesh_transform(vec3(scale3[0],scale3[1],scale3[2]),
                         vec3(rotate3[0],rotate3[1],rotate3[2]), vec3(translate3[0],translate3[1],translate3[2])));
        auto *out = new ref_scene_out;
        out->scene = flatten_scene(&source);
        camera c(source.lookfrom, source.lookat, vec3(0,1,0), vfov, aspect, aperture, source.focus_dist, t0, t1);
        out->camera = flatten_camera(c);
        return out;
    } catch(const std::exception& e) {
        fprintf(stderr, "ref_load_obj: %s\n", e.what());
        return nullptr;
    }
}

// Hand-built hitable scene of spheres in a reference bvh_node, flattened by the exporter.
void *ref_sphere_scene(long srand_seed) {
    srand48(srand_seed);
    static constant_texture paint(vec3(0.3,0.7,0.2)), lamp_color(vec3(4,3,2));
    static lambertian diffuse(&paint); static metal mirror(vec3(0.8,0.7,0.6),0.1);
    static dielectric glass(1.5); static diffuse_light lamp(&lamp_color);
    sphere a(vec3(-2,0,0),1,&diffuse), b(vec3(0,0,0),1,&glass), c(vec3(2,0,0),1,&mirror);
    sphere ground(vec3(0,-101,0),100,&diffuse), light(vec3(0,4,2),2,&lamp);
    hitable *objects[] = {&a,&b,&c,&ground,&light};
    bvh_node world(objects,5,0,1);
    auto *out = new ref_scene_out;
    out->scene = flatten_scene(&world);
    camera view(vec3(0,2,8),vec3(0,0,0),vec3(0,1,0),40,1.5,0.2,8,0.2,0.8);
    out->camera = flatten_camera(view);
    return out;
}

// Triangle-light fixture: a lit floor, a small emitter and an optional blocker.
void *ref_triangle_light_scene(int occluded) {
    constant_texture white(vec3(0.7,0.7,0.7)), emission(vec3(10,10,10)), black(vec3(0,0,0));
    lambertian diffuse(&white), blocker(&black); diffuse_light light(&emission);
    std::vector<triangle> faces;
    faces.emplace_back(vec3(-10,-10,0),vec3(10,-10,0),vec3(0,10,0),&diffuse);
    faces.emplace_back(vec3(-0.2,-0.2,2),vec3(0.2,-0.2,2),vec3(0,0.2,2),&light);
    if(occluded) faces.emplace_back(vec3(-10,-10,1),vec3(10,-10,1),vec3(0,10,1),&blocker);
    triangle_mesh mesh(std::move(faces));
    auto *out = new ref_scene_out;
    out->scene = flatten_scene(&mesh);
    out->camera = flatten_camera(::camera(vec3(0,0,0.1),vec3(0,0,0),vec3(0,1,0),40,19.0f/13,0,0.1,0,1));
    return out;
}

void ref_camera(const float *from, const float *at, const float *up, float vfov, float aspect, float aperture,
                float focus, float t0, float t1, flat_camera *out) {
    camera c(vec3(from[0],from[1],from[2]), vec3(at[0],at[1],at[2]), vec3(up[0],up[1],up[2]),
             vfov, aspect, aperture, focus, t0, t1);
    *out = flatten_camera(c);
}

void ref_scene_counts(void *h, int64_t *counts, int32_t *root_depth) {
    auto *o = static_cast<ref_scene_out*>(h);
    counts[0] = int64_t(o->scene.triangles.size()); counts[1] = int64_t(o->scene.spheres.size());
    counts[2] = int64_t(o->scene.nodes.size()); counts[3] = int64_t(o->scene.materials.size());
    counts[4] = int64_t(o->scene.textures.size()); counts[5] = int64_t(o->scene.pixels.size());
    root_depth[0] = o->scene.root; root_depth[1] = o->scene.max_depth;
}

void ref_scene_copy(void *h, flat_triangle *t, flat_sphere *s, flat_node *n, flat_material *m, flat_texture *x,
                    unsigned char *p, flat_camera *cam) {
    auto *o = static_cast<ref_scene_out*>(h);
    std::copy(o->scene.triangles.begin(), o->scene.triangles.end(), t);
    std::copy(o->scene.spheres.begin(), o->scene.spheres.end(), s);
    std::copy(o->scene.nodes.begin(), o->scene.nodes.end(), n);
    std::copy(o->scene.materials.begin(), o->scene.materials.end(), m);
    std::copy(o->scene.textures.begin(), o->scene.textures.end(), x);
    std::copy(o->scene.pixels.begin(), o->scene.pixels.end(), p);
    *cam = o->camera;
}

void ref_scene_free(void *h) { delete static_cast<ref_scene_out*>(h); }

// The reference's random-axis BVH (bvh_node, srand48-seeded) over
// [triangles..., spheres...] of a flat input, flattened by its exporter.
// Materials become the reference's material objects over constant, image
// and OBJ diffuse-map textures.
void *ref_build_flat(const ref_scene_in *in, long srand_seed) {
    try {
        auto *out = new ref_scene_out;
        static std::vector<std::unique_ptr<texture>> textures;
        static std::vector<std::unique_ptr<material>> materials;
        static std::vector<std::unique_ptr<hitable>> objects;
        static std::vector<std::vector<unsigned char>> images;
        textures.clear(); materials.clear(); objects.clear(); images.clear();
        std::vector<texture*> tex(size_t(in->texture_count), nullptr);
        for(int pass = 0; pass < 2; pass++) {
            for(int64_t i = 0; i < in->texture_count; i++) {
                const flat_texture& t = in->textures[i];
                if((pass == 0) != (t.type != FLAT_MAP)) continue;
                if(t.type == FLAT_CONSTANT) textures.push_back(std::make_unique<constant_texture>(t.color));
                else if(t.type == FLAT_IMAGE) {
                    images.emplace_back(in->pixels + t.byte_offset, in->pixels + t.byte_offset + size_t(t.width) * t.height * 3);
                    textures.push_back(std::make_unique<image_texture>(images.back().data(), t.width, t.height));
                } else {
                    tinyobj::texture_option_t opt{};
                    opt.scale[0] = t.scale.x(); opt.scale[1] = t.scale.y(); opt.scale[2] = 1;
                    opt.origin_offset[0] = t.offset.x(); opt.origin_offset[1] = t.offset.y(); opt.origin_offset[2] = 0;
                    opt.clamp = t.clamp != 0;
                    textures.push_back(std::make_unique<obj_detail::diffuse_map>(tex[size_t(t.image_id)], t.color, opt));
                }
                tex[size_t(i)] = textures.back().get();
            }
        }
        std::vector<material*> mats;
        for(int64_t i = 0; i < in->material_count; i++) {
            const flat_material& m = in->materials[i];
            if(m.type == FLAT_DIFFUSE) materials.push_back(std::make_unique<lambertian>(tex[size_t(m.texture_id)]));
            else if(m.type == FLAT_METAL) materials.push_back(std::make_unique<metal>(m.albedo, m.fuzz));
            else if(m.type == FLAT_GLASS) materials.push_back(std::make_unique<dielectric>(m.ref_idx));
            else materials.push_back(std::make_unique<diffuse_light>(tex[size_t(m.texture_id)]));
            mats.push_back(materials.back().get());
        }
        std::vector<hitable*> list;
        for(int64_t i = 0; i < in->triangle_count; i++) {
            const flat_triangle& f = in->triangles[i];
            auto t = std::make_unique<triangle>(f.v0, f.v1, f.v2, mats[size_t(f.material_id)]);
            t->n0 = f.n0; t->n1 = f.n1; t->n2 = f.n2; t->smooth = f.smooth != 0;
            t->uv0 = f.uv0; t->uv1 = f.uv1; t->uv2 = f.uv2;
            list.push_back(t.get());
            objects.push_back(std::move(t));
        }
        for(int64_t i = 0; i < in->sphere_count; i++) {
            const flat_sphere& s = in->spheres[i];
            auto o = std::make_unique<sphere>(s.center, s.radius, mats[size_t(s.material_id)]);
            list.push_back(o.get());
            objects.push_back(std::move(o));
        }
        srand48(srand_seed);
        bvh_node root(list.data(), int(list.size()), 0, 1);
        out->scene = flatten_scene(&root);
        out->camera = flat_camera{};
        return out;
    } catch(const std::exception& e) {
        fprintf(stderr, "ref_build_flat: %s\n", e.what());
        return nullptr;
    }
}

// stb_image as the reference loads textures (texture.h:65-72): 8-bit RGB.
int ref_load_image(const char *path, int *w, int *h, unsigned char *out, long capacity) {
    int channels = 0;
    unsigned char *data = stbi_load(path, w, h, &channels, STBI_rgb);
    if(!data) return -1;
    long n = long(*w) * long(*h) * 3;
    if(n <= capacity) std::memcpy(out, data, size_t(n));
    stbi_image_free(data);
    return n <= capacity ? 0 : 1;
}


}
