// pt_scene.cpp -- BVH builder emitting the flat transfer format, the camera
// basis, and seeded synthetic scenes (Cornell room, displaced icosphere,
// clustered triangle soup).  Host-only C++; no device code.
#include "pt_scene.h"
#include "pt_scene_impl.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string last_error;

} // namespace

void pts::set_error(const std::string& message) { last_error = message; }

pt_flat_texture pts::constant_texture(pt_vec3 color) {
    // flat_export.h:30: constant record defaults.
    pt_flat_texture t{};
    t.color = color; t.scale = v3(1, 1, 0); t.offset = v3(0, 0, 0);
    t.type = PT_CONSTANT; t.image_id = -1; t.clamp = 1;
    return t;
}

pt_flat_material pts::flat_material(int32_t type, pt_vec3 albedo, float fuzz, float ref_idx, int32_t texture) {
    pt_flat_material m{};
    m.albedo = albedo; m.fuzz = fuzz; m.ref_idx = ref_idx; m.type = type; m.texture_id = texture;
    return m;
}

namespace {

using namespace pts;

struct prim_ref {
    box3 bounds;
    float centroid[3];
    int32_t type, index;
};

class builder {
public:
    builder(pts_scene& out, int split) : scene(out), method(split) {}

    void run(const std::vector<pt_flat_triangle>& tris, const std::vector<pt_flat_sphere>& sph) {
        refs.reserve(tris.size() + sph.size());
        for(size_t i = 0; i < tris.size(); i++) add(triangle_bounds(tris[i]), PT_TRIANGLE, int32_t(i));
        for(size_t i = 0; i < sph.size(); i++) add(sphere_bounds(sph[i]), PT_SPHERE, int32_t(i));
        scene.nodes.clear(); scene.triangles.clear(); scene.spheres.clear();
        scene.max_depth = 0;
        if(refs.empty()) { scene.root = -1; return; }
        if(refs.size() > size_t(INT32_MAX / 2)) throw std::invalid_argument("too many primitives for 32-bit node indices");
        scene.nodes.reserve(2 * refs.size());
        scene.triangles.reserve(tris.size());
        scene.spheres.reserve(sph.size());
        src_tris = &tris; src_sph = &sph;
        box3 ignored;
        scene.root = build(0, refs.size(), 1, ignored);
    }

private:
    pts_scene& scene;
    int method;
    std::vector<prim_ref> refs;
    const std::vector<pt_flat_triangle> *src_tris = nullptr;
    const std::vector<pt_flat_sphere> *src_sph = nullptr;

    void add(const box3& b, int32_t type, int32_t index) {
        prim_ref r;
        r.bounds = b; r.type = type; r.index = index;
        for(int a = 0; a < 3; a++) r.centroid[a] = 0.5f * b.lo[a] + 0.5f * b.hi[a];
        refs.push_back(r);
    }

    static int ceil_log2(size_t n) { int d = 0; while((size_t(1) << d) < n) d++; return d; }

    // Returns the split position in [begin,end) after partitioning refs.
    size_t split(size_t begin, size_t end, int depth) {
        size_t n = end - begin;
        box3 cb; cb.empty();
        for(size_t i = begin; i < end; i++) cb.grow(refs[i].centroid);
        int axis = 0;
        float extent = cb.hi[0] - cb.lo[0];
        for(int a = 1; a < 3; a++) if(cb.hi[a] - cb.lo[a] > extent) { extent = cb.hi[a] - cb.lo[a]; axis = a; }
        bool median = method == PTS_SPLIT_MEDIAN || depth + 1 + ceil_log2(n) >= 62 || !(extent > 0) || n <= 2;
        if(!median) {
            const int bins = 16;
            double best = std::numeric_limits<double>::infinity();
            int best_axis = -1, best_bin = -1;
            for(int a = 0; a < 3; a++) {
                float lo = cb.lo[a], ext = cb.hi[a] - cb.lo[a];
                if(!(ext > 0)) continue;
                box3 bb[bins]; size_t cnt[bins] = {};
                for(int k = 0; k < bins; k++) bb[k].empty();
                float scale_k = bins / ext;
                for(size_t i = begin; i < end; i++) {
                    int k = int((refs[i].centroid[a] - lo) * scale_k);
                    k = std::min(bins - 1, std::max(0, k));
                    cnt[k]++; bb[k].grow(refs[i].bounds);
                }
                box3 right_box[bins]; size_t right_cnt[bins];
                box3 acc; acc.empty(); size_t c = 0;
                for(int k = bins - 1; k > 0; k--) { acc.grow(bb[k]); c += cnt[k]; right_box[k] = acc; right_cnt[k] = c; }
                acc.empty(); c = 0;
                for(int k = 0; k < bins - 1; k++) {
                    acc.grow(bb[k]); c += cnt[k];
                    size_t rc = right_cnt[k + 1];
                    if(c == 0 || rc == 0) continue;
                    double cost = acc.area() * double(c) + right_box[k + 1].area() * double(rc);
                    if(cost < best) { best = cost; best_axis = a; best_bin = k; }
                }
            }
            if(best_axis >= 0) {
                float lo = cb.lo[best_axis], ext = cb.hi[best_axis] - cb.lo[best_axis];
                float scale_k = bins / ext;
                auto mid = std::partition(refs.begin() + begin, refs.begin() + end, [&](const prim_ref& r) {
                    int k = int((r.centroid[best_axis] - lo) * scale_k);
                    k = std::min(bins - 1, std::max(0, k));
                    return k <= best_bin;
                });
                size_t m = size_t(mid - refs.begin());
                if(m > begin && m < end) return m;
            }
        }
        size_t m = begin + n / 2;
        std::nth_element(refs.begin() + begin, refs.begin() + m, refs.begin() + end,
                         [axis](const prim_ref& a, const prim_ref& b) {
                             if(a.centroid[axis] != b.centroid[axis]) return a.centroid[axis] < b.centroid[axis];
                             return a.index < b.index;
                         });
        return m;
    }

    int32_t build(size_t begin, size_t end, int depth, box3& out_bounds) {
        if(depth > 64) throw std::runtime_error("BVH exceeds the 64-level traversal stack");
        scene.max_depth = std::max(scene.max_depth, depth);
        int32_t id = int32_t(scene.nodes.size());
        scene.nodes.push_back(pt_flat_node{});
        if(end - begin == 1) {
            const prim_ref& r = refs[begin];
            pt_flat_node node{};
            node.type = r.type; node.left = -1; node.right = -1;
            if(r.type == PT_TRIANGLE) {
                node.primitive = int32_t(scene.triangles.size());
                scene.triangles.push_back((*src_tris)[size_t(r.index)]);
            } else {
                node.primitive = int32_t(scene.spheres.size());
                scene.spheres.push_back((*src_sph)[size_t(r.index)]);
            }
            node.minimum = v3(r.bounds.lo[0], r.bounds.lo[1], r.bounds.lo[2]);
            node.maximum = v3(r.bounds.hi[0], r.bounds.hi[1], r.bounds.hi[2]);
            scene.nodes[size_t(id)] = node;
            out_bounds = r.bounds;
            return id;
        }
        size_t mid = split(begin, end, depth);
        box3 lb, rb;
        int32_t left = build(begin, mid, depth + 1, lb);
        int32_t right = build(mid, end, depth + 1, rb);
        // Union with the reference's ffmin/ffmax comparisons (aabb.h:40-56).
        box3 b;
        for(int a = 0; a < 3; a++) {
            b.lo[a] = lb.lo[a] < rb.lo[a] ? lb.lo[a] : rb.lo[a];
            b.hi[a] = lb.hi[a] > rb.hi[a] ? lb.hi[a] : rb.hi[a];
        }
        pt_flat_node node{};
        node.type = PT_BRANCH; node.left = left; node.right = right; node.primitive = -1;
        node.minimum = v3(b.lo[0], b.lo[1], b.lo[2]);
        node.maximum = v3(b.hi[0], b.hi[1], b.hi[2]);
        scene.nodes[size_t(id)] = node;
        out_bounds = b;
        return id;
    }
};

void build_into(pts_scene& s, const std::vector<pt_flat_triangle>& tris, const std::vector<pt_flat_sphere>& sph, int split) {
    builder b(s, split);
    b.run(tris, sph);
}

int32_t add_constant(pts_scene& s, pt_vec3 color) {
    s.textures.push_back(constant_texture(color));
    return int32_t(s.textures.size() - 1);
}
int32_t add_material(pts_scene& s, int32_t type, pt_vec3 albedo, float fuzz, float ref_idx, int32_t texture) {
    s.materials.push_back(flat_material(type, albedo, fuzz, ref_idx, texture));
    return int32_t(s.materials.size() - 1);
}
int32_t diffuse(pts_scene& s, pt_vec3 c) { return add_material(s, PT_DIFFUSE, v3(0, 0, 0), 0, 1, add_constant(s, c)); }
int32_t light(pts_scene& s, pt_vec3 c) { return add_material(s, PT_LIGHT, v3(0, 0, 0), 0, 1, add_constant(s, c)); }

pt_flat_triangle flat_tri(pt_vec3 a, pt_vec3 b, pt_vec3 c, int32_t material) {
    pt_flat_triangle t{};
    t.v0 = a; t.v1 = b; t.v2 = c;
    t.n0 = t.n1 = t.n2 = v3(0, 0, 0);
    t.uv0 = v3(0, 0, 0); t.uv1 = v3(1, 0, 0); t.uv2 = v3(0, 1, 0);
    t.material_id = material; t.smooth = 0;
    return t;
}
// Quad a-b-c-d (counter-clockwise seen from the front) as two triangles.
void quad(std::vector<pt_flat_triangle>& out, pt_vec3 a, pt_vec3 b, pt_vec3 c, pt_vec3 d, int32_t m) {
    out.push_back(flat_tri(a, b, c, m));
    out.push_back(flat_tri(a, c, d, m));
}

struct seeded {
    std::mt19937_64 engine;
    explicit seeded(uint64_t s) : engine(s) {}
    float uniform(float lo, float hi) { return lo + (hi - lo) * float(double(engine() >> 11) * 0x1p-53); }
    // Three draws as a vector: z first, then y, then x.  Every multi-draw
    // expression of the synthetic scenes is written out in the order the
    // scenes were first generated with (right to left, as GCC evaluates
    // function arguments), since C++ leaves that order unspecified.
    pt_vec3 uniform3(float lo, float hi) {
        const float z = uniform(lo, hi), y = uniform(lo, hi);
        return v3(uniform(lo, hi), y, z);
    }
};

// Value noise on the integer lattice, hashed with the seed; smoothstep-trilinear.
struct value_noise {
    uint64_t seed;
    float lattice(int64_t x, int64_t y, int64_t z) const {
        uint64_t h = seed ^ (uint64_t(x) * 0x9E3779B97F4A7C15ull) ^ (uint64_t(y) * 0xC2B2AE3D27D4EB4Full) ^ (uint64_t(z) * 0x165667B19E3779F9ull);
        h ^= h >> 33; h *= 0xff51afd7ed558ccdull; h ^= h >> 33; h *= 0xc4ceb9fe1a85ec53ull; h ^= h >> 33;
        return float(double(h >> 11) * 0x1p-53) * 2.0f - 1.0f;
    }
    static float smooth(float t) { return t * t * (3.0f - 2.0f * t); }
    float sample(float x, float y, float z) const {
        float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
        int64_t ix = int64_t(fx), iy = int64_t(fy), iz = int64_t(fz);
        float tx = smooth(x - fx), ty = smooth(y - fy), tz = smooth(z - fz);
        float c[2][2][2];
        for(int a = 0; a < 2; a++) for(int b = 0; b < 2; b++) for(int d = 0; d < 2; d++) c[a][b][d] = lattice(ix + a, iy + b, iz + d);
        float x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx, x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
        float x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx, x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
        float y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
        return y0 + (y1 - y0) * tz;
    }
    // Three octaves, normalised to [-1,1].
    float octaves(pt_vec3 p) const {
        float sum = 0, amp = 1, freq = 2.5f, norm = 0;
        for(int o = 0; o < 3; o++) { sum += amp * sample(p.x * freq, p.y * freq, p.z * freq); norm += amp; amp *= 0.5f; freq *= 2; }
        return sum / norm;
    }
};

pt_flat_camera make_camera(pt_vec3 from, pt_vec3 at, pt_vec3 up, float vfov, float aspect, float aperture, float focus, float t0, float t1) {
    float f[3] = {from.x, from.y, from.z}, a[3] = {at.x, at.y, at.z}, u[3] = {up.x, up.y, up.z};
    pt_flat_camera c;
    pts_camera(f, a, u, vfov, aspect, aperture, focus, t0, t1, &c);
    return c;
}

} // namespace

void pts::build_scene(pts_scene& s, const std::vector<pt_flat_triangle>& tris, const std::vector<pt_flat_sphere>& sph,
                      int split) {
    build_into(s, tris, sph, split);
}

extern "C" {

const char *pt_scene_last_error(void) { return last_error.c_str(); }

const char *pts_warnings(const pts_scene *scene) { return scene ? scene->warnings.c_str() : ""; }

void pts_camera(const float lookfrom[3], const float lookat[3], const float vup[3], float vfov, float aspect,
                float aperture, float focus_dist, float time0, float time1, pt_flat_camera *out) {
    // src/core/camera.h:16-43, float arithmetic in the same order.
    pt_vec3 from = v3(lookfrom[0], lookfrom[1], lookfrom[2]), at = v3(lookat[0], lookat[1], lookat[2]);
    pt_vec3 up = v3(vup[0], vup[1], vup[2]);
    float lens_radius = aperture / 2;
    float theta = float(vfov * M_PI / 180);
    float half_height = std::tan(theta / 2);
    float half_width = aspect * half_height;
    pt_vec3 w = unit(from - at);
    pt_vec3 u = unit(cross(up, w));
    pt_vec3 v = cross(w, u);
    out->origin = from;
    out->lower_left_corner = from - (half_width * focus_dist) * u - (half_height * focus_dist) * v - focus_dist * w;
    out->horizontal = (2 * half_width * focus_dist) * u;
    out->vertical = (2 * half_height * focus_dist) * v;
    out->u = u; out->v = v;
    out->lens_radius = lens_radius; out->time0 = time0; out->time1 = time1;
}

pts_scene *pts_build(const pt_flat_triangle *triangles, int64_t triangle_count,
                     const pt_flat_sphere *spheres, int64_t sphere_count,
                     const pt_flat_material *materials, int64_t material_count,
                     const pt_flat_texture *textures, int64_t texture_count,
                     const uint8_t *pixels, int64_t pixel_bytes, int32_t split) {
    try {
        if(triangle_count < 0 || sphere_count < 0 || material_count < 0 || texture_count < 0 || pixel_bytes < 0)
            throw std::invalid_argument("negative array size");
        auto *s = new pts_scene;
        std::vector<pt_flat_triangle> tris(triangles, triangles + triangle_count);
        std::vector<pt_flat_sphere> sph(spheres, spheres + sphere_count);
        s->materials.assign(materials, materials + material_count);
        s->textures.assign(textures, textures + texture_count);
        s->pixels.assign(pixels, pixels + pixel_bytes);
        build_into(*s, tris, sph, split);
        return s;
    } catch(const std::exception& e) { last_error = e.what(); return nullptr; }
}

pts_scene *pts_cornell(uint64_t seed, int32_t width, int32_t height, pt_flat_camera *camera) {
    try {
        auto *s = new pts_scene;
        seeded rng(seed);
        int32_t white = diffuse(*s, v3(0.73f, 0.73f, 0.73f));
        int32_t red = diffuse(*s, v3(0.65f, 0.05f, 0.05f));
        int32_t green = diffuse(*s, v3(0.12f, 0.45f, 0.15f));
        int32_t lamp = light(*s, v3(15, 15, 15));
        std::vector<pt_flat_triangle> tris;
        // Room x,z in [-1,1], y in [0,2], open towards +z.
        quad(tris, v3(-1, 0, 1), v3(1, 0, 1), v3(1, 0, -1), v3(-1, 0, -1), white);   // floor
        quad(tris, v3(-1, 2, -1), v3(1, 2, -1), v3(1, 2, 1), v3(-1, 2, 1), white);   // ceiling
        quad(tris, v3(-1, 0, -1), v3(1, 0, -1), v3(1, 2, -1), v3(-1, 2, -1), white); // back
        quad(tris, v3(-1, 0, 1), v3(-1, 0, -1), v3(-1, 2, -1), v3(-1, 2, 1), red);   // left
        quad(tris, v3(1, 0, -1), v3(1, 0, 1), v3(1, 2, 1), v3(1, 2, -1), green);     // right
        // Ceiling light, 1e-3 below the ceiling to avoid coplanar ties, facing down.
        float ly = 2.0f - 0.001f;
        quad(tris, v3(-0.3f, ly, -0.3f), v3(-0.3f, ly, 0.3f), v3(0.3f, ly, 0.3f), v3(0.3f, ly, -0.3f), lamp);
        std::vector<pt_flat_sphere> sph;
        int32_t mats[3];
        mats[0] = diffuse(*s, rng.uniform3(0.1f, 0.9f));
        const pt_vec3 metal_albedo = rng.uniform3(0.5f, 1);
        const float metal_fuzz = rng.uniform(0, 0.3f);
        mats[1] = add_material(*s, PT_METAL, metal_albedo, metal_fuzz, 1, -1);
        mats[2] = add_material(*s, PT_GLASS, v3(0, 0, 0), 0, 1.5f, -1);
        for(int k = 0; k < 3; k++) {
            for(int attempt = 0; attempt < 10000; attempt++) {
                float r = rng.uniform(0.3f, 0.5f);
                float x = rng.uniform(-1 + r + 0.02f, 1 - r - 0.02f), z = rng.uniform(-1 + r + 0.02f, 1 - r - 0.02f);
                pt_vec3 c = v3(x, r + 0.001f, z);
                bool clear = true;
                for(const auto& o : sph) if(length(o.center - c) < o.radius + r + 0.01f) clear = false;
                if(clear || attempt == 9999) { sph.push_back(pt_flat_sphere{c, r, mats[k]}); break; }
            }
        }
        build_into(*s, tris, sph, PTS_SPLIT_SAH);
        if(camera) *camera = make_camera(v3(0, 1, 4.2f), v3(0, 1, 0), v3(0, 1, 0), 40, float(width) / float(height), 0, 4.2f, 0, 1);
        return s;
    } catch(const std::exception& e) { last_error = e.what(); return nullptr; }
}

pts_scene *pts_icosphere(int32_t frequency, uint64_t seed, int32_t width, int32_t height, pt_flat_camera *camera) {
    try {
        if(frequency < 1 || frequency > 2048) throw std::invalid_argument("icosphere frequency out of range");
        auto *s = new pts_scene;
        seeded rng(seed);
        value_noise noise{seed * 0x2545F4914F6CDD1Dull + 1};
        const float t = float((1.0 + std::sqrt(5.0)) / 2.0);
        const pt_vec3 corners[12] = {v3(-1, t, 0), v3(1, t, 0), v3(-1, -t, 0), v3(1, -t, 0), v3(0, -1, t), v3(0, 1, t),
                                     v3(0, -1, -t), v3(0, 1, -t), v3(t, 0, -1), v3(t, 0, 1), v3(-t, 0, -1), v3(-t, 0, 1)};
        const int faces[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                                  {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                                  {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
        const int n = frequency;
        // Shared vertices: key = sorted (corner, weight) pairs with non-zero weight.
        std::unordered_map<uint64_t, int32_t> index;
        index.reserve(size_t(10) * n * n + 16);
        std::vector<pt_vec3> pos;
        pos.reserve(size_t(10) * n * n + 16);
        auto vertex = [&](const int f[3], int i, int j) -> int32_t {
            int w[3] = {n - i - j, i, j};
            int c[3] = {f[0], f[1], f[2]};
            // order by corner index for a canonical sum
            for(int a = 0; a < 3; a++) for(int b = a + 1; b < 3; b++) if(c[b] < c[a]) { std::swap(c[a], c[b]); std::swap(w[a], w[b]); }
            // (corner, weight) pairs in corner order; weights <= 2048 fit 12 bits.
            uint64_t full = 0;
            for(int a = 0; a < 3; a++) if(w[a]) full = (full << 16) | (uint64_t(c[a]) << 12) | uint64_t(w[a]);
            auto it = index.find(full);
            if(it != index.end()) return it->second;
            pt_vec3 p = v3(0, 0, 0);
            for(int a = 0; a < 3; a++) if(w[a]) p = p + (float(w[a]) / float(n)) * corners[c[a]];
            pt_vec3 d = unit(p);
            float radius = 1.0f + 0.05f * noise.octaves(d);
            int32_t id = int32_t(pos.size());
            pos.push_back(radius * d);
            index.emplace(full, id);
            return id;
        };
        const int K = 4; // clusters per face edge
        std::vector<int32_t> cluster_material(static_cast<size_t>(20 * K * K));
        for(auto& m : cluster_material) {
            float pick = rng.uniform(0, 1);
            if(pick < 0.60f) m = diffuse(*s, rng.uniform3(0.1f, 0.9f));
            else if(pick < 0.85f) {
                const float fuzz = rng.uniform(0, 0.3f);
                m = add_material(*s, PT_METAL, rng.uniform3(0.5f, 1), fuzz, 1, -1);
            }
            else m = add_material(*s, PT_GLASS, v3(0, 0, 0), 0, 1.5f, -1);
        }
        struct face_ref { int32_t a, b, c, material; };
        std::vector<face_ref> mesh;
        mesh.reserve(size_t(20) * n * n);
        for(int f = 0; f < 20; f++) {
            for(int i = 0; i < n; i++) for(int j = 0; i + j < n; j++) {
                int cl = f * K * K + (i * K / n) * K + (j * K / n);
                int32_t m = cluster_material[size_t(cl)];
                mesh.push_back({vertex(faces[f], i, j), vertex(faces[f], i + 1, j), vertex(faces[f], i, j + 1), m});
                if(i + j < n - 1) mesh.push_back({vertex(faces[f], i + 1, j), vertex(faces[f], i + 1, j + 1), vertex(faces[f], i, j + 1), m});
            }
        }
        // Area-weighted vertex normals of the displaced mesh.
        std::vector<pt_vec3> nrm(pos.size(), v3(0, 0, 0));
        for(const auto& fr : mesh) {
            pt_vec3 fn = cross(pos[size_t(fr.b)] - pos[size_t(fr.a)], pos[size_t(fr.c)] - pos[size_t(fr.a)]);
            nrm[size_t(fr.a)] = nrm[size_t(fr.a)] + fn; nrm[size_t(fr.b)] = nrm[size_t(fr.b)] + fn; nrm[size_t(fr.c)] = nrm[size_t(fr.c)] + fn;
        }
        for(auto& v : nrm) v = unit(v);
        std::vector<pt_flat_triangle> tris;
        tris.reserve(mesh.size() + 10);
        for(const auto& fr : mesh) {
            pt_flat_triangle tr = flat_tri(pos[size_t(fr.a)], pos[size_t(fr.b)], pos[size_t(fr.c)], fr.material);
            tr.n0 = nrm[size_t(fr.a)]; tr.n1 = nrm[size_t(fr.b)]; tr.n2 = nrm[size_t(fr.c)];
            tr.smooth = 1;
            tris.push_back(tr);
        }
        int32_t floor_m = diffuse(*s, v3(0.5f, 0.5f, 0.5f));
        quad(tris, v3(-6, -1.15f, 6), v3(6, -1.15f, 6), v3(6, -1.15f, -6), v3(-6, -1.15f, -6), floor_m);
        // Four 1x1 lights around and above the sphere, each facing its centre.
        const pt_vec3 centres[4] = {v3(-2.2f, 2.4f, 1.6f), v3(2.2f, 2.0f, 1.8f), v3(-1.8f, 2.6f, -2.0f), v3(2.0f, 3.0f, -1.6f)};
        for(const auto& c : centres) {
            int32_t lm = light(*s, v3(0, 0, 0));
            float e = rng.uniform(5, 20);
            s->textures[size_t(s->materials[size_t(lm)].texture_id)].color = v3(e, e, e);
            pt_vec3 w = unit(v3(0, 0, 0) - c);
            pt_vec3 a = unit(cross(std::fabs(w.y) > 0.9f ? v3(1, 0, 0) : v3(0, 1, 0), w));
            pt_vec3 b = cross(w, a);
            pt_vec3 p0 = c - 0.5f * a - 0.5f * b, p1 = c + 0.5f * a - 0.5f * b, p2 = c + 0.5f * a + 0.5f * b, p3 = c - 0.5f * a + 0.5f * b;
            quad(tris, p0, p1, p2, p3, lm);
        }
        build_into(*s, tris, {}, PTS_SPLIT_SAH);
        if(camera) *camera = make_camera(v3(0.0f, 0.55f, 3.1f), v3(0, -0.05f, 0), v3(0, 1, 0), 40, float(width) / float(height), 0, 3.1f, 0, 1);
        return s;
    } catch(const std::exception& e) { last_error = e.what(); return nullptr; }
}

pts_scene *pts_soup(int64_t triangle_count, uint64_t seed, int32_t width, int32_t height, pt_flat_camera *camera) {
    try {
        if(triangle_count < 1 || triangle_count > (int64_t(1) << 28)) throw std::invalid_argument("soup size out of range");
        auto *s = new pts_scene;
        seeded rng(seed);
        std::normal_distribution<float> gauss(0.0f, 1.0f);
        int clusters = int(std::max<int64_t>(4, std::min<int64_t>(256, triangle_count / 4096)));
        std::vector<pt_vec3> centre(static_cast<size_t>(clusters));
        std::vector<float> spread(static_cast<size_t>(clusters));
        std::vector<int32_t> mat(static_cast<size_t>(clusters));
        for(int c = 0; c < clusters; c++) {
            const float cz = rng.uniform(-2, 2), cy = rng.uniform(-1, 1.5f), cx = rng.uniform(-2, 2);
            centre[size_t(c)] = v3(cx, cy, cz);
            spread[size_t(c)] = rng.uniform(0.1f, 0.5f);
            float pick = rng.uniform(0, 1);
            if(pick < 0.6f) mat[size_t(c)] = diffuse(*s, rng.uniform3(0.1f, 0.9f));
            else if(pick < 0.85f) {
                const float fuzz = rng.uniform(0, 0.3f);
                mat[size_t(c)] = add_material(*s, PT_METAL, rng.uniform3(0.5f, 1), fuzz, 1, -1);
            }
            else mat[size_t(c)] = add_material(*s, PT_GLASS, v3(0, 0, 0), 0, 1.5f, -1);
        }
        std::vector<pt_flat_triangle> tris;
        tris.reserve(size_t(triangle_count) + 8);
        float edge = 0.6f * std::cbrt(1.0f / float(triangle_count)) * 4.0f;
        for(int64_t i = 0; i < triangle_count; i++) {
            int c = int(rng.engine() % uint64_t(clusters));
            float sp = spread[size_t(c)];
            const float gz = gauss(rng.engine), gy = gauss(rng.engine), gx = gauss(rng.engine);
            const pt_vec3 p = centre[size_t(c)] + v3(sp * gx, sp * gy, sp * gz);
            const pt_vec3 da = rng.uniform3(-1, 1);
            const pt_vec3 a = p + v3(edge * da.x, edge * da.y, edge * da.z);
            const pt_vec3 db = rng.uniform3(-1, 1);
            const pt_vec3 b = p + v3(edge * db.x, edge * db.y, edge * db.z);
            tris.push_back(flat_tri(p, a, b, mat[size_t(c)]));
        }
        int32_t floor_m = diffuse(*s, v3(0.5f, 0.5f, 0.5f));
        quad(tris, v3(-8, -2, 8), v3(8, -2, 8), v3(8, -2, -8), v3(-8, -2, -8), floor_m);
        int32_t l1 = light(*s, v3(8, 8, 8)), l2 = light(*s, v3(4, 5, 6));
        quad(tris, v3(-1, 4, -1), v3(-1, 4, 1), v3(1, 4, 1), v3(1, 4, -1), l1);
        quad(tris, v3(3, 2, -1), v3(3, 2, 1), v3(3, 4, 1), v3(3, 4, -1), l2);
        build_into(*s, tris, {}, PTS_SPLIT_SAH);
        if(camera) *camera = make_camera(v3(0, 1, 7), v3(0, 0, 0), v3(0, 1, 0), 45, float(width) / float(height), 0, 7, 0, 1);
        return s;
    } catch(const std::exception& e) { last_error = e.what(); return nullptr; }
}

void pts_counts(const pts_scene *s, int64_t counts[6], int32_t root_depth[2]) {
    counts[0] = int64_t(s->triangles.size()); counts[1] = int64_t(s->spheres.size());
    counts[2] = int64_t(s->nodes.size()); counts[3] = int64_t(s->materials.size());
    counts[4] = int64_t(s->textures.size()); counts[5] = int64_t(s->pixels.size());
    root_depth[0] = s->root; root_depth[1] = s->max_depth;
}

void pts_copy(const pts_scene *s, pt_flat_triangle *t, pt_flat_sphere *sp, pt_flat_node *n, pt_flat_material *m,
              pt_flat_texture *x, uint8_t *p) {
    if(t) std::copy(s->triangles.begin(), s->triangles.end(), t);
    if(sp) std::copy(s->spheres.begin(), s->spheres.end(), sp);
    if(n) std::copy(s->nodes.begin(), s->nodes.end(), n);
    if(m) std::copy(s->materials.begin(), s->materials.end(), m);
    if(x) std::copy(s->textures.begin(), s->textures.end(), x);
    if(p) std::copy(s->pixels.begin(), s->pixels.end(), p);
}

void pts_desc(const pts_scene *s, pt_scene_desc *d) {
    d->triangles = s->triangles.data(); d->triangle_count = int64_t(s->triangles.size());
    d->spheres = s->spheres.data(); d->sphere_count = int64_t(s->spheres.size());
    d->nodes = s->nodes.data(); d->node_count = int64_t(s->nodes.size());
    d->materials = s->materials.data(); d->material_count = int64_t(s->materials.size());
    d->textures = s->textures.data(); d->texture_count = int64_t(s->textures.size());
    d->pixels = s->pixels.data(); d->pixel_bytes = int64_t(s->pixels.size());
    d->root = s->root; d->max_depth = s->max_depth;
}

void pts_free(pts_scene *s) { delete s; }

}
