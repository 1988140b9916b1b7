// pt_scene_impl.h -- internals shared by the host scene sources
// (pt_scene.cpp: builders and synthetic scenes; pt_obj.cpp: OBJ/MTL import,
// images, the reference-order builder).  Not part of the C ABI.
#ifndef PT_SCENE_IMPL_H
#define PT_SCENE_IMPL_H

#include <algorithm>
#include <cmath>
#include <limits>
#include <string>
#include <vector>

#include "pt_scene.h"

struct pts_scene {
    std::vector<pt_flat_triangle> triangles;
    std::vector<pt_flat_sphere> spheres;
    std::vector<pt_flat_node> nodes;
    std::vector<pt_flat_material> materials;
    std::vector<pt_flat_texture> textures;
    std::vector<uint8_t> pixels;
    int32_t root = -1, max_depth = 0;
    std::string warnings;
};

namespace pts {

void set_error(const std::string& message);

inline pt_vec3 v3(float x, float y, float z) { return pt_vec3{x, y, z}; }
inline pt_vec3 operator+(pt_vec3 a, pt_vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline pt_vec3 operator-(pt_vec3 a, pt_vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline pt_vec3 operator*(float t, pt_vec3 v) { return v3(t * v.x, t * v.y, t * v.z); }
inline float dot(pt_vec3 a, pt_vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline pt_vec3 cross(pt_vec3 a, pt_vec3 b) {
    return v3(a.y * b.z - a.z * b.y, -(a.x * b.z - a.z * b.x), a.x * b.y - a.y * b.x);
}
inline float length(pt_vec3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
inline pt_vec3 unit(pt_vec3 v) { float l = length(v); return v3(v.x / l, v.y / l, v.z / l); }
inline float axis_of(pt_vec3 v, int a) { return a == 0 ? v.x : a == 1 ? v.y : v.z; }

struct box3 {
    float lo[3], hi[3];
    void empty() {
        for(int a = 0; a < 3; a++) { lo[a] = std::numeric_limits<float>::infinity(); hi[a] = -lo[a]; }
    }
    void grow(const box3& b) {
        for(int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], b.lo[a]); hi[a] = std::max(hi[a], b.hi[a]); }
    }
    void grow(const float p[3]) {
        for(int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], p[a]); hi[a] = std::max(hi[a], p[a]); }
    }
    double area() const {
        double dx = std::max(0.0f, hi[0] - lo[0]), dy = std::max(0.0f, hi[1] - lo[1]), dz = std::max(0.0f, hi[2] - lo[2]);
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

// Leaf bounds exactly as the reference computes them.
inline box3 triangle_bounds(const pt_flat_triangle& t) {
    // src/geometry/triangle.h:96-109: min/max per axis, padded by 1e-4 and
    // pushed one ulp further outwards.
    box3 b;
    const float inf = std::numeric_limits<float>::infinity();
    for(int a = 0; a < 3; a++) {
        float x0 = axis_of(t.v0, a), x1 = axis_of(t.v1, a), x2 = axis_of(t.v2, a);
        float mn = x0 < (x1 < x2 ? x1 : x2) ? x0 : (x1 < x2 ? x1 : x2);
        float mx = x0 > (x1 > x2 ? x1 : x2) ? x0 : (x1 > x2 ? x1 : x2);
        b.lo[a] = std::nextafter(mn - 0.0001f, -inf);
        b.hi[a] = std::nextafter(mx + 0.0001f, inf);
    }
    return b;
}
inline box3 sphere_bounds(const pt_flat_sphere& s) {
    // src/geometry/sphere.h:69-75
    box3 b;
    for(int a = 0; a < 3; a++) { b.lo[a] = axis_of(s.center, a) - s.radius; b.hi[a] = axis_of(s.center, a) + s.radius; }
    return b;
}


pt_flat_texture constant_texture(pt_vec3 color);
void build_scene(pts_scene& s, const std::vector<pt_flat_triangle>& tris, const std::vector<pt_flat_sphere>& sph,
                 int split);
pt_flat_material flat_material(int32_t type, pt_vec3 albedo, float fuzz, float ref_idx, int32_t texture);

} // namespace pts

#endif
