// pt_layout.h -- host-side preparation of an uploaded scene: the boundary's
// validation (src/flat/flat_validate.h, src/renderer/render_validation.h),
// the emitter table (src/renderer/scene_view.h) and the traversal layouts the
// kernels read (binary branch records, triangle test records, shading
// records, the 4-wide hierarchy).  Host only; used by the runtime
// (pt_api.hip) and by the test suite's CPU model of the traversal.
#pragma once

#include <array>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include <thread>

#include "pt_device.h"
#include "pt_libm.h"

namespace pt {

// Failure of a boundary call: the status it returns and its message.
struct status_error : std::runtime_error {
    pt_status code;
    status_error(pt_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] inline void fail(pt_status c, const std::string& m) { throw status_error(c, m); }

// ---------------- validation (src/flat/flat_validate.h:10-83) ----------------
constexpr int stack_limit = 64; // flat_stack_capacity, src/flat/flat_data.h:52

inline bool finite3(const pt_vec3& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }
inline float axis(const pt_vec3& v, int a) { return a == 0 ? v.x : a == 1 ? v.y : v.z; }
inline bool valid_id(int32_t id, int64_t n) { return id >= 0 && int64_t(id) < n; }

inline void validate_scene(const pt_scene_desc& s) {
    auto require = [](bool ok) { if(!ok) fail(PT_ESCENE, "invalid flat scene data"); };
    require(s.triangle_count >= 0 && s.sphere_count >= 0 && s.node_count >= 0 && s.material_count >= 0 &&
            s.texture_count >= 0 && s.pixel_bytes >= 0);
    require((s.triangles || !s.triangle_count) && (s.spheres || !s.sphere_count) && (s.nodes || !s.node_count) &&
            (s.materials || !s.material_count) && (s.textures || !s.texture_count) && (s.pixels || !s.pixel_bytes));
    require(s.node_count < int64_t(INT32_MAX) && s.triangle_count < int64_t(INT32_MAX / 2) && s.sphere_count < int64_t(INT32_MAX / 2));
    for(int64_t i = 0; i < s.texture_count; i++) {
        const auto& t = s.textures[i];
        require(finite3(t.color) && finite3(t.scale) && finite3(t.offset));
        require(t.type >= PT_CONSTANT && t.type <= PT_MAP);
        if(t.type == PT_IMAGE) {
            require(t.width > 0 && t.height > 0 && t.byte_offset >= 0);
            uint64_t end = uint64_t(t.byte_offset) + 3 * uint64_t(t.width) * uint64_t(t.height);
            require(end <= uint64_t(s.pixel_bytes));
        } else if(t.type == PT_MAP) {
            require(valid_id(t.image_id, s.texture_count));
            require(s.textures[t.image_id].type == PT_IMAGE && (t.clamp == 0 || t.clamp == 1));
        }
    }
    for(int64_t i = 0; i < s.material_count; i++) {
        const auto& m = s.materials[i];
        require(m.type >= PT_DIFFUSE && m.type <= PT_LIGHT && finite3(m.albedo));
        require(std::isfinite(m.fuzz) && std::isfinite(m.ref_idx));
        if(m.type == PT_DIFFUSE || m.type == PT_LIGHT) require(valid_id(m.texture_id, s.texture_count));
        if(m.type == PT_GLASS) require(m.ref_idx > 0);
    }
    for(int64_t i = 0; i < s.triangle_count; i++) {
        const auto& t = s.triangles[i];
        require(valid_id(t.material_id, s.material_count) && (t.smooth == 0 || t.smooth == 1));
        require(finite3(t.v0) && finite3(t.v1) && finite3(t.v2) && finite3(t.n0) && finite3(t.n1) && finite3(t.n2));
        require(finite3(t.uv0) && finite3(t.uv1) && finite3(t.uv2));
    }
    for(int64_t i = 0; i < s.sphere_count; i++) {
        const auto& p = s.spheres[i];
        require(valid_id(p.material_id, s.material_count) && finite3(p.center) && std::isfinite(p.radius) && p.radius > 0);
    }
    if(s.root == -1) {
        require(s.node_count == 0 && s.triangle_count == 0 && s.sphere_count == 0 && s.max_depth == 0);
        return;
    }
    require(valid_id(s.root, s.node_count));
    std::vector<int> heights(size_t(s.node_count), 0);
    std::vector<char> active(size_t(s.node_count), 0);
    int64_t visited = 0;
    // Iterative post-order walk with the recursive checker's rules.
    struct frame { int32_t id; int depth; int stage; int left_height; };
    std::vector<frame> stack;
    int result = 0;
    auto enter = [&](int32_t id, int depth) -> bool {
        require(valid_id(id, s.node_count) && depth <= stack_limit && !active[size_t(id)]);
        if(heights[size_t(id)]) { require(depth + heights[size_t(id)] - 1 <= stack_limit); result = heights[size_t(id)]; return false; }
        active[size_t(id)] = 1;
        visited++;
        const auto& n = s.nodes[id];
        require(finite3(n.minimum) && finite3(n.maximum));
        for(int a = 0; a < 3; a++) require(axis(n.minimum, a) <= axis(n.maximum, a));
        stack.push_back({id, depth, 0, 0});
        return true;
    };
    if(enter(s.root, 1)) {
        while(!stack.empty()) {
            frame& f = stack.back();
            const auto& n = s.nodes[f.id];
            if(n.type == PT_BRANCH) {
                if(f.stage == 0) {
                    f.stage = 1;
                    int32_t child = n.left; int d = f.depth + 1;
                    if(enter(child, d)) continue;
                    // cached height returned in result
                }
                if(f.stage == 1) {
                    f.left_height = result;
                    f.stage = 2;
                    if(n.left != n.right) {
                        int32_t child = n.right; int d = f.depth + 1;
                        if(enter(child, d)) continue;
                    }
                }
                // stage 2: right height is in result (or equals left for an alias)
                int right_height = n.left == n.right ? f.left_height : result;
                int height = 1 + std::max(f.left_height, right_height);
                for(int32_t child : {n.left, n.right}) {
                    const auto& c = s.nodes[child];
                    for(int a = 0; a < 3; a++)
                        require(axis(n.minimum, a) <= axis(c.minimum, a) && axis(n.maximum, a) >= axis(c.maximum, a));
                }
                active[size_t(f.id)] = 0; heights[size_t(f.id)] = height; result = height;
                stack.pop_back();
            } else {
                if(n.type == PT_TRIANGLE) {
                    require(valid_id(n.primitive, s.triangle_count));
                    const auto& t = s.triangles[n.primitive];
                    for(const pt_vec3& p : {t.v0, t.v1, t.v2})
                        for(int a = 0; a < 3; a++) require(axis(p, a) >= axis(n.minimum, a) && axis(p, a) <= axis(n.maximum, a));
                } else if(n.type == PT_SPHERE) {
                    require(valid_id(n.primitive, s.sphere_count));
                    const auto& p = s.spheres[n.primitive];
                    for(int a = 0; a < 3; a++)
                        require(axis(p.center, a) - p.radius >= axis(n.minimum, a) && axis(p.center, a) + p.radius <= axis(n.maximum, a));
                } else require(false);
                active[size_t(f.id)] = 0; heights[size_t(f.id)] = 1; result = 1;
                stack.pop_back();
            }
        }
    }
    require(result == s.max_depth && visited == s.node_count);
}

// src/renderer/render_validation.h:7-22
inline size_t checked_pixel_count(int width, int height) {
    if(width <= 0 || height <= 0 || uint64_t(width) * uint64_t(height) > uint64_t(INT32_MAX))
        fail(PT_EINVAL, "image dimensions exceed supported pixel count");
    return size_t(width) * size_t(height);
}
inline void validate_mode(int32_t mode) {
    if(mode < PT_DEBUG_SHADED || mode > PT_DEBUG_INTERSECTIONS) fail(PT_EINVAL, "invalid render debug mode");
}
inline void validate_camera(const pt_flat_camera& c) {
    for(const pt_vec3& v : {c.origin, c.lower_left_corner, c.horizontal, c.vertical, c.u, c.v})
        if(!finite3(v)) fail(PT_EINVAL, "camera contains nonfinite values");
    if(!std::isfinite(c.lens_radius) || c.lens_radius < 0 || !std::isfinite(c.time0) || !std::isfinite(c.time1) || c.time1 < c.time0)
        fail(PT_EINVAL, "invalid camera lens/shutter");
}

// ---------------- light table (src/renderer/scene_view.h:10-28) ----------------
inline std::vector<light_rec> scene_lights(const pt_scene_desc& s) {
    std::vector<light_rec> lights;
    double area = 0;
    auto add = [&](int32_t type, int32_t id, double size) {
        if(!(size > 0) || !std::isfinite(size)) return;
        area += size;
        if(!std::isfinite(area) || area > FLT_MAX) fail(PT_EINVAL, "light area exceeds range");
        lights.push_back({type, id, float(area)});
    };
    for(int64_t i = 0; i < s.triangle_count; i++) {
        const auto& t = s.triangles[i];
        if(s.materials[t.material_id].type != PT_LIGHT) continue;
        pt_vec3 e1{t.v1.x - t.v0.x, t.v1.y - t.v0.y, t.v1.z - t.v0.z}, e2{t.v2.x - t.v0.x, t.v2.y - t.v0.y, t.v2.z - t.v0.z};
        pt_vec3 c{e1.y * e2.z - e1.z * e2.y, -(e1.x * e2.z - e1.z * e2.x), e1.x * e2.y - e1.y * e2.x};
        float len = std::sqrt(c.x * c.x + c.y * c.y + c.z * c.z);
        add(PT_TRIANGLE, int32_t(i), 0.5 * len);
    }
    for(int64_t i = 0; i < s.sphere_count; i++) {
        const auto& p = s.spheres[i];
        if(s.materials[p.material_id].type != PT_LIGHT) continue;
        add(PT_SPHERE, int32_t(i), 4 * M_PI * double(p.radius) * p.radius);
    }
    return lights;
}
// ---------------- host-libm power table (pt_libm.h) ----------------
// The table-domain inputs where the host libm's pow(double(x), 5.0) differs
// from pow5_rounded(x), ascending, with the host's values.
inline void build_pow5_table(std::vector<int32_t>& keys, std::vector<double>& values) {
    const unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int64_t low = pow5_table_low, high = pow5_table_high, span = high - low + 1;
    std::vector<std::vector<int32_t>> k_part(threads);
    std::vector<std::vector<double>> v_part(threads);
    std::vector<std::thread> pool;
    for(unsigned t = 0; t < threads; t++) {
        pool.emplace_back([&, t] {
            const int64_t begin = low + span * t / threads, end = low + span * (t + 1) / threads;
            for(int64_t k = begin; k < end; k++) {
                const float fk = float(k);
                if(int64_t(fk) != k) continue;  // k * 2^-24 is not a float
                const float x = fk * 0x1p-24f;
                const double host = std::pow(double(x), 5.0), exact = pow5_rounded(x);
                if(std::memcmp(&host, &exact, sizeof host) != 0) {
                    k_part[t].push_back(int32_t(k));
                    v_part[t].push_back(host);
                }
            }
        });
    }
    for(auto& th : pool) th.join();
    keys.clear();
    values.clear();
    for(unsigned t = 0; t < threads; t++) {
        keys.insert(keys.end(), k_part[t].begin(), k_part[t].end());
        values.insert(values.end(), v_part[t].begin(), v_part[t].end());
    }
}

// ---------------- traversal layouts ----------------
// Triangle test data: v0, e1 = v1-v0, e2 = v2-v0 and |e1 x e2| exactly as
// flat_intersect.h:9-12 computes them, plus 1e-7f*|e1 x e2| (:16).
inline std::vector<float4> triangle_test_records(const pt_scene_desc& s) {
    std::vector<float4> ti(size_t(3 * s.triangle_count));
    for(int64_t i = 0; i < s.triangle_count; i++) {
        const auto& t = s.triangles[i];
        float e1x = t.v1.x - t.v0.x, e1y = t.v1.y - t.v0.y, e1z = t.v1.z - t.v0.z;
        float e2x = t.v2.x - t.v0.x, e2y = t.v2.y - t.v0.y, e2z = t.v2.z - t.v0.z;
        float cx = e1y * e2z - e1z * e2y, cy = -(e1x * e2z - e1z * e2x), cz = e1x * e2y - e1y * e2x;
        float nl = std::sqrt(cx * cx + cy * cy + cz * cz);
        ti[size_t(3 * i)] = make_float4(t.v0.x, t.v0.y, t.v0.z, 0.0000001f * nl);
        ti[size_t(3 * i + 1)] = make_float4(e1x, e1y, e1z, nl);
        ti[size_t(3 * i + 2)] = make_float4(e2x, e2y, e2z, 0);
    }
    return ti;
}

// 4-wide layout for the shaded path.  Collapses the flat binary tree: every
// wide node takes up to four of the binary descendants (splitting the child of
// largest surface area first), and every subtree of at most wide_leaf_max
// leaves becomes a leaf cluster.  Leaf positions follow the depth-first,
// left-first leaf order of the flat tree, so a subtree's leaves are one
// contiguous range and a leaf's position is the order in which the
// reference's traversal would reach it (its tie rank).
//
// The boxes a wide node holds are boxes the reference itself tests
// (transport.h:61-83 tests branch boxes only):
// - a child that is a branch (a wide node or a cluster) keeps its own box;
// - a child that is a single leaf gets the box of its binary parent: the
//   reference never tests a leaf's box (which may be flat, for an
//   axis-aligned triangle, and then no ray would pass it), only the parent's
//   before it pushes the leaf;
// - a root that is a branch of at most wide_leaf_max leaves becomes a wide
//   node with that one cluster, so that the root's box is tested as the
//   reference tests it; a root that is a leaf is a cluster of one, untested,
//   as in the reference.
// Since the flat boxes nest (flat_validate.h:59-63) and rounding is monotone,
// a descendant's box test implies those of the binary branches between it
// and the wide node, so leaving those out changes no outcome.
struct wide_layout {
    std::vector<float4> nodes;   // 8 per node
    std::vector<float4> leaves;  // 3 per leaf position
    int32_t root_ref = empty_root;
    int depth = 0;
};

inline bool build_wide(const pt_scene_desc& s, const std::vector<float4>& ti, wide_layout& out) {
    if(s.root < 0) return false;
    const int64_t n = s.node_count;
    std::vector<int32_t> first(size_t(n), -1), count(size_t(n), 0);
    std::vector<int32_t> leaf_code;
    leaf_code.reserve(size_t(s.triangle_count + s.sphere_count));
    // Iterative depth-first walk, left before right; a node reached twice (a
    // shared subtree) disables the wide layout.
    struct item { int32_t node; bool exit; };
    std::vector<item> stack{{s.root, false}};
    while(!stack.empty()) {
        item it = stack.back(); stack.pop_back();
        const auto& nd = s.nodes[it.node];
        if(it.exit) { count[size_t(it.node)] = int32_t(leaf_code.size()) - first[size_t(it.node)]; continue; }
        if(first[size_t(it.node)] >= 0) return false;
        first[size_t(it.node)] = int32_t(leaf_code.size());
        if(nd.type == PT_BRANCH) {
            stack.push_back({it.node, true});
            if(nd.left != nd.right) stack.push_back({nd.right, false});
            stack.push_back({nd.left, false});
        } else {
            leaf_code.push_back(nd.primitive * 2 + (nd.type == PT_SPHERE ? 1 : 0));
            count[size_t(it.node)] = 1;
        }
    }
    if(leaf_code.size() >= (size_t(1) << 27)) return false;
    out.leaves.resize(3 * leaf_code.size());
    for(size_t k = 0; k < leaf_code.size(); k++) {
        int32_t code = leaf_code[k], prim = code >> 1;
        float cf; std::memcpy(&cf, &code, 4);
        if(code & 1) {
            const auto& p = s.spheres[prim];
            out.leaves[3 * k] = make_float4(p.center.x, p.center.y, p.center.z, p.radius);
            out.leaves[3 * k + 1] = make_float4(0, 0, 0, -1);
            out.leaves[3 * k + 2] = make_float4(0, 0, 0, cf);
        } else {
            out.leaves[3 * k] = ti[size_t(3 * prim)];
            out.leaves[3 * k + 1] = ti[size_t(3 * prim + 1)];
            float4 c = ti[size_t(3 * prim + 2)];
            out.leaves[3 * k + 2] = make_float4(c.x, c.y, c.z, cf);
        }
    }
    auto area = [&](int32_t node) {
        const auto& b = s.nodes[node];
        double dx = double(b.maximum.x) - b.minimum.x, dy = double(b.maximum.y) - b.minimum.y, dz = double(b.maximum.z) - b.minimum.z;
        return dx * dy + dy * dz + dz * dx;
    };
    // Which binary nodes become wide nodes: the cut that minimizes the
    // surface-area cost (a slot's work weighted by the surface area of the box
    // it is reached through), by dynamic programming over the binary tree,
    // children before parents.  F[x][k] is the least cost of covering the
    // subtree of x with at most k slots of one wide node; a slot is a single
    // leaf (reached through its binary parent's box), a cluster of at most
    // wide_leaf_max leaves, or a wide node of its own (slot_cost).
    std::vector<int32_t> parent(static_cast<size_t>(n), -1), post;
    post.reserve(size_t(n));
    {
        std::vector<std::pair<int32_t, bool>> walk{{s.root, false}};
        while(!walk.empty()) {
            auto [node, exit] = walk.back(); walk.pop_back();
            const auto& nd = s.nodes[node];
            if(exit || nd.type != PT_BRANCH) { post.push_back(node); continue; }
            walk.push_back({node, true});
            parent[size_t(nd.left)] = node;
            walk.push_back({nd.left, false});
            if(nd.right != nd.left) { parent[size_t(nd.right)] = node; walk.push_back({nd.right, false}); }
        }
    }
    constexpr double node_cost = 1.0, leaf_base = 0.3, leaf_each = 0.5;
    std::vector<std::array<double, 5>> F(static_cast<size_t>(n));
    std::vector<std::array<uint8_t, 5>> split(static_cast<size_t>(n));  // (i << 4) | j, or 0 for one slot
    std::vector<double> slot_cost(static_cast<size_t>(n));
    auto children = [&](int32_t x, int32_t& l, int32_t& r) {
        l = s.nodes[x].left; r = s.nodes[x].right;
        return l != r ? 2 : 1;
    };
    for(int32_t x : post) {
        const auto& nd = s.nodes[x];
        const size_t xi = size_t(x);
        if(nd.type != PT_BRANCH) {
            const int32_t p = parent[xi] >= 0 ? parent[xi] : x;
            slot_cost[xi] = area(p) * (leaf_base + leaf_each);
            F[xi].fill(slot_cost[xi]);
            split[xi].fill(0);
            continue;
        }
        int32_t l, r;
        const int m = children(x, l, r);
        // best cover of x's subtree below x with at most k slots (k >= 2)
        auto below = [&](int k, uint8_t& how) {
            if(m == 1) { how = uint8_t((k << 4) | 0); return F[size_t(l)][size_t(k)]; }
            double best = INFINITY;
            for(int i = 1; i < k; i++)
                for(int j = 1; i + j <= k; j++) {
                    const double c = F[size_t(l)][size_t(i)] + F[size_t(r)][size_t(j)];
                    if(c < best) { best = c; how = uint8_t((i << 4) | j); }
                }
            return best;
        };
        const int32_t leaves = count[xi];
        uint8_t how = 0;
        if(leaves <= wide_leaf_max) slot_cost[xi] = area(x) * (leaf_base + leaf_each * leaves);
        else slot_cost[xi] = area(x) * node_cost + below(4, how);
        F[xi][1] = slot_cost[xi];
        split[xi][1] = 0;
        for(int k = 2; k <= 4; k++) {
            const double c = below(k, how);
            if(c < slot_cost[xi]) { F[xi][size_t(k)] = c; split[xi][size_t(k)] = how; }
            else { F[xi][size_t(k)] = slot_cost[xi]; split[xi][size_t(k)] = 0; }
        }
    }
    // A cluster of count <= wide_leaf_max leaves from depth-first position first; note that
    // the first leaf alone encodes as -1, a valid reference.
    auto cluster_ref = [&](int32_t node) { return ~((first[size_t(node)] << 3) | (count[size_t(node)] - 1)); };
    bool failed = false;
    auto expandable = [&](int32_t node) { return s.nodes[node].type == PT_BRANCH && count[size_t(node)] > wide_leaf_max; };
    struct kid { int32_t node, parent; };
    std::function<int32_t(int32_t, int)> make;
    // One wide node over `kids`: each kid's box (its parent's for a leaf) and reference.
    auto emit = [&](const std::vector<kid>& kids, int depth) -> int32_t {
        out.depth = std::max(out.depth, depth);
        size_t w = out.nodes.size() / 8;
        out.nodes.resize(out.nodes.size() + 8);
        float lo[3][4], hi[3][4];
        int32_t ref[4];
        for(int k = 0; k < 4; k++) {
            for(int a = 0; a < 3; a++) { lo[a][k] = INFINITY; hi[a][k] = -INFINITY; }
            ref[k] = wide_empty;
        }
        for(size_t k = 0; k < kids.size(); k++) {
            const int32_t boxed = s.nodes[kids[k].node].type == PT_BRANCH ? kids[k].node : kids[k].parent;
            const auto& c = s.nodes[boxed];
            lo[0][k] = c.minimum.x; lo[1][k] = c.minimum.y; lo[2][k] = c.minimum.z;
            hi[0][k] = c.maximum.x; hi[1][k] = c.maximum.y; hi[2][k] = c.maximum.z;
            ref[k] = make(kids[k].node, depth + 1);
            if(failed) return 0;
        }
        float4 *dst = &out.nodes[8 * w];
        for(int a = 0; a < 3; a++) {
            dst[a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
            dst[3 + a] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
        }
        float rf[4];
        std::memcpy(rf, ref, sizeof rf);
        dst[6] = make_float4(rf[0], rf[1], rf[2], rf[3]);
        dst[7] = make_float4(0, 0, 0, 0);
        return int32_t(w);
    };
    // The slots covering the subtree of y with at most k of them (y's binary
    // parent p), left to right.
    std::function<void(int32_t, int32_t, int, std::vector<kid>&)> cover = [&](int32_t y, int32_t p, int k,
                                                                             std::vector<kid>& kids) {
        const uint8_t how = k >= 2 ? split[size_t(y)][size_t(k)] : 0;
        if(how == 0) { kids.push_back({y, p}); return; }
        int32_t l, r;
        const int m = children(y, l, r);
        if(m == 1) { cover(l, y, how >> 4, kids); return; }
        cover(l, y, how >> 4, kids);
        cover(r, y, how & 15, kids);
    };
    make = [&](int32_t b, int depth) -> int32_t {
        if(!expandable(b)) return cluster_ref(b);
        std::vector<kid> kids;
        int32_t l, r;
        if(children(b, l, r) == 1) cover(l, b, 4, kids);
        else {
            // the split of b's subtree into at most 4 slots that slot_cost[b] was computed with
            double best = INFINITY;
            int bi = 1, bj = 1;
            for(int i = 1; i < 4; i++)
                for(int j = 1; i + j <= 4; j++) {
                    const double c = F[size_t(l)][size_t(i)] + F[size_t(r)][size_t(j)];
                    if(c < best) { best = c; bi = i; bj = j; }
                }
            cover(l, b, bi, kids);
            cover(r, b, bj, kids);
        }
        if(kids.empty() || kids.size() > 4) {  // cannot happen: at most 4 slots by construction
            failed = true;
            return 0;
        }
        return emit(kids, depth);
    };
    out.depth = 1;
    if(s.nodes[s.root].type != PT_BRANCH) out.root_ref = cluster_ref(s.root);
    else if(!expandable(s.root)) out.root_ref = emit({{s.root, s.root}}, 1);
    else out.root_ref = make(s.root, 1);
    return !failed;
}

// Everything the kernels read of one scene besides the transfer arrays
// themselves, built on the host at upload.
struct scene_layout {
    std::vector<int32_t> branch_id;   // compact branch index of each flat node, -1 for leaves
    std::vector<float4> bnodes;       // 2 per branch: {lo.x,lo.y,lo.z,hi.x} {hi.y,hi.z,left_ref,right_ref}
    std::vector<float4> tri_isect;    // 3 per triangle (triangle_test_records)
    std::vector<float4> sph_isect;    // 1 per sphere: {center, radius}
    std::vector<float4> shade_tri;    // 5 per triangle (scene_dev::shade_tri)
    std::vector<light_rec> lights;
    wide_layout wide;
    bool has_wide = false;
    int32_t root_ref = empty_root;
    int stack_capacity = 2;           // binary traversal stack entries
    int wide_stack_capacity = 0;      // 4-wide stack entries (0 without the wide layout)

    // Child reference of flat node `node` in the binary layout.
    int32_t ref_of(const pt_scene_desc& s, int32_t node) const {
        const auto& n = s.nodes[node];
        if(n.type == PT_BRANCH) return branch_id[size_t(node)];
        return leaf_ref(n.primitive, n.type == PT_SPHERE);
    }
};

inline scene_layout build_layout(const pt_scene_desc& s, bool with_wide) {
    scene_layout L;
    // Binary layout: branch nodes renumbered in array order, children encoded
    // as branch indices or leaf references.
    L.branch_id.assign(size_t(s.node_count), -1);
    int64_t branches = 0;
    for(int64_t i = 0; i < s.node_count; i++) if(s.nodes[i].type == PT_BRANCH) L.branch_id[size_t(i)] = int32_t(branches++);
    L.bnodes.resize(size_t(2 * branches));
    for(int64_t i = 0; i < s.node_count; i++) {
        const auto& n = s.nodes[i];
        if(n.type != PT_BRANCH) continue;
        size_t b = size_t(L.branch_id[size_t(i)]);
        int32_t l = L.ref_of(s, n.left), r = L.ref_of(s, n.right);
        float lf, rf;
        std::memcpy(&lf, &l, 4); std::memcpy(&rf, &r, 4);
        L.bnodes[2 * b] = make_float4(n.minimum.x, n.minimum.y, n.minimum.z, n.maximum.x);
        L.bnodes[2 * b + 1] = make_float4(n.maximum.y, n.maximum.z, lf, rf);
    }
    L.tri_isect = triangle_test_records(s);
    // Shading records (see scene_dev::shade_tri).  The outward normal is
    // cross(e1,e2) scaled by float(1.0/|e1 x e2|) as vec3::operator/= does
    // (src/core/vec3.h:152-158, flat_intersect.h:28).
    L.shade_tri.resize(size_t(5 * s.triangle_count));
    for(int64_t i = 0; i < s.triangle_count; i++) {
        const auto& t = s.triangles[i];
        const float4 B = L.tri_isect[size_t(3 * i + 1)], C = L.tri_isect[size_t(3 * i + 2)];
        float cx = B.y * C.z - B.z * C.y, cy = -(B.x * C.z - B.z * C.x), cz = B.x * C.y - B.y * C.x;
        float k = float(1.0 / double(B.w));
        float mf, sf;
        std::memcpy(&mf, &t.material_id, 4);
        std::memcpy(&sf, &t.smooth, 4);
        L.shade_tri[size_t(5 * i)] = make_float4(cx * k, cy * k, cz * k, mf);
        L.shade_tri[size_t(5 * i + 1)] = make_float4(t.n0.x, t.n0.y, t.n0.z, sf);
        L.shade_tri[size_t(5 * i + 2)] = make_float4(t.n1.x, t.n1.y, t.n1.z, t.uv0.x);
        L.shade_tri[size_t(5 * i + 3)] = make_float4(t.n2.x, t.n2.y, t.n2.z, t.uv0.y);
        L.shade_tri[size_t(5 * i + 4)] = make_float4(t.uv1.x, t.uv1.y, t.uv2.x, t.uv2.y);
    }
    L.sph_isect.resize(size_t(s.sphere_count));
    for(int64_t i = 0; i < s.sphere_count; i++) {
        const auto& p = s.spheres[i];
        L.sph_isect[size_t(i)] = make_float4(p.center.x, p.center.y, p.center.z, p.radius);
    }
    L.lights = scene_lights(s);
    L.root_ref = s.root < 0 ? empty_root : L.ref_of(s, s.root);
    L.stack_capacity = std::max(2, std::min(stack_limit, int(s.max_depth)));
    if(with_wide && build_wide(s, L.tri_isect, L.wide) && L.wide.nodes.size() / 8 <= wide_node_limit) {
        L.has_wide = true;
        L.wide_stack_capacity = std::max(3 * L.wide.depth + 1, L.stack_capacity);
    }
    return L;
}

} // namespace pt
