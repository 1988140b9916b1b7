// wide_model.hip -- CPU model of the shaded path's traversal (test
// infrastructure, not part of the product).
//
// Builds the layouts exactly as the renderer's upload does (pt_layout.h) and
// runs the traversal source the kernels run (pt_traverse.h) on the host, one
// ray at a time, so that the 4-wide nearest-first traversal, its
// reference-order fallback and its stack occupancy can be examined without a
// GPU.  Built by tests/native/build_model.py with the renderer's arithmetic
// flags (no contraction; host IEEE division and square root).
#include <cstdio>
#include <cstring>
#include <vector>

#include "pt_layout.h"
#include "pt_traverse.h"

namespace {

// Host wide stack: records the deepest entry written; an entry beyond the
// capacity the renderer sizes its stack for is counted, never written.
struct model_stack {
    std::vector<int2> e;
    int capacity = 0;
    int max_used = 0;
    int overflow = 0;
    void store(int k, int32_t ref, float entry) {
        if(k < 0 || k >= capacity) { overflow++; return; }
        e[size_t(k)] = make_int2(ref, pt::f2i(entry));
        if(k + 1 > max_used) max_used = k + 1;
    }
    int2 load(int k) const {
        if(k < 0 || k >= capacity) return make_int2(pt::empty_root, 0);
        return e[size_t(k)];
    }
};

pt::scene_dev host_view(const pt_scene_desc& s, const pt::scene_layout& L) {
    pt::scene_dev v{};
    v.bnodes = L.bnodes.data(); v.tri_isect = L.tri_isect.data(); v.sph_isect = L.sph_isect.data();
    v.shade_tri = L.shade_tri.data();
    v.root_ref = L.root_ref;
    v.triangles = s.triangles; v.spheres = s.spheres; v.nodes = s.nodes; v.node_count = int32_t(s.node_count);
    v.materials = s.materials; v.textures = s.textures; v.pixels = s.pixels;
    v.stack_capacity = L.stack_capacity;
    v.wnodes = L.has_wide ? L.wide.nodes.data() : nullptr;
    v.leaves = L.has_wide ? L.wide.leaves.data() : nullptr;
    v.wroot_ref = L.has_wide ? L.wide.root_ref : pt::empty_root;
    v.wide_stack_capacity = L.wide_stack_capacity;
    return v;
}

}  // namespace

extern "C" {

// Layout facts: [0] wide depth, [1] wide stack capacity, [2] binary stack
// capacity, [3] wide nodes, [4] leaf positions, [5] whether the wide layout
// exists.  Returns 0, or -1 with a message in `error` for an invalid scene.
int wm_layout_info(const pt_scene_desc *scene, int32_t *info, char *error, size_t capacity) {
    try {
        pt::validate_scene(*scene);
        const pt::scene_layout L = pt::build_layout(*scene, true);
        info[0] = L.has_wide ? L.wide.depth : 0;
        info[1] = L.wide_stack_capacity;
        info[2] = L.stack_capacity;
        info[3] = int32_t(L.wide.nodes.size() / 8);
        info[4] = int32_t(L.wide.leaves.size() / 3);
        info[5] = L.has_wide ? 1 : 0;
        return 0;
    } catch(const std::exception& e) {
        std::snprintf(error, capacity, "%s", e.what());
        return -1;
    }
}

// Closest (any_hit = 0) or any (any_hit = 1) hit of each ray on
// [0.001, high[i]).  mode 0: the renderer's 4-wide traversal with its
// reference-order fallback; mode 1: the same traversal without the fallback
// (window 0); mode 2: the reference-order traversal over the binary layout.
// Per ray: hit[4] = {t bits, u bits, v bits, code} (code -1 for no hit) and
// stats[4] = {wide nodes or binary branches visited, primitives tested,
// deepest stack entry + 1, retraced}.  Returns the number of stack
// overflows (writes beyond the renderer's capacity), or -1 on an invalid scene.
int64_t wm_trace(const pt_scene_desc *scene, const pt_ray *rays, const float *high, int64_t count, int32_t mode,
                 int32_t any_hit, int32_t *hit, int32_t *stats) {
    try {
        pt::validate_scene(*scene);
    } catch(const std::exception&) {
        return -1;
    }
    const pt::scene_layout L = pt::build_layout(*scene, true);
    const pt::scene_dev v = host_view(*scene, L);
    int64_t overflows = 0;
    model_stack stk;
    stk.capacity = mode == 2 ? L.stack_capacity : L.wide_stack_capacity;
    stk.e.assign(size_t(std::max(1, stk.capacity)), make_int2(0, 0));
    for(int64_t i = 0; i < count; i++) {
        const pt_ray& ray = rays[i];
        const pt::ray_in r = pt::setup_ray(ray.origin.x, ray.origin.y, ray.origin.z, ray.direction.x,
                                           ray.direction.y, ray.direction.z);
        pt::best_hit b{0, 0, 0, -1};
        int boxes = 0, prims = 0, spheres = 0, error = 0;
        bool found = false, retraced = false;
        stk.max_used = 0;
        stk.overflow = 0;
        if(mode == 2 || !L.has_wide) {
            pt::wide_as_binary<model_stack> bs{stk};
            found = any_hit ? pt::traverse_bvh<true, true>(v, L.root_ref, r, 0.001f, high[i], bs, b, boxes, prims, spheres, error)
                            : pt::traverse_bvh<false, true>(v, L.root_ref, r, 0.001f, high[i], bs, b, boxes, prims, spheres, error);
        } else {
            pt::wide_query q;
            q.r = r;
            pt::wide_begin(L.wide.root_ref, q, any_hit != 0, high[i]);
            if(mode == 0) while(!pt::wide_step<true>(v, q, stk, boxes, prims, spheres)) {}
            else while(!pt::wide_step<true, model_stack, 0>(v, q, stk, boxes, prims, spheres)) {}
            found = q.found;
            b = q.best;
            if(pt::needs_retrace(q)) {
                retraced = true;
                found = pt::exact_retrace(v, L.root_ref, r, stk, b, error);
            }
        }
        overflows += stk.overflow + error;
        if(!found) b = pt::best_hit{0, 0, 0, -1};
        hit[4 * i] = pt::f2i(b.t); hit[4 * i + 1] = pt::f2i(b.u); hit[4 * i + 2] = pt::f2i(b.v);
        hit[4 * i + 3] = found ? b.code : -1;
        stats[4 * i] = boxes; stats[4 * i + 1] = prims; stats[4 * i + 2] = stk.max_used; stats[4 * i + 3] = retraced;
    }
    return overflows;
}

// The Fresnel power of pt_libm.h: pow5_rounded(x[i]) into exact[i], and the
// host-libm value the renderer uses (table-corrected) into libm[i].  Returns
// the number of table entries (built once).
int64_t wm_pow5(const float *x, int64_t count, double *exact, double *libm) {
    static std::vector<int32_t> keys;
    static std::vector<double> values;
    static bool built = false;
    if(!built) { pt::build_pow5_table(keys, values); built = true; }
    const pt::pow5_table t{keys.data(), values.data(), int32_t(keys.size())};
    for(int64_t i = 0; i < count; i++) {
        exact[i] = pt::pow5_rounded(x[i]);
        libm[i] = pt::libm_pow5(t, x[i]);
    }
    return int64_t(keys.size());
}

}  // extern "C"
