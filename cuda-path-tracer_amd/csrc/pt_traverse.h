// pt_traverse.h -- ray set-up, the slab, triangle and sphere tests and the
// BVH traversals of the path-tracing core.  Every function here is compiled
// for the device (pt_kernels.hip) and for the host, where the test suite's
// CPU model of the traversal (tests/native/wide_model.hip) runs the same
// source on the same layouts, so that traversal order, stack occupancy and
// the exact-order fallback can be examined off the GPU.
//
// Arithmetic follows the reference's host evaluation order exactly; both
// builds use IEEE division and square root without contraction.
#pragma once

#include <cfloat>
#include <cmath>
#include <stdint.h>

#include "pt_device.h"

#define PT_HD __host__ __device__ __forceinline__

namespace pt {

PT_HD int32_t f2i(float x) { return __builtin_bit_cast(int32_t, x); }
PT_HD float i2f(int32_t x) { return __builtin_bit_cast(float, x); }
PT_HD uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
PT_HD float u2f(uint32_t x) { return __builtin_bit_cast(float, x); }

// ---------------- vec3 arithmetic (src/core/vec3.h:49-162) ----------------
struct v3 { float x, y, z; };
PT_HD v3 mk(float x, float y, float z) { return v3{x, y, z}; }
PT_HD v3 mk(const pt_vec3& a) { return v3{a.x, a.y, a.z}; }
PT_HD v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
PT_HD v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
PT_HD v3 mul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
// operator*(float,vec3) / operator*(vec3,float): t*e[i] (vec3.h:93-103)
PT_HD v3 scale(float t, v3 v) { return mk(t * v.x, t * v.y, t * v.z); }
// operator/(vec3,float): per-component division (vec3.h:97-99)
PT_HD v3 divs(v3 v, float t) { return mk(v.x / t, v.y / t, v.z / t); }
PT_HD v3 neg(v3 v) { return mk(-v.x, -v.y, -v.z); }
PT_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// vec3.h:109-115 (negated middle difference kept for signed zeros)
PT_HD v3 cross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, -(a.x * b.z - a.z * b.x), a.x * b.y - a.y * b.x); }
PT_HD float length(v3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
PT_HD float squared_length(v3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
// operator/=(float): multiply by float(1.0/t) computed in double (vec3.h:152-158)
PT_HD v3 div_assign(v3 v, float t) {
    float k = float(1.0 / double(t));
    return mk(v.x * k, v.y * k, v.z * k);
}
PT_HD v3 unit_vector(v3 v) { return divs(v, length(v)); }

// ---------------- rays ----------------
struct ray_in {
    float ox, oy, oz, dx, dy, dz;
    float ix, iy, iz;   // 1.0f/d per axis (the value trace_box recomputes per node)
    float dlen;         // r.direction().length() (flat_intersect.h:13)
};
PT_HD ray_in setup_ray(float ox, float oy, float oz, float dx, float dy, float dz) {
    ray_in r;
    r.ox = ox; r.oy = oy; r.oz = oz; r.dx = dx; r.dy = dy; r.dz = dz;
    r.ix = 1.0f / dx; r.iy = 1.0f / dy; r.iz = 1.0f / dz;
    r.dlen = sqrtf(dx * dx + dy * dy + dz * dz);
    return r;
}

// nextafterf(x, toward) for finite x != toward.
PT_HD float next_toward(float x, float toward) {
    if(!(x < toward) && !(x > toward)) return toward;
    uint32_t b = f2u(x);
    if(x == 0.0f) return x < toward ? u2f(1u) : u2f(0x80000001u);
    bool up = x < toward;
    bool positive = !(b >> 31);
    b = (up == positive) ? b + 1 : b - 1;
    return u2f(b);
}

// ---------------- slab test (src/renderer/transport.h:14-31) ----------------
// All three axes are evaluated with the reference's ternaries; low only grows
// and high only shrinks, so the reference's early exits cannot change the result.
PT_HD bool slab(float mn, float mx, float o, float d, float inv, float& low, float& high, bool& ok) {
    if(d == 0) { ok = ok && !(o < mn || o > mx); return ok; }
    float a = (mn - o) * inv;
    float b = (mx - o) * inv;
    if(inv < 0) { float s = a; a = b; b = s; }
    low = a > low ? a : low;
    high = b < high ? b : high;
    return ok;
}
PT_HD bool box_test(const float4& A, const float4& B, const ray_in& r, float low, float high) {
    bool ok = true;
    slab(A.x, A.w, r.ox, r.dx, r.ix, low, high, ok);
    slab(A.y, B.x, r.oy, r.dy, r.iy, low, high, ok);
    slab(A.z, B.y, r.oz, r.dz, r.iz, low, high, ok);
    return ok && high > low;
}

// ---------------- Moller-Trumbore (src/flat/flat_intersect.h:8-27) ----------------
// A = {v0, 1e-7f*|n|}, B = {e1, |n|}, C = {e2, -}, with e1 = v1-v0, e2 = v2-v0, n = e1 x e2.
PT_HD bool tri_test(const float4& A, const float4& B, const float4& C, const ray_in& r, float t_min, float t_max,
                    float& t, float& u, float& v) {
    if(!(B.w > 0) || !(r.dlen > 0)) return false;
    v3 d = mk(r.dx, r.dy, r.dz), e1 = mk(B.x, B.y, B.z), e2 = mk(C.x, C.y, C.z);
    v3 pvec = cross(d, e2);
    float det = dot(e1, pvec);
    if(fabsf(det) <= A.w * r.dlen) return false;
    float inv_det = 1.0f / det;
    v3 tvec = sub(mk(r.ox, r.oy, r.oz), mk(A.x, A.y, A.z));
    u = dot(tvec, pvec) * inv_det;
    if(!(u >= 0 && u <= 1)) return false;
    v3 qvec = cross(tvec, e1);
    v = dot(d, qvec) * inv_det;
    if(!(v >= 0 && u + v <= 1)) return false;
    t = dot(e2, qvec) * inv_det;
    return t > t_min && t < t_max;
}

// ---------------- sphere (src/flat/flat_intersect.h:49-58) ----------------
PT_HD bool sphere_test(const float4& S, const ray_in& r, float t_min, float t_max, float& t) {
    v3 d = mk(r.dx, r.dy, r.dz);
    v3 oc = sub(mk(r.ox, r.oy, r.oz), mk(S.x, S.y, S.z));
    float a = dot(d, d);
    float b = dot(oc, d);
    float c = dot(oc, oc) - S.w * S.w;
    float disc = b * b - a * c;
    if(!(disc > 0)) return false;
    float root = sqrtf(disc);
    t = (-b - root) / a;
    if(!(t < t_max && t > t_min)) t = (-b + root) / a;
    return t < t_max && t > t_min;
}

struct best_hit { float t, u, v; int32_t code; };

// ---------------- closest / any hit (src/renderer/transport.h:43-85) ----------------
// Depth-first, pop then test, right pushed before left, pruning limit
// nextafterf(closest, high), later leaf wins exact ties: the reference's
// order, bit for bit.  `root` is a reference into the binary layout
// (scene_dev::bnodes); STACK provides put(k, ref) and get(k).
template<bool ANY_HIT, bool STATS, class STACK>
PT_HD bool traverse_bvh(const scene_dev& s, int32_t root, const ray_in& r, float low, float high, STACK& stk,
                        best_hit& best, int& boxes, int& prims, int& spheres, int& error) {
    if(root == empty_root) return false;
    bool found = false;
    float closest = high;
    int count = 0;
    stk.put(0, root); count = 1;
    const int cap = s.stack_capacity;
    while(count) {
        --count;
        int32_t ref = stk.get(count);
        float limit = found ? next_toward(closest, high) : high;
        if(ref >= 0) {
            if(STATS) boxes++;
            float4 A = s.bnodes[2 * ref], B = s.bnodes[2 * ref + 1];
            if(!box_test(A, B, r, low, limit)) continue;
            int32_t left = f2i(B.z), right = f2i(B.w);
            int needed = left == right ? 1 : 2;
            if(count + needed > cap) { error = 1; return false; }
            if(left != right) { stk.put(count, right); count++; }
            stk.put(count, left); count++;
        } else {
            if(STATS) prims++;
            int32_t code = ~ref;
            int32_t prim = code >> 1;
            float t, u = 0, v = 0;
            bool hit;
            if(STATS && (code & 1)) spheres++;
            if(code & 1) hit = sphere_test(s.sph_isect[prim], r, low, limit, t);
            else hit = tri_test(s.tri_isect[3 * prim], s.tri_isect[3 * prim + 1], s.tri_isect[3 * prim + 2], r, low, limit, t, u, v);
            if(hit && (!found || t <= closest)) {
                found = true; closest = t;
                best.t = t; best.u = u; best.v = v; best.code = code;
                if(ANY_HIT) return true;
            }
        }
    }
    return found;
}

// Brute mode (transport.h:48-60): every leaf in node-array order, no pruning.
template<bool ANY_HIT>
PT_HD bool traverse_brute(const scene_dev& s, const ray_in& r, float low, float high, best_hit& best, int& prims,
                          int& spheres) {
    if(s.root_ref == empty_root) return false;
    bool found = false;
    float closest = high;
    for(int32_t i = 0; i < s.node_count; i++) {
        const pt_flat_node& n = s.nodes[i];
        if(n.type == PT_BRANCH) continue;
        prims++;
        if(n.type == PT_SPHERE) spheres++;
        float t, u = 0, v = 0;
        bool hit;
        int32_t code = n.primitive * 2 + (n.type == PT_SPHERE ? 1 : 0);
        if(n.type == PT_SPHERE) hit = sphere_test(s.sph_isect[n.primitive], r, low, high, t);
        else hit = tri_test(s.tri_isect[3 * n.primitive], s.tri_isect[3 * n.primitive + 1], s.tri_isect[3 * n.primitive + 2], r, low, high, t, u, v);
        if(hit && (!found || t <= closest)) {
            found = true; closest = t;
            best.t = t; best.u = u; best.v = v; best.code = code;
            if(ANY_HIT) return true;
        }
    }
    return found;
}

// ---------------- 4-wide nearest-first traversal (shaded path) ----------------
// The four child boxes of a wide node (SoA: child k in component k), with
// trace_box's arithmetic (transport.h:14-31): per axis a = (near - o) * (1/d)
// and b = (far - o) * (1/d), where the near plane is the lower one for
// 1/d >= 0 and the upper one otherwise (the caller loads the planes in that
// order, so no selects are needed here); low = a > low ? a : low and
// high = b < high ? b : high in x, y, z order, the box hit when high > low;
// returns the entry distance low.  fmaxf/fminf give what those ternaries
// give, since both skip a NaN plane distance, and they let the compiler use
// 3-input min/max.  One form serves every ray, d == 0 included, where
// trace_box instead rejects the box if o < min or o > max and leaves low and
// high alone: 1/d is then +-inf, so with o outside the slab one plane
// distance is +inf as the near one or -inf as the far one (a miss), with o
// strictly inside they are -inf and +inf (no bound), and with o on a plane
// that distance is 0 * inf = NaN (skipped).  (min - o is never 0 for
// min != o: denormals are kept.)  The sign of a zero distance cannot
// matter: low starts at 0.001 and a zero high misses either way.
PT_HD void slab_entry4(const float4& NX, const float4& NY, const float4& NZ, const float4& FX, const float4& FY,
                       const float4& FZ, const ray_in& r, float low, float high, bool h[4], float e[4]) {
    const float nx[4] = {NX.x, NX.y, NX.z, NX.w}, ny[4] = {NY.x, NY.y, NY.z, NY.w}, nz[4] = {NZ.x, NZ.y, NZ.z, NZ.w};
    const float fx[4] = {FX.x, FX.y, FX.z, FX.w}, fy[4] = {FY.x, FY.y, FY.z, FY.w}, fz[4] = {FZ.x, FZ.y, FZ.z, FZ.w};
#pragma unroll
    for(int k = 0; k < 4; k++) {
        const float ax = (nx[k] - r.ox) * r.ix, bx = (fx[k] - r.ox) * r.ix;
        const float ay = (ny[k] - r.oy) * r.iy, by = (fy[k] - r.oy) * r.iy;
        const float az = (nz[k] - r.oz) * r.iz, bz = (fz[k] - r.oz) * r.iz;
        const float lo = fmaxf(fmaxf(fmaxf(low, ax), ay), az);
        const float hi = fminf(fminf(fminf(high, bx), by), bz);
        e[k] = lo;
        h[k] = hi > lo;
    }
}

// Order-sensitive cases.  The reference visits depth-first, the 4-wide
// traversal nearest-first.  Both find every hit at the final distance whose
// boxes all have rounded entry distances no larger than that distance, and
// both return the latest of those in depth-first order, so they agree unless
// some primitive hit at the final distance lies in front of the entry
// distance of one of its boxes, as rounding can make it for a triangle lying
// on a face of its box hit where it meets another at exactly the same
// distance.  Whether the reference finds such a hit depends on what it found
// before it reached the box.  A closest-hit query of the 4-wide traversal
// therefore stops and is traced again in the reference's order
// (exact_retrace) when
// - it hits a primitive in front of the entry distance of the box it reached
//   it through (the deepest of its boxes, since nested boxes have ordered
//   entry distances), or
// - it would prune a box whose entry distance lies less than
//   wide_window_ulps units in the last place above the pruning limit: boxes
//   are tested against a limit widened by that much, and the query stops
//   where one inside the band would be dropped.
// Any-hit (shadow) queries need no such care: whether some primitive is hit
// below the upper bound does not depend on the order.
constexpr uint32_t wide_window_ulps = 4;

// Closest (or any) hit over the 4-wide layout.  Children are visited nearest
// first; a stacked child is dropped on pop when its entry distance is not
// below the current limit, which is the reference's pop-time box test.
// Among hits at the same distance the leaf of larger depth-first rank wins,
// the reference's tie rule (transport.h:67-68,79).  STACK provides
// store(k, ref, entry) and load(k) -> int2 {ref, entry bits}.
//
// The query state lives in wide_query so that a traversal can be cut into
// steps (one wide node or one leaf cluster, then the pop): a lane whose query
// has finished takes a new one while the rest of its wave still traverses.
struct wide_query {
    ray_in r;
    // limit: the pruning limit, the query's upper bound until the first hit,
    // then nextafterf(closest, upper bound), which for 0 < closest < upper
    // bound is the next float above closest.  Boxes are tested against the
    // widened limit (widened()).  entry: the entry distance of the box of
    // the node or cluster being visited (-inf for a root without a box).
    // best.t is the closest hit so far, the upper bound before the first.
    // wlimit: widened(limit), kept beside it (it changes only with a hit).
    float limit, wlimit, entry;
    int32_t cur, count, best_rank;
    best_hit best;
    bool found, any;
};

// q.cur of a query that stopped at an order-sensitive case, to be traced
// again in the reference's order.  (Kept in an integer rather than a flag of
// its own: a per-lane flag carried through the divergent traversal loop
// costs scalar mask instructions at every join.)
constexpr int32_t wide_retrace = INT32_MIN + 2;
PT_HD bool needs_retrace(const wide_query& q) { return q.cur == wide_retrace; }

PT_HD void wide_begin(int32_t root, wide_query& q, bool any, float high) {
    q.limit = high; q.wlimit = high; q.entry = -INFINITY; q.cur = root; q.count = 0; q.best_rank = -1;
    q.best.t = high; q.best.code = -1;
    q.found = false; q.any = any;
}

// The widened limit: wide_window_ulps above the pruning limit once a closest
// hit is known (an any-hit query ends at its first hit), at most +inf.
template<uint32_t WINDOW>
PT_HD float widened(const wide_query& q, float limit) {
    const uint32_t w = f2u(limit) + WINDOW;
    return q.found ? u2f(w < 0x7f800000u ? w : 0x7f800000u) : limit;
}

// Returns true when the query has finished (or stopped for a retrace).
// STATS counts wide nodes visited in `boxes` and leaf records tested in
// `prims` (spheres among them in `spheres`).
// WINDOW: the band in ulps; 0 switches the reference-order fallback off
// altogether (the CPU model runs it so to show what the fallback changes).
template<bool STATS, class STACK, uint32_t WINDOW = wide_window_ulps>
PT_HD bool wide_step(const scene_dev& s, wide_query& q, STACK& stk, int& boxes, int& prims, int& spheres) {
    const float low = 0.001f;
    const ray_in& r = q.r;
    float limit = q.limit;
    if(q.cur >= 0) {
        // The node's record at byte offset q.cur * 128 (32 bits: at most
        // wide_node_limit nodes): lo.x lo.y lo.z hi.x hi.y hi.z ref, 16 bytes
        // each.  Per axis the near plane is lo, or hi (48 bytes on) where
        // 1/d < 0, and the far plane the other one.
        const char *nb = reinterpret_cast<const char *>(s.wnodes);
        const uint32_t at = uint32_t(q.cur) << 7;
        const uint32_t sx = r.ix < 0 ? 48u : 0u, sy = r.iy < 0 ? 48u : 0u, sz = r.iz < 0 ? 48u : 0u;
        auto plane = [nb](uint32_t offset) { return *reinterpret_cast<const float4 *>(nb + offset); };
        const float4 NX = plane(at + sx), NY = plane(at + 16u + sy), NZ = plane(at + 32u + sz);
        const float4 FX = plane(at + (sx ^ 48u)), FY = plane(at + 16u + (sy ^ 48u)), FZ = plane(at + 32u + (sz ^ 48u));
        const float4 RF = plane(at + 96u);
        const int32_t r0 = f2i(RF.x), r1 = f2i(RF.y), r2 = f2i(RF.z), r3 = f2i(RF.w);
        float e[4];
        int32_t c[4];
        bool hh[4];
        slab_entry4(NX, NY, NZ, FX, FY, FZ, r, low, q.wlimit, hh, e);
        const bool h0 = (r0 != wide_empty) & hh[0], h1 = (r1 != wide_empty) & hh[1];
        const bool h2 = (r2 != wide_empty) & hh[2], h3 = (r3 != wide_empty) & hh[3];
        if(STATS) boxes++;  // wide nodes visited (each a 128-byte record)
        if(!h0) e[0] = INFINITY;
        if(!h1) e[1] = INFINITY;
        if(!h2) e[2] = INFINITY;
        if(!h3) e[3] = INFINITY;
        c[0] = r0; c[1] = r1; c[2] = r2; c[3] = r3;
        const int nh = int(h0) + int(h1) + int(h2) + int(h3);
        if(nh) {
            // nearest first (any-hit queries too: their order does not change the
            // answer); swaps as selects, so the network has no branches
#define PT_CSWAP(i, j) { const bool sw = e[j] < e[i]; const float ei = sw ? e[j] : e[i], ej = sw ? e[i] : e[j]; \
                               const int32_t ci = sw ? c[j] : c[i], cj = sw ? c[i] : c[j]; \
                               e[i] = ei; e[j] = ej; c[i] = ci; c[j] = cj; }
            PT_CSWAP(0, 1) PT_CSWAP(2, 3) PT_CSWAP(0, 2) PT_CSWAP(1, 3) PT_CSWAP(1, 2)
#undef PT_CSWAP
            // The nearest child passed only the widened test: an order-sensitive
            // prune, and the query stops for a retrace.
            const bool band = !(e[0] < limit);
            // No overflow test: at a node of depth d (the root at 1) the stack
            // holds at most the 3 unvisited siblings of each of its d - 1
            // ancestors, so after these pushes count <= 3d, within the
            // capacity 3 * (deepest node's depth) + 1 the stack is sized for.
            // c[nh-1] .. c[1] at count .. count + nh - 2 (c[k] at count + nh - 1 - k)
            if(nh > 1) stk.store(q.count + nh - 2, c[1], e[1]);
            if(nh > 2) stk.store(q.count + nh - 3, c[2], e[2]);
            if(nh > 3) stk.store(q.count, c[3], e[3]);
            q.count += nh - 1;
            q.cur = band ? wide_retrace : c[0];
            q.entry = e[0];
            return band;
        }
    } else {
        const int32_t v = ~q.cur;
        const int32_t first = v >> 3, n = (v & 7) + 1;
        // A cluster holds at most two leaves: both records are loaded before
        // the first test, so that their latencies overlap.
        static_assert(wide_leaf_max <= 2, "the leaf step loads at most two leaves");
        const float4 *L = s.leaves + 3 * size_t(first);
        float4 rec[2][3];
        rec[0][0] = L[0]; rec[0][1] = L[1]; rec[0][2] = L[2];
        if(n > 1) { rec[1][0] = L[3]; rec[1][1] = L[4]; rec[1][2] = L[5]; }
#pragma unroll
        for(int j = 0; j < 2; j++) {
            if(j >= n) break;
            const int32_t k = first + j;
            const float4 A = rec[j][0], B = rec[j][1], C = rec[j][2];
            float t, u = 0, vv = 0;
            bool hit;
            if(B.w < 0) { hit = sphere_test(A, r, low, limit, t); if(STATS) spheres++; }
            else hit = tri_test(A, B, C, r, low, limit, t, u, vv);
            if(STATS) prims++;
            // A hit in front of its own box's entry distance: the reference
            // may find it or not depending on the order (see above).
            const bool front = WINDOW && hit && !q.any && t < q.entry;
            if(front) q.cur = wide_retrace;
            hit = hit && !front;
            // A hit has t < limit, so t <= closest (before the first hit closest
            // is the upper bound and best_rank -1): the later leaf wins a tie.
            if(hit && (t < q.best.t || k > q.best_rank)) {
                q.found = true; q.best_rank = k;
                q.best.t = t; q.best.u = u; q.best.v = vv; q.best.code = f2i(C.w);
                if(q.any) return true;
                // next_toward(t, upper bound): a hit has 0.001 < t < limit
                limit = u2f(f2u(t) + 1u);
                q.limit = limit;
                q.wlimit = widened<WINDOW>(q, limit);
            }
        }
    }
    if(q.cur == wide_retrace) return true;
    for(;;) {
        if(q.count == 0) return true;
        --q.count;
        const int2 top = stk.load(q.count);
        // Entries were pushed below the widened limit of their time; before
        // the first hit the limit is the upper bound, so only entries pushed
        // before a hit, or inside the widened band, can fail this test.
        const float entry = i2f(top.y);
        if(entry < limit) {
            q.cur = top.x;
            q.entry = entry;
            return false;
        }
        if(entry < q.wlimit) {
            q.cur = wide_retrace;
            return true;
        }
    }
}

// Binary-stack view of a wide stack, for the reference-order traversal.
template<class W>
struct wide_as_binary {
    W& w;
    PT_HD void put(int k, int32_t ref) { w.store(k, ref, 0.0f); }
    PT_HD int32_t get(int k) const { return w.load(k).x; }
};

// A closest-hit query that stopped at an order-sensitive prune, traced again
// from the start in the reference's order over the binary layout (`broot`:
// the query's root there).  The wide stack serves as the binary stack; it is
// sized for both.
template<class STACK>
PT_HD bool exact_retrace(const scene_dev& s, int32_t broot, const ray_in& r, STACK& stk, best_hit& best, int& error) {
    int boxes = 0, prims = 0, spheres = 0;
    wide_as_binary<STACK> bs{stk};
    return traverse_bvh<false, false>(s, broot, r, 0.001f, FLT_MAX, bs, best, boxes, prims, spheres, error);
}
// The same, counting the binary traversal's branch and leaf pops.
template<class STACK>
PT_HD bool exact_retrace_counted(const scene_dev& s, int32_t broot, const ray_in& r, STACK& stk, best_hit& best,
                                 int& error, int& boxes, int& prims, int& spheres) {
    wide_as_binary<STACK> bs{stk};
    return traverse_bvh<false, true>(s, broot, r, 0.001f, FLT_MAX, bs, best, boxes, prims, spheres, error);
}

// A whole 4-wide query in one call (wavefront trace kernel, drain kernels),
// the reference-order retrace included.
template<bool ANY_HIT, bool STATS, class STACK, uint32_t WINDOW = wide_window_ulps>
PT_HD bool traverse_wide(const scene_dev& s, int32_t root, int32_t broot, const ray_in& r, float high, STACK& stk,
                         best_hit& best, int& boxes, int& prims, int& spheres, int& error) {
    wide_query q;
    q.r = r;
    wide_begin(root, q, ANY_HIT, high);
    while(!wide_step<STATS, STACK, WINDOW>(s, q, stk, boxes, prims, spheres)) {}
    if(needs_retrace(q)) return exact_retrace(s, broot, r, stk, best, error);
    if(q.found) best = q.best;
    return q.found;
}

} // namespace pt
