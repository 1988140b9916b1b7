// pt_kernels.hip -- CDNA4 kernels of the path-tracing core.
//
// Wavefront structure (one live path per pixel, so each pixel's PCG stream
// is consumed in the same order as the reference's sequential sample loop,
// src/cuda/cuda_renderer.cu:70-85):
//   start    : camera rays for every pixel, identity extension queue
//   trace    : persistent kernel; closest hit for the extension queue and
//              any-hit for the shadow queue (one launch), LDS traversal stacks
//   bookkeep : one thread; ray counters and queue resets
//   shade    : hit reconstruction, MIS/NEE, BSDF sampling, accumulation and
//              path regeneration; wave-ballot compaction into the next queues
// Primary-only diagnostic views use one pixel-parallel kernel.
//
// Arithmetic follows the reference's host evaluation order exactly and the
// library is compiled with -ffp-contract=off: see the citation on each helper.
#include <cfloat>
#include <cmath>
#include <stdexcept>

#include "pt_device.h"
#include "pt_traverse.h"
#include "pt_libm.h"

namespace pt {
namespace {

constexpr float PI_F = float(M_PI);

// ---------------- PCG-XSH-RR 64/32 (src/renderer/random.h:8-23) ----------------
struct rng_t {
    uint64_t state, increment;
    __device__ __forceinline__ uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + increment;
        uint32_t shifted = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t rotation = uint32_t(old >> 59u);
        return (shifted >> rotation) | (shifted << ((-rotation) & 31u));
    }
    __device__ __forceinline__ uint32_t next24() { return next() >> 8; }
    __device__ __forceinline__ float uniform() { return float(next() >> 8) * 0x1p-24f; }
};
__device__ __forceinline__ rng_t pixel_rng(uint64_t state, uint64_t pixel) { return rng_t{state, (pixel << 1u) | 1u}; }
__device__ __forceinline__ uint64_t seed_state(uint64_t value, uint64_t pixel) {
    rng_t r{0, (pixel << 1u) | 1u};
    r.next(); r.state += value; r.next();
    return r.state;
}
// random.h:25-33
__device__ v3 sample_ball(rng_t& r) {
    v3 p;
    do {
        float x = r.uniform(), y = r.uniform(), z = r.uniform();
        p = sub(scale(2.0f, mk(x, y, z)), mk(1, 1, 1));
    } while(squared_length(p) >= 1.0f);
    return p;
}
// random.h:35-42
__device__ v3 sample_disk(rng_t& r) {
    v3 p;
    do {
        float x = r.uniform(), y = r.uniform();
        p = sub(scale(2.0f, mk(x, y, 0)), mk(1, 1, 0));
    } while(squared_length(p) >= 1.0f);
    return p;
}

// Traversal stack of the wide path: the first wide_lds_entries entries of a
// lane live in LDS (entry k at lds[k*stride]), deeper ones in a global
// overflow region laid out so that a wave's accesses are coalesced.  STRIDE
// is the block size when the kernel fixes it (the LDS address is then a
// shift), 0 when it is known only at run time.
template<int STRIDE>
struct wide_stack_t {
    int2 *lds;
    int stride;
    int2 *spill;
    size_t spill_stride;
    uint32_t lane_id;   // grid-wide thread index (grids stay below 2^32 threads)
    __device__ __forceinline__ int lds_stride() const { return STRIDE ? STRIDE : stride; }
    __device__ __forceinline__ void store(int k, int32_t ref, float entry) {
        if(k < wide_lds_entries) {
            // one 64-bit LDS store (kept apart from the global one below, which
            // the compiler would otherwise merge into a generic-address store)
            *reinterpret_cast<unsigned long long *>(lds + k * lds_stride()) =
                (unsigned long long)uint32_t(ref) | ((unsigned long long)__float_as_uint(entry) << 32);
        } else {
            int2 *slot = spill + size_t(k - wide_lds_entries) * spill_stride + lane_id;
            slot->x = ref;
            slot->y = __float_as_int(entry);
        }
    }
    __device__ __forceinline__ int2 load(int k) const {
        if(k < wide_lds_entries) {
            const unsigned long long v = *reinterpret_cast<const unsigned long long *>(lds + k * lds_stride());
            return make_int2(int32_t(uint32_t(v)), int32_t(uint32_t(v >> 32)));
        }
        return spill[size_t(k - wide_lds_entries) * spill_stride + lane_id];
    }
};

using wide_stack = wide_stack_t<0>;

// Binary traversal stack in LDS: entry k of a lane at p[k*stride].
struct lds_stack {
    int32_t *p;
    int stride;
    __device__ __forceinline__ void put(int k, int32_t v) { p[k * stride] = v; }
    __device__ __forceinline__ int32_t get(int k) const { return p[k * stride]; }
};

template<bool ANY_HIT, bool STATS>
__device__ __forceinline__ bool trace(const scene_dev& s, bool brute, const ray_in& r, float low, float high,
                                      int32_t *stk, int stride, best_hit& best, int& boxes, int& prims, int& error) {
    int spheres = 0;
    if(brute) return traverse_brute<ANY_HIT>(s, r, low, high, best, prims, spheres);
    lds_stack st{stk, stride};
    return traverse_bvh<ANY_HIT, STATS>(s, s.root_ref, r, low, high, st, best, boxes, prims, spheres, error);
}
template<bool ANY_HIT>
__device__ __forceinline__ bool trace_counted(const scene_dev& s, bool brute, const ray_in& r, float low, float high,
                                              int32_t *stk, int stride, best_hit& best, int& boxes, int& prims,
                                              int& spheres, int& error) {
    if(brute) return traverse_brute<ANY_HIT>(s, r, low, high, best, prims, spheres);
    lds_stack st{stk, stride};
    return traverse_bvh<ANY_HIT, true>(s, s.root_ref, r, low, high, st, best, boxes, prims, spheres, error);
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
    for(int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// ---------------- hit record (flat_intersect.h:28-42,59-64; transport.h:39) ----------------
struct hit_rec { float t, u, v; v3 p, normal; int32_t material_id, primitive_type, primitive_id; };

// src/geometry/sphere_uv.h:7-13: {u, v}.  atan2/asin are evaluated in double
// and rounded, so u,v can differ from a float libm's result in the last bit.
// (Returned by value: out-parameters of an out-of-line call would keep the
// caller's hit record in private memory.)
__device__ __noinline__ float2 sphere_uv(v3 p) {
    float phi = float(atan2(double(p.z), double(p.x)));
    float theta = float(asin(double(p.y)));
    return make_float2(float(1 - (double(phi) + M_PI) / (2 * M_PI)), float((double(theta) + M_PI / 2) / M_PI));
}

__device__ hit_rec make_hit(const scene_dev& s, v3 o, v3 d, const best_hit& b) {
    hit_rec h;
    int32_t prim = b.code >> 1;
    h.t = b.t;
    h.p = add(o, scale(b.t, d));
    h.primitive_id = prim;
    if(b.code & 1) {
        const pt_flat_sphere& sp = s.spheres[prim];
        v3 c = mk(sp.center);
        h.normal = divs(sub(h.p, c), sp.radius);
        const float2 uv = sphere_uv(h.normal);
        h.u = uv.x; h.v = uv.y;
        h.material_id = sp.material_id;
        h.primitive_type = PT_SPHERE;
    } else {
        // Shading record: {unit outward normal, material} {n0, smooth}
        // {n1, uv0.x} {n2, uv0.y} {uv1.xy, uv2.xy}; the outward normal is
        // cross(e1,e2) /= |e1 x e2| precomputed with the same arithmetic.
        const float4 *sr = s.shade_tri + 5 * size_t(prim);
        const float4 G = sr[0], N0 = sr[1], UV = sr[4];
        const v3 outward = mk(G.x, G.y, G.z);
        v3 normal = outward;
        const float u = b.u, v = b.v;
        if(__float_as_int(N0.w)) {
            const float4 N1 = sr[2], N2 = sr[3];
            float w = 1 - u - v;
            v3 interp = add(add(scale(w, mk(N0.x, N0.y, N0.z)), scale(u, mk(N1.x, N1.y, N1.z))), scale(v, mk(N2.x, N2.y, N2.z)));
            float ls = squared_length(interp);
            if(ls > 0 && isfinite(ls)) {
                normal = unit_vector(interp);
                if(dot(normal, outward) < 0) normal = neg(normal);
                if(dot(normal, outward) == 0) normal = outward;
            }
        }
        h.normal = normal;
        const float w = 1 - u - v;
        const float uv0x = sr[2].w, uv0y = sr[3].w;
        h.u = w * uv0x + u * UV.x + v * UV.z;
        h.v = w * uv0y + u * UV.y + v * UV.w;
        h.material_id = __float_as_int(G.w);
        h.primitive_type = PT_TRIANGLE;
    }
    return h;
}

// ---------------- textures (src/renderer/transport.h:87-111) ----------------
__device__ v3 trace_image(const scene_dev& s, const pt_flat_texture& tex, float u, float v) {
    u = fminf(1.0f, fmaxf(0.0f, u));
    v = fminf(1.0f, fmaxf(0.0f, v));
    int i = int(u * float(tex.width));
    int j = int(double((1 - v) * float(tex.height)) - 0.001);
    if(i < 0) i = 0;
    if(j < 0) j = 0;
    if(i > tex.width - 1) i = tex.width - 1;
    if(j > tex.height - 1) j = tex.height - 1;
    size_t off = size_t(tex.byte_offset) + 3 * size_t(i) + 3 * size_t(tex.width) * size_t(j);
    float red = float(int(s.pixels[off]) / 255.0);
    float green = float(int(s.pixels[off + 1]) / 255.0);
    float blue = float(int(s.pixels[off + 2]) / 255.0);
    return mk(red, green, blue);
}
__device__ v3 trace_texture(const scene_dev& s, int32_t id, float u, float v) {
    const pt_flat_texture& tex = s.textures[id];
    if(tex.type == PT_CONSTANT) return mk(tex.color);
    if(tex.type == PT_IMAGE) return trace_image(s, tex, u, v);
    u = u * tex.scale.x + tex.offset.x;
    v = v * tex.scale.y + tex.offset.y;
    if(!tex.clamp) { u -= floorf(u); v -= floorf(v); }
    return mul(mk(tex.color), trace_image(s, s.textures[tex.image_id], u, v));
}

// ---------------- light sampling (src/renderer/light_sampling.h) ----------------
// light_sampling.h:10-16
__device__ __forceinline__ float mis_weight(float a, float b) {
    float sc = fmaxf(a, b);
    if(!(sc > 0)) return 0;
    a /= sc; b /= sc;
    return a * a / (a * a + b * b);
}
// cosf/sinf of angle = 2*float(pi)*(k*2^-24) (light_sampling.h:20,47) as
// the host libm computes them: double evaluation rounded to float, replaced
// by the stored host value for the k where the two differ.  The double
// evaluation: angle = j/128 + l with |l| <= 1/256 (exact in double), cos and
// sin of j/128 from trig_base (805 angles cover [0, 2*pi)), cos l and sin l
// by their Taylor polynomials to l^8 and l^9 (truncation below 2^-100), and
// the angle-addition formulas.
__device__ double2 trig_base[805];
__global__ void trig_base_kernel() {
    const int j = int(blockIdx.x * blockDim.x + threadIdx.x);
    if(j < 805) trig_base[j] = make_double2(cos(double(j) * 0x1p-7), sin(double(j) * 0x1p-7));
}
__device__ __noinline__ float2 device_cos_sin(uint32_t k) {
    const float angle = (2 * PI_F) * (float(k) * 0x1p-24f);
    const double a = double(angle);
    const int j = int(a * 128.0 + 0.5);
    const double l = a - double(j) * 0x1p-7;
    const double l2 = l * l;
    double c = fma(l2, 1.0 / 40320, -1.0 / 720);
    c = fma(c, l2, 1.0 / 24);
    c = fma(c, l2, -0.5);
    c = fma(c, l2, 1.0);
    double sn = fma(l2, 1.0 / 362880, -1.0 / 5040);
    sn = fma(sn, l2, 1.0 / 120);
    sn = fma(sn, l2, -1.0 / 6);
    sn = fma(sn, l2, 1.0);
    sn = sn * l;
    const double2 b = trig_base[j];
    return make_float2(float(fma(b.x, c, -(b.y * sn))), float(fma(b.y, c, b.x * sn)));
}
__device__ __forceinline__ float2 libm_cos_sin(const scene_dev& s, uint32_t k) {
    const unsigned long long word = s.trig_bits[k >> 6];
    const unsigned long long bit = 1ull << (k & 63u);
    if(word & bit) return s.trig_exceptions[s.trig_rank[k >> 6] + uint32_t(__popcll(word & (bit - 1ull)))];
    return device_cos_sin(k);
}

// light_sampling.h:18-25.
__device__ __forceinline__ v3 cosine_direction(const scene_dev& s, v3 normal, rng_t& r) {
    float u = r.uniform();
    uint32_t k = r.next24();
    float radius = sqrtf(u);
    float2 cs = libm_cos_sin(s, k);
    v3 helper = fabsf(normal.x) > 0.9f ? mk(0, 1, 0) : mk(1, 0, 0);
    v3 tangent = unit_vector(cross(helper, normal));
    v3 bitangent = cross(normal, tangent);
    return add(add(scale(radius * cs.x, tangent), scale(radius * cs.y, bitangent)), scale(sqrtf(1 - u), normal));
}
struct light_sample { v3 point, normal; float u, v; int32_t material; };
// The emitters a path samples: the scene's, or one view's range of a batch.
struct light_set { const light_rec *lights; int32_t count; float area; };
__device__ __forceinline__ light_set scene_lights(const scene_dev& s) { return light_set{s.lights, s.light_count, s.light_area}; }
// light_sampling.h:27-54
__device__ light_sample sample_light(const scene_dev& s, const light_set& ls, rng_t& r) {
    float choice = r.uniform() * ls.area;
    int low = 0, high = ls.count - 1;
    while(low < high) {
        int middle = low + (high - low) / 2;
        if(choice < ls.lights[middle].cumulative_area) high = middle;
        else low = middle + 1;
    }
    const light_rec light = ls.lights[low];
    float u = r.uniform();
    uint32_t kv = r.next24();
    float v = float(kv) * 0x1p-24f;
    light_sample res;
    if(light.type == PT_TRIANGLE) {
        const pt_flat_triangle& t = s.triangles[light.primitive];
        float root = sqrtf(u), a = 1 - root, b = root * (1 - v), c = root * v;
        res.point = add(add(scale(a, mk(t.v0)), scale(b, mk(t.v1))), scale(c, mk(t.v2)));
        res.normal = unit_vector(cross(sub(mk(t.v1), mk(t.v0)), sub(mk(t.v2), mk(t.v0))));
        v3 uv = add(add(scale(a, mk(t.uv0)), scale(b, mk(t.uv1))), scale(c, mk(t.uv2)));
        res.u = uv.x; res.v = uv.y; res.material = t.material_id;
    } else {
        const pt_flat_sphere& sp = s.spheres[light.primitive];
        float z = 1 - 2 * u, radius = sqrtf(fmaxf(0, 1 - z * z));
        float2 cs = libm_cos_sin(s, kv);
        res.normal = mk(radius * cs.x, z, radius * cs.y);
        res.point = add(mk(sp.center), scale(fabsf(sp.radius), res.normal));
        if(sp.radius < 0) res.normal = neg(res.normal);
        const float2 uv = sphere_uv(res.normal);
        res.u = uv.x; res.v = uv.y;
        res.material = sp.material_id;
    }
    return res;
}
// light_sampling.h:56-64
__device__ __forceinline__ float light_direction_pdf(float area, v3 from, v3 point, v3 normal) {
    v3 delta = sub(point, from);
    float d2 = squared_length(delta);
    if(!(area > 0) || !(d2 > 0)) return 0;
    float cosine = fabsf(dot(normal, delta)) / sqrtf(d2);
    if(!(cosine > 0)) return 0;
    float pdf = d2 / (cosine * area);
    return isfinite(pdf) ? pdf : 0;
}
// light_sampling.h:66-73
__device__ __forceinline__ float hit_light_pdf(const scene_dev& s, float light_area, v3 from, const hit_rec& h) {
    v3 normal = h.normal;
    if(h.primitive_type == PT_TRIANGLE) {
        const pt_flat_triangle& t = s.triangles[h.primitive_id];
        normal = unit_vector(cross(sub(mk(t.v1), mk(t.v0)), sub(mk(t.v2), mk(t.v0))));
    }
    return light_direction_pdf(light_area, from, h.p, normal);
}

// ---------------- BSDF sampling (src/renderer/transport.h:113-155) ----------------
__device__ __forceinline__ v3 reflect(v3 v, v3 n) { return sub(v, scale(2 * dot(v, n), n)); }

// r0 + (1-r0)*pow(1-cosine, 5) in double, rounded to float, with pow the host
// libm's (transport.h:147; the host build promotes pow(float,int) to double);
// out of line so its registers do not count against every shading path.
__device__ __noinline__ float schlick_double(const int32_t *keys, const double *values, int32_t count, float r0,
                                             float cosine) {
    const pow5_table table{keys, values, count};
    return float(double(r0) + double(1 - r0) * libm_pow5(table, 1 - cosine));
}

__device__ bool scatter(const scene_dev& s, const pt_flat_material& mat, v3 in_d, const hit_rec& rec, rng_t& r,
                        v3& attenuation, v3& out_d) {
    if(mat.type == PT_DIFFUSE) {
        v3 target = add(add(rec.p, rec.normal), sample_ball(r));
        out_d = sub(target, rec.p);
        attenuation = trace_texture(s, mat.texture_id, rec.u, rec.v);
        return true;
    }
    if(mat.type == PT_METAL) {
        v3 reflected = reflect(unit_vector(in_d), rec.normal);
        out_d = add(reflected, scale(mat.fuzz, sample_ball(r)));
        attenuation = mk(mat.albedo);
        return dot(out_d, rec.normal) > 0;
    }
    if(mat.type == PT_GLASS) {
        v3 outward, refracted = mk(0, 0, 0);
        v3 reflected = reflect(in_d, rec.normal);
        float ratio, cosine;
        if(dot(in_d, rec.normal) > 0) {
            outward = neg(rec.normal); ratio = mat.ref_idx;
            cosine = mat.ref_idx * dot(in_d, rec.normal) / length(in_d);
        } else {
            outward = rec.normal; ratio = 1.0f / mat.ref_idx;
            cosine = -dot(in_d, rec.normal) / length(in_d);
        }
        v3 uv = unit_vector(in_d);
        float dt = dot(uv, outward);
        float discriminant = 1.0f - ratio * ratio * (1 - dt * dt);
        float probability = 1;
        if(discriminant > 0) {
            refracted = sub(scale(ratio, sub(uv, scale(dt, outward))), scale(sqrtf(discriminant), outward));
            float r0 = (1 - mat.ref_idx) / (1 + mat.ref_idx);
            r0 *= r0;
            // The host build promotes pow(float,int) to double (transport.h:147).
            probability = schlick_double(s.pow5_keys, s.pow5_values, s.pow5_count, r0, cosine);
        }
        out_d = r.uniform() < probability ? reflected : refracted;
        attenuation = mk(1, 1, 1);
        return true;
    }
    return false;
}

// ---------------- camera (transport.h:157-162, 232-237) ----------------
__device__ __forceinline__ void camera_ray(const pt_flat_camera& c, int x, int y, int width, int height, rng_t& r,
                                           v3& o, v3& d, float& time) {
    float u = (float(x) + r.uniform()) / float(width);
    float v = (float(y) + r.uniform()) / float(height);
    v3 disk = scale(c.lens_radius, sample_disk(r));
    v3 offset = add(scale(disk.x, mk(c.u)), scale(disk.y, mk(c.v)));
    time = c.time0 + r.uniform() * (c.time1 - c.time0);
    o = add(mk(c.origin), offset);
    d = sub(sub(add(add(mk(c.lower_left_corner), scale(u, mk(c.horizontal))), scale(v, mk(c.vertical))), mk(c.origin)), offset);
}

// transport.h:164-167
__device__ __forceinline__ v3 id_color(int32_t id) {
    uint32_t bits = (uint32_t(id) + 1u) * 2654435761u;
    return mk(float(bits & 255) / 255.0f, float((bits >> 8) & 255) / 255.0f, float((bits >> 16) & 255) / 255.0f);
}

__device__ __forceinline__ float cost_log2(const scene_dev& s, int count) {
    // log2f(1.0f + float(count)) of the host libm, from the per-scene table.
    if(count >= 0 && count < s.log2_count) return s.log2_table[count];
    return float(log2(double(1.0f + float(count))));
}

// Block-aggregated append to up to three queues: one atomic per queue per
// block (a single counter shared by the whole grid is the contended
// resource), order kept within the block.  Every thread of the block must
// call it; it synchronises the block.
constexpr int max_block_waves = 16;
constexpr int trace_chunk = 64;  // items per work claim of a trace wave (larger claims starve small launches)
__device__ __forceinline__ void block_append3(bool p0, bool p1, bool p2, int32_t value,
                                              int32_t *q0, uint32_t *c0, int32_t *q1, uint32_t *c1,
                                              int32_t *q2, uint32_t *c2) {
    __shared__ uint32_t counts[3][max_block_waves];
    __shared__ uint32_t bases[3][max_block_waves];
    const int lane = __lane_id(), wave = int(threadIdx.x) >> 6, waves = int(blockDim.x) >> 6;
    const uint64_t m0 = __ballot(p0), m1 = __ballot(p1), m2 = __ballot(p2);
    if(lane == 0) {
        counts[0][wave] = uint32_t(__popcll(m0));
        counts[1][wave] = uint32_t(__popcll(m1));
        counts[2][wave] = uint32_t(__popcll(m2));
    }
    __syncthreads();
    if(threadIdx.x < 3 && (threadIdx.x != 2 || q2)) {
        const int q = int(threadIdx.x);
        uint32_t total = 0;
        for(int w = 0; w < waves; w++) total += counts[q][w];
        uint32_t *counter = q == 0 ? c0 : q == 1 ? c1 : c2;
        uint32_t running = total ? atomicAdd(counter, total) : 0;
        for(int w = 0; w < waves; w++) { bases[q][w] = running; running += counts[q][w]; }
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    if(p0) q0[bases[0][wave] + uint32_t(__popcll(m0 & below))] = value;
    if(p1) q1[bases[1][wave] + uint32_t(__popcll(m1 & below))] = value;
    if(p2) q2[bases[2][wave] + uint32_t(__popcll(m2 & below))] = value;
    __syncthreads();
}

// Wave-aggregated append: one atomic per wave, order kept within the wave.
__device__ __forceinline__ void append(bool pred, int32_t value, int32_t *queue, uint32_t *counter) {
    uint64_t mask = __ballot(pred);
    if(mask == 0) return;
    int lane = __lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if(lane == leader) base = atomicAdd(counter, uint32_t(__popcll(mask)));
    base = __shfl(base, leader);
    if(pred) queue[base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)))] = value;
}

// ======================================================================
// Kernels
// ======================================================================

// cuda_renderer.cu:64-68 (initialize_pixels) plus the failure flag reset.
__global__ void init_pixels_kernel(float *sum, uint64_t *rng, int64_t count, uint64_t seed, loop_ctl *ctl) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if(i == 0) { ctl->failure = 0; ctl->rays = 0; }
    if(i >= count) return;
    sum[3 * i] = 0; sum[3 * i + 1] = 0; sum[3 * i + 2] = 0;
    rng[i] = seed_state(seed, uint64_t(i));
}

// The trace_result of a closest-hit query (cuda_renderer.cu:116-119): the hit
// record, or the empty record with hit = 0.
__device__ pt_trace_result trace_result(const scene_dev& s, const ray_in& r, bool hit, const best_hit& b, int error) {
    pt_trace_result res;
    res.rec.t = 0; res.rec.u = 0; res.rec.v = 0;
    res.rec.p = pt_vec3{0, 0, 0}; res.rec.normal = pt_vec3{0, 0, 0};
    res.rec.material_id = -1; res.rec.primitive_type = -1; res.rec.primitive_id = -1;
    if(hit && !error) {
        hit_rec h = make_hit(s, mk(r.ox, r.oy, r.oz), mk(r.dx, r.dy, r.dz), b);
        res.rec.t = h.t; res.rec.u = h.u; res.rec.v = h.v;
        res.rec.p = pt_vec3{h.p.x, h.p.y, h.p.z}; res.rec.normal = pt_vec3{h.normal.x, h.normal.y, h.normal.z};
        res.rec.material_id = h.material_id; res.rec.primitive_type = h.primitive_type; res.rec.primitive_id = h.primitive_id;
    }
    res.hit = (hit && !error) ? 1 : 0;
    res.error = error;
    return res;
}

// cuda_renderer.cu:111-120 (query_rays), with optional trace_stats output.
__global__ __launch_bounds__(256) void query_kernel(scene_dev s, const pt_ray *rays, pt_trace_result *out,
                                                    int32_t *boxes_out, int32_t *prims_out, int32_t count, int brute) {
    extern __shared__ int32_t lds[];
    int32_t i = int32_t(blockIdx.x) * int32_t(blockDim.x) + int32_t(threadIdx.x);
    if(i >= count) return;
    pt_ray ray = rays[i];
    ray_in r = setup_ray(ray.origin.x, ray.origin.y, ray.origin.z, ray.direction.x, ray.direction.y, ray.direction.z);
    best_hit b; int boxes = 0, prims = 0, error = 0;
    bool hit = trace<false, true>(s, brute != 0, r, 0.001f, FLT_MAX, lds + threadIdx.x, blockDim.x, b, boxes, prims, error);
    out[i] = trace_result(s, r, hit, b, error);
    if(boxes_out) boxes_out[i] = boxes;
    if(prims_out) prims_out[i] = prims;
}

// The same closest-hit query through the shaded path's traversal: 4-wide,
// nearest first, with its reference-order retrace (verification entry
// pt_query_rays_wide; its results equal query_kernel's).  retraced_out[i] is
// 1 where the retrace ran.  `spill`: wide stack entries beyond LDS.
__global__ __launch_bounds__(256) void wide_query_kernel(scene_dev s, const pt_ray *rays, pt_trace_result *out,
                                                         int32_t *retraced_out, int2 *spill, int32_t count) {
    extern __shared__ int32_t lds[];
    const int32_t i = int32_t(blockIdx.x) * int32_t(blockDim.x) + int32_t(threadIdx.x);
    if(i >= count) return;
    wide_stack wstk{reinterpret_cast<int2 *>(lds) + threadIdx.x, int(blockDim.x), spill,
                    size_t(gridDim.x) * blockDim.x, uint32_t(i)};
    const pt_ray ray = rays[i];
    wide_query q;
    q.r = setup_ray(ray.origin.x, ray.origin.y, ray.origin.z, ray.direction.x, ray.direction.y, ray.direction.z);
    wide_begin(s.wroot_ref, q, false, FLT_MAX);
    int boxes = 0, prims = 0, spheres = 0, error = 0;
    while(!wide_step<false>(s, q, wstk, boxes, prims, spheres)) {}
    if(needs_retrace(q)) q.found = exact_retrace(s, s.root_ref, q.r, wstk, q.best, error);
    out[i] = trace_result(s, q.r, q.found, q.best, error);
    retraced_out[i] = needs_retrace(q) ? 1 : 0;
}

// Primary-only diagnostic views (transport.h:179-187): pixel-parallel, each
// thread runs its pixel's samples in order.
__global__ void debug_kernel(scene_dev s, frame_params f, float *sum, uint64_t *rngs, int32_t samples, int32_t *failure) {
    extern __shared__ int32_t lds[];
    int tid = int(threadIdx.y * blockDim.x + threadIdx.x);
    int stride = int(blockDim.x * blockDim.y);
    int x = int(blockIdx.x * blockDim.x + threadIdx.x);
    int y = int(blockIdx.y * blockDim.y + threadIdx.y);
    if(x >= f.width || y >= f.height) return;
    size_t i = size_t(y) * size_t(f.width) + size_t(x);
    rng_t r = pixel_rng(rngs[i], i);
    v3 acc = mk(sum[3 * i], sum[3 * i + 1], sum[3 * i + 2]);
    int error = 0;
    for(int k = 0; k < samples; k++) {
        v3 o, d; float time;
        camera_ray(f.camera, x, y, f.width, f.height, r, o, d, time);
        ray_in ray = setup_ray(o.x, o.y, o.z, d.x, d.y, d.z);
        best_hit b; int boxes = 0, prims = 0;
        bool hit = trace<false, true>(s, f.brute != 0, ray, 0.001f, FLT_MAX, lds + tid, stride, b, boxes, prims, error);
        v3 c;
        if(f.debug == PT_DEBUG_BVH || f.debug == PT_DEBUG_INTERSECTIONS) {
            float cost = cost_log2(s, f.debug == PT_DEBUG_BVH ? boxes : prims) / 10.0f;
            cost = fminf(1, cost);
            c = mk(cost, cost * cost, 1 - cost);
        } else if(!hit || error) {
            c = mk(0, 0, 0);
        } else {
            hit_rec h = make_hit(s, o, d, b);
            if(f.debug == PT_DEBUG_NORMALS) c = scale(0.5f, add(h.normal, mk(1, 1, 1)));
            else if(f.debug == PT_DEBUG_TRIANGLE) c = h.primitive_type == PT_TRIANGLE ? id_color(h.primitive_id) : mk(float(0.3), float(0.3), float(0.3));
            else c = id_color(h.material_id);
        }
        acc = add(acc, c);
    }
    if(error) atomicExch(failure, error);
    sum[3 * i] = acc.x; sum[3 * i + 1] = acc.y; sum[3 * i + 2] = acc.z;
    rngs[i] = r.state;
}

// Start of an add_samples call: the first camera ray of every pixel of one
// partition.  Pixels are dealt to `parts` partitions in chunks of
// `chunk` consecutive pixels: chunk c belongs to partition c % parts.
// ---------------- per-pixel view (single frame or batch) ----------------
// A single render reads its frame parameters and the scene's root and lights;
// a batch looks up the pixel's view.
template<bool BATCH>
__device__ __forceinline__ uint64_t local_index(const batch_dev& b, int32_t pix) {
    if constexpr(BATCH) return uint64_t(int64_t(pix) - b.views[b.pixel_view[pix]].pixel_offset);
    else return uint64_t(pix);
}
template<bool BATCH>
__device__ __forceinline__ int32_t query_root(const scene_dev& s, const batch_dev& b, int32_t pix) {
    if constexpr(BATCH) return b.views[b.pixel_view[pix]].wroot_ref;
    else return s.wroot_ref;
}
// The same scene's root in the binary layout (reference-order retraces).
template<bool BATCH>
__device__ __forceinline__ int32_t query_broot(const scene_dev& s, const batch_dev& b, int32_t pix) {
    if constexpr(BATCH) return b.views[b.pixel_view[pix]].broot_ref;
    else return s.root_ref;
}
template<bool BATCH>
__device__ __forceinline__ light_set pixel_lights(const scene_dev& s, const batch_dev& b, int32_t pix) {
    if constexpr(BATCH) {
        const view_rec& v = b.views[b.pixel_view[pix]];
        return light_set{s.lights + v.light_offset, v.light_count, v.light_area};
    } else {
        return scene_lights(s);
    }
}
template<bool BATCH>
__device__ __forceinline__ void pixel_camera_ray(const frame_params& f, const batch_dev& b, int32_t pix, rng_t& r, v3& o,
                                                 v3& d, float& time) {
    if constexpr(BATCH) {
        const view_rec& v = b.views[b.pixel_view[pix]];
        const int64_t local = int64_t(pix) - v.pixel_offset;
        camera_ray(v.camera, int(local % v.width), int(local / v.width), v.width, v.height, r, o, d, time);
    } else {
        camera_ray(f.camera, int(pix % f.width), int(pix / f.width), f.width, f.height, r, o, d, time);
    }
}

// Batch set-up: zero sums, PCG state of each pixel from its view's seed and
// its index within the view (cuda_renderer.cu:64-68 per view).
__global__ void batch_init_kernel(batch_dev b, float *sum, uint64_t *rng, int64_t count, loop_ctl *ctl) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if(i == 0) { ctl->failure = 0; ctl->rays = 0; }
    if(i >= count) return;
    const view_rec& v = b.views[b.pixel_view[i]];
    sum[3 * i] = 0; sum[3 * i + 1] = 0; sum[3 * i + 2] = 0;
    rng[i] = seed_state(v.seed, uint64_t(i - v.pixel_offset));
}

template<bool BATCH>
__global__ void start_kernel(scene_dev s, frame_params f, paths_dev p, loop_ctl *ctl, int32_t samples,
                             int part, int parts, int chunk, int64_t count, int64_t n, batch_dev bd) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if(j == 0) {
        ctl->n_ext[0] = uint32_t(count); ctl->n_ext[1] = 0; ctl->n_shadow = 0; ctl->n_finish = 0; ctl->work = 0;
        ctl->n_saved[0] = 0; ctl->n_saved[1] = 0; ctl->n_waiting = 0;
        ctl->iterations = 0; ctl->ext_rays = 0; ctl->shadow_rays = 0;
        ctl->box_tests = 0; ctl->triangle_tests = 0; ctl->sphere_tests = 0; ctl->retraces = 0;
        ctl->hits = 0; ctl->retrace_branches = 0; ctl->retrace_leaves = 0;
    }
    if(j >= count) return;
    const int64_t i = ((j / chunk) * parts + part) * chunk + j % chunk;
    if(i >= n) return;
    rng_t r = pixel_rng(p.rng[i], local_index<BATCH>(bd, int32_t(i)));
    v3 o, d; float time;
    pixel_camera_ray<BATCH>(f, bd, int32_t(i), r, o, d, time);
    p.ray_o[i] = make_float4(o.x, o.y, o.z, time);
    p.ray_d[i] = make_float4(d.x, d.y, d.z, 0);
    p.thr[i] = make_float4(1, 1, 1, 0);
    p.rad[i] = make_float4(0, 0, 0, 0);
    p.ctr[i] = make_int2(0, samples);
    p.rng[i] = r.state;
    p.sh_d[i].w = 0;
    p.queue[0][j] = int32_t(i);
}

__global__ void clear_ctl_kernel(loop_ctl *ctl) { ctl->failure = 0; ctl->rays = 0; }

template<bool COUNT>
__global__ __launch_bounds__(256) void trace_kernel(scene_dev s, frame_params f, paths_dev p, loop_ctl *ctl, int parity) {
    extern __shared__ int32_t lds[];
    const uint32_t n_ext = ctl->n_ext[parity];
    const uint32_t total = n_ext + ctl->n_shadow;
    const int32_t *q = p.queue[parity];
    const int lane = __lane_id();
    int32_t *stk = lds + threadIdx.x;
    const int stride = int(blockDim.x);
    const bool brute = f.brute != 0;
    const bool wide = !brute && s.wroot_ref != empty_root && !f.exact_traversal;
    wide_stack wstk{reinterpret_cast<int2 *>(lds) + threadIdx.x, stride, p.spill,
                    size_t(gridDim.x) * blockDim.x, blockIdx.x * blockDim.x + threadIdx.x};
    int error = 0;
    unsigned long long n_box = 0, n_leaf = 0, n_sph = 0;
    // Each wave claims trace_chunk consecutive items per atomic and works
    // through them 64 at a time.
    uint32_t chunk_base = 0, chunk_left = 0;
    for(;;) {
        if(chunk_left == 0) {
            uint32_t b = 0;
            if(lane == 0) b = atomicAdd(&ctl->work, uint32_t(trace_chunk));
            chunk_base = __shfl(b, 0);
            chunk_left = trace_chunk;
            if(chunk_base >= total) break;
        }
        const uint32_t base = chunk_base;
        chunk_base += 64; chunk_left -= 64;
        if(base >= total) { chunk_left = 0; continue; }
        uint32_t item = base + uint32_t(lane);
        int boxes = 0, prims = 0, spheres = 0;
        if(item < n_ext) {
            int32_t pix = q[item];
            float4 o = p.ray_o[pix], d = p.ray_d[pix];
            ray_in r = setup_ray(o.x, o.y, o.z, d.x, d.y, d.z);
            best_hit b;
            bool hit;
            if(wide) hit = traverse_wide<false, COUNT>(s, s.wroot_ref, s.root_ref, r, FLT_MAX, wstk, b, boxes, prims, spheres, error);
            else hit = COUNT ? trace_counted<false>(s, brute, r, 0.001f, FLT_MAX, stk, stride, b, boxes, prims, spheres, error)
                             : trace<false, false>(s, brute, r, 0.001f, FLT_MAX, stk, stride, b, boxes, prims, error);
            p.hit[pix] = hit ? make_int4(__float_as_int(b.t), __float_as_int(b.u), __float_as_int(b.v), b.code)
                             : make_int4(0, 0, 0, -1);
        } else if(item < total) {
            int32_t pix = p.shadow_queue[item - n_ext];
            float4 o = p.sh_o[pix], d = p.sh_d[pix];
            ray_in r = setup_ray(o.x, o.y, o.z, d.x, d.y, d.z);
            best_hit b;
            bool blocked;
            if(wide) blocked = traverse_wide<true, COUNT>(s, s.wroot_ref, s.root_ref, r, o.w, wstk, b, boxes, prims, spheres, error);
            else blocked = COUNT ? trace_counted<true>(s, brute, r, 0.001f, o.w, stk, stride, b, boxes, prims, spheres, error)
                                 : trace<true, false>(s, brute, r, 0.001f, o.w, stk, stride, b, boxes, prims, error);
            if(!blocked) {
                float4 c = p.sh_c[pix], ra = p.rad[pix];
                ra.x = ra.x + c.x; ra.y = ra.y + c.y; ra.z = ra.z + c.z;
                p.rad[pix] = ra;
            }
        }
        if(COUNT) { n_box += unsigned(boxes); n_leaf += unsigned(prims); n_sph += unsigned(spheres); }
    }
    if(error) atomicExch(&ctl->failure, 1);
    if(COUNT) {
        n_box = wave_sum(n_box); n_leaf = wave_sum(n_leaf); n_sph = wave_sum(n_sph);
        if(lane == 0) {
            atomicAdd(&ctl->box_tests, n_box);
            atomicAdd(&ctl->triangle_tests, n_leaf - n_sph);
            atomicAdd(&ctl->sphere_tests, n_sph);
        }
    }
}

// Trace stage on the wide layout with persistent lanes: whenever at least
// `refill` lanes of a wave are idle, the wave claims that many new queries
// (one atomic) and hands them to its idle lanes, so a wave does not wait for
// its slowest query before taking new work.
//
// Handing on.  Half of a launch used to pass after its queues ran dry,
// waiting for a few queries of hundreds of steps (rays grazing the surface)
// in waves with most lanes idle.  Once the queues are empty, a wave traces
// at most `hand_steps` more steps, then saves the queries still running
// (node, stack, closest hit so far) into saved[parity ^ 1] and ends; w of
// the query's direction is set to 1.  The next launch resumes them first,
// packed into full waves.  Meanwhile the shade stage
// queues the paths of handed-on queries again as ~pixel, unshaded, and the
// trace stage skips such entries; a path is shaded once both of its queries
// are done, so every result, and the order in which a path's radiance
// terms are added, is unchanged.
template<bool COUNT, bool BATCH>
__global__ __launch_bounds__(256) void trace_lanes_kernel(scene_dev s, frame_params f, paths_dev p, loop_ctl *ctl,
                                                          int parity, int refill, int hand_steps, batch_dev bd) {
    extern __shared__ int32_t lds[];
    const uint32_t n_resume = min(ctl->n_saved[parity], uint32_t(p.save_capacity));
    const uint32_t n_ext = ctl->n_ext[parity];
    const uint32_t ext_end = n_resume + n_ext;
    const uint32_t total = ext_end + ctl->n_shadow;
    const int32_t *queue = p.queue[parity];
    const saved_queries in = p.saved[parity], out = p.saved[parity ^ 1];
    const int lane = __lane_id();
    wide_stack wstk{reinterpret_cast<int2 *>(lds) + threadIdx.x, int(blockDim.x), p.spill,
                    size_t(gridDim.x) * blockDim.x, blockIdx.x * blockDim.x + threadIdx.x};
    int error = 0;
    unsigned long long n_box = 0, n_leaf = 0, n_sph = 0, n_hits = 0, n_rbox = 0, n_rleaf = 0;
    int boxes = 0, prims = 0, spheres = 0;
    wide_query q;
    bool active = false, exhausted = false, handed = false;
    int tail_left = hand_steps;  // wave-uniform
    int32_t pix = 0;
    for(;;) {
        const uint64_t idle = __ballot(!active);
        const int nidle = __popcll(idle);
        if(!exhausted && nidle >= refill) {
            uint32_t b = 0;
            if(lane == 0) b = atomicAdd(&ctl->work, uint32_t(nidle));
            b = __shfl(b, 0);
            if(b >= total) exhausted = true;
            else if(!active) {
                const uint32_t item = b + uint32_t(__popcll(idle & ((1ull << lane) - 1ull)));
                if(item < n_resume) {
                    const int4 h = in.head[item];
                    const bool any = h.x < 0;
                    pix = any ? ~h.x : h.x;
                    const float4 o = any ? p.sh_o[pix] : p.ray_o[pix], d = any ? p.sh_d[pix] : p.ray_d[pix];
                    q.r = setup_ray(o.x, o.y, o.z, d.x, d.y, d.z);
                    const float high = any ? o.w : FLT_MAX;
                    wide_begin(query_root<BATCH>(s, bd, pix), q, any, high);
                    // The saved stack ends with the node the query was at and its entry distance.
                    q.count = h.z - 1; q.best_rank = h.w; q.found = h.w >= 0;
                    if(q.found) {
                        const int4 bb = in.best[item];
                        q.best.t = __int_as_float(bb.x); q.best.u = __int_as_float(bb.y);
                        q.best.v = __int_as_float(bb.z); q.best.code = bb.w;
                        q.limit = next_toward(q.best.t, high);
                        q.wlimit = widened<wide_window_ulps>(q, q.limit);
                    }
                    const int2 *src = in.stack + size_t(item) * size_t(p.save_stride);
                    for(int k = 0; k < q.count; k++) { const int2 e = src[k]; wstk.store(k, e.x, __int_as_float(e.y)); }
                    q.cur = src[q.count].x; q.entry = __int_as_float(src[q.count].y);
                    active = true;
                } else if(item < ext_end) {
                    const int32_t v = queue[item - n_resume];
                    if(v >= 0) {  // ~pixel: a path waiting for a handed-on query, nothing to trace
                        pix = v;
                        const float4 o = p.ray_o[pix], d = p.ray_d[pix];
                        q.r = setup_ray(o.x, o.y, o.z, d.x, d.y, d.z);
                        wide_begin(query_root<BATCH>(s, bd, pix), q, false, FLT_MAX);
                        active = true;
                    }
                } else if(item < total) {
                    pix = p.shadow_queue[item - ext_end];
                    const float4 o = p.sh_o[pix], d = p.sh_d[pix];
                    q.r = setup_ray(o.x, o.y, o.z, d.x, d.y, d.z);
                    wide_begin(query_root<BATCH>(s, bd, pix), q, true, o.w);
                    active = true;
                }
            }
        }
        if(__ballot(active) == 0) {
            if(exhausted) break;
            continue;
        }
        if(exhausted && hand_steps > 0 && !handed) {
            if(tail_left > 0) {
                tail_left--;
            } else {
                handed = true;
                const uint64_t hand = __ballot(active);
                uint32_t base = 0;
                if(lane == 0) base = atomicAdd(&ctl->n_saved[parity ^ 1], uint32_t(__popcll(hand)));
                base = __shfl(base, 0);
                const uint32_t slot = base + uint32_t(__popcll(hand & ((1ull << lane) - 1ull)));
                if(active && slot < uint32_t(p.save_capacity)) {  // beyond the capacity a query runs on
                    out.head[slot] = make_int4(q.any ? ~pix : pix, q.cur, q.count + 1, q.found ? q.best_rank : -1);
                    out.best[slot] = make_int4(__float_as_int(q.best.t), __float_as_int(q.best.u),
                                               __float_as_int(q.best.v), q.best.code);
                    int2 *dst = out.stack + size_t(slot) * size_t(p.save_stride);
                    for(int k = 0; k < q.count; k++) dst[k] = wstk.load(k);
                    dst[q.count] = make_int2(q.cur, __float_as_int(q.entry));
                    if(q.any) p.sh_d[pix].w = 1.0f;
                    else p.ray_d[pix].w = 1.0f;
                    active = false;
                }
            }
        }
        if(active && wide_step<COUNT>(s, q, wstk, boxes, prims, spheres)) {
            active = false;
            if(needs_retrace(q)) {
                if(COUNT) {
                    int rb = 0, rp = 0, rs = 0;
                    q.found = exact_retrace_counted(s, query_broot<BATCH>(s, bd, pix), q.r, wstk, q.best, error, rb, rp, rs);
                    n_rbox += unsigned(rb); n_rleaf += unsigned(rp);
                } else {
                    q.found = exact_retrace(s, query_broot<BATCH>(s, bd, pix), q.r, wstk, q.best, error);
                }
                atomicAdd(&ctl->retraces, 1ull);
            }
            if(COUNT && !q.any && q.found) n_hits++;
            if(!q.any) {
                p.hit[pix] = q.found ? make_int4(__float_as_int(q.best.t), __float_as_int(q.best.u),
                                                 __float_as_int(q.best.v), q.best.code)
                                     : make_int4(0, 0, 0, -1);
                p.ray_d[pix].w = 0;
            } else {
                if(!q.found) {
                    float4 c = p.sh_c[pix], ra = p.rad[pix];
                    ra.x = ra.x + c.x; ra.y = ra.y + c.y; ra.z = ra.z + c.z;
                    p.rad[pix] = ra;
                }
                p.sh_d[pix].w = 0;
            }
            if(COUNT) { n_box += unsigned(boxes); n_leaf += unsigned(prims); n_sph += unsigned(spheres); boxes = prims = spheres = 0; }
        }
    }
    if(error) atomicExch(&ctl->failure, 1);
    if(COUNT) {
        n_box = wave_sum(n_box); n_leaf = wave_sum(n_leaf); n_sph = wave_sum(n_sph);
        n_hits = wave_sum(n_hits); n_rbox = wave_sum(n_rbox); n_rleaf = wave_sum(n_rleaf);
        if(lane == 0) {
            atomicAdd(&ctl->box_tests, n_box);
            atomicAdd(&ctl->triangle_tests, n_leaf - n_sph);
            atomicAdd(&ctl->sphere_tests, n_sph);
            atomicAdd(&ctl->hits, n_hits);
            atomicAdd(&ctl->retrace_branches, n_rbox);
            atomicAdd(&ctl->retrace_leaves, n_rleaf);
        }
    }
}

__global__ void bookkeep_kernel(loop_ctl *ctl, int parity) {
    uint32_t e = ctl->n_ext[parity], sh = ctl->n_shadow;
    // ~pixel entries were not traced; resumed queries were counted when they started.
    const uint32_t w = ctl->n_waiting;
    if(e && ctl->log && ctl->iterations < ctl->log_capacity) {
        ctl->log[2 * ctl->iterations] = e - w;
        ctl->log[2 * ctl->iterations + 1] = sh;
    }
    ctl->rays += (unsigned long long)(e - w) + sh;
    ctl->ext_rays += e - w; ctl->shadow_rays += sh;
    ctl->n_waiting = 0;
    ctl->n_saved[parity] = 0;
    ctl->n_shadow = 0;
    ctl->n_finish = 0;
    ctl->n_ext[parity ^ 1] = 0;
    ctl->work = 0;
    if(e) ctl->iterations++;
}

__global__ void add_rays_kernel(loop_ctl *ctl, unsigned long long rays) { ctl->rays += rays; }

// State of one path between bounces (the locals of trace_color,
// transport.h:171-173, plus the current ray and the samples still to take).
struct path_state {
    v3 o, d;
    float time;
    v3 throughput, radiance;
    float previous_pdf;
    int depth, left;
};
struct shadow_query { v3 o, d; float t_max; v3 contribution; };

__device__ __forceinline__ path_state load_path(const paths_dev& p, int32_t pix) {
    const float4 ro = p.ray_o[pix], rd = p.ray_d[pix], th = p.thr[pix], ra = p.rad[pix];
    const int2 ct = p.ctr[pix];
    path_state st;
    st.o = mk(ro.x, ro.y, ro.z); st.d = mk(rd.x, rd.y, rd.z); st.time = ro.w;
    st.throughput = mk(th.x, th.y, th.z); st.previous_pdf = th.w;
    st.radiance = mk(ra.x, ra.y, ra.z);
    st.depth = ct.x; st.left = ct.y;
    return st;
}
__device__ __forceinline__ void store_path(const paths_dev& p, int32_t pix, const path_state& st) {
    p.ray_o[pix] = make_float4(st.o.x, st.o.y, st.o.z, st.time);
    p.ray_d[pix] = make_float4(st.d.x, st.d.y, st.d.z, 0);
    p.thr[pix] = make_float4(st.throughput.x, st.throughput.y, st.throughput.z, st.previous_pdf);
    p.rad[pix] = make_float4(st.radiance.x, st.radiance.y, st.radiance.z, 0);
    p.ctr[pix] = make_int2(st.depth, st.left);
}

// First-hit feature buffers for the denoiser (no counterpart in the
// reference): the surface albedo and shading normal of the camera ray's hit,
// summed per pixel.  Reads no random numbers, so radiance is unchanged.
__device__ void record_aov(const scene_dev& s, const hit_rec& rec, const pt_flat_material& mat, float *aov) {
    v3 albedo;
    if(mat.type == PT_DIFFUSE) albedo = trace_texture(s, mat.texture_id, rec.u, rec.v);
    else if(mat.type == PT_METAL) albedo = mk(mat.albedo);
    else if(mat.type == PT_GLASS) albedo = mk(1, 1, 1);
    else {
        v3 e = trace_texture(s, mat.texture_id, rec.u, rec.v);
        albedo = mk(fminf(e.x, 1.0f), fminf(e.y, 1.0f), fminf(e.z, 1.0f));
    }
    aov[0] = aov[0] + albedo.x; aov[1] = aov[1] + albedo.y; aov[2] = aov[2] + albedo.z;
    aov[3] = aov[3] + rec.normal.x; aov[4] = aov[4] + rec.normal.y; aov[5] = aov[5] + rec.normal.z;
}

// One iteration of trace_color's loop after its closest-hit query
// (transport.h:184-227).  Returns true when the path continues with st.o/st.d
// as the next ray, false when the sample has ended with st.radiance.  A next
// event estimation query is returned in `sq` (has_shadow); its contribution
// must be added to st.radiance if unblocked, before the next bounce.
__device__ bool shade_bounce(const scene_dev& s, const frame_params& f, const light_set& ls, path_state& st, bool hit,
                             const best_hit& b, rng_t& r, shadow_query& sq, bool& has_shadow, float *aov) {
    has_shadow = false;
    if(!hit) return false;
    hit_rec rec = make_hit(s, st.o, st.d, b);
    const pt_flat_material mat = s.materials[rec.material_id];
    if(aov && st.depth == 0) record_aov(s, rec, mat, aov);
    if(mat.type == PT_LIGHT) {
        float weight = st.previous_pdf > 0 ? mis_weight(st.previous_pdf, hit_light_pdf(s, ls.area, st.o, rec)) : 1;
        st.radiance = add(st.radiance, mul(scale(weight, st.throughput), trace_texture(s, mat.texture_id, rec.u, rec.v)));
        return false;
    }
    if(st.depth == f.max_depth) return false;
    if(f.direct_light && mat.type == PT_DIFFUSE) {
        v3 attenuation = trace_texture(s, mat.texture_id, rec.u, rec.v);
        if(ls.count) {
            light_sample light = sample_light(s, ls, r);
            v3 delta = sub(light.point, rec.p);
            float distance = length(delta);
            if(distance > 0.002f && isfinite(distance)) {
                v3 direction = divs(delta, distance);
                float cosine = fmaxf(0, dot(rec.normal, direction));
                float pdf = light_direction_pdf(ls.area, rec.p, light.point, light.normal);
                if(cosine > 0 && pdf > 0) {
                    const pt_flat_material emitter = s.materials[light.material];
                    v3 emission = trace_texture(s, emitter.texture_id, light.u, light.v);
                    float bsdf_pdf = cosine / PI_F;
                    float w = bsdf_pdf * mis_weight(pdf, bsdf_pdf) / pdf;
                    sq.o = rec.p; sq.d = direction; sq.t_max = distance - 0.001f;
                    sq.contribution = scale(w, mul(mul(st.throughput, attenuation), emission));
                    has_shadow = true;
                }
            }
        }
        v3 direction = cosine_direction(s, rec.normal, r);
        st.previous_pdf = fmaxf(0, dot(rec.normal, direction)) / PI_F;
        st.throughput = mul(st.throughput, attenuation);
        st.o = rec.p; st.d = direction;
        st.depth++;
        return true;
    }
    st.previous_pdf = 0;
    v3 attenuation, out_d;
    if(!scatter(s, mat, st.d, rec, r, attenuation, out_d)) return false;
    st.throughput = mul(st.throughput, attenuation);
    st.o = rec.p; st.d = out_d;
    st.depth++;
    return true;
}

// One bounce for every path in the extension queue.
template<bool BATCH>
__global__ __launch_bounds__(1024) void shade_kernel(scene_dev s, frame_params f, paths_dev p, loop_ctl *ctl, int parity,
                                                     batch_dev bd) {
    const uint32_t n = ctl->n_ext[parity];
    const int32_t *q = p.queue[parity];
    int32_t *qn = p.queue[parity ^ 1];
    const uint32_t step = gridDim.x * blockDim.x;
    for(uint32_t base = blockIdx.x * blockDim.x; base < n; base += step) {
        uint32_t item = base + threadIdx.x;
        bool cont = false, shadow = false, done = false;
        int32_t pix = 0;
        // A path whose extension or shadow query was handed on waits, queued
        // again as ~pixel and not shaded, until both of its queries are done.
        bool wait = false;
        if(item < n) {
            const int32_t v = q[item];
            pix = v >= 0 ? v : ~v;
            wait = p.ray_d[pix].w != 0 || p.sh_d[pix].w != 0;
            if(wait) { cont = true; atomicAdd(&ctl->n_waiting, 1u); }
        }
        if(item < n && !wait) {
            const int4 hb = p.hit[pix];
            path_state st = load_path(p, pix);
            rng_t r = pixel_rng(p.rng[pix], local_index<BATCH>(bd, pix));
            const best_hit b{__int_as_float(hb.x), __int_as_float(hb.y), __int_as_float(hb.z), hb.w};
            shadow_query sq;
            cont = shade_bounce(s, f, pixel_lights<BATCH>(s, bd, pix), st, hb.w >= 0, b, r, sq, shadow,
                                p.aov ? p.aov + 6 * size_t(pix) : nullptr);
            if(shadow) {
                p.sh_o[pix] = make_float4(sq.o.x, sq.o.y, sq.o.z, sq.t_max);
                p.sh_d[pix] = make_float4(sq.d.x, sq.d.y, sq.d.z, 0);
                p.sh_c[pix] = make_float4(sq.contribution.x, sq.contribution.y, sq.contribution.z, 0);
            }
            p.rng[pix] = r.state;
            if(cont) store_path(p, pix, st);
            else p.rad[pix] = make_float4(st.radiance.x, st.radiance.y, st.radiance.z, 0);
            done = !cont;
        }
        block_append3(cont, shadow, done, wait ? ~pix : pix, qn, &ctl->n_ext[parity ^ 1], p.shadow_queue, &ctl->n_shadow,
                      p.finish_queue, &ctl->n_finish);
    }
}

// One query of a pixel of the drain kernels: 4-wide (root, with broot for a
// reference-order retrace) or the reference-order / brute traversal.
template<bool ANY_HIT>
__device__ __forceinline__ bool trace_query(const scene_dev& s, int32_t root, int32_t broot, bool brute, bool wide,
                                            const ray_in& r, float high, int32_t *stk, int stride, wide_stack& wstk,
                                            best_hit& b, int& error) {
    int boxes = 0, prims = 0, spheres = 0;
    if(wide) return traverse_wide<ANY_HIT, false>(s, root, broot, r, high, wstk, b, boxes, prims, spheres, error);
    if(brute) return traverse_brute<ANY_HIT>(s, r, 0.001f, high, b, prims, spheres);
    lds_stack st{stk, stride};
    return traverse_bvh<ANY_HIT, false>(s, broot, r, 0.001f, high, st, b, boxes, prims, spheres, error);
}

// Tail of an add_samples call, part 1: the shadow queries still pending
// from the last wavefront bounce.
template<bool BATCH>
__global__ __launch_bounds__(256) void drain_shadow_kernel(scene_dev s, frame_params f, paths_dev p, loop_ctl *ctl,
                                                           int parity, batch_dev bd) {
    extern __shared__ int32_t lds[];
    const uint32_t n = ctl->n_shadow;
    const bool brute = f.brute != 0;
    const bool wide = !brute && s.wroot_ref != empty_root && !f.exact_traversal;
    const int stride = int(blockDim.x);
    wide_stack wstk{reinterpret_cast<int2 *>(lds) + threadIdx.x, stride, p.spill,
                    size_t(gridDim.x) * blockDim.x, blockIdx.x * blockDim.x + threadIdx.x};
    int32_t *stk = lds + threadIdx.x;
    int error = 0;
    const uint32_t step = gridDim.x * blockDim.x;
    for(uint32_t item = blockIdx.x * blockDim.x + threadIdx.x; item < n; item += step) {
        const int32_t pix = p.shadow_queue[item];
        const float4 o = p.sh_o[pix], d = p.sh_d[pix];
        best_hit b;
        if(!trace_query<true>(s, query_root<BATCH>(s, bd, pix), query_broot<BATCH>(s, bd, pix), brute, wide, setup_ray(o.x, o.y, o.z, d.x, d.y, d.z), o.w,
                              stk, stride, wstk, b, error)) {
            float4 c = p.sh_c[pix], ra = p.rad[pix];
            ra.x = ra.x + c.x; ra.y = ra.y + c.y; ra.z = ra.z + c.z;
            p.rad[pix] = ra;
        }
    }
    // Queries the last trace launch handed on: shadow queries are traced here
    // from the start (the extension ones by drain_kernel, whose queue holds
    // their paths as ~pixel).  Not counted again.
    const uint32_t ns = min(ctl->n_saved[parity], uint32_t(p.save_capacity));
    for(uint32_t item = blockIdx.x * blockDim.x + threadIdx.x; item < ns; item += step) {
        const int32_t v = p.saved[parity].head[item].x;
        if(v >= 0) { p.ray_d[v].w = 0; continue; }
        const int32_t pix = ~v;
        const float4 o = p.sh_o[pix], d = p.sh_d[pix];
        best_hit b;
        if(!trace_query<true>(s, query_root<BATCH>(s, bd, pix), query_broot<BATCH>(s, bd, pix), brute, wide, setup_ray(o.x, o.y, o.z, d.x, d.y, d.z), o.w,
                              stk, stride, wstk, b, error)) {
            float4 c = p.sh_c[pix], ra = p.rad[pix];
            ra.x = ra.x + c.x; ra.y = ra.y + c.y; ra.z = ra.z + c.z;
            p.rad[pix] = ra;
        }
        p.sh_d[pix].w = 0;
    }
    if(error) atomicExch(&ctl->failure, 1);
    if(blockIdx.x == 0 && threadIdx.x == 0) {
        // drain_kernel traces the ~pixel entries' extension queries again.
        const uint32_t w = ctl->n_waiting;
        ctl->rays = ctl->rays + n - w; ctl->shadow_rays += n; ctl->ext_rays -= w;
        ctl->n_waiting = 0;
    }
}

// Tail of an add_samples call, part 2: each remaining pixel finishes all of
// its samples in one thread -- the megakernel order of the reference
// (cuda_renderer.cu:70-85) with the same per-bounce code, so every pixel
// consumes its PCG stream and accumulates exactly as in the wavefront loop.
// Used once few paths are live, where a bounce iteration costs as much as
// its slowest query whatever the queue length.
template<bool BATCH>
__global__ __launch_bounds__(256) void drain_kernel(scene_dev s, frame_params f, paths_dev p, loop_ctl *ctl, int parity,
                                                    batch_dev bd) {
    extern __shared__ int32_t lds[];
    const uint32_t n = ctl->n_ext[parity];
    const int32_t *q = p.queue[parity];
    const bool brute = f.brute != 0;
    const bool wide = !brute && s.wroot_ref != empty_root && !f.exact_traversal;
    const int stride = int(blockDim.x);
    wide_stack wstk{reinterpret_cast<int2 *>(lds) + threadIdx.x, stride, p.spill,
                    size_t(gridDim.x) * blockDim.x, blockIdx.x * blockDim.x + threadIdx.x};
    int32_t *stk = lds + threadIdx.x;
    int error = 0;
    unsigned long long ext = 0, shadow = 0;
    const uint32_t step = gridDim.x * blockDim.x;
    for(uint32_t item = blockIdx.x * blockDim.x + threadIdx.x; item < n; item += step) {
        const int32_t pix = q[item] >= 0 ? q[item] : ~q[item];
        path_state st = load_path(p, pix);
        p.ray_d[pix].w = 0;
        rng_t r = pixel_rng(p.rng[pix], local_index<BATCH>(bd, pix));
        const int32_t root = query_root<BATCH>(s, bd, pix), broot = query_broot<BATCH>(s, bd, pix);
        const light_set ls = pixel_lights<BATCH>(s, bd, pix);
        float *sum = p.sum + 3 * size_t(pix);
        for(;;) {
            best_hit b;
            const bool hit = trace_query<false>(s, root, broot, brute, wide, setup_ray(st.o.x, st.o.y, st.o.z, st.d.x, st.d.y, st.d.z),
                                                FLT_MAX, stk, stride, wstk, b, error);
            ext++;
            shadow_query sq;
            bool has_shadow;
            const bool cont = shade_bounce(s, f, ls, st, hit, b, r, sq, has_shadow, p.aov ? p.aov + 6 * size_t(pix) : nullptr);
            if(has_shadow) {
                shadow++;
                best_hit bb;
                if(!trace_query<true>(s, root, broot, brute, wide, setup_ray(sq.o.x, sq.o.y, sq.o.z, sq.d.x, sq.d.y, sq.d.z),
                                      sq.t_max, stk, stride, wstk, bb, error))
                    st.radiance = add(st.radiance, sq.contribution);
            }
            if(cont) continue;
            sum[0] = sum[0] + st.radiance.x; sum[1] = sum[1] + st.radiance.y; sum[2] = sum[2] + st.radiance.z;
            if(--st.left <= 0) break;
            pixel_camera_ray<BATCH>(f, bd, pix, r, st.o, st.d, st.time);
            st.throughput = mk(1, 1, 1); st.radiance = mk(0, 0, 0); st.previous_pdf = 0; st.depth = 0;
        }
        p.rng[pix] = r.state;
        p.ctr[pix] = make_int2(0, 0);
    }
    if(error) atomicExch(&ctl->failure, 1);
    ext = wave_sum(ext); shadow = wave_sum(shadow);
    if(__lane_id() == 0 && (ext || shadow)) {
        atomicAdd(&ctl->rays, ext + shadow);
        atomicAdd(&ctl->ext_rays, ext);
        atomicAdd(&ctl->shadow_rays, shadow);
    }
}

// Path lanes: every remaining pixel of an add_samples call on a lane of its
// own (the renderer's default for all pixels; the tail of the wavefront loop
// otherwise).  Each lane runs a small state machine -- trace its current
// query one step (wide node or leaf cluster) at a time, shade, next bounce or
// sample -- and takes the next pixel when its pixel is done.  A wave shades
// the lanes that are ready together, once at least `shade_batch` of them wait
// or no lane is traversing, so the shading code runs with many lanes active
// and no lane waits for the slowest traversal of its wave (as in
// drain_kernel, where each bounce costs the longest of 64 queries).  Path
// state stays in registers: no queues, no per-bounce launches.  Per pixel the
// work and its order are those of drain_kernel and of the wavefront loop:
// extension query, shade_bounce, shadow query (its contribution added if
// unblocked), then the next bounce or sample.
// 4 waves per SIMD (at most 128 VGPRs, a few spilled): 2150 against 1868
// Mrays/s at the compiler's own 140 VGPRs and 3 waves; 5 waves spill 106
// VGPRs and fall to 1613.  A tracing lane takes up to `trace_steps` steps
// before the wave looks again for lanes to shade and idle lanes to refill
// (1080p, 64 spp: 1, 4 and 8 steps give 2440, 2554 and 2563 Mrays/s), each
// step on the kind of work most tracing lanes are at (see below).
//
// Longest first (MODE).  A lane that takes an expensive pixel near the end of
// the queue keeps the GPU waiting for all of that pixel's samples, since a
// pixel's samples run one after the other.  A launch over every pixel from a
// sample boundary may therefore run in two passes: PATHS_FIRST runs the first
// `first_samples` samples of each pixel and writes the pixel's cost (the loop
// passes its lane spent on it) at its queue position; the queue is reordered
// by cost, most expensive first, in groups of consecutive positions (summed
// cost, each group keeping its pixels' order); PATHS_REST then starts each
// pixel at its next sample.  A pixel's samples keep their order and its PCG
// stream its continuity, so the results equal a single pass.  1080p, 64 spp,
// 1M triangles: 2170 Mrays/s in one pass; 4 first samples and groups of 1,
// 16 and 64 pixels: 2348, 2378 and 2329; groups of 256 or more fall below
// one pass (the ordering is then too coarse for the tail).
template<bool BATCH, int MODE>
__global__ __launch_bounds__(path_lane_block) __attribute__((amdgpu_waves_per_eu(4))) void path_lanes_kernel(scene_dev s, frame_params f, paths_dev p, loop_ctl *ctl,
                                                          int parity, const int32_t *queue, int shade_batch,
                                                          int trace_steps, int first_samples, uint32_t *costs,
                                                          batch_dev bd) {
    extern __shared__ int32_t lds[];
    const uint32_t n = ctl->n_ext[parity];
    const int lane = __lane_id();
    wide_stack_t<path_lane_block> wstk{reinterpret_cast<int2 *>(lds) + threadIdx.x, int(blockDim.x), p.spill,
                    size_t(gridDim.x) * blockDim.x, blockIdx.x * blockDim.x + threadIdx.x};
    // The running sum of the lane's pixel, [3][path_lane_block] after the
    // stack's LDS entries: read when the pixel is taken, written back when
    // it is left, so that a sample's accumulation does not go through HBM
    // (the additions and their order are those of sum += radiance).
    float *acc = reinterpret_cast<float *>(reinterpret_cast<int2 *>(lds) + wide_lds_entries * path_lane_block) +
                 threadIdx.x;
    int error = 0, boxes = 0, prims = 0, spheres = 0;
    uint32_t ext = 0, shadow = 0;   // queries this lane began, flushed to the totals before 2^31
    auto flush_rays = [&]() {
        atomicAdd(&ctl->rays, (unsigned long long)ext + shadow);
        atomicAdd(&ctl->ext_rays, (unsigned long long)ext);
        atomicAdd(&ctl->shadow_rays, (unsigned long long)shadow);
        ext = shadow = 0;
    };
    enum : int { IDLE = 0, TRACE = 1, SHADE = 2, RETRACE = 3 };
    int phase = IDLE;
    bool exhausted = false, cont_after = false;
    int32_t pix = 0, root = 0;
    path_state st;
    rng_t r{0, 1};
    v3 contribution = mk(0, 0, 0);
    wide_query q;
    // PATHS_FIRST: loop passes spent on the pixel, its queue position, and the
    // sample count at which the pass stops with it.
    uint32_t cost = 0, claim = 0;
    int stop_left = 0;
    auto begin_ext = [&]() {
        q.r = setup_ray(st.o.x, st.o.y, st.o.z, st.d.x, st.d.y, st.d.z);
        wide_begin(root, q, false, FLT_MAX);
        if(++ext >= (1u << 30)) flush_rays();
        phase = TRACE;
    };
    // After a bounce's queries: the next bounce, the next sample, or done.
    auto next_bounce = [&](bool cont) {
        if(cont) { begin_ext(); return; }
        acc[0] = acc[0] + st.radiance.x;
        acc[path_lane_block] = acc[path_lane_block] + st.radiance.y;
        acc[2 * path_lane_block] = acc[2 * path_lane_block] + st.radiance.z;
        if(--st.left <= 0 || (MODE == PATHS_FIRST && st.left <= stop_left)) {
            // Done, or (first pass) stopped at a sample boundary: the next
            // sample's camera ray is drawn when the pixel is taken again.
            float *sum = p.sum + 3 * size_t(pix);
            sum[0] = acc[0]; sum[1] = acc[path_lane_block]; sum[2] = acc[2 * path_lane_block];
            p.rng[pix] = r.state;
            if constexpr(MODE != PATHS_FRESH) p.ctr[pix] = make_int2(0, st.left > 0 ? st.left : 0);
            if constexpr(MODE == PATHS_FIRST) costs[claim] = st.left > 0 ? cost : 0u;
            phase = IDLE;
            return;
        }
        pixel_camera_ray<BATCH>(f, bd, pix, r, st.o, st.d, st.time);
        st.throughput = mk(1, 1, 1); st.radiance = mk(0, 0, 0); st.previous_pdf = 0; st.depth = 0;
        begin_ext();
    };
    for(;;) {
        const uint64_t idle = __ballot(phase == IDLE);
        if(!exhausted && idle) {
            uint32_t b = 0;
            if(lane == 0) b = atomicAdd(&ctl->work, uint32_t(__popcll(idle)));
            b = __shfl(b, 0);
            if(b >= n) exhausted = true;
            else if(phase == IDLE) {
                const uint32_t item = b + uint32_t(__popcll(idle & ((1ull << lane) - 1ull)));
                if(item < n) {
                    if constexpr(MODE == PATHS_FRESH) {
                        pix = int32_t(item);
                    } else {
                        const int32_t v = queue[item];
                        pix = v >= 0 ? v : ~v;
                    }
                    r = pixel_rng(p.rng[pix], local_index<BATCH>(bd, pix));
                    root = query_root<BATCH>(s, bd, pix);
                    {
                        const float *sum = p.sum + 3 * size_t(pix);
                        acc[0] = sum[0]; acc[path_lane_block] = sum[1]; acc[2 * path_lane_block] = sum[2];
                    }
                    if constexpr(MODE == PATHS_REST || MODE == PATHS_FRESH) {
                        st.left = MODE == PATHS_FRESH ? first_samples : p.ctr[pix].y;
                        if(st.left > 0) {
                            pixel_camera_ray<BATCH>(f, bd, pix, r, st.o, st.d, st.time);
                            st.throughput = mk(1, 1, 1); st.radiance = mk(0, 0, 0); st.previous_pdf = 0; st.depth = 0;
                            begin_ext();
                        }
                    } else {
                        st = load_path(p, pix);
                        p.ray_d[pix].w = 0;
                        if constexpr(MODE == PATHS_FIRST) { claim = item; cost = 0; stop_left = st.left - first_samples; }
                        begin_ext();
                    }
                }
            }
        }
        if constexpr(MODE == PATHS_FIRST) cost += phase != IDLE ? 1u : 0u;
        // Rare: closest-hit queries that stopped at an order-sensitive case
        // are traced again in the reference's order, here rather than in the
        // traversal steps below, so that the binary traversal's registers do
        // not add to theirs.  (Inline: a function call anywhere in this
        // kernel measured 30% slower, even where it never ran.)
        if(__ballot(phase == RETRACE)) {
            if(phase == RETRACE) {
                q.found = exact_retrace(s, query_broot<BATCH>(s, bd, pix), q.r, wstk, q.best, error);
                phase = SHADE;
                atomicAdd(&ctl->retraces, 1ull);
            }
        }
        const uint64_t tracing = __ballot(phase == TRACE), ready = __ballot(phase == SHADE);
        if((tracing | ready) == 0) {
            if(exhausted) break;
            continue;
        }
        if(ready && (__popcll(ready) >= shade_batch || tracing == 0)) {
            if(phase == SHADE) {
                shadow_query sq;
                bool has_shadow;
                const bool cont = shade_bounce(s, f, pixel_lights<BATCH>(s, bd, pix), st, q.found, q.best, r, sq, has_shadow,
                                               p.aov ? p.aov + 6 * size_t(pix) : nullptr);
                if(has_shadow) {
                    if(++shadow >= (1u << 30)) flush_rays();
                    contribution = sq.contribution;
                    cont_after = cont;
                    q.r = setup_ray(sq.o.x, sq.o.y, sq.o.z, sq.d.x, sq.d.y, sq.d.z);
                    wide_begin(root, q, true, sq.t_max);
                    phase = TRACE;
                } else {
                    next_bounce(cont);
                }
            }
        } else {
            // Up to trace_steps traversal steps per pass; a lane whose shadow
            // query ends and whose next extension query begins keeps going.
            // Each step takes the kind of work (wide node or leaf cluster) that
            // most tracing lanes of the wave are at; the others wait a step, so
            // a step issues one of the two code paths, not both.
            for(int k = 0; k < trace_steps; k++) {
                const uint64_t tr = __ballot(phase == TRACE);
                if(!tr) break;
                const uint64_t at_leaf = __ballot(phase == TRACE && q.cur < 0);
                const bool leaves = 2 * __popcll(at_leaf) >= __popcll(tr);
                if(phase != TRACE || (q.cur < 0) != leaves) continue;
                if(!wide_step<false>(s, q, wstk, boxes, prims, spheres)) continue;
                if(needs_retrace(q)) {
                    phase = RETRACE;
                } else if(!q.any) {
                    phase = SHADE;
                } else {
                    if(!q.found) st.radiance = add(st.radiance, contribution);
                    next_bounce(cont_after);
                }
            }
        }
    }
    if(error) atomicExch(&ctl->failure, 1);
    const unsigned long long ext_total = wave_sum((unsigned long long)ext), shadow_total = wave_sum((unsigned long long)shadow);
    if(lane == 0 && (ext_total || shadow_total)) {
        atomicAdd(&ctl->rays, ext_total + shadow_total);
        atomicAdd(&ctl->ext_rays, ext_total);
        atomicAdd(&ctl->shadow_rays, shadow_total);
    }
}

// End of a sample (cuda_renderer.cu:81): accumulated += radiance in sample
// order, then the pixel's next camera ray while samples remain.
template<bool BATCH>
__global__ __launch_bounds__(1024) void regen_kernel(frame_params f, paths_dev p, loop_ctl *ctl, int parity, batch_dev bd) {
    const uint32_t n = ctl->n_finish;
    int32_t *qn = p.queue[parity ^ 1];
    const uint32_t step = gridDim.x * blockDim.x;
    for(uint32_t base = blockIdx.x * blockDim.x; base < n; base += step) {
        uint32_t item = base + threadIdx.x;
        bool cont = false;
        int32_t pix = 0;
        if(item < n) {
            pix = p.finish_queue[item];
            float4 ra = p.rad[pix];
            float *sp = p.sum + 3 * size_t(pix);
            sp[0] = sp[0] + ra.x; sp[1] = sp[1] + ra.y; sp[2] = sp[2] + ra.z;
            int left = p.ctr[pix].y - 1;
            if(left > 0) {
                rng_t r = pixel_rng(p.rng[pix], local_index<BATCH>(bd, pix));
                v3 o, d; float time;
                pixel_camera_ray<BATCH>(f, bd, pix, r, o, d, time);
                p.rng[pix] = r.state;
                p.ray_o[pix] = make_float4(o.x, o.y, o.z, time);
                p.ray_d[pix] = make_float4(d.x, d.y, d.z, 0);
                p.thr[pix] = make_float4(1, 1, 1, 0);
                p.rad[pix] = make_float4(0, 0, 0, 0);
                cont = true;
            }
            p.ctr[pix] = make_int2(0, left);
        }
        block_append3(cont, false, false, pix, qn, &ctl->n_ext[parity ^ 1], nullptr, nullptr, nullptr, nullptr);
    }
}

// cuda_renderer.cu:99-109 (pack_pixels)
__global__ void pack_kernel(const float *sum, uint8_t *rgba, int64_t count, int32_t samples) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if(i >= count) return;
    float c[3];
    for(int k = 0; k < 3; k++) c[k] = samples ? sum[3 * i + k] / float(samples) : 0.0f;
    for(int k = 0; k < 3; k++) {
        float value = c[k];
        float gamma = value > 0 && value < 1 ? sqrtf(value) : 0;
        rgba[4 * i + k] = uint8_t(!(value > 0) ? 0 : value >= 1 ? 255 : int(255.99 * double(gamma)));
    }
    rgba[4 * i + 3] = 255;
}

// image(): sum * float(1.0 / samples) per component (vec3.h:152-158).
__global__ void scale_kernel(const float *sum, float *out, int64_t count, float k) {
    for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x)
        out[i] = sum[i] * k;
}

} // namespace

__global__ void trig_values_kernel(float2 *out) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k < (1u << 24)) out[k] = device_cos_sin(k);
}

// ---------------- launchers ----------------
void launch_trig_values(float2 *out, hipStream_t stream) {
    hipLaunchKernelGGL(trig_base_kernel, dim3(4), dim3(256), 0, stream);
    hipLaunchKernelGGL(trig_values_kernel, dim3((1u << 24) / 256), dim3(256), 0, stream, out);
}
void launch_query(const scene_dev& s, const pt_ray *rays, pt_trace_result *out, int32_t *boxes, int32_t *prims,
                  int32_t count, bool brute, hipStream_t stream) {
    const int block = 256;
    unsigned grid = unsigned((count + block - 1) / block);
    hipLaunchKernelGGL(query_kernel, dim3(grid), dim3(block), stack_bytes(s.stack_capacity, block), stream,
                       s, rays, out, boxes, prims, count, brute ? 1 : 0);
}
void launch_wide_query(const scene_dev& s, const pt_ray *rays, pt_trace_result *out, int32_t *retraced, int2 *spill,
                       int32_t count, int block, hipStream_t stream) {
    const unsigned grid = unsigned((count + block - 1) / block);
    if(grid) hipLaunchKernelGGL(wide_query_kernel, dim3(grid), dim3(unsigned(block)), wide_stack_bytes(s.wide_stack_capacity, block),
                                stream, s, rays, out, retraced, spill, count);
}
void launch_init_pixels(float *sum, uint64_t *rng, int64_t count, uint64_t seed, loop_ctl *ctl, hipStream_t stream) {
    unsigned grid = unsigned((count + 255) / 256);
    hipLaunchKernelGGL(init_pixels_kernel, dim3(grid ? grid : 1), dim3(256), 0, stream, sum, rng, count, seed, ctl);
}
void launch_debug(const scene_dev& s, const frame_params& f, float *sum, uint64_t *rng, int32_t samples,
                  int32_t *failure, dim3 block, hipStream_t stream) {
    dim3 grid((unsigned(f.width) + block.x - 1) / block.x, (unsigned(f.height) + block.y - 1) / block.y);
    hipLaunchKernelGGL(debug_kernel, grid, block, stack_bytes(s.stack_capacity, int(block.x * block.y)), stream,
                       s, f, sum, rng, samples, failure);
}
void launch_start(const scene_dev& s, const frame_params& f, const paths_dev& p, loop_ctl *ctl, int32_t samples,
                  int part, int parts, int chunk, int64_t count, hipStream_t stream, const batch_dev *batch) {
    unsigned grid = unsigned((count + 255) / 256);
    const int64_t n = batch ? count : int64_t(f.width) * f.height;
    if(batch)
        hipLaunchKernelGGL(start_kernel<true>, dim3(grid ? grid : 1), dim3(256), 0, stream, s, f, p, ctl, samples, part,
                           parts, chunk, count, n, *batch);
    else
        hipLaunchKernelGGL(start_kernel<false>, dim3(grid ? grid : 1), dim3(256), 0, stream, s, f, p, ctl, samples, part,
                           parts, chunk, count, n, batch_dev{});
}
void launch_batch_init(const batch_dev& b, float *sum, uint64_t *rng, int64_t count, loop_ctl *ctl, hipStream_t stream) {
    unsigned grid = unsigned((count + 255) / 256);
    hipLaunchKernelGGL(batch_init_kernel, dim3(grid ? grid : 1), dim3(256), 0, stream, b, sum, rng, count, ctl);
}
void launch_clear_ctl(loop_ctl *ctl, hipStream_t stream) {
    hipLaunchKernelGGL(clear_ctl_kernel, dim3(1), dim3(1), 0, stream, ctl);
}
void launch_trace(const scene_dev& s, const frame_params& f, const paths_dev& p, loop_ctl *ctl, int parity,
                  int grid, int block, bool count_tests, hipStream_t stream) {
    size_t lds = trace_lds_bytes(s, f, block);
    if(count_tests)
        hipLaunchKernelGGL(trace_kernel<true>, dim3(unsigned(grid)), dim3(unsigned(block)), lds, stream, s, f, p, ctl, parity);
    else
        hipLaunchKernelGGL(trace_kernel<false>, dim3(unsigned(grid)), dim3(unsigned(block)), lds, stream, s, f, p, ctl, parity);
}
void launch_trace_lanes(const scene_dev& s, const frame_params& f, const paths_dev& p, loop_ctl *ctl, int parity,
                        int grid, int block, int refill, int hand_steps, bool count_tests, hipStream_t stream,
                        const batch_dev *batch) {
    size_t lds = wide_stack_bytes(s.wide_stack_capacity, block);
    const dim3 g{unsigned(grid)}, b{unsigned(block)};
    // The counting pass runs every query to its end within its launch; nothing
    // is handed on without save buffers either.
    const int tail = count_tests || p.save_capacity <= 0 ? -1 : hand_steps;
    if(batch) {
        if(count_tests) hipLaunchKernelGGL((trace_lanes_kernel<true, true>), g, b, lds, stream, s, f, p, ctl, parity, refill, tail, *batch);
        else hipLaunchKernelGGL((trace_lanes_kernel<false, true>), g, b, lds, stream, s, f, p, ctl, parity, refill, tail, *batch);
    } else {
        if(count_tests)
            hipLaunchKernelGGL((trace_lanes_kernel<true, false>), g, b, lds, stream, s, f, p, ctl, parity, refill, tail, batch_dev{});
        else
            hipLaunchKernelGGL((trace_lanes_kernel<false, false>), g, b, lds, stream, s, f, p, ctl, parity, refill, tail, batch_dev{});
    }
}
void launch_bookkeep(loop_ctl *ctl, int parity, hipStream_t stream) {
    hipLaunchKernelGGL(bookkeep_kernel, dim3(1), dim3(1), 0, stream, ctl, parity);
}
void launch_shade(const scene_dev& s, const frame_params& f, const paths_dev& p, loop_ctl *ctl, int parity,
                  int grid, int block, hipStream_t stream, const batch_dev *batch) {
    if(batch)
        hipLaunchKernelGGL(shade_kernel<true>, dim3(unsigned(grid)), dim3(unsigned(block)), 0, stream, s, f, p, ctl, parity,
                           *batch);
    else
        hipLaunchKernelGGL(shade_kernel<false>, dim3(unsigned(grid)), dim3(unsigned(block)), 0, stream, s, f, p, ctl, parity,
                           batch_dev{});
}
void launch_regen(const frame_params& f, const paths_dev& p, loop_ctl *ctl, int parity, int grid, int block,
                  hipStream_t stream, const batch_dev *batch) {
    if(batch)
        hipLaunchKernelGGL(regen_kernel<true>, dim3(unsigned(grid)), dim3(unsigned(block)), 0, stream, f, p, ctl, parity,
                           *batch);
    else
        hipLaunchKernelGGL(regen_kernel<false>, dim3(unsigned(grid)), dim3(unsigned(block)), 0, stream, f, p, ctl, parity,
                           batch_dev{});
}
void launch_drain(const scene_dev& s, const frame_params& f, const paths_dev& p, loop_ctl *ctl, int parity,
                  uint32_t shadows, uint32_t saved, uint32_t paths, int block, path_lanes lanes, hipStream_t stream,
                  const batch_dev *batch) {
    size_t lds = trace_lds_bytes(s, f, block);
    const batch_dev bd = batch ? *batch : batch_dev{};
    saved = std::min(saved, uint32_t(std::max(0, p.save_capacity)));
    if(shadows || saved) {
        unsigned grid = unsigned((std::max(shadows, saved) + uint32_t(block) - 1) / uint32_t(block));
        if(batch) hipLaunchKernelGGL(drain_shadow_kernel<true>, dim3(grid), dim3(unsigned(block)), lds, stream, s, f, p, ctl, parity, bd);
        else hipLaunchKernelGGL(drain_shadow_kernel<false>, dim3(grid), dim3(unsigned(block)), lds, stream, s, f, p, ctl, parity, bd);
    }
    if(paths) {
        unsigned grid = unsigned((paths + uint32_t(block) - 1) / uint32_t(block));
        if(wide_traversal(s, f)) {  // lanes on their own state machines, taking path after path
            launch_path_lanes(s, f, p, ctl, parity, p.queue[parity], paths, block, lanes, path_pass{}, stream, batch);
        } else if(batch) {
            hipLaunchKernelGGL(drain_kernel<true>, dim3(grid), dim3(unsigned(block)), lds, stream, s, f, p, ctl, parity, bd);
        } else {
            hipLaunchKernelGGL(drain_kernel<false>, dim3(grid), dim3(unsigned(block)), lds, stream, s, f, p, ctl, parity, bd);
        }
    }
}
template<int MODE>
static void path_lanes_mode(const scene_dev& s, const frame_params& f, const paths_dev& p, loop_ctl *ctl, int parity,
                            const int32_t *queue, int shade_batch, int trace_steps, const path_pass& pass, dim3 grid,
                            dim3 block, size_t lds, hipStream_t stream, const batch_dev *batch) {
    if constexpr(MODE == PATHS_FRESH) {  // single renderers only
        hipLaunchKernelGGL((path_lanes_kernel<false, MODE>), grid, block, lds, stream, s, f, p, ctl, parity, queue,
                           shade_batch, trace_steps, pass.first_samples, pass.costs, batch_dev{});
        return;
    }
    if(batch)
        hipLaunchKernelGGL((path_lanes_kernel<true, MODE>), grid, block, lds, stream, s, f, p, ctl, parity, queue,
                           shade_batch, trace_steps, pass.first_samples, pass.costs, *batch);
    else
        hipLaunchKernelGGL((path_lanes_kernel<false, MODE>), grid, block, lds, stream, s, f, p, ctl, parity, queue,
                           shade_batch, trace_steps, pass.first_samples, pass.costs, batch_dev{});
}
void launch_path_lanes(const scene_dev& s, const frame_params& f, const paths_dev& p, loop_ctl *ctl, int parity,
                       const int32_t *queue, uint32_t paths, int block, path_lanes lanes, path_pass pass,
                       hipStream_t stream, const batch_dev *batch) {
    // The kernel keeps wide_lds_entries stack entries per lane in LDS (the
    // binary stack of a retrace uses the same entries).
    if(!wide_traversal(s, f)) throw std::logic_error("path lanes need the 4-wide traversal");
    if(block != path_lane_block) throw std::logic_error("path lanes run blocks of path_lane_block threads");
    if(pass.mode == PATHS_FIRST && (!pass.costs || pass.first_samples <= 0))
        throw std::logic_error("first path-lane pass without a cost buffer");
    if(pass.mode == PATHS_FRESH && (batch || pass.first_samples <= 0))
        throw std::logic_error("fresh path-lane pass without a sample count, or on a batch");
    if(!paths) return;
    // wide_lds_entries stack entries and the pixel's running sum per lane
    const size_t lds = size_t(block) * (wide_lds_entries * sizeof(int2) + 3 * sizeof(float));
    const unsigned grid = unsigned((paths + uint32_t(block) - 1) / uint32_t(block));
    const unsigned g = lanes.blocks > 0 ? std::min(grid, unsigned(lanes.blocks)) : grid;
    const int batch_lanes = std::max(1, std::min(64, lanes.shade_batch));
    const int steps = std::max(1, lanes.trace_steps);
    const dim3 gd{g}, bd{unsigned(block)};
    switch(pass.mode) {
    case PATHS_FIRST: path_lanes_mode<PATHS_FIRST>(s, f, p, ctl, parity, queue, batch_lanes, steps, pass, gd, bd, lds, stream, batch); break;
    case PATHS_REST: path_lanes_mode<PATHS_REST>(s, f, p, ctl, parity, queue, batch_lanes, steps, pass, gd, bd, lds, stream, batch); break;
    case PATHS_FRESH: path_lanes_mode<PATHS_FRESH>(s, f, p, ctl, parity, queue, batch_lanes, steps, pass, gd, bd, lds, stream, batch); break;
    default: path_lanes_mode<PATHS_WHOLE>(s, f, p, ctl, parity, queue, batch_lanes, steps, pass, gd, bd, lds, stream, batch); break;
    }
}
// 8-bit key of a group's summed cost: exponent and top three mantissa bits of
// float(cost), 1 for cost <= 1 and +8 per doubling, at most 255.
__device__ __forceinline__ uint8_t cost_level(uint64_t cost) {
    const uint32_t level = (__float_as_uint(float(cost > 1 ? cost : 1)) >> 20) - 1015u;
    return uint8_t(level < 255u ? level : 255u);
}
__global__ void group_keys_kernel(const uint32_t *costs, uint32_t n, int group, uint8_t *keys, uint32_t *ids) {
    const uint32_t groups = (n + uint32_t(group) - 1) / uint32_t(group);
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if(g >= groups) return;
    const uint32_t begin = g * uint32_t(group), end = min(n, begin + uint32_t(group));
    uint64_t sum = 0;
    for(uint32_t i = begin; i < end; i++) sum += costs[i];
    keys[g] = end - begin < uint32_t(group) ? uint8_t(0) : cost_level(sum);
    ids[g] = g;
}
__global__ void group_gather_kernel(const int32_t *queue, const uint32_t *ids, uint32_t n, int group, int32_t *out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if(k >= n) return;
    out[k] = queue[ids[k / uint32_t(group)] * uint32_t(group) + k % uint32_t(group)];
}
void launch_group_keys(const uint32_t *costs, uint32_t n, int group, uint8_t *keys, uint32_t *ids, hipStream_t stream) {
    const uint32_t groups = (n + uint32_t(group) - 1) / uint32_t(group);
    if(groups) hipLaunchKernelGGL(group_keys_kernel, dim3((groups + 255) / 256), dim3(256), 0, stream, costs, n, group, keys, ids);
}
void launch_group_gather(const int32_t *queue, const uint32_t *ids, uint32_t n, int group, int32_t *out,
                         hipStream_t stream) {
    if(n) hipLaunchKernelGGL(group_gather_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, queue, ids, n, group, out);
}
__global__ void zero_work_kernel(loop_ctl *ctl) { ctl->work = 0; }
// A control block cleared for one launch over `n` items (a kernel rather than a
// memset, so that the call can be captured into a HIP graph).
__global__ void fresh_ctl_kernel(loop_ctl *ctl, uint32_t n) {
    loop_ctl c{};
    c.n_ext[0] = n;
    *ctl = c;
}
void launch_fresh_ctl(loop_ctl *ctl, uint32_t n, hipStream_t stream) {
    hipLaunchKernelGGL(fresh_ctl_kernel, dim3(1), dim3(1), 0, stream, ctl, n);
}
void launch_zero_work(loop_ctl *ctl, hipStream_t stream) {
    hipLaunchKernelGGL(zero_work_kernel, dim3(1), dim3(1), 0, stream, ctl);
}
void launch_pack(const float *sum, uint8_t *rgba, int64_t count, int32_t samples, hipStream_t stream) {
    unsigned grid = unsigned((count + 255) / 256);
    hipLaunchKernelGGL(pack_kernel, dim3(grid ? grid : 1), dim3(256), 0, stream, sum, rgba, count, samples);
}
void launch_scale(const float *sum, float *out, int64_t count, float k, hipStream_t stream) {
    const int64_t blocks = std::min<int64_t>(std::max<int64_t>(1, (count + 255) / 256), 65536);
    hipLaunchKernelGGL(scale_kernel, dim3(unsigned(blocks)), dim3(256), 0, stream, sum, out, count, k);
}
void launch_add_rays(loop_ctl *ctl, unsigned long long rays, hipStream_t stream) {
    hipLaunchKernelGGL(add_rays_kernel, dim3(1), dim3(1), 0, stream, ctl, rays);
}

} // namespace pt
