// pt_lbvh.hip -- linear BVH builder on the GPU, emitting the reference's flat
// hierarchy format (src/flat/flat_data.h: one primitive per leaf, branch
// boxes enclosing their children, depth at most flat_stack_capacity).  It is
// the fast-build option beside the host binned-SAH and reference-order
// builders (include/pt_scene.h): milliseconds instead of seconds at a million
// triangles, at a lower tree quality.
//
// Method (Karras, "Maximizing Parallelism in the Construction of BVHs,
// Octrees, and k-d Trees", HPG 2012):
//   1. leaf boxes by the reference's rules (triangle: min/max per axis padded
//      by 1e-4 and one ulp outwards, src/geometry/triangle.h:96-109; sphere:
//      center -+ radius, src/geometry/sphere.h:69-75) and box centroids;
//   2. the centroids' bounds, 63-bit Morton codes of the centroids
//      (21 bits per axis), a stable radix sort of (code, primitive);
//   3. every internal node independently from the sorted codes: the range
//      of leaves it covers and its split, by common-prefix lengths (equal
//      codes ordered by position);
//   4. branch boxes bottom up, one launch per level: each internal node's
//      depth from its parent links, then the deepest level first, every node
//      the union of its two children's boxes.  Kernel boundaries order the
//      levels; a climb with per-node arrival counters would need a device
//      scope release per node, which on this chip writes back an L2 per
//      fence (8 ms instead of a fraction of one at a million primitives).
// Internal node k is flat node k, leaf k of the sorted order flat node
// N - 1 + k; the root is node 0.  The output is the same for every run.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "pt_layout.h"

namespace pt {
namespace {

void check(hipError_t e, const char *what) {
    if(e != hipSuccess) fail(PT_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define LC(x) check((x), #x)

template<class T> struct buffer {
    T *p = nullptr;
    explicit buffer(size_t n) { if(n) LC(hipMalloc(reinterpret_cast<void **>(&p), n * sizeof(T))); }
    ~buffer() { if(p) (void)hipFree(p); }
    buffer(const buffer&) = delete;
    buffer& operator=(const buffer&) = delete;
};

// nextafterf(x, -inf) / nextafterf(x, +inf) for finite x.
__device__ inline float next_down(float x) {
    if(x == 0.0f) return -__uint_as_float(1u);
    const uint32_t b = __float_as_uint(x);
    return __uint_as_float(x > 0.0f ? b - 1u : b + 1u);
}
__device__ inline float next_up(float x) {
    if(x == 0.0f) return __uint_as_float(1u);
    const uint32_t b = __float_as_uint(x);
    return __uint_as_float(x > 0.0f ? b + 1u : b - 1u);
}

// Leaf boxes and centroids; primitive p < nt is triangle p, else sphere p - nt.
__global__ void leaf_boxes_kernel(const pt_flat_triangle *tris, int64_t nt, const pt_flat_sphere *sph, int64_t ns,
                                  float4 *lo, float4 *hi, float4 *centroid) {
    const int64_t n = nt + ns;
    for(int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < n; p += int64_t(gridDim.x) * blockDim.x) {
        float l[3], h[3];
        if(p < nt) {
            const pt_flat_triangle& t = tris[p];
            const float x0[3] = {t.v0.x, t.v0.y, t.v0.z}, x1[3] = {t.v1.x, t.v1.y, t.v1.z};
            const float x2[3] = {t.v2.x, t.v2.y, t.v2.z};
            for(int a = 0; a < 3; a++) {
                const float m12 = x1[a] < x2[a] ? x1[a] : x2[a], M12 = x1[a] > x2[a] ? x1[a] : x2[a];
                const float mn = x0[a] < m12 ? x0[a] : m12, mx = x0[a] > M12 ? x0[a] : M12;
                l[a] = next_down(mn - 0.0001f);
                h[a] = next_up(mx + 0.0001f);
            }
        } else {
            const pt_flat_sphere& s = sph[p - nt];
            const float c[3] = {s.center.x, s.center.y, s.center.z};
            for(int a = 0; a < 3; a++) { l[a] = c[a] - s.radius; h[a] = c[a] + s.radius; }
        }
        lo[p] = make_float4(l[0], l[1], l[2], 0);
        hi[p] = make_float4(h[0], h[1], h[2], 0);
        centroid[p] = make_float4(0.5f * l[0] + 0.5f * h[0], 0.5f * l[1] + 0.5f * h[1], 0.5f * l[2] + 0.5f * h[2], 0);
    }
}

// Order-preserving unsigned image of a finite float (for atomic min/max).
__device__ inline uint32_t ordered(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float unordered(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// bounds[0..2] = ordered min per axis, bounds[3..5] = ordered max (pre-set to
// 0xffffffff and 0).  Each block reduces in LDS and adds one atomic per
// word, so the atomics on the six words stay few.  256 threads per block.
__global__ void __launch_bounds__(256) centroid_bounds_kernel(const float4 *centroid, int64_t n, uint32_t *bounds) {
    uint32_t v[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0, 0, 0};
    for(int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < n; p += int64_t(gridDim.x) * blockDim.x) {
        const float4 c = centroid[p];
        const uint32_t o[3] = {ordered(c.x), ordered(c.y), ordered(c.z)};
        for(int a = 0; a < 3; a++) {
            v[a] = o[a] < v[a] ? o[a] : v[a];
            v[3 + a] = o[a] > v[3 + a] ? o[a] : v[3 + a];
        }
    }
    __shared__ uint32_t red[6][256];
    for(int a = 0; a < 6; a++) red[a][threadIdx.x] = v[a];
    __syncthreads();
    for(int half = 128; half > 0; half >>= 1) {
        if(int(threadIdx.x) < half) {
            for(int a = 0; a < 3; a++) {
                const uint32_t m = red[a][threadIdx.x + half], M = red[3 + a][threadIdx.x + half];
                if(m < red[a][threadIdx.x]) red[a][threadIdx.x] = m;
                if(M > red[3 + a][threadIdx.x]) red[3 + a][threadIdx.x] = M;
            }
        }
        __syncthreads();
    }
    if(threadIdx.x < 3) atomicMin(&bounds[threadIdx.x], red[threadIdx.x][0]);
    else if(threadIdx.x < 6) atomicMax(&bounds[threadIdx.x], red[threadIdx.x][0]);
}

// 21 bits of x spread to every third bit of a 63-bit word.
__device__ inline uint64_t spread21(uint64_t x) {
    x &= 0x1fffffull;
    x = (x | (x << 32)) & 0x001f00000000ffffull;
    x = (x | (x << 16)) & 0x001f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__global__ void morton_kernel(const float4 *centroid, int64_t n, const uint32_t *bounds, uint64_t *codes, int32_t *ids) {
    const float lo[3] = {unordered(bounds[0]), unordered(bounds[1]), unordered(bounds[2])};
    const float hi[3] = {unordered(bounds[3]), unordered(bounds[4]), unordered(bounds[5])};
    for(int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < n; p += int64_t(gridDim.x) * blockDim.x) {
        const float4 c = centroid[p];
        const float v[3] = {c.x, c.y, c.z};
        uint64_t q[3];
        for(int a = 0; a < 3; a++) {
            const float ext = hi[a] - lo[a];
            float x = ext > 0.0f ? (v[a] - lo[a]) / ext : 0.0f;
            x = fminf(fmaxf(x, 0.0f), 1.0f);
            const uint32_t k = uint32_t(fminf(x * 2097152.0f, 2097151.0f));
            q[a] = k;
        }
        codes[p] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
        ids[p] = int32_t(p);
    }
}

// Common prefix length of the sorted keys at positions i and j, extended by
// the positions' own prefix where the keys are equal; -1 outside [0, n).
__device__ inline int delta(const uint64_t *keys, int64_t n, int64_t i, int64_t j) {
    if(j < 0 || j >= n) return -1;
    const uint64_t a = keys[i], b = keys[j];
    if(a != b) return __clzll(a ^ b);
    return 64 + __clzll(uint64_t(i) ^ uint64_t(j));
}

// Internal node i of n leaves (n >= 2): children and parents in flat
// numbering (internal k -> k, leaf k -> n - 1 + k).  The root (node 0) is
// no node's child; its parent entry is set to -1 here, which ends every
// climb of refit_kernel.
__global__ void internal_nodes_kernel(const uint64_t *keys, int64_t n, int32_t *left, int32_t *right, int32_t *parent) {
    for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n - 1; i += int64_t(gridDim.x) * blockDim.x) {
        if(i == 0) parent[0] = -1;
        const int d = delta(keys, n, i, i + 1) > delta(keys, n, i, i - 1) ? 1 : -1;
        const int dmin = delta(keys, n, i, i - d);
        int64_t lmax = 2;
        while(delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
        int64_t l = 0;
        for(int64_t t = lmax >> 1; t >= 1; t >>= 1)
            if(delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
        const int64_t j = i + l * d;
        const int dnode = delta(keys, n, i, j);
        int64_t s = 0, t = l;
        do {
            t = (t + 1) >> 1;
            if(delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        } while(t > 1);
        const int64_t split = i + s * d + (d < 0 ? -1 : 0);
        const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
        const int32_t lc = int32_t(lo == split ? n - 1 + split : split);
        const int32_t rc = int32_t(hi == split + 1 ? n - 1 + split + 1 : split + 1);
        left[i] = lc;
        right[i] = rc;
        parent[lc] = int32_t(i);
        parent[rc] = int32_t(i);
    }
}

// A radix tree over 63-bit codes extended by 31-bit positions is at most
// 63 + 31 levels deep; a longer parent chain is broken.
constexpr int max_climb = 96;

// Leaf records: sorted leaf k is flat node n - 1 + k.
__global__ void leaves_kernel(const int32_t *ids, int64_t n, int64_t nt, const float4 *lo, const float4 *hi,
                              pt_flat_node *nodes) {
    for(int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += int64_t(gridDim.x) * blockDim.x) {
        const int32_t p = ids[k];
        const float4 l = lo[p], h = hi[p];
        pt_flat_node leaf;
        leaf.minimum = pt_vec3{l.x, l.y, l.z};
        leaf.maximum = pt_vec3{h.x, h.y, h.z};
        leaf.type = p < nt ? PT_TRIANGLE : PT_SPHERE;
        leaf.left = -1;
        leaf.right = -1;
        leaf.primitive = p < nt ? p : int32_t(p - nt);
        nodes[n - 1 + k] = leaf;
    }
}

// Depth of every internal node (root 0) by following its parent links to
// the root, and the largest depth into *deepest (pre-set to 0); a chain
// longer than max_climb sets *deepest to max_climb.  256 threads per block.
__global__ void __launch_bounds__(256) depth_kernel(const int32_t *parent, int64_t n, int32_t *depth,
                                                    uint32_t *deepest) {
    uint32_t most = 0;
    for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n - 1; i += int64_t(gridDim.x) * blockDim.x) {
        int32_t d = 0;
        for(int32_t q = parent[i]; q >= 0 && d < max_climb; q = parent[q]) d++;
        depth[i] = d;
        most = uint32_t(d) > most ? uint32_t(d) : most;
    }
    __shared__ uint32_t red[256];
    red[threadIdx.x] = most;
    __syncthreads();
    for(int half = 128; half > 0; half >>= 1) {
        if(int(threadIdx.x) < half && red[threadIdx.x + half] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + half];
        __syncthreads();
    }
    if(threadIdx.x == 0) atomicMax(deepest, red[0]);
}

// Branch records of the internal nodes at one depth; the nodes below were
// written by the launches for the deeper levels (or leaves_kernel) earlier
// on the same stream.  The box is the reference's union (aabb.h:40-56,
// ffmin/ffmax of left and right).
__global__ void level_kernel(const int32_t *depth, int64_t n, int32_t level, const int32_t *left, const int32_t *right,
                             pt_flat_node *nodes) {
    for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n - 1; i += int64_t(gridDim.x) * blockDim.x) {
        if(depth[i] != level) continue;
        const int32_t l = left[i], r = right[i];
        const pt_flat_node a = nodes[l], b = nodes[r];
        pt_flat_node br;
        br.minimum = pt_vec3{a.minimum.x < b.minimum.x ? a.minimum.x : b.minimum.x,
                             a.minimum.y < b.minimum.y ? a.minimum.y : b.minimum.y,
                             a.minimum.z < b.minimum.z ? a.minimum.z : b.minimum.z};
        br.maximum = pt_vec3{a.maximum.x > b.maximum.x ? a.maximum.x : b.maximum.x,
                             a.maximum.y > b.maximum.y ? a.maximum.y : b.maximum.y,
                             a.maximum.z > b.maximum.z ? a.maximum.z : b.maximum.z};
        br.type = PT_BRANCH;
        br.left = l;
        br.right = r;
        br.primitive = -1;
        nodes[i] = br;
    }
}

unsigned grid_of(int64_t n) { return unsigned(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536))); }

} // namespace

// The flat hierarchy of nt triangles and ns spheres (n = nt + ns >= 1):
// 2n - 1 nodes into nodes_out (host memory), root 0; the device time of the
// build in *milliseconds.  Inputs are validated by the caller.
void build_lbvh(const pt_flat_triangle *triangles, int64_t nt, const pt_flat_sphere *spheres, int64_t ns,
                pt_flat_node *nodes_out, float *milliseconds) {
    const int64_t n = nt + ns;
    const size_t count = size_t(2 * n - 1);
    hipStream_t stream;
    LC(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    struct stream_guard { hipStream_t s; ~stream_guard() { (void)hipStreamDestroy(s); } } sg{stream};
    struct event_guard {
        hipEvent_t e = nullptr;
        event_guard() { LC(hipEventCreate(&e)); }
        ~event_guard() { (void)hipEventDestroy(e); }
    } g0, g1, g2, g3;
    // Device time: e0..e1 (boxes, codes, sort, splits, depths, leaves) and
    // e2..e3 (branch levels), without the host's read of the level count.
    hipEvent_t e0 = g0.e, e1 = g1.e, e2 = g2.e, e3 = g3.e;

    buffer<pt_flat_triangle> tris{size_t(nt)};
    buffer<pt_flat_sphere> sph{size_t(ns)};
    if(nt) LC(hipMemcpyAsync(tris.p, triangles, size_t(nt) * sizeof(pt_flat_triangle), hipMemcpyHostToDevice, stream));
    if(ns) LC(hipMemcpyAsync(sph.p, spheres, size_t(ns) * sizeof(pt_flat_sphere), hipMemcpyHostToDevice, stream));
    buffer<float4> lo{size_t(n)}, hi{size_t(n)}, centroid{size_t(n)};
    buffer<uint32_t> bounds{7};
    buffer<uint64_t> keys{size_t(n)}, keys_sorted{size_t(n)};
    buffer<int32_t> ids{size_t(n)}, ids_sorted{size_t(n)};
    buffer<int32_t> left{size_t(n)}, right{size_t(n)}, parent{count};
    buffer<int32_t> depth{size_t(n - 1)};
    buffer<pt_flat_node> nodes{count};
    size_t temp_bytes = 0;
    LC(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys.p, keys_sorted.p, ids.p, ids_sorted.p, int(n), 0, 63,
                                          stream));
    buffer<unsigned char> temp{temp_bytes};
    // bounds: ordered centroid minima and maxima, then the deepest internal node.
    const uint32_t init[7] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0, 0, 0, 0};
    LC(hipMemcpyAsync(bounds.p, init, sizeof init, hipMemcpyHostToDevice, stream));

    LC(hipEventRecord(e0, stream));
    hipLaunchKernelGGL(leaf_boxes_kernel, dim3(grid_of(n)), dim3(256), 0, stream, tris.p, nt, sph.p, ns, lo.p, hi.p,
                       centroid.p);
    hipLaunchKernelGGL(centroid_bounds_kernel, dim3(std::min(grid_of(n), 512u)), dim3(256), 0, stream, centroid.p, n,
                       bounds.p);
    hipLaunchKernelGGL(morton_kernel, dim3(grid_of(n)), dim3(256), 0, stream, centroid.p, n, bounds.p, keys.p, ids.p);
    LC(hipcub::DeviceRadixSort::SortPairs(temp.p, temp_bytes, keys.p, keys_sorted.p, ids.p, ids_sorted.p, int(n), 0, 63,
                                          stream));
    if(n > 1) {
        hipLaunchKernelGGL(internal_nodes_kernel, dim3(grid_of(n - 1)), dim3(256), 0, stream, keys_sorted.p, n, left.p,
                           right.p, parent.p);
        hipLaunchKernelGGL(depth_kernel, dim3(std::min(grid_of(n - 1), 1024u)), dim3(256), 0, stream, parent.p, n,
                           depth.p, bounds.p + 6);
    }
    hipLaunchKernelGGL(leaves_kernel, dim3(grid_of(n)), dim3(256), 0, stream, ids_sorted.p, n, nt, lo.p, hi.p, nodes.p);
    LC(hipGetLastError());
    LC(hipEventRecord(e1, stream));
    // The number of levels decides the launches below.
    uint32_t deepest = 0;
    if(n > 1) {
        LC(hipMemcpyAsync(&deepest, bounds.p + 6, sizeof deepest, hipMemcpyDeviceToHost, stream));
        LC(hipStreamSynchronize(stream));
        if(deepest >= uint32_t(max_climb)) fail(PT_EDEVICE, "linear BVH build: broken parent chain");
    }
    LC(hipEventRecord(e2, stream));
    for(int level = int(deepest); n > 1 && level >= 0; level--)
        hipLaunchKernelGGL(level_kernel, dim3(grid_of(n - 1)), dim3(256), 0, stream, depth.p, n, level, left.p, right.p,
                           nodes.p);
    LC(hipGetLastError());
    LC(hipEventRecord(e3, stream));
    LC(hipMemcpyAsync(nodes_out, nodes.p, count * sizeof(pt_flat_node), hipMemcpyDeviceToHost, stream));
    LC(hipStreamSynchronize(stream));
    float first = 0, second = 0;
    LC(hipEventElapsedTime(&first, e0, e1));
    LC(hipEventElapsedTime(&second, e2, e3));
    if(milliseconds) *milliseconds = first + second;
}

} // namespace pt
