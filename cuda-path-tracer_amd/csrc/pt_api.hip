// pt_api.hip -- host runtime of the path-tracing core and the C ABI of
// include/pt_render.h.  It replaces the host half of the reference's
// src/cuda/cuda_renderer.cu: validation, scene upload, the per-renderer
// accumulation state and the add_samples bounce loop.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include <hipcub/hipcub.hpp>
#include "pt_device.h"
#include "pt_layout.h"
#include "pt_libm.h"

namespace {

thread_local std::string g_last_error;

using pt::status_error;
using pt::fail;
using pt::validate_scene;
using pt::validate_camera;
using pt::validate_mode;
using pt::checked_pixel_count;
using pt::finite3;

void hip_check(hipError_t e, const char *what) {
    if(e != hipSuccess) fail(PT_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIP_CHECK(call) hip_check((call), #call)

template<class F> pt_status guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return PT_OK;
    } catch(const status_error& e) {
        g_last_error = e.what();
        return e.code;
    } catch(const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return PT_ENOMEM;
    } catch(const std::exception& e) {
        g_last_error = e.what();
        return PT_EDEVICE;
    }
}

template<class T> class dev_buf {
public:
    dev_buf() = default;
    explicit dev_buf(size_t n) { alloc(n); }
    ~dev_buf() { release(); }
    dev_buf(const dev_buf&) = delete;
    dev_buf& operator=(const dev_buf&) = delete;
    void alloc(size_t n) {
        release();
        if(n > std::numeric_limits<size_t>::max() / sizeof(T)) fail(PT_EINVAL, "device allocation size overflow");
        count = n;
        if(n) HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&ptr), n * sizeof(T)));
    }
    void upload(const T *src, size_t n) {
        alloc(n);
        if(n) HIP_CHECK(hipMemcpy(ptr, src, n * sizeof(T), hipMemcpyHostToDevice));
    }
    void release() { if(ptr) (void)hipFree(ptr); ptr = nullptr; count = 0; }
    T *get() const { return ptr; }
    size_t size() const { return count; }
private:
    T *ptr = nullptr;
    size_t count = 0;
};

// ---------------- host-libm tables ----------------
// cos/sin of every angle cosine_direction and sphere-light sampling can form:
// 2*float(M_PI)*v with v = k*2^-24 (light_sampling.h:20,47), with the host
// libm, as the reference's host build computes them.
constexpr size_t trig_entries = size_t(1) << 24;

const std::vector<float2>& host_trig_table() {
    static std::once_flag once;
    static std::vector<float2> table;
    std::call_once(once, [] {
        table.resize(trig_entries);
        unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for(unsigned t = 0; t < threads; t++) {
            pool.emplace_back([t, threads] {
                const float two_pi = 2 * float(M_PI);
                size_t chunk = trig_entries / threads;
                size_t begin = t * chunk, end = t + 1 == threads ? trig_entries : begin + chunk;
                for(size_t k = begin; k < end; k++) {
                    float v = float(uint32_t(k)) * 0x1p-24f;
                    float angle = two_pi * v;
                    table[k] = make_float2(cosf(angle), sinf(angle));
                }
            });
        }
        for(auto& th : pool) th.join();
    });
    return table;
}

struct trig_tables { const unsigned long long *bits; const uint32_t *rank; const float2 *exceptions; };

// Per device, once per process: compare the device's double-rounded cos/sin
// with the host libm for all 2^24 angles and keep only the differences.
trig_tables device_trig_tables() {
    static std::mutex lock;
    static std::map<int, trig_tables> cache;
    int device = 0;
    HIP_CHECK(hipGetDevice(&device));
    std::lock_guard<std::mutex> guard(lock);
    auto it = cache.find(device);
    if(it != cache.end()) return it->second;
    const auto& host = host_trig_table();
    std::vector<float2> dev(trig_entries);
    {
        dev_buf<float2> tmp(trig_entries);
        pt::launch_trig_values(tmp.get(), nullptr);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(dev.data(), tmp.get(), trig_entries * sizeof(float2), hipMemcpyDeviceToHost));
    }
    const size_t words = trig_entries / 64;
    std::vector<unsigned long long> bits(words, 0);
    std::vector<uint32_t> rank(words, 0);
    std::vector<float2> exceptions;
    for(size_t w = 0; w < words; w++) {
        rank[w] = uint32_t(exceptions.size());
        for(size_t b = 0; b < 64; b++) {
            size_t k = 64 * w + b;
            const float2 h = host[k], d = dev[k];
            if(std::memcmp(&h.x, &d.x, 4) != 0 || std::memcmp(&h.y, &d.y, 4) != 0) {
                bits[w] |= 1ull << b;
                exceptions.push_back(h);
            }
        }
    }
    if(exceptions.empty()) exceptions.push_back(make_float2(0, 0));
    unsigned long long *db = nullptr;
    uint32_t *dr = nullptr;
    float2 *de = nullptr;
    HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&db), words * sizeof(unsigned long long)));
    HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&dr), words * sizeof(uint32_t)));
    HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&de), exceptions.size() * sizeof(float2)));
    HIP_CHECK(hipMemcpy(db, bits.data(), words * sizeof(unsigned long long), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(dr, rank.data(), words * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(de, exceptions.data(), exceptions.size() * sizeof(float2), hipMemcpyHostToDevice));
    trig_tables t{db, dr, de};  // kept for the life of the process
    cache[device] = t;
    return t;
}

// x^5 of the host libm's pow(double, double) over the table domain of
// pt_libm.h, kept where it differs from the device's correctly rounded value
// (about 1 input in 2000 with glibc 2.35); per device, once per process.
struct pow5_device { const int32_t *keys; const double *values; int32_t count; };
pow5_device device_pow5_table() {
    static std::mutex lock;
    static std::map<int, pow5_device> cache;
    static std::vector<int32_t> keys;
    static std::vector<double> values;
    static std::once_flag once;
    std::call_once(once, [] { pt::build_pow5_table(keys, values); });
    int device = 0;
    HIP_CHECK(hipGetDevice(&device));
    std::lock_guard<std::mutex> guard(lock);
    auto it = cache.find(device);
    if(it != cache.end()) return it->second;
    int32_t *dk = nullptr;
    double *dv = nullptr;
    const size_t n = std::max<size_t>(1, keys.size());
    HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&dk), n * sizeof(int32_t)));
    HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&dv), n * sizeof(double)));
    if(!keys.empty()) {
        HIP_CHECK(hipMemcpy(dk, keys.data(), keys.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(dv, values.data(), values.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    pow5_device d{dk, dv, int32_t(keys.size())};  // kept for the life of the process
    cache[device] = d;
    return d;
}

void require_device() {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if(e != hipSuccess) fail(PT_EDEVICE, std::string("HIP device unavailable: ") + hipGetErrorString(e));
    if(count < 1) fail(PT_EDEVICE, "HIP device unavailable: no HIP devices");
}

// ---------------- uploaded scene (cuda_renderer.cu:36-54) ----------------
struct device_scene {
    dev_buf<float4> bnodes, tri_isect, sph_isect, shade_tri;
    dev_buf<pt_flat_triangle> triangles;
    dev_buf<pt_flat_sphere> spheres;
    dev_buf<pt_flat_node> nodes;
    dev_buf<pt_flat_material> materials;
    dev_buf<pt_flat_texture> textures;
    dev_buf<uint8_t> pixels;
    dev_buf<pt::light_rec> lights;
    dev_buf<float> log2_table;
    dev_buf<float4> wnodes, leaves;
    pt::scene_dev view{};
    int64_t branch_count = 0;
    std::vector<int32_t> branch_ids;   // compact branch index of each flat node (binary layout), -1 for leaves

    device_scene(const pt_scene_desc& s, bool with_tables, bool with_wide = false, bool with_log2 = true) {
        const pt::scene_layout L = pt::build_layout(s, with_wide);
        branch_ids = L.branch_id;
        branch_count = int64_t(L.bnodes.size() / 2);
        bnodes.upload(L.bnodes.data(), L.bnodes.size());
        tri_isect.upload(L.tri_isect.data(), L.tri_isect.size());
        sph_isect.upload(L.sph_isect.data(), L.sph_isect.size());
        shade_tri.upload(L.shade_tri.data(), L.shade_tri.size());
        triangles.upload(s.triangles, size_t(s.triangle_count));
        spheres.upload(s.spheres, size_t(s.sphere_count));
        nodes.upload(s.nodes, size_t(s.node_count));
        materials.upload(s.materials, size_t(s.material_count));
        textures.upload(s.textures, size_t(s.texture_count));
        pixels.upload(s.pixels, size_t(s.pixel_bytes));
        lights.upload(L.lights.data(), L.lights.size());
        view.bnodes = bnodes.get(); view.tri_isect = tri_isect.get(); view.sph_isect = sph_isect.get();
        view.shade_tri = shade_tri.get();
        view.root_ref = L.root_ref;
        view.triangles = triangles.get(); view.spheres = spheres.get(); view.nodes = nodes.get();
        view.node_count = int32_t(s.node_count);
        view.materials = materials.get(); view.textures = textures.get(); view.pixels = pixels.get();
        view.lights = lights.get(); view.light_count = int32_t(L.lights.size());
        view.light_area = L.lights.empty() ? 0 : L.lights.back().cumulative_area;
        view.stack_capacity = L.stack_capacity;
        view.trig_bits = nullptr; view.trig_rank = nullptr; view.trig_exceptions = nullptr;
        view.log2_table = nullptr; view.log2_count = 0;
        view.wnodes = nullptr; view.leaves = nullptr; view.wroot_ref = pt::empty_root; view.wide_stack_capacity = 0;
        if(L.has_wide) {
            wnodes.upload(L.wide.nodes.data(), L.wide.nodes.size());
            leaves.upload(L.wide.leaves.data(), L.wide.leaves.size());
            view.wnodes = wnodes.get(); view.leaves = leaves.get();
            view.wroot_ref = L.wide.root_ref;
            view.wide_stack_capacity = L.wide_stack_capacity;
        }
        view.pow5_keys = nullptr; view.pow5_values = nullptr; view.pow5_count = 0;
        if(with_tables) {
            const trig_tables t = device_trig_tables();
            view.trig_bits = t.bits; view.trig_rank = t.rank; view.trig_exceptions = t.exceptions;
            const pow5_device p = device_pow5_table();
            view.pow5_keys = p.keys; view.pow5_values = p.values; view.pow5_count = p.count;
        }
        if(with_tables && with_log2) {
            // log2f(1+n) of the host libm for every count a traversal can reach.
            size_t n = size_t(std::min<int64_t>(s.node_count + 2, int64_t(1) << 26));
            std::vector<float> lg(n);
            for(size_t k = 0; k < n; k++) lg[k] = log2f(1.0f + float(k));
            log2_table.upload(lg.data(), lg.size());
            view.log2_table = log2_table.get();
            view.log2_count = int32_t(n);
        }
    }
};

int device_cu_count() {
    int device = 0;
    HIP_CHECK(hipGetDevice(&device));
    hipDeviceProp_t props{};
    HIP_CHECK(hipGetDeviceProperties(&props, device));
    return std::max(1, props.multiProcessorCount);
}

} // namespace

// ---------------- renderer (cuda_renderer.cu:142-247) ----------------
// One independent bounce loop over a subset of the pixels, with its own
// queues, control block and stream.  Pixels never interact (one PCG stream
// and one writer each), so partitions run concurrently and one partition's
// launch tail overlaps another's work.
struct partition {
    int index = 0;
    int64_t count = 0;                   // pixels in this partition
    hipStream_t stream = nullptr;
    dev_buf<pt::loop_ctl> ctl;
    pt::loop_ctl *ctl_host = nullptr;    // pinned mirror
    dev_buf<int32_t> queue0, queue1, shadow_queue, finish_queue;
    dev_buf<int2> spill;
    dev_buf<uint32_t> log;
    dev_buf<int4> save_head[2], save_best[2];   // queries handed on between trace launches
    dev_buf<int2> save_stack[2];
    // split path-lane launches: pixel costs, group keys and indices, sort space
    dev_buf<uint32_t> costs, group_ids, sorted_ids;
    dev_buf<uint8_t> group_keys, sorted_keys, sort_temp;
    pt::paths_dev paths{};
    std::vector<hipEvent_t> events;      // 5 per iteration of a check interval
    hipEvent_t checked = nullptr, drain0 = nullptr, drain1 = nullptr, finished = nullptr;
    int parity = 0;
    bool done = false, drained = false;
    ~partition() {
        if(stream) (void)hipStreamSynchronize(stream);
        for(auto e : events) (void)hipEventDestroy(e);
        for(auto e : {checked, drain0, drain1, finished}) if(e) (void)hipEventDestroy(e);
        if(ctl_host) (void)hipHostFree(ctl_host);
        if(stream) (void)hipStreamDestroy(stream);
    }
    // Buffers for the queries a trace launch hands on to the next: up to
    // `capacity` per launch, `stride` stack entries each.
    void ensure_saves(int64_t capacity, int stride) {
        if(paths.save_capacity >= capacity && paths.save_stride == stride) return;
        for(int k = 0; k < 2; k++) {
            save_head[k].alloc(size_t(capacity)); save_best[k].alloc(size_t(capacity));
            save_stack[k].alloc(size_t(capacity) * size_t(std::max(1, stride)));
            paths.saved[k] = pt::saved_queries{save_head[k].get(), save_best[k].get(), save_stack[k].get()};
        }
        paths.save_capacity = int32_t(capacity);
        paths.save_stride = stride;
    }
    // Every path of the partition on path lanes, right after launch_start
    // (each pixel at the start of its first sample, queue0 in pixel order).
    // With more than `first` samples per pixel, more pixels than lanes in
    // flight (so that lanes take pixels from a queue) and the 4-wide
    // traversal, in two passes: `first` samples of every pixel, then the rest
    // with the queue reordered by the first samples' cost, most expensive
    // first, in groups of `group` consecutive pixels, so that no expensive
    // pixel starts near the end of the launch.  Returns whether it split.
    bool paths_from_start(const pt::scene_dev& view, const pt::frame_params& frame, int32_t samples, int32_t first,
                          int32_t group, pt::path_lanes lanes, int block, int cus, hipStream_t s,
                          const pt::batch_dev *batch = nullptr) {
        const uint32_t n = uint32_t(count);
        const uint64_t grid = (uint64_t(n) + uint64_t(block) - 1) / uint64_t(block);
        const uint64_t launched = (lanes.blocks > 0 ? std::min<uint64_t>(grid, uint64_t(lanes.blocks)) : grid) * uint64_t(block);
        const uint64_t in_flight = std::min<uint64_t>(launched, uint64_t(cus) * uint64_t(pt::path_lanes_per_cu));
        if(first <= 0 || samples <= first || uint64_t(n) <= in_flight || !pt::wide_traversal(view, frame)) {
            pt::launch_drain(view, frame, paths, ctl.get(), 0, 0, 0, n, block, lanes, s, batch);
            return false;
        }
        group = std::max(1, group);
        const uint32_t groups = (n + uint32_t(group) - 1) / uint32_t(group);
        if(costs.size() < n) costs.alloc(n);
        if(group_ids.size() < groups) {
            group_ids.alloc(groups); sorted_ids.alloc(groups); group_keys.alloc(groups); sorted_keys.alloc(groups);
        }
        size_t temp = 0;
        HIP_CHECK(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, temp, group_keys.get(), sorted_keys.get(),
                                                                group_ids.get(), sorted_ids.get(), int(groups), 0, 8, s));
        if(sort_temp.size() < temp) sort_temp.alloc(temp);
        pt::path_pass pass;
        pass.mode = pt::PATHS_FIRST;
        pass.first_samples = first;
        pass.costs = costs.get();
        pt::launch_path_lanes(view, frame, paths, ctl.get(), 0, queue0.get(), n, block, lanes, pass, s, batch);
        pt::launch_zero_work(ctl.get(), s);
        pt::launch_group_keys(costs.get(), n, group, group_keys.get(), group_ids.get(), s);
        HIP_CHECK(hipcub::DeviceRadixSort::SortPairsDescending(sort_temp.get(), temp, group_keys.get(), sorted_keys.get(),
                                                                group_ids.get(), sorted_ids.get(), int(groups), 0, 8, s));
        pt::launch_group_gather(queue0.get(), sorted_ids.get(), n, group, queue1.get(), s);
        pt::launch_path_lanes(view, frame, paths, ctl.get(), 0, queue1.get(), n, block, lanes, pt::path_pass{pt::PATHS_REST},
                              s, batch);
        return true;
    }
    void read_ctl_async() { read_ctl_async(stream); }
    void read_ctl_async(hipStream_t s) {
        HIP_CHECK(hipMemcpyAsync(ctl_host, ctl.get(), sizeof(pt::loop_ctl), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipEventRecord(checked, s));
    }
};

constexpr int partition_chunk = 4096;  // pixels per chunk dealt round-robin to partitions
constexpr int64_t max_saved_queries = 262144;  // queries one trace launch can hand on to the next
constexpr int32_t all_paths = 1 << 30;         // drain threshold: every path on path lanes
constexpr uint32_t log_capacity = 1u << 16;

// ---------------- allocation-free rendering ----------------
// Scratch of pt_renderer_add_samples_noalloc, for callers that cannot allocate
// device memory while they render: one array per GPU, held by the process
// from the library's load until it ends, shared by every such call on that
// GPU (calls that overlap on one GPU race).  Layout: the launch's control
// block (work cursor and counters) in the first scratch_ctl_bytes, then the
// 4-wide traversal stack entries that do not fit in LDS.  Kernels take its
// address as found by hipGetSymbolAddress (static_scratch), never the symbol.
constexpr size_t static_scratch_bytes = size_t(64) << 20;
constexpr size_t scratch_ctl_bytes = 256;
static_assert(sizeof(pt::loop_ctl) <= scratch_ctl_bytes, "control block fits its slot");
static __device__ __align__(256) unsigned char scratch[static_scratch_bytes];

namespace {
std::mutex scratch_lock;
std::map<int, unsigned char *> scratch_address;  // per device

// The device address of this GPU's scratch array: resolved once per device,
// at renderer creation (set-up) or else at the first use on the device.
unsigned char *static_scratch() {
    int device = 0;
    HIP_CHECK(hipGetDevice(&device));
    std::lock_guard<std::mutex> guard(scratch_lock);
    auto it = scratch_address.find(device);
    if(it != scratch_address.end()) return it->second;
    void *ptr = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if(!ptr) fail(PT_EDEVICE, "static scratch array has no device address");
    scratch_address[device] = static_cast<unsigned char *>(ptr);
    return static_cast<unsigned char *>(ptr);
}
} // namespace

struct pt_renderer {
    std::unique_ptr<device_scene> scene;
    pt::frame_params frame{};
    uint64_t seed = 0;
    int32_t samples = 0;
    uint64_t rays = 0;
    float milliseconds = 0;
    dim3 block = dim3(16, 16);
    hipStream_t stream = nullptr;        // reset, debug views, packing
    hipEvent_t start = nullptr, stop = nullptr;
    size_t pixels = 0;
    dev_buf<float> sum;
    dev_buf<uint64_t> rng;
    dev_buf<uint8_t> rgba;
    mutable dev_buf<float> image_buf;    // image(): the scaled sums, allocated at the first read
    dev_buf<pt::loop_ctl> ctl;           // debug views' counters
    pt::loop_ctl *ctl_host = nullptr;
    // wavefront state (shaded mode only), indexed by pixel and shared by the partitions
    dev_buf<float4> ray_o, ray_d, thr, rad, sh_o, sh_d, sh_c;
    dev_buf<int2> ctr;
    dev_buf<int4> hit;
    std::vector<std::unique_ptr<partition>> parts;
    dev_buf<float> aov;                  // first-hit albedo and normal sums (6 per pixel), when enabled
    pt_tuning tuning{128, 256, 16, 8, 0, 0, 0, all_paths, 48, 1, 16, 32, 12, 4, 4, 8};
    int cus = 1;
    pt_launch_stats last{};
    uint64_t rays_noalloc = 0;           // rays of add_samples_noalloc calls since reset (not in the device counters)

    ~pt_renderer() {
        parts.clear();
        if(stream) (void)hipStreamSynchronize(stream);
        if(start) (void)hipEventDestroy(start);
        if(stop) (void)hipEventDestroy(stop);
        if(ctl_host) (void)hipHostFree(ctl_host);
        if(stream) (void)hipStreamDestroy(stream);
    }

    void reset(const pt_flat_camera& camera) {
        validate_camera(camera);
        pt::launch_init_pixels(sum.get(), rng.get(), int64_t(pixels), seed, ctl.get(), stream);
        for(auto& part : parts) pt::launch_clear_ctl(part->ctl.get(), stream);
        if(aov.size()) HIP_CHECK(hipMemsetAsync(aov.get(), 0, aov.size() * sizeof(float), stream));
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(stream));
        frame.camera = camera; samples = 0; milliseconds = 0; rays = 0; rays_noalloc = 0;
    }

    // Largest block of at most tuning.trace_block threads whose traversal
    // stacks fit the 64 KiB of LDS a workgroup may hold; the wide traversal is
    // switched off if even 64 threads do not fit.
    pt::path_lanes path_lanes() const {
        return pt::path_lanes{tuning.path_shade_batch, tuning.path_blocks_per_cu > 0 ? cus * tuning.path_blocks_per_cu : 0,
                              tuning.path_trace_steps};
    }
    int trace_block() {
        frame.exact_traversal = tuning.exact_traversal;
        int block = tuning.trace_block;
        while(block > 64 && pt::trace_lds_bytes(scene->view, frame, block) > 65536) block -= 64;
        if(pt::trace_lds_bytes(scene->view, frame, block) > 65536) frame.exact_traversal = 1;
        return block;
    }

    void read_ctl() {
        HIP_CHECK(hipMemcpyAsync(ctl_host, ctl.get(), sizeof(pt::loop_ctl), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
    }

    // (Re)build the partitions when their number changes.
    void ensure_partitions(int count) {
        count = std::max(1, std::min(count, 8));
        if(int(parts.size()) == count) return;
        HIP_CHECK(hipDeviceSynchronize());
        std::vector<uint64_t> rays_before;
        for(auto& part : parts) {
            HIP_CHECK(hipMemcpy(part->ctl_host, part->ctl.get(), sizeof(pt::loop_ctl), hipMemcpyDeviceToHost));
            rays_before.push_back(part->ctl_host->rays);
        }
        uint64_t carried = 0;
        for(uint64_t r : rays_before) carried += r;
        parts.clear();
        const int64_t chunks = (int64_t(pixels) + partition_chunk - 1) / partition_chunk;
        for(int k = 0; k < count; k++) {
            auto part = std::make_unique<partition>();
            part->index = k;
            for(int64_t c = k; c < chunks; c += count)
                part->count += std::min<int64_t>(partition_chunk, int64_t(pixels) - c * partition_chunk);
            HIP_CHECK(hipStreamCreateWithFlags(&part->stream, hipStreamNonBlocking));
            for(auto *e : {&part->checked, &part->drain0, &part->drain1, &part->finished}) HIP_CHECK(hipEventCreate(e));
            part->ctl.alloc(1);
            HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&part->ctl_host), sizeof(pt::loop_ctl), hipHostMallocDefault));
            const size_t n = size_t(std::max<int64_t>(1, part->count));
            part->queue0.alloc(n); part->queue1.alloc(n); part->shadow_queue.alloc(n); part->finish_queue.alloc(n);
            part->log.alloc(2 * size_t(log_capacity));
            pt::loop_ctl init{};
            init.log = part->log.get();
            init.log_capacity = log_capacity;
            if(k == 0) init.rays = carried;
            HIP_CHECK(hipMemcpy(part->ctl.get(), &init, sizeof init, hipMemcpyHostToDevice));
            *part->ctl_host = init;
            auto& p = part->paths;
            p.ray_o = ray_o.get(); p.ray_d = ray_d.get(); p.thr = thr.get(); p.rad = rad.get();
            p.ctr = ctr.get(); p.hit = hit.get();
            p.sh_o = sh_o.get(); p.sh_d = sh_d.get(); p.sh_c = sh_c.get();
            p.rng = rng.get(); p.sum = sum.get(); p.aov = aov.get();
            p.queue[0] = part->queue0.get(); p.queue[1] = part->queue1.get();
            p.shadow_queue = part->shadow_queue.get(); p.finish_queue = part->finish_queue.get();
            parts.push_back(std::move(part));
        }
    }

    void add_samples(int32_t count) {
        if(count <= 0 || count > INT32_MAX - samples) fail(PT_EINVAL, "invalid accumulated sample count");
        pt_launch_stats st{};
        if(frame.debug != PT_DEBUG_SHADED) {
            HIP_CHECK(hipEventRecord(start, stream));
            pt::launch_debug(scene->view, frame, sum.get(), rng.get(), count, &ctl.get()->failure, block, stream);
            HIP_CHECK(hipGetLastError());
            pt::launch_add_rays(ctl.get(), (unsigned long long)pixels * (unsigned long long)count, stream);
            HIP_CHECK(hipEventRecord(stop, stream));
            read_ctl();
            st.iterations = 1;
            st.extend_rays = uint64_t(pixels) * uint64_t(count);
            HIP_CHECK(hipEventSynchronize(stop));
            if(ctl_host->failure) fail(PT_EOVERFLOW, "HIP traversal stack overflow");
            rays = ctl_host->rays + rays_noalloc;
        } else {
            shaded_samples(count, st);
        }
        float elapsed = 0;
        HIP_CHECK(hipEventElapsedTime(&elapsed, start, stop));
        milliseconds += elapsed;
        samples += count;
        last = st;
    }

    // add_samples with every scratch buffer in the static array at `base`
    // (static_scratch()): no device or host allocation.  One path-lane launch
    // over all pixels, each lane taking pixel indices and running `count`
    // samples of each from its PCG state and sum (PATHS_FRESH), with as many
    // resident blocks as the array's stack region holds; or the diagnostic
    // kernel.  Results equal add_samples' (each pixel's samples run in order
    // on its own PCG stream whatever the schedule).
    void add_samples_noalloc(int32_t count, unsigned char *base) {
        if(count <= 0 || count > INT32_MAX - samples) fail(PT_EINVAL, "invalid accumulated sample count");
        pt::loop_ctl *c = reinterpret_cast<pt::loop_ctl *>(base);
        int2 *spill = reinterpret_cast<int2 *>(base + scratch_ctl_bytes);
        pt_launch_stats st{};
        if(frame.debug != PT_DEBUG_SHADED) {
            HIP_CHECK(hipEventRecord(start, stream));
            pt::launch_fresh_ctl(c, 0, stream);
            pt::launch_debug(scene->view, frame, sum.get(), rng.get(), count, &c->failure, block, stream);
            HIP_CHECK(hipEventRecord(stop, stream));
        } else {
            frame.exact_traversal = tuning.exact_traversal;
            if(!pt::wide_traversal(scene->view, frame))
                fail(PT_EINVAL, "allocation-free rendering of shaded frames needs the 4-wide traversal "
                                "(not brute-force or exact-traversal renderers)");
            // Resident blocks: as many as the stack region holds (at the deepest
            // hierarchy validation allows, 64 levels, 181 entries per lane, 181
            // blocks still fit), at most the renderer's path-lane tuning.
            const size_t entries = size_t(pt::wide_spill_entries(scene->view.wide_stack_capacity));
            const size_t per_block = size_t(pt::path_lane_block) * entries * sizeof(int2);
            const size_t fit = per_block ? (static_scratch_bytes - scratch_ctl_bytes) / per_block : size_t(INT32_MAX);
            if(fit == 0) fail(PT_EINVAL, "traversal stack exceeds the static scratch array");
            size_t blocks = fit;
            if(tuning.path_blocks_per_cu > 0) blocks = std::min(blocks, size_t(cus) * size_t(tuning.path_blocks_per_cu));
            blocks = std::min(blocks, size_t(INT32_MAX));
            const pt::path_lanes lanes{tuning.path_shade_batch, int(blocks), tuning.path_trace_steps};
            pt::paths_dev p{};
            p.sum = sum.get(); p.rng = rng.get(); p.aov = aov.size() ? aov.get() : nullptr; p.spill = spill;
            pt::path_pass pass;
            pass.mode = pt::PATHS_FRESH;
            pass.first_samples = count;
            HIP_CHECK(hipEventRecord(start, stream));
            pt::launch_fresh_ctl(c, uint32_t(pixels), stream);
            pt::launch_path_lanes(scene->view, frame, p, c, 0, nullptr, uint32_t(pixels), pt::path_lane_block, lanes,
                                  pass, stream);
            HIP_CHECK(hipEventRecord(stop, stream));
        }
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(ctl_host, c, sizeof(pt::loop_ctl), hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        if(ctl_host->failure) fail(PT_EOVERFLOW, "HIP traversal stack overflow");
        uint64_t call_rays;
        if(frame.debug != PT_DEBUG_SHADED) {
            call_rays = uint64_t(pixels) * uint64_t(count);
            st.iterations = 1;
            st.extend_rays = call_rays;
        } else {
            call_rays = ctl_host->rays;
            st.extend_rays = ctl_host->ext_rays;
            st.shadow_rays = ctl_host->shadow_rays;
            st.retraced_queries = ctl_host->retraces;
            st.drained_paths = pixels;
        }
        float elapsed = 0;
        HIP_CHECK(hipEventElapsedTime(&elapsed, start, stop));
        milliseconds += elapsed;
        samples += count;
        rays += call_rays;
        rays_noalloc += call_rays;
        last = st;
    }

    void shaded_samples(int32_t count, pt_launch_stats& st) {
        ensure_partitions(tuning.partitions);
        const int nparts = int(parts.size());
        const int trace_threads = trace_block();
        const bool use_lanes = tuning.lane_refill > 0 && !frame.brute && !frame.exact_traversal &&
                               scene->view.wroot_ref != pt::empty_root;
        const int trace_grid = cus * std::max(1, tuning.blocks_per_cu);
        const int drain_block = 256;
        const int64_t drain_threshold = tuning.drain_threshold > 0 ? std::max<int64_t>(1, tuning.drain_threshold / nparts) : 0;
        int64_t largest = 0;
        for(auto& part : parts) largest = std::max(largest, part->count);
        const int64_t drain_paths = std::min(drain_threshold, largest);  // most paths one drain launch can take
        size_t drain_threads = drain_paths > 0 ? size_t((drain_paths + drain_block - 1) / drain_block * drain_block) : 0;
        if(path_lanes().blocks > 0) drain_threads = std::min(drain_threads, size_t(path_lanes().blocks) * size_t(drain_block));
        const size_t spill_need = std::max(size_t(trace_grid) * size_t(trace_threads), drain_threads) *
                                  size_t(pt::wide_spill_entries(scene->view.wide_stack_capacity));
        const int interval = std::max(1, tuning.check_interval);
        HIP_CHECK(hipEventRecord(start, stream));
        for(auto& part : parts) {
            if(part->spill.size() < spill_need) { part->spill.alloc(spill_need); part->paths.spill = part->spill.get(); }
            if(use_lanes) part->ensure_saves(std::min<int64_t>(max_saved_queries, 2 * std::max<int64_t>(1, part->count)),
                                             scene->view.wide_stack_capacity);
            if(tuning.time_kernels && part->events.size() < size_t(5 * interval)) {
                for(auto e : part->events) (void)hipEventDestroy(e);
                part->events.assign(size_t(5 * interval), nullptr);
                for(auto& e : part->events) HIP_CHECK(hipEventCreate(&e));
            }
            HIP_CHECK(hipStreamWaitEvent(part->stream, start, 0));
            part->parity = 0;
            part->done = part->count == 0;
            part->drained = false;
            pt::launch_start(scene->view, frame, part->paths, part->ctl.get(), count, part->index, nparts,
                             partition_chunk, part->count, part->stream);
            if(!part->done && drain_threshold && part->count <= drain_threshold) {
                // Few enough paths for the tail kernel from the start: no wavefront bounces.
                if(tuning.time_kernels) HIP_CHECK(hipEventRecord(part->drain0, part->stream));
                if(part->paths_from_start(scene->view, frame, count, tuning.path_first_samples, tuning.path_sort_group,
                                          path_lanes(), drain_block, cus, part->stream))
                    st.split_launches++;
                if(tuning.time_kernels) HIP_CHECK(hipEventRecord(part->drain1, part->stream));
                st.drained_paths += uint64_t(part->count);
                part->read_ctl_async();
                part->done = part->drained = true;
            }
        }
        HIP_CHECK(hipGetLastError());
        auto launch_iterations = [&](partition& part) {
            pt::loop_ctl *c = part.ctl.get();
            for(int k = 0; k < interval; k++) {
                const int par = part.parity;
                const int shade_grid = int(std::max<size_t>(1, std::min<size_t>(
                    (size_t(part.count) + size_t(tuning.shade_block) - 1) / size_t(tuning.shade_block), size_t(cus) * 16)));
                if(tuning.time_kernels) HIP_CHECK(hipEventRecord(part.events[size_t(5 * k)], part.stream));
                if(use_lanes)
                    pt::launch_trace_lanes(scene->view, frame, part.paths, c, par, trace_grid, trace_threads,
                                           tuning.lane_refill, tuning.handoff_steps, tuning.count_tests != 0, part.stream);
                else
                    pt::launch_trace(scene->view, frame, part.paths, c, par, trace_grid, trace_threads,
                                     tuning.count_tests != 0, part.stream);
                if(tuning.time_kernels) HIP_CHECK(hipEventRecord(part.events[size_t(5 * k + 1)], part.stream));
                pt::launch_bookkeep(c, par, part.stream);
                if(tuning.time_kernels) HIP_CHECK(hipEventRecord(part.events[size_t(5 * k + 2)], part.stream));
                pt::launch_shade(scene->view, frame, part.paths, c, par, shade_grid, tuning.shade_block, part.stream);
                if(tuning.time_kernels) HIP_CHECK(hipEventRecord(part.events[size_t(5 * k + 3)], part.stream));
                pt::launch_regen(frame, part.paths, c, par, shade_grid, tuning.shade_block, part.stream);
                if(tuning.time_kernels) HIP_CHECK(hipEventRecord(part.events[size_t(5 * k + 4)], part.stream));
                part.parity ^= 1;
                st.trace_launches++;
            }
            HIP_CHECK(hipGetLastError());
            part.read_ctl_async();
        };
        for(;;) {
            bool any = false;
            for(auto& part : parts) if(!part->done) { launch_iterations(*part); any = true; }
            if(!any) break;
            for(auto& part : parts) {
                if(part->done) continue;
                HIP_CHECK(hipEventSynchronize(part->checked));
                if(tuning.time_kernels) {
                    for(int k = 0; k < interval; k++) {
                        float a = 0, b = 0, c = 0;
                        HIP_CHECK(hipEventElapsedTime(&a, part->events[size_t(5 * k)], part->events[size_t(5 * k + 1)]));
                        HIP_CHECK(hipEventElapsedTime(&b, part->events[size_t(5 * k + 2)], part->events[size_t(5 * k + 3)]));
                        HIP_CHECK(hipEventElapsedTime(&c, part->events[size_t(5 * k + 3)], part->events[size_t(5 * k + 4)]));
                        st.trace_ms += a; st.shade_ms += b; st.regen_ms += c;
                    }
                }
                const pt::loop_ctl& h = *part->ctl_host;
                const uint32_t live = h.n_ext[part->parity];
                if(live == 0 || h.failure) { part->done = true; continue; }
                if(drain_threshold && int64_t(live) <= drain_threshold) {
                    if(tuning.time_kernels) HIP_CHECK(hipEventRecord(part->drain0, part->stream));
                    pt::launch_drain(scene->view, frame, part->paths, part->ctl.get(), part->parity, h.n_shadow,
                                     h.n_saved[part->parity], live, drain_block, path_lanes(), part->stream);
                    HIP_CHECK(hipGetLastError());
                    if(tuning.time_kernels) HIP_CHECK(hipEventRecord(part->drain1, part->stream));
                    st.drained_paths += live;
                    part->read_ctl_async();
                    part->done = part->drained = true;
                }
            }
        }
        for(auto& part : parts) {
            HIP_CHECK(hipEventRecord(part->finished, part->stream));
            HIP_CHECK(hipStreamWaitEvent(stream, part->finished, 0));
        }
        HIP_CHECK(hipEventRecord(stop, stream));
        HIP_CHECK(hipEventSynchronize(stop));
        uint64_t total_rays = 0;
        int failure = 0;
        for(auto& part : parts) {
            HIP_CHECK(hipEventSynchronize(part->checked));
            const pt::loop_ctl& h = *part->ctl_host;
            st.iterations = std::max(st.iterations, int32_t(h.iterations));
            st.extend_rays += h.ext_rays;
            st.shadow_rays += h.shadow_rays;
            st.box_tests += h.box_tests;
            st.triangle_tests += h.triangle_tests;
            st.sphere_tests += h.sphere_tests;
            st.retraced_queries += h.retraces;
            st.hit_queries += h.hits;
            st.retrace_box_tests += h.retrace_branches;
            st.retrace_primitive_tests += h.retrace_leaves;
            total_rays += h.rays;
            failure |= h.failure;
            if(tuning.time_kernels && part->drained) {
                float d = 0;
                HIP_CHECK(hipEventElapsedTime(&d, part->drain0, part->drain1));
                st.drain_ms += d;
            }
        }
        if(failure) fail(PT_EOVERFLOW, "HIP traversal stack overflow");
        rays = total_rays + rays_noalloc;
    }
};

bool pt::renderer_frame(const pt_renderer *r, pt::frame_view *out) {
    if(!r || !out) return false;
    out->sum = r->sum.get();
    out->aov = r->aov.size() ? r->aov.get() : nullptr;
    out->width = r->frame.width;
    out->height = r->frame.height;
    out->samples = r->samples;
    out->stream = r->stream;
    return true;
}

namespace {

void allocate_state(pt_renderer& r, bool wavefront) {
    size_t n = r.pixels;
    r.sum.alloc(3 * n);
    r.rng.alloc(n);
    r.rgba.alloc(4 * n);
    r.ctl.alloc(1);
    HIP_CHECK(hipMemset(r.ctl.get(), 0, sizeof(pt::loop_ctl)));
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&r.ctl_host), sizeof(pt::loop_ctl), hipHostMallocDefault));
    std::memset(r.ctl_host, 0, sizeof(pt::loop_ctl));
    if(!wavefront) return;
    r.ray_o.alloc(n); r.ray_d.alloc(n); r.thr.alloc(n); r.rad.alloc(n);
    r.sh_o.alloc(n); r.sh_d.alloc(n); r.sh_c.alloc(n);
    r.ctr.alloc(n); r.hit.alloc(n);
    r.ensure_partitions(r.tuning.partitions);
}

void check_grid(const pt_renderer& r, unsigned bx, unsigned by) {
    int device = 0;
    HIP_CHECK(hipGetDevice(&device));
    hipDeviceProp_t props{};
    HIP_CHECK(hipGetDeviceProperties(&props, device));
    unsigned gx = (unsigned(r.frame.width) + bx - 1) / bx, gy = (unsigned(r.frame.height) + by - 1) / by;
    if(gx > unsigned(props.maxGridSize[0]) || gy > unsigned(props.maxGridSize[1]))
        fail(PT_EINVAL, "image exceeds HIP grid limits");
}

void query(const pt_scene_desc *scene, const pt_ray *rays, size_t count, int32_t brute, pt_trace_result *results,
           int32_t *boxes, int32_t *prims) {
    if(!scene) fail(PT_EINVAL, "null scene");
    validate_scene(*scene);
    require_device();
    if(count == 0) return;
    if(count > size_t(INT32_MAX)) fail(PT_EINVAL, "too many validation rays");
    if(!rays || !results) fail(PT_EINVAL, "null ray or result buffer");
    device_scene uploaded(*scene, false);
    dev_buf<pt_ray> input; input.upload(rays, count);
    dev_buf<pt_trace_result> output(count);
    dev_buf<int32_t> b(boxes ? count : 0), p(prims ? count : 0);
    pt::launch_query(uploaded.view, input.get(), output.get(), b.get(), p.get(), int32_t(count), brute != 0, nullptr);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipDeviceSynchronize());
    HIP_CHECK(hipMemcpy(results, output.get(), count * sizeof(pt_trace_result), hipMemcpyDeviceToHost));
    if(boxes) HIP_CHECK(hipMemcpy(boxes, b.get(), count * sizeof(int32_t), hipMemcpyDeviceToHost));
    if(prims) HIP_CHECK(hipMemcpy(prims, p.get(), count * sizeof(int32_t), hipMemcpyDeviceToHost));
}

} // namespace

// ---------------- batched rendering (no counterpart in the reference) ----------------
// Many views of many scenes rendered by one wavefront bounce loop: the
// scenes are merged into one scene set (arrays concatenated, indices
// rebased), each keeping its own 4-wide hierarchy and light range; the loop
// is captured once into a HIP graph and replayed until the tail kernel takes
// over.  Every view's image equals the one a single renderer makes of it.
namespace {

struct merged_scenes {
    std::vector<pt_flat_triangle> triangles;
    std::vector<pt_flat_sphere> spheres;
    std::vector<pt_flat_node> nodes;
    std::vector<pt_flat_material> materials;
    std::vector<pt_flat_texture> textures;
    std::vector<uint8_t> pixels;
    std::vector<int64_t> tri_offset, sph_offset, node_offset;
    pt_scene_desc desc() const {
        pt_scene_desc d{};
        d.triangles = triangles.data(); d.triangle_count = int64_t(triangles.size());
        d.spheres = spheres.data(); d.sphere_count = int64_t(spheres.size());
        d.nodes = nodes.data(); d.node_count = int64_t(nodes.size());
        d.materials = materials.data(); d.material_count = int64_t(materials.size());
        d.textures = textures.data(); d.texture_count = int64_t(textures.size());
        d.pixels = pixels.data(); d.pixel_bytes = int64_t(pixels.size());
        d.root = -1; d.max_depth = 1;
        return d;
    }
};

merged_scenes merge_scenes(const pt_scene_desc *scenes, int32_t count) {
    merged_scenes m;
    for(int32_t k = 0; k < count; k++) {
        const pt_scene_desc& s = scenes[k];
        const int64_t to = int64_t(m.triangles.size()), so = int64_t(m.spheres.size()), no = int64_t(m.nodes.size());
        const int64_t mo = int64_t(m.materials.size()), xo = int64_t(m.textures.size()), po = int64_t(m.pixels.size());
        if(to + s.triangle_count > INT32_MAX / 2 || so + s.sphere_count > INT32_MAX / 2 || no + s.node_count > INT32_MAX ||
           po + s.pixel_bytes > INT32_MAX)
            fail(PT_EINVAL, "batch exceeds 32-bit indices");
        m.tri_offset.push_back(to);
        m.sph_offset.push_back(so);
        m.node_offset.push_back(no);
        for(int64_t i = 0; i < s.triangle_count; i++) {
            pt_flat_triangle t = s.triangles[i];
            t.material_id += int32_t(mo);
            m.triangles.push_back(t);
        }
        for(int64_t i = 0; i < s.sphere_count; i++) {
            pt_flat_sphere p = s.spheres[i];
            p.material_id += int32_t(mo);
            m.spheres.push_back(p);
        }
        for(int64_t i = 0; i < s.node_count; i++) {
            pt_flat_node n = s.nodes[i];
            if(n.type == PT_BRANCH) { n.left += int32_t(no); n.right += int32_t(no); }
            else n.primitive += int32_t(n.type == PT_SPHERE ? so : to);
            m.nodes.push_back(n);
        }
        for(int64_t i = 0; i < s.material_count; i++) {
            pt_flat_material x = s.materials[i];
            if(x.texture_id >= 0) x.texture_id += int32_t(xo);
            m.materials.push_back(x);
        }
        for(int64_t i = 0; i < s.texture_count; i++) {
            pt_flat_texture x = s.textures[i];
            if(x.image_id >= 0) x.image_id += int32_t(xo);
            if(x.type == PT_IMAGE) x.byte_offset += int32_t(po);
            m.textures.push_back(x);
        }
        m.pixels.insert(m.pixels.end(), s.pixels, s.pixels + s.pixel_bytes);
    }
    return m;
}

// A wide-hierarchy reference moved into the concatenated layout.
int32_t rebase_ref(int32_t ref, int32_t node_offset, int32_t leaf_offset) {
    if(ref >= 0) return ref + node_offset;
    if(ref == pt::wide_empty) return ref;
    const int32_t v = ~ref;
    return ~((((v >> 3) + leaf_offset) << 3) | (v & 7));
}

} // namespace

struct pt_batch {
    std::unique_ptr<device_scene> scene;
    dev_buf<float4> wnodes, leaves;
    dev_buf<pt::light_rec> lights;
    dev_buf<pt::view_rec> views;
    dev_buf<uint16_t> pixel_view;
    std::vector<pt::view_rec> host_views;
    pt::batch_dev bd{};
    pt::frame_params frame{};
    size_t pixels = 0;
    dev_buf<float> sum;
    dev_buf<uint64_t> rng;
    dev_buf<float4> ray_o, ray_d, thr, rad, sh_o, sh_d, sh_c;
    dev_buf<int2> ctr;
    dev_buf<int4> hit;
    partition part;
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, stop = nullptr;
    // One path-lane pass in queue order (path_first_samples -1): the views'
    // pixels stay together, and so do the hierarchies they read, which keeps
    // the L2 useful for a batch of large scenes (64 soups of 10k-2M
    // triangles: 1520 against 1430 Mrays/s with two cost-ordered passes).
    // Shading batches of 24 lanes: the batch's waves wait less for shading
    // partners, which its memory-bound traversal prefers.
    pt_tuning tuning{128, 256, 16, 8, 0, 0, 0, all_paths, 48, 1, 16, 24, 12, -1, 4, 8};
    int cus = 1;
    int32_t samples = 0;
    uint64_t rays = 0;
    float milliseconds = 0;
    hipGraphExec_t graph = nullptr;
    int graph_iterations = 0, graph_trace_block = 0, graph_blocks = 0, graph_shade_block = 0, graph_refill = 0, graph_tail = 0;
    const int2 *graph_spill = nullptr;     // buffers the captured kernels point to
    const int4 *graph_saved = nullptr;
    pt_launch_stats last{};

    ~pt_batch() {
        if(stream) (void)hipStreamSynchronize(stream);
        if(graph) (void)hipGraphExecDestroy(graph);
        if(start) (void)hipEventDestroy(start);
        if(stop) (void)hipEventDestroy(stop);
        if(stream) (void)hipStreamDestroy(stream);
    }

    pt::path_lanes path_lanes() const {
        return pt::path_lanes{tuning.path_shade_batch, tuning.path_blocks_per_cu > 0 ? cus * tuning.path_blocks_per_cu : 0,
                              tuning.path_trace_steps};
    }
    int trace_block() {
        int block = tuning.trace_block;
        while(block > 64 && pt::trace_lds_bytes(scene->view, frame, block) > 65536) block -= 64;
        if(pt::trace_lds_bytes(scene->view, frame, block) > 65536) fail(PT_EINVAL, "batch hierarchy too deep for the LDS stack");
        return block;
    }

    void reset() {
        HIP_CHECK(hipStreamSynchronize(stream));
        pt::launch_batch_init(bd, sum.get(), rng.get(), int64_t(pixels), part.ctl.get(), stream);
        pt::launch_clear_ctl(part.ctl.get(), stream);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(stream));
        samples = 0; rays = 0; milliseconds = 0;
    }

    // The bounce loop as one graph: trace, bookkeep, shade, regenerate,
    // `iterations` times (even, so each replay ends on parity 0).
    void capture(int iterations, int block, int blocks, hipStream_t s) {
        // The graph holds its kernels' arguments by value: any buffer they
        // point to that was reallocated since (the traversal spill region
        // grows with the drain threshold and launch shapes) needs a new graph.
        if(graph && graph_iterations == iterations && graph_trace_block == block && graph_blocks == blocks &&
           graph_shade_block == tuning.shade_block && graph_refill == tuning.lane_refill &&
           graph_tail == tuning.handoff_steps && graph_spill == part.paths.spill &&
           graph_saved == part.paths.saved[0].head)
            return;
        if(graph) { HIP_CHECK(hipGraphExecDestroy(graph)); graph = nullptr; }
        const int shade_grid = int(std::max<size_t>(1, std::min<size_t>(
            (pixels + size_t(tuning.shade_block) - 1) / size_t(tuning.shade_block), size_t(cus) * 16)));
        hipGraph_t g = nullptr;
        HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for(int k = 0; k < iterations; k++) {
            const int par = k & 1;
            pt::launch_trace_lanes(scene->view, frame, part.paths, part.ctl.get(), par, blocks, block, tuning.lane_refill,
                                   tuning.handoff_steps, false, s, &bd);
            pt::launch_bookkeep(part.ctl.get(), par, s);
            pt::launch_shade(scene->view, frame, part.paths, part.ctl.get(), par, shade_grid, tuning.shade_block, s, &bd);
            pt::launch_regen(frame, part.paths, part.ctl.get(), par, shade_grid, tuning.shade_block, s, &bd);
        }
        HIP_CHECK(hipStreamEndCapture(s, &g));
        hipError_t e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIP_CHECK(e);
        graph_iterations = iterations; graph_trace_block = block; graph_blocks = blocks;
        graph_shade_block = tuning.shade_block; graph_refill = tuning.lane_refill; graph_tail = tuning.handoff_steps;
        graph_spill = part.paths.spill; graph_saved = part.paths.saved[0].head;
    }

    void add_samples(int32_t count) {
        if(count <= 0 || count > INT32_MAX - samples) fail(PT_EINVAL, "invalid accumulated sample count");
        pt_launch_stats st{};
        const int block = trace_block();
        const int blocks = cus * std::max(1, tuning.blocks_per_cu);
        const int drain_block = 256;
        const int64_t threshold = tuning.drain_threshold > 0 ? tuning.drain_threshold
                                : tuning.drain_threshold < 0 ? 0 : std::max<int64_t>(4096, int64_t(pixels) / 8);
        const int64_t drain_paths = std::min<int64_t>(threshold, int64_t(pixels));  // most paths one drain launch can take
        size_t drain_threads = drain_paths > 0 ? size_t((drain_paths + drain_block - 1) / drain_block * drain_block) : 0;
        if(path_lanes().blocks > 0) drain_threads = std::min(drain_threads, size_t(path_lanes().blocks) * size_t(drain_block));
        const size_t spill_need = std::max(size_t(blocks) * size_t(block), drain_threads) *
                                  size_t(pt::wide_spill_entries(scene->view.wide_stack_capacity));
        if(part.spill.size() < spill_need) { part.spill.alloc(spill_need); part.paths.spill = part.spill.get(); }
        part.ensure_saves(std::min<int64_t>(max_saved_queries, 2 * int64_t(std::max<size_t>(1, pixels))),
                          scene->view.wide_stack_capacity);
        const int iterations = 2 * std::max(1, tuning.check_interval / 2);
        capture(iterations, block, blocks, stream);
        HIP_CHECK(hipEventRecord(start, stream));
        pt::launch_start(scene->view, frame, part.paths, part.ctl.get(), count, 0, 1, partition_chunk, int64_t(pixels),
                         stream, &bd);
        HIP_CHECK(hipGetLastError());
        bool drained = false;
        if(threshold && int64_t(pixels) <= threshold) {
            // Few enough paths for the tail kernel from the start: no wavefront bounces.
            if(part.paths_from_start(scene->view, frame, count, tuning.path_first_samples, tuning.path_sort_group,
                                     path_lanes(), drain_block, cus, stream, &bd))
                st.split_launches++;
            HIP_CHECK(hipGetLastError());
            st.drained_paths = pixels;
            drained = true;
        }
        for(; !drained;) {
            HIP_CHECK(hipGraphLaunch(graph, stream));
            part.read_ctl_async(stream);
            st.trace_launches += iterations;
            HIP_CHECK(hipEventSynchronize(part.checked));
            const pt::loop_ctl& h = *part.ctl_host;
            if(h.failure) break;
            const uint32_t live = h.n_ext[0];
            if(live == 0) break;
            if(threshold && int64_t(live) <= threshold) {
                pt::launch_drain(scene->view, frame, part.paths, part.ctl.get(), 0, h.n_shadow, h.n_saved[0], live,
                                 drain_block, path_lanes(), stream, &bd);
                HIP_CHECK(hipGetLastError());
                st.drained_paths = live;
                drained = true;
                break;
            }
        }
        HIP_CHECK(hipEventRecord(stop, stream));
        part.read_ctl_async(stream);
        HIP_CHECK(hipEventSynchronize(part.checked));
        (void)drained;
        const pt::loop_ctl& h = *part.ctl_host;
        if(h.failure) fail(PT_EOVERFLOW, "HIP traversal stack overflow");
        st.iterations = int32_t(h.iterations);
        st.extend_rays = h.ext_rays;
        st.shadow_rays = h.shadow_rays;
        st.retraced_queries = h.retraces;
        float elapsed = 0;
        HIP_CHECK(hipEventElapsedTime(&elapsed, start, stop));
        milliseconds += elapsed;
        rays = h.rays;
        samples += count;
        last = st;
    }
};

namespace {

pt_batch *create_batch(const pt_scene_desc *scenes, int32_t scene_count, const pt_batch_view *views, int32_t view_count,
                       int32_t direct_light, int32_t max_path_depth) {
    if(!scenes || !views || scene_count <= 0 || view_count <= 0) fail(PT_EINVAL, "empty batch");
    if(view_count > 65535) fail(PT_EINVAL, "too many views in one batch");
    if(max_path_depth < 0) fail(PT_EINVAL, "invalid path depth");
    for(int32_t k = 0; k < scene_count; k++) {
        validate_scene(scenes[k]);
        if(scenes[k].root < 0) fail(PT_EINVAL, "batch scenes must not be empty");
    }
    size_t total = 0;
    for(int32_t v = 0; v < view_count; v++) {
        if(views[v].scene < 0 || views[v].scene >= scene_count) fail(PT_EINVAL, "view scene index out of range");
        validate_camera(views[v].camera);
        total += checked_pixel_count(views[v].width, views[v].height);
        if(total > size_t(INT32_MAX)) fail(PT_EINVAL, "batch exceeds 2^31 pixels");
    }
    require_device();
    auto b = std::make_unique<pt_batch>();
    b->pixels = total;
    HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreate(&b->start));
    HIP_CHECK(hipEventCreate(&b->stop));
    b->cus = device_cu_count();
    b->frame.max_depth = max_path_depth ? max_path_depth : 50;
    b->frame.direct_light = direct_light ? 1 : 0;
    b->frame.debug = PT_DEBUG_SHADED;
    b->frame.width = views[0].width; b->frame.height = views[0].height; b->frame.camera = views[0].camera;

    merged_scenes m = merge_scenes(scenes, scene_count);
    const pt_scene_desc md = m.desc();
    b->scene = std::make_unique<device_scene>(md, true, false, false);

    // Per-scene 4-wide hierarchies and light ranges, concatenated.
    std::vector<float4> wn, lv;
    std::vector<pt::light_rec> lights;
    const size_t nscenes = static_cast<size_t>(scene_count);
    std::vector<int32_t> roots(nscenes, pt::empty_root), broots(nscenes, pt::empty_root), light_offset(nscenes, 0),
        light_count(nscenes, 0);
    int binary_capacity = 2;
    std::vector<float> light_area(nscenes, 0.0f);
    int depth = 1;
    for(int32_t k = 0; k < scene_count; k++) {
        const pt_scene_desc& s = scenes[k];
        std::vector<float4> ti = pt::triangle_test_records(s);
        pt::wide_layout w;
        if(s.root >= 0) {
            if(!pt::build_wide(s, ti, w)) fail(PT_EINVAL, "batch scenes need tree hierarchies (no shared subtrees)");
            const int32_t node_offset = int32_t(wn.size() / 8), leaf_offset = int32_t(lv.size() / 3);
            if(lv.size() / 3 + w.leaves.size() / 3 >= (size_t(1) << 27)) fail(PT_EINVAL, "batch exceeds 2^27 leaves");
            if(wn.size() / 8 + w.nodes.size() / 8 > pt::wide_node_limit) fail(PT_EINVAL, "batch exceeds 2^25 wide nodes");
            for(size_t i = 0; i < w.nodes.size() / 8; i++) {
                float4 *r = &w.nodes[8 * i + 6];
                int32_t ref[4];
                std::memcpy(ref, r, sizeof ref);
                for(int j = 0; j < 4; j++) ref[j] = rebase_ref(ref[j], node_offset, leaf_offset);
                std::memcpy(r, ref, sizeof ref);
            }
            for(size_t i = 0; i < w.leaves.size() / 3; i++) {
                float *cf = &w.leaves[3 * i + 2].w;
                int32_t code;
                std::memcpy(&code, cf, 4);
                const int32_t prim = code >> 1, sphere = code & 1;
                code = int32_t((prim + (sphere ? m.sph_offset[size_t(k)] : m.tri_offset[size_t(k)])) * 2 + sphere);
                std::memcpy(cf, &code, 4);
            }
            roots[size_t(k)] = rebase_ref(w.root_ref, node_offset, leaf_offset);
            // The scene's root in the merged binary layout, for reference-order retraces.
            const int32_t merged_root = int32_t(m.node_offset[size_t(k)]) + s.root;
            const pt_flat_node& rn = md.nodes[merged_root];
            broots[size_t(k)] = rn.type == PT_BRANCH ? b->scene->branch_ids[size_t(merged_root)]
                                                     : pt::leaf_ref(rn.primitive, rn.type == PT_SPHERE);
            binary_capacity = std::max(binary_capacity, std::min(pt::stack_limit, int(s.max_depth)));
            wn.insert(wn.end(), w.nodes.begin(), w.nodes.end());
            lv.insert(lv.end(), w.leaves.begin(), w.leaves.end());
            depth = std::max(depth, w.depth);
        }
        std::vector<pt::light_rec> lk = pt::scene_lights(s);
        light_offset[size_t(k)] = int32_t(lights.size());
        light_count[size_t(k)] = int32_t(lk.size());
        light_area[size_t(k)] = lk.empty() ? 0.0f : lk.back().cumulative_area;
        for(auto l : lk) {
            l.primitive += int32_t(l.type == PT_SPHERE ? m.sph_offset[size_t(k)] : m.tri_offset[size_t(k)]);
            lights.push_back(l);
        }
    }
    b->wnodes.upload(wn.data(), wn.size());
    b->leaves.upload(lv.data(), lv.size());
    b->lights.upload(lights.data(), lights.size());
    pt::scene_dev& sv = b->scene->view;
    sv.wnodes = b->wnodes.get(); sv.leaves = b->leaves.get();
    // Marks the wide layout as present; each query starts at its view's root.
    sv.wroot_ref = roots[0];
    sv.stack_capacity = binary_capacity;
    sv.wide_stack_capacity = std::max(3 * depth + 1, binary_capacity);
    sv.lights = b->lights.get(); sv.light_count = int32_t(lights.size());
    sv.light_area = 0;

    // Views and the pixel -> view map.
    std::vector<uint16_t> pv(total);
    int64_t offset = 0;
    for(int32_t v = 0; v < view_count; v++) {
        pt::view_rec r{};
        const pt_batch_view& in = views[v];
        r.camera = in.camera; r.width = in.width; r.height = in.height;
        r.pixel_offset = offset; r.seed = in.seed;
        r.wroot_ref = roots[size_t(in.scene)];
        r.broot_ref = broots[size_t(in.scene)];
        r.light_offset = light_offset[size_t(in.scene)]; r.light_count = light_count[size_t(in.scene)];
        r.light_area = light_area[size_t(in.scene)];
        b->host_views.push_back(r);
        const int64_t n = int64_t(in.width) * in.height;
        std::fill(pv.begin() + offset, pv.begin() + offset + n, uint16_t(v));
        offset += n;
    }
    b->views.upload(b->host_views.data(), b->host_views.size());
    b->pixel_view.upload(pv.data(), pv.size());
    b->bd.views = b->views.get();
    b->bd.pixel_view = b->pixel_view.get();

    // Path state for every pixel of every view; one set of queues.
    const size_t n = total;
    b->sum.alloc(3 * n); b->rng.alloc(n);
    b->ray_o.alloc(n); b->ray_d.alloc(n); b->thr.alloc(n); b->rad.alloc(n);
    b->sh_o.alloc(n); b->sh_d.alloc(n); b->sh_c.alloc(n); b->ctr.alloc(n); b->hit.alloc(n);
    partition& part = b->part;
    part.count = int64_t(n);
    part.stream = nullptr;
    for(auto *e : {&part.checked, &part.drain0, &part.drain1, &part.finished}) HIP_CHECK(hipEventCreate(e));
    part.ctl.alloc(1);
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&part.ctl_host), sizeof(pt::loop_ctl), hipHostMallocDefault));
    part.queue0.alloc(n); part.queue1.alloc(n); part.shadow_queue.alloc(n); part.finish_queue.alloc(n);
    part.log.alloc(2 * size_t(log_capacity));
    pt::loop_ctl init{};
    init.log = part.log.get();
    init.log_capacity = log_capacity;
    HIP_CHECK(hipMemcpy(part.ctl.get(), &init, sizeof init, hipMemcpyHostToDevice));
    auto& p = part.paths;
    p.ray_o = b->ray_o.get(); p.ray_d = b->ray_d.get(); p.thr = b->thr.get(); p.rad = b->rad.get();
    p.ctr = b->ctr.get(); p.hit = b->hit.get();
    p.sh_o = b->sh_o.get(); p.sh_d = b->sh_d.get(); p.sh_c = b->sh_c.get();
    p.rng = b->rng.get(); p.sum = b->sum.get(); p.aov = nullptr;
    p.queue[0] = part.queue0.get(); p.queue[1] = part.queue1.get();
    p.shadow_queue = part.shadow_queue.get(); p.finish_queue = part.finish_queue.get();
    b->reset();
    return b.release();
}

// Tuning dimensions the autotuner sweeps, in order.
struct tune_axis { int32_t pt_tuning::*field; std::vector<int32_t> values; };

} // namespace

extern "C" {

const char *pt_last_error(void) { return g_last_error.c_str(); }

pt_status pt_device_available(char *description, size_t capacity) {
    std::string text;
    pt_status st = PT_OK;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if(e != hipSuccess) { text = hipGetErrorString(e); st = PT_EDEVICE; }
    else if(count < 1) { text = "no HIP devices"; st = PT_EDEVICE; }
    else {
        int current = 0;
        hipDeviceProp_t props{};
        e = hipGetDevice(&current);
        if(e == hipSuccess) e = hipGetDeviceProperties(&props, current);
        if(e != hipSuccess) { text = hipGetErrorString(e); st = PT_EDEVICE; }
        else text = std::string(props.name) + " (" + props.gcnArchName + ")";
    }
    if(description && capacity) {
        size_t n = std::min(capacity - 1, text.size());
        std::memcpy(description, text.data(), n);
        description[n] = 0;
    }
    g_last_error = st == PT_OK ? "" : text;
    return st;
}

pt_status pt_host_alloc(size_t bytes, void **out) {
    return guarded([&] {
        if(!out) fail(PT_EINVAL, "null output pointer");
        *out = nullptr;
        if(bytes == 0) fail(PT_EINVAL, "zero-byte host buffer");
        HIP_CHECK(hipHostMalloc(out, bytes, hipHostMallocDefault));
    });
}

pt_status pt_host_free(void *p) {
    return guarded([&] { if(p) HIP_CHECK(hipHostFree(p)); });
}

pt_status pt_build_bvh(const pt_flat_triangle *triangles, int64_t triangle_count, const pt_flat_sphere *spheres,
                       int64_t sphere_count, pt_flat_node *nodes, int64_t node_capacity, int32_t *root,
                       int32_t *max_depth, float *milliseconds) {
    return guarded([&] {
        if(!root || !max_depth) fail(PT_EINVAL, "null output pointer");
        *root = -1;
        *max_depth = 0;
        if(milliseconds) *milliseconds = 0;
        if(triangle_count < 0 || sphere_count < 0 || (triangle_count && !triangles) || (sphere_count && !spheres))
            fail(PT_EINVAL, "invalid primitive arrays");
        const int64_t n = triangle_count + sphere_count;
        if(n > int64_t(INT32_MAX / 2)) fail(PT_EINVAL, "too many primitives for 32-bit node indices");
        // The leaf-box rules of flat_validate.h:66-77 need finite vertices and
        // finite, positive radii.
        for(int64_t i = 0; i < triangle_count; i++) {
            const pt_flat_triangle& t = triangles[i];
            if(!finite3(t.v0) || !finite3(t.v1) || !finite3(t.v2)) fail(PT_ESCENE, "non-finite triangle vertex");
        }
        for(int64_t i = 0; i < sphere_count; i++) {
            const pt_flat_sphere& s = spheres[i];
            if(!finite3(s.center) || !std::isfinite(s.radius) || !(s.radius > 0)) fail(PT_ESCENE, "invalid sphere");
        }
        if(n == 0) return;
        if(!nodes || node_capacity < 2 * n - 1) fail(PT_EINVAL, "node array holds fewer than 2n - 1 nodes");
        require_device();
        pt::build_lbvh(triangles, triangle_count, spheres, sphere_count, nodes, milliseconds);
        // Depth of the hierarchy (the traversal stack bound, flat_data.h:52).
        const int64_t count = 2 * n - 1;
        std::vector<std::pair<int32_t, int>> stack{{0, 1}};
        int depth = 0;
        int64_t visited = 0;
        while(!stack.empty()) {
            const auto [id, d] = stack.back();
            stack.pop_back();
            if(id < 0 || id >= count || ++visited > count) fail(PT_EDEVICE, "linear BVH build produced a broken tree");
            depth = std::max(depth, d);
            if(nodes[id].type == PT_BRANCH) {
                stack.push_back({nodes[id].right, d + 1});
                stack.push_back({nodes[id].left, d + 1});
            }
        }
        if(visited != count) fail(PT_EDEVICE, "linear BVH build produced a broken tree");
        if(depth > pt::stack_limit)
            fail(PT_ESCENE, "linear BVH deeper than " + std::to_string(pt::stack_limit) +
                                " levels; build this scene with the host builder (pts_build)");
        *root = 0;
        *max_depth = depth;
    });
}

pt_status pt_renderer_create(const pt_scene_desc *scene, const pt_flat_camera *camera, const pt_settings *settings,
                             pt_renderer **out) {
    return guarded([&] {
        if(!out) fail(PT_EINVAL, "null output handle");
        *out = nullptr;
        if(!scene || !camera || !settings) fail(PT_EINVAL, "null argument");
        // cuda_renderer.cu:177-179 order: scene, camera, mode, size, device.
        validate_scene(*scene);
        validate_camera(*camera);
        validate_mode(settings->debug);
        size_t pixels = checked_pixel_count(settings->width, settings->height);
        if(settings->max_path_depth < 0) fail(PT_EINVAL, "invalid path depth");
        require_device();
        auto r = std::make_unique<pt_renderer>();
        r->pixels = pixels;
        r->seed = settings->seed;
        r->frame.width = settings->width; r->frame.height = settings->height;
        r->frame.max_depth = settings->max_path_depth ? settings->max_path_depth : 50;
        r->frame.direct_light = settings->direct_light ? 1 : 0;
        r->frame.brute = settings->brute ? 1 : 0;
        r->frame.debug = settings->debug;
        r->frame.camera = *camera;
        check_grid(*r, 16, 16);
        HIP_CHECK(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreate(&r->start));
        HIP_CHECK(hipEventCreate(&r->stop));
        r->cus = device_cu_count();
        (void)static_scratch();  // resolve this GPU's allocation-free scratch at set-up
        r->scene = std::make_unique<device_scene>(*scene, true, settings->debug == PT_DEBUG_SHADED);
        allocate_state(*r, settings->debug == PT_DEBUG_SHADED);
        r->reset(*camera);
        *out = r.release();
    });
}

pt_status pt_renderer_destroy(pt_renderer *renderer) {
    return guarded([&] { delete renderer; });
}

pt_status pt_renderer_reset(pt_renderer *renderer, const pt_flat_camera *camera) {
    return guarded([&] {
        if(!renderer || !camera) fail(PT_EINVAL, "null argument");
        renderer->reset(*camera);
    });
}

pt_status pt_renderer_add_samples(pt_renderer *renderer, int32_t count) {
    return guarded([&] {
        if(!renderer) fail(PT_EINVAL, "null renderer");
        renderer->add_samples(count);
    });
}

pt_status pt_renderer_add_samples_noalloc(pt_renderer *renderer, int32_t count) {
    return guarded([&] {
        if(!renderer) fail(PT_EINVAL, "null renderer");
        renderer->add_samples_noalloc(count, static_scratch());
    });
}

pt_status pt_renderer_set_block_shape(pt_renderer *renderer, int32_t width, int32_t height) {
    return guarded([&] {
        if(!renderer) fail(PT_EINVAL, "null renderer");
        if(!((width == 8 && height == 8) || (width == 16 && height == 8) || (width == 16 && height == 16) || (width == 32 && height == 8)))
            fail(PT_EINVAL, "HIP block must be 8x8, 16x8, 16x16 or 32x8");
        if(renderer->block.x == unsigned(width) && renderer->block.y == unsigned(height)) return;
        check_grid(*renderer, unsigned(width), unsigned(height));
        renderer->block = dim3(unsigned(width), unsigned(height));
    });
}

pt_status pt_renderer_read_image(const pt_renderer *renderer, float *rgb, size_t float_count) {
    return guarded([&] {
        if(!renderer || !rgb) fail(PT_EINVAL, "null argument");
        if(float_count < 3 * renderer->pixels) fail(PT_EINVAL, "image buffer is too small");
        const size_t count = 3 * renderer->pixels;
        const float *src = renderer->sum.get();
        if(renderer->samples) {
            // value /= float(samples): multiply by float(1.0/samples) (vec3.h:152-158), on the device
            const float k = float(1.0 / double(float(renderer->samples)));
            if(renderer->image_buf.size() < count) renderer->image_buf.alloc(count);
            pt::launch_scale(src, renderer->image_buf.get(), int64_t(count), k, renderer->stream);
            HIP_CHECK(hipGetLastError());
            src = renderer->image_buf.get();
        }
        HIP_CHECK(hipMemcpyAsync(rgb, src, count * sizeof(float), hipMemcpyDeviceToHost, renderer->stream));
        HIP_CHECK(hipStreamSynchronize(renderer->stream));
    });
}

pt_status pt_renderer_write_rgba_device(const pt_renderer *renderer, void *device_rgba, size_t bytes) {
    return guarded([&] {
        if(!renderer) fail(PT_EINVAL, "null renderer");
        if(!device_rgba || bytes < 4 * renderer->pixels) fail(PT_EINVAL, "HIP RGBA destination is too small");
        pt::launch_pack(renderer->sum.get(), static_cast<uint8_t *>(device_rgba), int64_t(renderer->pixels), renderer->samples, renderer->stream);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(renderer->stream));
    });
}

pt_status pt_renderer_read_rgba(const pt_renderer *renderer, uint8_t *rgba, size_t bytes) {
    return guarded([&] {
        if(!renderer || !rgba) fail(PT_EINVAL, "null argument");
        if(bytes < 4 * renderer->pixels) fail(PT_EINVAL, "RGBA buffer is too small");
        pt::launch_pack(renderer->sum.get(), renderer->rgba.get(), int64_t(renderer->pixels), renderer->samples, renderer->stream);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(renderer->stream));
        HIP_CHECK(hipMemcpy(rgba, renderer->rgba.get(), 4 * renderer->pixels, hipMemcpyDeviceToHost));
    });
}

pt_status pt_renderer_stats(const pt_renderer *renderer, int32_t *samples, float *kernel_ms, uint64_t *rays) {
    return guarded([&] {
        if(!renderer) fail(PT_EINVAL, "null renderer");
        if(samples) *samples = renderer->samples;
        if(kernel_ms) *kernel_ms = renderer->milliseconds;
        if(rays) *rays = renderer->rays;
    });
}

pt_status pt_query_rays(const pt_scene_desc *scene, const pt_ray *rays, size_t count, int32_t brute, pt_trace_result *results) {
    return guarded([&] { query(scene, rays, count, brute, results, nullptr, nullptr); });
}

pt_status pt_query_rays_stats(const pt_scene_desc *scene, const pt_ray *rays, size_t count, int32_t brute,
                              pt_trace_result *results, int32_t *boxes, int32_t *primitives) {
    return guarded([&] { query(scene, rays, count, brute, results, boxes, primitives); });
}

pt_status pt_query_rays_wide(const pt_scene_desc *scene, const pt_ray *rays, size_t count, pt_trace_result *results,
                             int32_t *retraced) {
    return guarded([&] {
        if(!scene) fail(PT_EINVAL, "null scene");
        validate_scene(*scene);
        require_device();
        if(count == 0) return;
        if(count > size_t(INT32_MAX)) fail(PT_EINVAL, "too many validation rays");
        if(!rays || !results) fail(PT_EINVAL, "null ray or result buffer");
        device_scene uploaded(*scene, false, true);
        dev_buf<pt_ray> input; input.upload(rays, count);
        dev_buf<pt_trace_result> output(count);
        dev_buf<int32_t> flags(count);
        if(uploaded.view.wroot_ref == pt::empty_root) {
            // No 4-wide layout (an empty scene or shared subtrees): the shaded
            // path then runs the reference-order traversal too.
            pt::launch_query(uploaded.view, input.get(), output.get(), nullptr, nullptr, int32_t(count), false, nullptr);
            HIP_CHECK(hipMemset(flags.get(), 0, count * sizeof(int32_t)));
        } else {
            const int block = 256;
            const size_t threads = (count + block - 1) / block * block;
            dev_buf<int2> spill(std::max<size_t>(1, threads * size_t(pt::wide_spill_entries(uploaded.view.wide_stack_capacity))));
            pt::launch_wide_query(uploaded.view, input.get(), output.get(), flags.get(), spill.get(), int32_t(count), block, nullptr);
        }
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(results, output.get(), count * sizeof(pt_trace_result), hipMemcpyDeviceToHost));
        if(retraced) HIP_CHECK(hipMemcpy(retraced, flags.get(), count * sizeof(int32_t), hipMemcpyDeviceToHost));
    });
}

pt_status pt_renderer_read_state(const pt_renderer *renderer, float *sum, uint64_t *rng_state, size_t pixels) {
    return guarded([&] {
        if(!renderer) fail(PT_EINVAL, "null renderer");
        if(pixels != renderer->pixels) fail(PT_EINVAL, "pixel count mismatch");
        if(sum) HIP_CHECK(hipMemcpy(sum, renderer->sum.get(), 3 * pixels * sizeof(float), hipMemcpyDeviceToHost));
        if(rng_state) HIP_CHECK(hipMemcpy(rng_state, renderer->rng.get(), pixels * sizeof(uint64_t), hipMemcpyDeviceToHost));
    });
}

pt_status pt_renderer_write_state(pt_renderer *renderer, const float *sum, const uint64_t *rng_state, size_t pixels,
                                  int32_t samples, uint64_t rays) {
    return guarded([&] {
        if(!renderer || !sum || !rng_state) fail(PT_EINVAL, "null argument");
        if(pixels != renderer->pixels) fail(PT_EINVAL, "pixel count mismatch");
        if(samples < 0) fail(PT_EINVAL, "invalid accumulated sample count");
        HIP_CHECK(hipMemcpy(renderer->sum.get(), sum, 3 * pixels * sizeof(float), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(renderer->rng.get(), rng_state, pixels * sizeof(uint64_t), hipMemcpyHostToDevice));
        pt::loop_ctl c{};
        c.rays = rays;
        HIP_CHECK(hipMemcpy(renderer->ctl.get(), &c, sizeof(c), hipMemcpyHostToDevice));
        for(auto& part : renderer->parts) {
            pt::loop_ctl pc{};
            pc.rays = part->index == 0 ? rays : 0;
            pc.log = part->log.get();
            pc.log_capacity = log_capacity;
            HIP_CHECK(hipMemcpy(part->ctl.get(), &pc, sizeof(pc), hipMemcpyHostToDevice));
        }
        renderer->samples = samples; renderer->rays = rays; renderer->rays_noalloc = 0;
    });
}

pt_status pt_renderer_last_launch_stats(const pt_renderer *renderer, pt_launch_stats *stats) {
    return guarded([&] {
        if(!renderer || !stats) fail(PT_EINVAL, "null argument");
        *stats = renderer->last;
    });
}

pt_status pt_renderer_iteration_log(const pt_renderer *renderer, uint32_t *extension, uint32_t *shadow,
                                    size_t capacity, int32_t *count) {
    return guarded([&] {
        if(!renderer || !count) fail(PT_EINVAL, "null argument");
        *count = 0;
        if(renderer->frame.debug != PT_DEBUG_SHADED) return;
        // Partitions in order, each with its own iterations.
        size_t written = 0;
        int32_t total = 0;
        for(const auto& part : renderer->parts) {
            HIP_CHECK(hipMemcpy(part->ctl_host, part->ctl.get(), sizeof(pt::loop_ctl), hipMemcpyDeviceToHost));
            const size_t n = std::min<size_t>(part->ctl_host->iterations, log_capacity);
            total += int32_t(n);
            const size_t m = std::min(n, capacity - std::min(capacity, written));
            if(m) {
                std::vector<uint32_t> buf(2 * m);
                HIP_CHECK(hipMemcpy(buf.data(), part->log.get(), buf.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
                for(size_t i = 0; i < m; i++) {
                    if(extension) extension[written + i] = buf[2 * i];
                    if(shadow) shadow[written + i] = buf[2 * i + 1];
                }
                written += m;
            }
        }
        *count = total;
    });
}

pt_status pt_renderer_enable_aovs(pt_renderer *renderer, int32_t enable) {
    return guarded([&] {
        if(!renderer) fail(PT_EINVAL, "null renderer");
        if(renderer->frame.debug != PT_DEBUG_SHADED) fail(PT_EINVAL, "feature buffers need the shaded mode");
        if(renderer->samples != 0) fail(PT_EINVAL, "feature buffers must be enabled before samples are added");
        HIP_CHECK(hipStreamSynchronize(renderer->stream));
        if(enable) {
            if(!renderer->aov.size()) renderer->aov.alloc(6 * renderer->pixels);
            HIP_CHECK(hipMemset(renderer->aov.get(), 0, renderer->aov.size() * sizeof(float)));
        } else {
            renderer->aov.release();
        }
        for(auto& part : renderer->parts) part->paths.aov = renderer->aov.get();
    });
}

pt_status pt_renderer_read_aovs(const pt_renderer *renderer, float *albedo, float *normal, size_t count) {
    return guarded([&] {
        if(!renderer) fail(PT_EINVAL, "null renderer");
        if(!renderer->aov.size()) fail(PT_EINVAL, "feature buffers are not enabled");
        if(count != 3 * renderer->pixels) fail(PT_EINVAL, "feature buffer size mismatch");
        if(renderer->samples <= 0) fail(PT_EINVAL, "no samples accumulated");
        std::vector<float> host(renderer->aov.size());
        HIP_CHECK(hipMemcpy(host.data(), renderer->aov.get(), host.size() * sizeof(float), hipMemcpyDeviceToHost));
        const float k = float(1.0 / double(renderer->samples));
        for(size_t i = 0; i < renderer->pixels; i++) {
            for(int c = 0; c < 3; c++) {
                if(albedo) albedo[3 * i + size_t(c)] = host[6 * i + size_t(c)] * k;
                if(normal) normal[3 * i + size_t(c)] = host[6 * i + 3 + size_t(c)] * k;
            }
        }
    });
}

pt_status pt_renderer_set_tuning(pt_renderer *renderer, const pt_tuning *tuning) {
    return guarded([&] {
        if(!renderer || !tuning) fail(PT_EINVAL, "null argument");
        pt_tuning t = renderer->tuning;
        auto block_ok = [](int32_t b) { return b >= 64 && b <= 256 && b % 64 == 0; };
        if(tuning->trace_block) { if(!block_ok(tuning->trace_block)) fail(PT_EINVAL, "trace block must be 64..256, multiple of 64"); t.trace_block = tuning->trace_block; }
        if(tuning->shade_block) {
            if(!(tuning->shade_block >= 64 && tuning->shade_block <= 1024 && tuning->shade_block % 64 == 0))
                fail(PT_EINVAL, "shade block must be 64..1024, multiple of 64");
            t.shade_block = tuning->shade_block;
        }
        if(tuning->blocks_per_cu) { if(tuning->blocks_per_cu < 1 || tuning->blocks_per_cu > 64) fail(PT_EINVAL, "blocks per CU must be 1..64"); t.blocks_per_cu = tuning->blocks_per_cu; }
        if(tuning->check_interval) { if(tuning->check_interval < 1 || tuning->check_interval > 4096) fail(PT_EINVAL, "check interval must be 1..4096"); t.check_interval = tuning->check_interval; }
        t.time_kernels = tuning->time_kernels ? 1 : 0;
        t.count_tests = tuning->count_tests ? 1 : 0;
        t.exact_traversal = tuning->exact_traversal ? 1 : 0;
        if(tuning->drain_threshold) t.drain_threshold = tuning->drain_threshold < 0 ? -1 : tuning->drain_threshold;
        if(tuning->lane_refill) {
            if(tuning->lane_refill > 64) fail(PT_EINVAL, "lane refill must be 1..64 or -1");
            t.lane_refill = tuning->lane_refill < 0 ? -1 : tuning->lane_refill;
        }
        if(tuning->partitions) {
            if(tuning->partitions < 1 || tuning->partitions > 8) fail(PT_EINVAL, "partitions must be 1..8");
            t.partitions = tuning->partitions;
        }
        if(tuning->handoff_steps) {
            if(tuning->handoff_steps > 65536) fail(PT_EINVAL, "handoff steps must be 1..65536 or -1");
            t.handoff_steps = tuning->handoff_steps < 0 ? -1 : tuning->handoff_steps;
        }
        if(tuning->path_shade_batch) {
            if(tuning->path_shade_batch < 1 || tuning->path_shade_batch > 64) fail(PT_EINVAL, "path shade batch must be 1..64");
            t.path_shade_batch = tuning->path_shade_batch;
        }
        if(tuning->path_blocks_per_cu) {
            if(tuning->path_blocks_per_cu > 64) fail(PT_EINVAL, "path blocks per CU must be 1..64 or -1");
            t.path_blocks_per_cu = tuning->path_blocks_per_cu < 0 ? -1 : tuning->path_blocks_per_cu;
        }
        if(tuning->path_first_samples) {
            if(tuning->path_first_samples > 65536) fail(PT_EINVAL, "path first samples must be 1..65536 or -1");
            t.path_first_samples = tuning->path_first_samples < 0 ? -1 : tuning->path_first_samples;
        }
        if(tuning->path_sort_group) {
            if(tuning->path_sort_group < 1 || tuning->path_sort_group > 4096) fail(PT_EINVAL, "path sort group must be 1..4096");
            t.path_sort_group = tuning->path_sort_group;
        }
        if(tuning->path_trace_steps) {
            if(tuning->path_trace_steps < 1 || tuning->path_trace_steps > 64) fail(PT_EINVAL, "path trace steps must be 1..64");
            t.path_trace_steps = tuning->path_trace_steps;
        }
        renderer->tuning = t;
    });
}


pt_status pt_batch_create(const pt_scene_desc *scenes, int32_t scene_count, const pt_batch_view *views,
                          int32_t view_count, int32_t direct_light, int32_t max_path_depth, pt_batch **out) {
    return guarded([&] {
        if(!out) fail(PT_EINVAL, "null output handle");
        *out = nullptr;
        *out = create_batch(scenes, scene_count, views, view_count, direct_light, max_path_depth);
    });
}

pt_status pt_batch_destroy(pt_batch *batch) {
    return guarded([&] { delete batch; });
}

pt_status pt_batch_reset(pt_batch *batch) {
    return guarded([&] {
        if(!batch) fail(PT_EINVAL, "null batch");
        batch->reset();
    });
}

pt_status pt_batch_add_samples(pt_batch *batch, int32_t count) {
    return guarded([&] {
        if(!batch) fail(PT_EINVAL, "null batch");
        batch->add_samples(count);
    });
}

pt_status pt_batch_read_image(const pt_batch *batch, int32_t view, float *rgb, size_t count) {
    return guarded([&] {
        if(!batch || !rgb) fail(PT_EINVAL, "null argument");
        if(view < 0 || size_t(view) >= batch->host_views.size()) fail(PT_EINVAL, "view index out of range");
        const pt::view_rec& v = batch->host_views[size_t(view)];
        const size_t n = size_t(v.width) * size_t(v.height);
        if(count != 3 * n) fail(PT_EINVAL, "image buffer size mismatch");
        if(batch->samples <= 0) fail(PT_EINVAL, "no samples accumulated");
        HIP_CHECK(hipMemcpy(rgb, batch->sum.get() + 3 * size_t(v.pixel_offset), 3 * n * sizeof(float), hipMemcpyDeviceToHost));
        // image(): sum times float(1.0 / samples) (cuda_renderer.cu:230-234).
        const float k = float(1.0 / double(batch->samples));
        for(size_t i = 0; i < 3 * n; i++) rgb[i] *= k;
    });
}

pt_status pt_batch_stats(const pt_batch *batch, int32_t *samples, float *milliseconds, uint64_t *rays) {
    return guarded([&] {
        if(!batch) fail(PT_EINVAL, "null batch");
        if(samples) *samples = batch->samples;
        if(milliseconds) *milliseconds = batch->milliseconds;
        if(rays) *rays = batch->rays;
    });
}

pt_status pt_batch_last_launch_stats(const pt_batch *batch, pt_launch_stats *stats) {
    return guarded([&] {
        if(!batch || !stats) fail(PT_EINVAL, "null argument");
        *stats = batch->last;
    });
}

pt_status pt_batch_get_tuning(const pt_batch *batch, pt_tuning *tuning) {
    return guarded([&] {
        if(!batch || !tuning) fail(PT_EINVAL, "null argument");
        *tuning = batch->tuning;
    });
}

pt_status pt_batch_set_tuning(pt_batch *batch, const pt_tuning *tuning) {
    return guarded([&] {
        if(!batch || !tuning) fail(PT_EINVAL, "null argument");
        pt_tuning t = batch->tuning;
        auto block_ok = [](int32_t b, int32_t hi) { return b >= 64 && b <= hi && b % 64 == 0; };
        if(tuning->trace_block) { if(!block_ok(tuning->trace_block, 256)) fail(PT_EINVAL, "trace block must be 64..256, multiple of 64"); t.trace_block = tuning->trace_block; }
        if(tuning->shade_block) { if(!block_ok(tuning->shade_block, 1024)) fail(PT_EINVAL, "shade block must be 64..1024, multiple of 64"); t.shade_block = tuning->shade_block; }
        if(tuning->blocks_per_cu) { if(tuning->blocks_per_cu < 1 || tuning->blocks_per_cu > 64) fail(PT_EINVAL, "blocks per CU must be 1..64"); t.blocks_per_cu = tuning->blocks_per_cu; }
        if(tuning->check_interval) { if(tuning->check_interval < 1 || tuning->check_interval > 64) fail(PT_EINVAL, "check interval must be 1..64"); t.check_interval = tuning->check_interval; }
        if(tuning->drain_threshold) t.drain_threshold = tuning->drain_threshold < 0 ? -1 : tuning->drain_threshold;
        if(tuning->lane_refill) { if(tuning->lane_refill < 1 || tuning->lane_refill > 64) fail(PT_EINVAL, "lane refill must be 1..64"); t.lane_refill = tuning->lane_refill; }
        if(tuning->handoff_steps) {
            if(tuning->handoff_steps > 65536) fail(PT_EINVAL, "handoff steps must be 1..65536 or -1");
            t.handoff_steps = tuning->handoff_steps < 0 ? -1 : tuning->handoff_steps;
        }
        if(tuning->path_shade_batch) {
            if(tuning->path_shade_batch < 1 || tuning->path_shade_batch > 64) fail(PT_EINVAL, "path shade batch must be 1..64");
            t.path_shade_batch = tuning->path_shade_batch;
        }
        if(tuning->path_blocks_per_cu) {
            if(tuning->path_blocks_per_cu > 64) fail(PT_EINVAL, "path blocks per CU must be 1..64 or -1");
            t.path_blocks_per_cu = tuning->path_blocks_per_cu < 0 ? -1 : tuning->path_blocks_per_cu;
        }
        if(tuning->path_first_samples) {
            if(tuning->path_first_samples > 65536) fail(PT_EINVAL, "path first samples must be 1..65536 or -1");
            t.path_first_samples = tuning->path_first_samples < 0 ? -1 : tuning->path_first_samples;
        }
        if(tuning->path_sort_group) {
            if(tuning->path_sort_group < 1 || tuning->path_sort_group > 4096) fail(PT_EINVAL, "path sort group must be 1..4096");
            t.path_sort_group = tuning->path_sort_group;
        }
        if(tuning->path_trace_steps) {
            if(tuning->path_trace_steps < 1 || tuning->path_trace_steps > 64) fail(PT_EINVAL, "path trace steps must be 1..64");
            t.path_trace_steps = tuning->path_trace_steps;
        }
        batch->tuning = t;
    });
}

pt_status pt_batch_autotune(pt_batch *batch, int32_t probe_samples, int32_t passes, pt_tuning *best, float *best_mrays,
                            pt_batch_trial *trials, int32_t trial_capacity, int32_t *trial_count) {
    return guarded([&] {
        if(!batch) fail(PT_EINVAL, "null batch");
        if(probe_samples <= 0 || passes <= 0) fail(PT_EINVAL, "invalid probe");
        const int32_t px = int32_t(std::min<size_t>(batch->pixels, size_t(INT32_MAX)));
        const std::vector<tune_axis> axes = {
            {&pt_tuning::trace_block, {64, 128, 192, 256}},
            {&pt_tuning::blocks_per_cu, {8, 16, 24, 32}},
            {&pt_tuning::lane_refill, {16, 32, 48, 64}},
            {&pt_tuning::handoff_steps, {-1, 8, 16, 32}},
            {&pt_tuning::shade_block, {256, 512, 1024}},
            {&pt_tuning::drain_threshold, {std::max(1024, px / 32), std::max(2048, px / 16), std::max(4096, px / 8), std::max(8192, px / 4), all_paths}},
            {&pt_tuning::path_shade_batch, {16, 24, 32}},
            {&pt_tuning::path_trace_steps, {1, 2, 4, 8}},
            {&pt_tuning::check_interval, {2, 4, 8, 16}},
        };
        int32_t count = 0;
        auto measure = [&](const pt_tuning& t) {
            batch->tuning = t;
            batch->reset();
            batch->add_samples(probe_samples);
            const float rate = float(double(batch->rays) / (double(batch->milliseconds) * 1e3));
            if(trials && count < trial_capacity) trials[count] = pt_batch_trial{t, rate};
            count++;
            return rate;
        };
        pt_tuning current = batch->tuning;
        if(current.drain_threshold == 0) current.drain_threshold = std::max(4096, px / 8);
        float current_rate = measure(current);
        for(int32_t pass = 0; pass < passes; pass++) {
            for(const tune_axis& axis : axes) {
                for(int32_t value : axis.values) {
                    if(current.*axis.field == value) continue;
                    pt_tuning t = current;
                    t.*axis.field = value;
                    const float rate = measure(t);
                    if(rate > current_rate) { current = t; current_rate = rate; }
                }
            }
        }
        batch->tuning = current;
        batch->reset();
        if(best) *best = current;
        if(best_mrays) *best_mrays = current_rate;
        if(trial_count) *trial_count = count;
    });
}

}
