// cuda_renderer_hip.cpp -- drop-in implementation of the reference renderer's
// src/cuda/cuda_renderer.h over the MI355X core's C ABI (include/pt_render.h).
//
// A maintainer of the reference builds this file in place of
// src/cuda/cuda_renderer.cu and links libptrace.so; the header and every
// caller (src/app/main.cpp, src/display/*, tests/cuda_tests.cpp,
// tests/lighting_tests.cpp) stay unchanged.  Status codes become the
// reference's exceptions: PT_EINVAL -> std::invalid_argument, anything else
// -> std::runtime_error.
#include <stdexcept>
#include <string>
#include <type_traits>

#include "cuda/cuda_renderer.h"
#include "pt_render.h"

static_assert(sizeof(flat_triangle) == sizeof(pt_flat_triangle) && sizeof(flat_sphere) == sizeof(pt_flat_sphere) &&
              sizeof(flat_node) == sizeof(pt_flat_node) && sizeof(flat_material) == sizeof(pt_flat_material) &&
              sizeof(flat_texture) == sizeof(pt_flat_texture) && sizeof(flat_camera) == sizeof(pt_flat_camera) &&
              sizeof(trace_result) == sizeof(pt_trace_result) && sizeof(ray) == sizeof(pt_ray) &&
              sizeof(vec3) == 3 * sizeof(float), "transfer records must match the C ABI byte for byte");
static_assert(std::is_standard_layout<vec3>::value && std::is_standard_layout<ray>::value, "plain records");

namespace {

void check(pt_status status) {
    if(status == PT_OK) return;
    std::string message = pt_last_error();
    if(status == PT_EINVAL) throw std::invalid_argument(message);
    throw std::runtime_error(message);
}

pt_scene_desc describe(const flat_scene& s) {
    pt_scene_desc d;
    d.triangles = reinterpret_cast<const pt_flat_triangle *>(s.triangles.data());
    d.triangle_count = int64_t(s.triangles.size());
    d.spheres = reinterpret_cast<const pt_flat_sphere *>(s.spheres.data());
    d.sphere_count = int64_t(s.spheres.size());
    d.nodes = reinterpret_cast<const pt_flat_node *>(s.nodes.data());
    d.node_count = int64_t(s.nodes.size());
    d.materials = reinterpret_cast<const pt_flat_material *>(s.materials.data());
    d.material_count = int64_t(s.materials.size());
    d.textures = reinterpret_cast<const pt_flat_texture *>(s.textures.data());
    d.texture_count = int64_t(s.textures.size());
    d.pixels = s.pixels.data();
    d.pixel_bytes = int64_t(s.pixels.size());
    d.root = s.root;
    d.max_depth = s.max_depth;
    return d;
}

const pt_flat_camera *as_camera(const flat_camera& c) { return reinterpret_cast<const pt_flat_camera *>(&c); }

} // namespace

bool cuda_device_available(std::string& description) {
    char text[512] = {0};
    pt_status status = pt_device_available(text, sizeof text);
    description = text;
    return status == PT_OK;
}

struct cuda_renderer::implementation {
    pt_renderer *handle = nullptr;
    size_t pixels = 0;
    ~implementation() { if(handle) pt_renderer_destroy(handle); }
};

cuda_renderer::cuda_renderer(const flat_scene& scene, const flat_camera& camera, int width, int height,
                             uint64_t seed, bool brute, render_debug debug, bool direct_light)
    : state(std::make_unique<implementation>()) {
    pt_scene_desc desc = describe(scene);
    pt_settings settings{width, height, seed, brute ? 1 : 0, int32_t(debug), direct_light ? 1 : 0, 0};
    check(pt_renderer_create(&desc, as_camera(camera), &settings, &state->handle));
    state->pixels = size_t(width) * size_t(height);
}

cuda_renderer::~cuda_renderer() = default;

void cuda_renderer::reset(const flat_camera& camera) { check(pt_renderer_reset(state->handle, as_camera(camera))); }

void cuda_renderer::add_samples(int count) { check(pt_renderer_add_samples(state->handle, count)); }

void cuda_renderer::set_block_shape(int width, int height) {
    check(pt_renderer_set_block_shape(state->handle, width, height));
}

std::vector<vec3> cuda_renderer::image() const {
    std::vector<vec3> out(state->pixels);
    check(pt_renderer_read_image(state->handle, reinterpret_cast<float *>(out.data()), 3 * out.size()));
    return out;
}

std::vector<unsigned char> cuda_renderer::rgba_image() const {
    std::vector<unsigned char> out(4 * state->pixels);
    check(pt_renderer_read_rgba(state->handle, out.data(), out.size()));
    return out;
}

void cuda_renderer::write_rgba(unsigned char *device_pixels, size_t bytes) const {
    check(pt_renderer_write_rgba_device(state->handle, device_pixels, bytes));
}

int cuda_renderer::sample_count() const {
    int32_t samples = 0;
    check(pt_renderer_stats(state->handle, &samples, nullptr, nullptr));
    return samples;
}

float cuda_renderer::kernel_milliseconds() const {
    float ms = 0;
    check(pt_renderer_stats(state->handle, nullptr, &ms, nullptr));
    return ms;
}

uint64_t cuda_renderer::rays_traced() const {
    uint64_t rays = 0;
    check(pt_renderer_stats(state->handle, nullptr, nullptr, &rays));
    return rays;
}

std::vector<trace_result> cuda_query_rays(const flat_scene& scene, const std::vector<ray>& rays, bool brute) {
    pt_scene_desc desc = describe(scene);
    std::vector<trace_result> out(rays.size());
    check(pt_query_rays(&desc, reinterpret_cast<const pt_ray *>(rays.data()), rays.size(), brute ? 1 : 0,
                        reinterpret_cast<pt_trace_result *>(out.data())));
    return out;
}
