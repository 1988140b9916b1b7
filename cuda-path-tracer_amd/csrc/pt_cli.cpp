// pt_cli.cpp -- the command-line renderer, a drop-in for the reference's
// ray_tracer application (src/app/main.cpp, src/app/render_options.h:49-107).
//
// Same flags, defaults, messages and exit codes (0 ok, 1 bad arguments or
// failure, 77 no device), same scene staging (srand48(seed); obj_scene;
// flatten_scene) and the same ASCII PPM on stdout.  Additions:
//   --backend hip      ("cuda" is accepted as the reference's name for it)
//   --bvh reference|sah|device  hierarchy of the imported mesh (default:
//                        reference, which reproduces the reference's output
//                        exactly; device: the linear BVH built on the GPU,
//                        pt_build_bvh, over the SAH import's arrays, and for
//                        the built-in scenes too)
//   --max-depth N      path length limit (default 50, the reference's)
//   --output FILE      write the PPM to FILE instead of stdout
//   --scene obj|cornell|icosphere  built-in synthetic scenes
//   --frequency N      icosphere subdivision (20 N^2 triangles)
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <unistd.h>
#include <vector>

#include "pt_render.h"
#include "pt_scene.h"

namespace {

struct options {
    int width = 500, height = 500, samples = 100, seed = 7, batch_samples = 16;
    int debug = PT_DEBUG_SHADED;
    std::string obj_file;
    std::string backend = "hip";
    float scale[3] = {1, 1, 1}, rotate[3] = {0, 0, 0}, translate[3] = {0, 0, 0};
    bool help = false, direct_light = true;
    std::string bvh = "reference", scene = "obj", output = "-";
    int max_depth = 0, frequency = 64;
};

// render_options.h:30-38: base-10 integer in [minimum, INT_MAX], nothing trailing.
int parse_integer(const char *value, int minimum) {
    char *end;
    errno = 0;
    long parsed = std::strtol(value, &end, 10);
    if(errno || end == value || *end != '\0' || parsed < minimum || parsed > INT_MAX)
        throw std::invalid_argument(std::string("invalid integer: ") + value);
    return int(parsed);
}

// render_options.h:40-47
float parse_float(const char *value) {
    char *end;
    errno = 0;
    float parsed = std::strtof(value, &end);
    if(errno || end == value || *end != '\0' || !std::isfinite(parsed))
        throw std::invalid_argument(std::string("invalid finite number: ") + value);
    return parsed;
}

// The asset directory the default OBJ file is read from, fixed when the CLI
// is built (render_options.h:19 uses the reference's RAY_TRACER_ASSET_DIR);
// the build sets it from its PT_ASSET_DIR environment variable.
#ifndef PT_ASSET_DIR
#define PT_ASSET_DIR "assets"
#endif

options parse_options(int argc, char **argv) {
    options o;
    o.obj_file = PT_ASSET_DIR "/models/blender_sphere/sphere.obj";
    bool has_transform = false, has_obj = false;
    for(int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto value = [&]() -> const char * {
            if(i + 1 >= argc) throw std::invalid_argument("missing value for " + a);
            return argv[++i];
        };
        if(a == "--help") o.help = true;
        else if(a == "--lighting") {
            if(i + 1 >= argc) throw std::invalid_argument("missing value for --lighting");
            std::string v = argv[++i];
            if(v != "classic" && v != "mis") throw std::invalid_argument("lighting must be classic or mis");
            o.direct_light = v == "mis";
        } else if(a == "--normals") {
            o.debug = PT_DEBUG_NORMALS;
        } else if(a == "--obj-scale" || a == "--obj-rotate" || a == "--obj-translate") {
            if(i + 3 >= argc) throw std::invalid_argument("expected three values for " + a);
            float *dst = a == "--obj-scale" ? o.scale : a == "--obj-rotate" ? o.rotate : o.translate;
            for(int axis = 0; axis < 3; axis++) dst[axis] = parse_float(argv[++i]);
            has_transform = true;
        } else if(a == "--obj") { o.obj_file = value(); has_obj = true; }
        else if(a == "--backend") o.backend = value();
        else if(a == "--debug") {
            std::string mode = value();
            if(mode == "shaded") o.debug = PT_DEBUG_SHADED;
            else if(mode == "normals") o.debug = PT_DEBUG_NORMALS;
            else if(mode == "triangle-id") o.debug = PT_DEBUG_TRIANGLE;
            else if(mode == "material-id") o.debug = PT_DEBUG_MATERIAL;
            else if(mode == "bvh") o.debug = PT_DEBUG_BVH;
            else if(mode == "intersections") o.debug = PT_DEBUG_INTERSECTIONS;
            else throw std::invalid_argument("unknown debug mode: " + mode);
        } else if(a == "--batch-samples") o.batch_samples = parse_integer(value(), 1);
        else if(a == "--width") o.width = parse_integer(value(), 1);
        else if(a == "--height") o.height = parse_integer(value(), 1);
        else if(a == "--samples") o.samples = parse_integer(value(), 1);
        else if(a == "--seed") o.seed = parse_integer(value(), 0);
        else if(a == "--bvh") {
            o.bvh = value();
            if(o.bvh != "reference" && o.bvh != "sah" && o.bvh != "device")
                throw std::invalid_argument("bvh must be reference, sah or device");
        } else if(a == "--max-depth") o.max_depth = parse_integer(value(), 1);
        else if(a == "--output") o.output = value();
        else if(a == "--scene") {
            o.scene = value();
            if(o.scene != "obj" && o.scene != "cornell" && o.scene != "icosphere")
                throw std::invalid_argument("scene must be obj, cornell or icosphere");
        } else if(a == "--frequency") o.frequency = parse_integer(value(), 1);
        else throw std::invalid_argument("unknown option: " + a);
    }
    if(o.obj_file.empty()) throw std::invalid_argument("OBJ filename must not be empty");
    if(has_transform && !has_obj) throw std::invalid_argument("OBJ transforms require --obj FILE");
    for(int i = 0; i < 3; i++)
        if(!std::isfinite(o.scale[i]) || o.scale[i] == 0 || !std::isfinite(o.rotate[i]) || !std::isfinite(o.translate[i]))
            throw std::invalid_argument("mesh transform requires finite values and nonzero scale");
    if(o.backend != "hip" && o.backend != "cuda") throw std::invalid_argument("HIP is the only rendering backend");
    return o;
}

void check(pt_status status) {
    if(status == PT_OK) return;
    std::string message = pt_last_error();
    if(status == PT_EINVAL) throw std::invalid_argument(message);
    throw std::runtime_error(message);
}

const char *usage =
    "Usage: pt_render [--obj FILE] [--width N] [--height N]\n"
    "  [--samples N] [--batch-samples N] [--seed N]\n"
    "  [--obj-scale X Y Z] [--obj-rotate X Y Z] [--obj-translate X Y Z]\n"
    "  [--debug shaded|normals|triangle-id|material-id|bvh|intersections] [--normals]\n"
    "  [--lighting classic|mis] (default: mis)\n"
    "  [--backend hip] [--bvh reference|sah|device] [--max-depth N] [--output FILE]\n"
    "  [--scene obj|cornell|icosphere] [--frequency N]\n"
    "Defaults: HIP, Blender sphere, 500x500, 100 samples/pixel, seed 7.\n"
    "PPM output goes to stdout. OBJ transforms apply scale, XYZ rotation in degrees, then translation.\n";

} // namespace

int main(int argc, char **argv) {
    try {
        options o = parse_options(argc, argv);
        if(o.help) { std::cout << usage; return 0; }
        char device[512] = {0};
        if(pt_device_available(device, sizeof device) != PT_OK) {
            std::cerr << "HIP device unavailable: " << device << '\n';
            return 77;
        }
        const float aspect = float(o.width) / o.height;
        pt_flat_camera camera;
        pts_scene *scene = nullptr;
        std::string warnings;
        if(o.scene == "obj") {
            pts_obj_options opt;
            std::memcpy(opt.scale, o.scale, sizeof opt.scale);
            std::memcpy(opt.rotate, o.rotate, sizeof opt.rotate);
            std::memcpy(opt.translate, o.translate, sizeof opt.translate);
            opt.aspect = aspect;
            opt.split = o.bvh == "reference" ? PTS_SPLIT_REFERENCE : PTS_SPLIT_SAH;
            opt.seed = o.seed;
            pts_obj_info info;
            scene = pts_load_obj(o.obj_file.c_str(), &opt, &info);
            if(!scene) throw std::runtime_error(pt_scene_last_error());
            const float up[3] = {0, 1, 0};
            const float from[3] = {info.lookfrom.x, info.lookfrom.y, info.lookfrom.z};
            const float at[3] = {info.lookat.x, info.lookat.y, info.lookat.z};
            pts_camera(from, at, up, 40, aspect, 0, info.focus_dist, 0, 1, &camera);
            warnings = pts_warnings(scene);
        } else if(o.scene == "cornell") {
            scene = pts_cornell(uint64_t(o.seed), o.width, o.height, &camera);
        } else {
            scene = pts_icosphere(o.frequency, uint64_t(o.seed), o.width, o.height, &camera);
        }
        if(!scene) throw std::runtime_error(pt_scene_last_error());
        pt_scene_desc desc;
        pts_desc(scene, &desc);
        std::vector<pt_flat_node> device_nodes;  // the device-built hierarchy, while desc points at it
        if(o.bvh == "device" && desc.triangle_count + desc.sphere_count > 0) {
            device_nodes.resize(size_t(2 * (desc.triangle_count + desc.sphere_count) - 1));
            int32_t root = -1, depth = 0;
            const pt_status built = pt_build_bvh(desc.triangles, desc.triangle_count, desc.spheres, desc.sphere_count,
                                                 device_nodes.data(), int64_t(device_nodes.size()), &root, &depth,
                                                 nullptr);
            if(built != PT_OK) {
                pts_free(scene);
                check(built);
            }
            desc.nodes = device_nodes.data();
            desc.node_count = int64_t(device_nodes.size());
            desc.root = root;
            desc.max_depth = depth;
        }
        pt_settings settings{o.width, o.height, uint64_t(o.seed), 0, o.debug, o.direct_light ? 1 : 0, o.max_depth};
        pt_renderer *renderer = nullptr;
        pt_status status = pt_renderer_create(&desc, &camera, &settings, &renderer);
        const int64_t triangles = desc.triangle_count, nodes = desc.node_count;
        pts_free(scene);
        check(status);
        std::cerr << "HIP: " << device << "\nLoaded " << triangles << " triangles, " << nodes << " BVH nodes\n" << warnings;
        int32_t count = 0;
        try {
            while(count < o.samples) {
                check(pt_renderer_add_samples(renderer, std::min(o.batch_samples, o.samples - count)));
                check(pt_renderer_stats(renderer, &count, nullptr, nullptr));
            }
            std::vector<float> pixels(size_t(o.width) * size_t(o.height) * 3);
            check(pt_renderer_read_image(renderer, pixels.data(), pixels.size()));
            pt_renderer_destroy(renderer);
            renderer = nullptr;
            std::cerr << "Rendered " << o.width << "x" << o.height << ", " << count << " samples/pixel\n";
            if(pts_write_ppm(o.output.c_str(), pixels.data(), o.width, o.height) != 0)
                throw std::runtime_error("could not write image output");
        } catch(...) {
            if(renderer) pt_renderer_destroy(renderer);
            throw;
        }
    } catch(const std::exception& error) {
        std::cerr << error.what() << "\nUse --help for usage.\n";
        return 1;
    }
    return 0;
}
