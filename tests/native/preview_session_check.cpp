// preview_session<cuda_renderer> on the drop-in binding.
//
// The reference's viewer drives the GPU renderer through its preview session
// (src/display/viewer.cpp:36, src/display/preview_session.h:18-104), whose
// cuda_renderer overloads read the device-packed RGBA image and the kernel
// time (preview_session.h:13-15).  This program instantiates that session on
// the binding over libptrace.so (built by oracle/build_ref.sh from the
// reference's headers, which only this container holds) and walks it through
// what a viewer does: idle frames, camera moves, focus loss, minimize and
// restore, resizes and debug-mode changes.  After each step the session's
// pixels must equal a fresh cuda_renderer's of the same camera, size, mode
// and sample count byte for byte, and the reference's own host renderer
// (tests/support/portable_renderer.h) to within 1 per channel (the bar of
// the reference's interop test, tests/cuda_gl_tests.cpp:29-31); the session
// must report the frame's rays and a positive kernel time.
#include <cstdlib>
#include <iostream>
#include <string>

#include "display/preview_session.h"
#include "support/portable_renderer.h"

namespace {

int checks = 0, failures = 0;

void check(bool ok, const std::string& what) {
    checks++;
    if(!ok) {
        failures++;
        if(failures < 20) std::cerr << "FAIL: " << what << "\n";
    }
}

bool within_one(const std::vector<unsigned char>& a, const std::vector<unsigned char>& b) {
    if(a.size() != b.size()) return false;
    for(size_t i = 0; i < a.size(); i++)
        if(std::abs(int(a[i]) - int(b[i])) > 1) return false;
    return true;
}

}  // namespace

int main() {
    std::string device;
    if(!cuda_device_available(device)) {
        std::cout << "no device: " << device << "\n";
        return 77;
    }
    srand48(7);
    obj_scene source(RAY_TRACER_ASSET_DIR "/models/blender_sphere/sphere.obj", 19.0f / 13);
    flat_scene scene = flatten_scene(&source);
    camera_controller camera(source.lookfrom, source.lookat, 40, 19.0f / 13, 0, source.focus_dist, 0, 1, 1, 0.1f, true);
    for(bool direct : {false, true}) {
        const std::string tag = direct ? " (MIS)" : " (classic)";
        preview_session<cuda_renderer> session(scene, camera, 19, 13, 7, 2, false, DEBUG_SHADED, direct);
        auto compare_fresh = [&](const std::string& step) {
            cuda_renderer fresh(scene, flatten_camera(session.camera_frame()), session.image_width(),
                                session.image_height(), 7, false, session.debug_mode(), direct);
            fresh.add_samples(session.sample_count());
            check(session.rgba() == fresh.rgba_image(), step + ": session pixels equal a fresh renderer's" + tag);
            portable_renderer host(scene, flatten_camera(session.camera_frame()), session.image_width(),
                                   session.image_height(), 7, false, session.debug_mode(), direct);
            host.add_samples(session.sample_count());
            check(within_one(session.rgba(), display_pixels(host.image())),
                  step + ": session pixels within 1 of the reference host renderer's" + tag);
        };
        check(session.sample_count() == 0 && session.rgba().size() == 19 * 13 * 4, "new session is empty" + tag);
        check(session.render_frame({}, 0) && session.render_frame({}, 0.1), "idle frames render" + tag);
        check(session.sample_count() == 4, "idle frames accumulate" + tag);
        check(session.rays_last_frame() > 0, "frame reports its rays" + tag);
        check(session.kernel_frame_milliseconds().has_value() && *session.kernel_frame_milliseconds() > 0,
              "frame reports the binding's kernel time" + tag);
        compare_fresh("idle");
        camera_input input;
        input.right = 1;
        input.mouse_dx = 50;
        session.render_frame(input, 0.5);
        check(session.sample_count() == 2, "camera move restarts the accumulation" + tag);
        compare_fresh("move");
        session.set_focused(false);
        check(!session.render_frame(input, 10), "unfocused frames do not render" + tag);
        session.set_focused(true);
        session.render_frame(input, 10);
        check(session.sample_count() == 4, "resumed frame discards stale input" + tag);
        compare_fresh("resume");
        check(!session.resize(0, 0) && session.paused(), "zero-size framebuffer suspends" + tag);
        check(!session.resize(19, 13) && !session.paused(), "restore keeps the framebuffer" + tag);
        check(session.resize(23, 17) && session.sample_count() == 0, "resize starts a new accumulation" + tag);
        session.render_frame({}, 0.1);
        compare_fresh("resize");
        for(auto mode : {DEBUG_NORMALS, DEBUG_TRIANGLE, DEBUG_MATERIAL, DEBUG_BVH, DEBUG_INTERSECTIONS, DEBUG_SHADED}) {
            check(session.set_debug(mode) && session.sample_count() == 0, "debug mode change resets" + tag);
            session.render_frame({}, 0);
            compare_fresh("debug mode " + std::to_string(int(mode)));
        }
        session.set_batch_samples(3);
        const int before = session.sample_count();
        session.render_frame({}, 0);
        check(session.sample_count() == before + 3, "batch size change keeps the view" + tag);
        compare_fresh("batch size");
    }
    std::cout << checks << " preview session checks on " << device << ", " << failures << " failures\n";
    return failures ? 1 : 0;
}
