// fuzz_image.cpp -- libFuzzer target for the texture decoders (PNG, PNM,
// JPEG) behind pts_load_image (test infrastructure, built with AddressSanitizer
// and UndefinedBehaviorSanitizer by tests/native/build_sanitized.py).
//
// Each input is written to a file and loaded through the public entry point,
// which must either decode it (width and height positive, width * height * 3
// bytes of pixels) or return -1 with a message; any memory error, undefined
// behaviour, crash or oversized allocation is a failure.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <unistd.h>

#include "pt_scene.h"

namespace {
char scratch_path[4096];
}  // namespace

extern "C" int LLVMFuzzerInitialize(int *, char ***) {
    const char *dir = std::getenv("PT_FUZZ_DIR");
    std::snprintf(scratch_path, sizeof scratch_path, "%s/fuzz_image_%d", dir ? dir : "/tmp", int(getpid()));
    return 0;
}

extern "C" int LLVMFuzzerTestOneInput(const uint8_t *data, size_t size) {
    const char *path = scratch_path;
    FILE *f = std::fopen(path, "wb");
    if(!f) std::abort();
    if(size && std::fwrite(data, 1, size, f) != size) std::abort();
    std::fclose(f);
    int32_t w = 0, h = 0;
    uint8_t *rgb = nullptr;
    if(pts_load_image(path, &w, &h, &rgb) == 0) {
        if(w <= 0 || h <= 0 || !rgb) {
            std::fprintf(stderr, "pts_load_image succeeded with %d x %d pixels at %p\n", w, h, static_cast<void *>(rgb));
            std::abort();
        }
        // touch every byte, so that a short buffer shows up under ASan
        unsigned sum = 0;
        for(size_t i = 0; i < size_t(w) * size_t(h) * 3; i++) sum += rgb[i];
        if(sum == 0xFFFFFFFFu) std::fputc(' ', stderr);
        pts_free_image(rgb);
    } else if(!pt_scene_last_error()[0]) {  // a failure must say why
        std::fprintf(stderr, "pts_load_image failed without a message\n");
        std::abort();
    }
    return 0;
}
