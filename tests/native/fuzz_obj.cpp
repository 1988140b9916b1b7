// fuzz_obj.cpp -- libFuzzer target for the OBJ/MTL import behind pts_load_obj
// (test infrastructure, built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/native/build_sanitized.py).
//
// An input is up to three parts separated by NUL bytes: OBJ text, MTL text and
// the bytes of an image file.  They are written as m.obj (which starts with
// "mtllib m.mtl"), m.mtl and t.png in a scratch directory, so that a mutated
// material can name the texture, and the OBJ is imported through the public
// entry point with both hierarchy builders.  The import must either succeed
// with a valid flat scene or return NULL with a message; any memory error,
// undefined behaviour or crash is a failure.  The scene check is the oracle's
// restatement of the reference's flat-scene validation (oracle/pt_oracle.c,
// flat_validate.h:10-83).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <unistd.h>

#include "pt_oracle.h"
#include "pt_scene.h"

namespace {
char scratch_dir[4000];
void write_file(const char *name, const void *data, size_t size) {
    char path[4096];
    std::snprintf(path, sizeof path, "%s/%s", scratch_dir, name);
    FILE *f = std::fopen(path, "wb");
    if(!f) std::abort();
    if(size && std::fwrite(data, 1, size, f) != size) std::abort();
    std::fclose(f);
}
}  // namespace

extern "C" int LLVMFuzzerInitialize(int *, char ***) {
    const char *dir = std::getenv("PT_FUZZ_DIR");
    std::snprintf(scratch_dir, sizeof scratch_dir, "%s/fuzz_obj_%d", dir ? dir : "/tmp", int(getpid()));
    mkdir(scratch_dir, 0700);
    return 0;
}

extern "C" int LLVMFuzzerTestOneInput(const uint8_t *data, size_t size) {
    const uint8_t *end = data + size, *a = data;
    const uint8_t *b = static_cast<const uint8_t *>(std::memchr(a, 0, size_t(end - a)));
    const uint8_t *obj_end = b ? b : end;
    const uint8_t *mtl = b ? b + 1 : end;
    const uint8_t *c = static_cast<const uint8_t *>(std::memchr(mtl, 0, size_t(end - mtl)));
    const uint8_t *mtl_end = c ? c : end;
    const uint8_t *img = c ? c + 1 : end;
    std::string obj = "mtllib m.mtl\n";
    obj.append(reinterpret_cast<const char *>(a), size_t(obj_end - a));
    write_file("m.obj", obj.data(), obj.size());
    write_file("m.mtl", mtl, size_t(mtl_end - mtl));
    write_file("t.png", img, size_t(end - img));
    for(int32_t split : {PTS_SPLIT_SAH, PTS_SPLIT_REFERENCE}) {
        pts_obj_options opt{{1, 1, 1}, {0, 30, 0}, {0, 0, 0}, 1.5f, split, 7};
        pts_obj_info info;
        char obj_path[4096];
        std::snprintf(obj_path, sizeof obj_path, "%s/m.obj", scratch_dir);
        pts_scene *s = pts_load_obj(obj_path, &opt, &info);
        if(!s) {
            if(!pt_scene_last_error()[0]) {  // a failure must say why
                std::fprintf(stderr, "pts_load_obj failed without a message (split %d)\n", split);
                std::abort();
            }
            continue;
        }
        pt_scene_desc desc;
        pts_desc(s, &desc);
        // pt_scene_desc and po_scene have the same layout (flat_data.h:47-56)
        static_assert(sizeof(pt_scene_desc) == sizeof(po_scene), "scene views differ");
        po_scene view;
        std::memcpy(&view, &desc, sizeof view);
        if(po_validate_scene(&view) != 0) {  // the importer's output is a valid flat scene
            std::fprintf(stderr, "pts_load_obj returned an invalid flat scene (split %d)\n", split);
            std::abort();
        }
        (void)pts_warnings(s);
        pts_free(s);
    }
    return 0;
}
