"""Host builds under AddressSanitizer and UndefinedBehaviorSanitizer (test
infrastructure), all with the ROCm toolchain's clang so that one sanitizer
runtime serves them:

- libptscene_asan.so: the scene library (BVH builders, OBJ/MTL import, PNG,
  PNM and JPEG decoding, PPM output), run from Python under LD_PRELOAD of the
  runtime (runtime_library());
- liboracle_asan.so: the CPU oracle, with its own arithmetic flags;
- libwidemodel_asan.so: the CPU model of the traversal (build_model.py);
- fuzz_image, fuzz_obj: libFuzzer targets for the texture decoders and the
  OBJ/MTL import (fuzz_image.cpp, fuzz_obj.cpp).

``python tests/native/build_sanitized.py`` builds them all.
"""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "cuda-path-tracer_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")
ORACLE = os.path.join(ROOT, "oracle")
LLVM = os.environ.get("PT_LLVM_BIN", "/opt/rocm/llvm/bin")
# With this toolchain AddressSanitizer registers string literals twice (and
# reports them as misaligned duplicates) when UndefinedBehaviorSanitizer's
# runtime checks are built non-recoverable or linked into a libFuzzer binary.
# So the libraries keep recoverable checks and halt at run time
# (UBSAN_OPTIONS=halt_on_error=1, set by the tests), and the fuzzers compile
# the checks as traps, which libFuzzer reports as crashes with their stack.
SANITIZE = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-g"]
FUZZ_SANITIZE = ["-fsanitize=address,undefined", "-fsanitize-trap=undefined", "-fno-omit-frame-pointer", "-g"]
RUN_ENV = {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=1", "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}
SCENE_SOURCES = [os.path.join(CSRC, n) for n in ("pt_scene.cpp", "pt_obj.cpp", "pt_image.cpp")]


def runtime_library():
    """The clang ASan runtime a Python process preloads to run the *_asan.so libraries."""
    found = sorted(glob.glob(os.path.join(os.path.dirname(LLVM), "lib", "clang", "*", "lib", "linux",
                                          "libclang_rt.asan-x86_64.so")))
    if not found:
        raise RuntimeError("clang AddressSanitizer runtime not found under " + os.path.dirname(LLVM))
    return found[-1]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError("sanitized build failed: " + " ".join(cmd[:3]))


def _link(out, cmd):
    tmp = out + ".tmp"
    _run(cmd + ["-o", tmp])
    os.replace(tmp, out)
    return out


def build_scene():
    return _link(os.path.join(HERE, "libptscene_asan.so"),
                 [os.path.join(LLVM, "clang++"), "-O1", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                  *SANITIZE, "-shared-libsan", "-I", INCLUDE, *SCENE_SOURCES])


def build_oracle():
    return _link(os.path.join(HERE, "liboracle_asan.so"),
                 [os.path.join(LLVM, "clang"), "-O1", "-std=c11", "-ffp-contract=off", "-fPIC", "-shared",
                  *SANITIZE, "-shared-libsan", os.path.join(ORACLE, "pt_oracle.c"), "-lm", "-lpthread"])


def build_fuzzer(name):
    """A libFuzzer target: the harness plus the scene sources (and, for the OBJ
    target, the oracle's scene validation), each compiled with coverage
    instrumentation."""
    objs = []
    cxx = [os.path.join(LLVM, "clang++"), "-O1", "-std=c++17", "-ffp-contract=off", "-fsanitize=fuzzer-no-link",
           *FUZZ_SANITIZE, "-I", INCLUDE, "-I", ORACLE, "-c"]
    units = [(os.path.join(HERE, name + ".cpp"), cxx)]
    units += [(src, cxx) for src in SCENE_SOURCES]
    if name == "fuzz_obj":
        units.append((os.path.join(ORACLE, "pt_oracle.c"),
                      [os.path.join(LLVM, "clang"), "-O1", "-std=c11", "-ffp-contract=off",
                       "-fsanitize=fuzzer-no-link", *FUZZ_SANITIZE, "-c"]))
    for k, (src, cmd) in enumerate(units):
        obj = os.path.join(HERE, f"_{name}_{k}.o")
        _run(cmd + [src, "-o", obj])
        objs.append(obj)
    try:
        return _link(os.path.join(HERE, name), [os.path.join(LLVM, "clang++"), "-fsanitize=fuzzer", *FUZZ_SANITIZE,
                                                *objs, "-lpthread"])
    finally:
        for obj in objs:
            os.remove(obj)


def build_all():
    import build_model
    jobs = [build_scene, build_oracle, lambda: build_fuzzer("fuzz_image"), lambda: build_fuzzer("fuzz_obj"),
            lambda: build_model.build(sanitize=True)]
    with ThreadPoolExecutor(len(jobs)) as ex:
        return list(ex.map(lambda f: f(), jobs))


if __name__ == "__main__":
    sys.path.insert(0, HERE)
    for path in build_all():
        print("built", path)
