"""Build tests/native/libwidemodel.so, the CPU model of the traversal (test
infrastructure).  It compiles the renderer's own traversal and layout headers
(cuda-path-tracer_amd/csrc/pt_traverse.h, pt_layout.h) for the host with the
renderer's arithmetic flags: no floating-point contraction, IEEE division and
square root, denormals kept.

``python tests/native/build_model.py [--sanitize]``: with --sanitize the host
code is built with AddressSanitizer and UndefinedBehaviorSanitizer into
libwidemodel_asan.so (run it under LD_PRELOAD of the compiler's ASan runtime).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "cuda-path-tracer_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build(sanitize=False):
    name = "libwidemodel_asan.so" if sanitize else "libwidemodel.so"
    out = os.path.join(HERE, name)
    tmp = out + ".tmp"
    flags = ["--offload-arch=gfx950", "-O2" if sanitize else "-O3", "-std=c++17", "-fPIC", "-shared",
             "-ffp-contract=off", "-Wall", "-Wno-unused-function", "-I", INCLUDE, "-I", CSRC]
    if sanitize:
        # UndefinedBehaviorSanitizer halts at run time (UBSAN_OPTIONS=halt_on_error=1)
        flags += ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-g"]
    proc = subprocess.run([HIPCC, *flags, os.path.join(HERE, "wide_model.hip"), "-o", tmp],
                          stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    lines = [l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if proc.returncode != 0:
        sys.stderr.write("\n".join(lines) + "\n")
        raise RuntimeError("building the traversal model failed")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print("built", build(sanitize="--sanitize" in sys.argv))
