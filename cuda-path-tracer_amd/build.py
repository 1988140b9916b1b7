"""Build the native libraries of the path-tracing core in-tree.

libptrace.so  -- HIP kernels + C ABI (include/pt_render.h, include/pt_denoise.h), gfx950 only.
libptscene.so -- host-only BVH builders, OBJ/MTL and image import, PPM output and
                 synthetic scenes (include/pt_scene.h).
pt_render     -- the command-line renderer (the reference application's flags).

Run as ``python cuda-path-tracer_amd/build.py``.  Every HIP translation unit is
compiled without FP contraction (the reference builds with --fmad=false,
CMakeLists.txt:19-20), with correctly rounded fp32 division and square root,
with fp32 denormals kept, and with packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, cwd=None):
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 target feature; that
    # warning is expected and dropped from the log.
    lines = [l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if proc.returncode != 0:
        sys.stderr.write("\n".join(lines) + "\n")
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return lines


def build_ptrace():
    out = os.path.join(HERE, "libptrace.so")
    objs = []
    for name in ("pt_kernels.hip", "pt_api.hip", "pt_denoise.hip", "pt_lbvh.hip"):
        obj = os.path.join(HERE, "_build", name + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        _run([HIPCC, *HIP_FLAGS, "-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, name), "-o", obj])
        objs.append(obj)
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp, "-lpthread"])
    os.replace(tmp, out)
    return out


def build_scene():
    out = os.path.join(HERE, "libptscene.so")
    tmp = out + ".tmp"
    sources = [os.path.join(CSRC, n) for n in ("pt_scene.cpp", "pt_obj.cpp", "pt_image.cpp")]
    _run(["g++", "-O3", "-DNDEBUG", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-Wall", "-Wextra",
          "-Wno-unused-function", "-I", INCLUDE, *sources, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_cli():
    """pt_render: the command-line renderer over both libraries.  Without
    --obj it loads models/blender_sphere/sphere.obj from the asset directory
    named by PT_ASSET_DIR at build time ("assets", relative to the working
    directory, when unset), as the reference's RAY_TRACER_ASSET_DIR."""
    out = os.path.join(HERE, "pt_render")
    tmp = out + ".tmp"
    asset_dir = os.environ.get("PT_ASSET_DIR", "assets")
    if '"' in asset_dir or "\\" in asset_dir:
        raise ValueError("PT_ASSET_DIR must not contain quotes or backslashes")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", INCLUDE, f'-DPT_ASSET_DIR="{asset_dir}"',
          os.path.join(CSRC, "pt_cli.cpp"),
          "-L", HERE, "-lptscene", "-lptrace", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_all():
    return [build_scene(), build_ptrace(), build_cli()]


if __name__ == "__main__":
    for path in build_all():
        print("built", path)
