#!/usr/bin/env bash
# Build oracle/_ref/ from the reference sources where they lie under
# /root/reference (test infrastructure; outputs stay out of git):
#   libptref.so          oracle/ref_shim.cpp over the reference's host path
#   cuda_tests           the reference's tests/cuda_tests.cpp and
#   cuda_lighting_tests  tests/lighting_tests.cpp (RAY_TRACER_HAS_CUDA=1),
#                        linked against this repository's drop-in binding
#                        (cuda-path-tracer_amd/binding/cuda_renderer_hip.cpp)
#                        and libptrace.so, with the reference's test assets
#                        read from tests/golden/assets.
#   preview_session_check  this repository's tests/native/preview_session_check.cpp:
#                        the reference's preview_session<cuda_renderer>
#                        (src/display/preview_session.h) on the same binding.
# Flags mirror the reference Release host build: -O3, no -march, no FMA
# contraction (CMakeLists.txt:19-20 builds device code with --fmad=false).
set -euo pipefail
REF=${PT_REFERENCE_ROOT:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(dirname "$HERE")"
PKG="$ROOT/cuda-path-tracer_amd"
if [ ! -d "$REF/src" ]; then
    echo "reference tree not found at $REF; skipping oracle/_ref" >&2
    exit 0
fi
mkdir -p "$HERE/_ref"
COMMON=(-O3 -DNDEBUG -std=c++17 -ffp-contract=off -I "$REF/src" -I "$REF/tests" -isystem "$REF/third_party")
# The vendored loaders (image and OBJ import) are compiled once per build.
g++ "${COMMON[@]}" -fPIC -c "$REF/third_party/stb_image.cpp" -o "$HERE/_ref/stb_image.o"
g++ "${COMMON[@]}" -fPIC -c "$REF/third_party/tinyobjloader/implementation.cpp" -o "$HERE/_ref/tinyobj.o"
LOADERS=("$HERE/_ref/stb_image.o" "$HERE/_ref/tinyobj.o")
g++ "${COMMON[@]}" -fPIC -shared -DRAY_TRACER_ASSET_DIR="\"$REF/assets\"" \
    "$HERE/ref_shim.cpp" "${LOADERS[@]}" -o "$HERE/_ref/libptref.so.tmp" -lpthread
mv "$HERE/_ref/libptref.so.tmp" "$HERE/_ref/libptref.so"
echo "built $HERE/_ref/libptref.so"
if [ -f "$PKG/libptrace.so" ]; then
    ASSETS="$ROOT/tests/golden/assets"
    LINK=(-I "$ROOT/include" "$PKG/binding/cuda_renderer_hip.cpp" "${LOADERS[@]}" -L "$PKG" -lptrace
          "-Wl,-rpath,\$ORIGIN/../../cuda-path-tracer_amd" -lpthread)
    g++ "${COMMON[@]}" -DRAY_TRACER_ASSET_DIR="\"$ASSETS\"" "$REF/tests/cuda_tests.cpp" "${LINK[@]}" \
        -o "$HERE/_ref/cuda_tests"
    g++ "${COMMON[@]}" -DRAY_TRACER_ASSET_DIR="\"$ASSETS\"" -DRAY_TRACER_HAS_CUDA=1 "$REF/tests/lighting_tests.cpp" \
        "${LINK[@]}" -o "$HERE/_ref/cuda_lighting_tests"
    g++ "${COMMON[@]}" -DRAY_TRACER_ASSET_DIR="\"$ASSETS\"" "$ROOT/tests/native/preview_session_check.cpp" \
        "${LINK[@]}" -o "$HERE/_ref/preview_session_check"
    echo "built $HERE/_ref/cuda_tests $HERE/_ref/cuda_lighting_tests $HERE/_ref/preview_session_check"
fi
