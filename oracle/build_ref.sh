#!/usr/bin/env bash
# Build oracle/_ref/libptref.so: oracle/ref_shim.cpp compiled against the
# reference's own host headers where they lie under /root/reference.
# Test infrastructure only; the output stays out of git (oracle/_ref/).
# Flags mirror the reference Release host build: -O3, no -march, no FMA
# contraction (CMakeLists.txt:19-20 builds device code with --fmad=false).
set -euo pipefail
REF=${PT_REFERENCE_ROOT:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
if [ ! -d "$REF/src" ]; then
    echo "reference tree not found at $REF; skipping oracle/_ref" >&2
    exit 0
fi
mkdir -p "$HERE/_ref"
g++ -O3 -DNDEBUG -std=c++17 -ffp-contract=off -fPIC -shared \
    -I "$REF/src" -I "$REF/tests" -isystem "$REF/third_party" \
    -DRAY_TRACER_ASSET_DIR="\"$REF/assets\"" \
    "$HERE/ref_shim.cpp" "$REF/third_party/stb_image.cpp" \
    "$REF/third_party/tinyobjloader/implementation.cpp" -o "$HERE/_ref/libptref.so.tmp" -lpthread
mv "$HERE/_ref/libptref.so.tmp" "$HERE/_ref/libptref.so"
echo "built $HERE/_ref/libptref.so"
