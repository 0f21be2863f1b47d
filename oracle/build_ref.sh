#!/usr/bin/env bash
# Compiles the reference libvship (single translation unit src/VshipLib.cpp,
# with its vendored VapourSynth headers) for gfx950 into oracle/_ref/.
# Used only as a checker and for side-by-side timing in tests; nothing from
# the reference tree is copied into this repository, only this build output,
# which .gitignore keeps out of history.
set -euo pipefail
REF=${VSHIP_REFERENCE_TREE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$REF/src/VshipLib.cpp" ]; then
    echo "reference tree not present at $REF; skipping reference build" >&2
    exit 0
fi
mkdir -p "$OUT"
/opt/rocm/bin/hipcc -std=c++17 -O3 -Wno-ignored-attributes -Wno-unused-result \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    -I "$REF/include" --offload-arch=gfx950 -shared -fPIC \
    "$REF/src/VshipLib.cpp" -o "$OUT/libvship_ref.so.tmp"
mv "$OUT/libvship_ref.so.tmp" "$OUT/libvship_ref.so"
echo "built $OUT/libvship_ref.so"
