#!/usr/bin/env bash
# Runs the CPU test suite (pytest -m "not gpu") against host code built with
# AddressSanitizer + UndefinedBehaviorSanitizer: the C ABI and engine host code
# of libvship_amd.so (hipcc, sanitizers on the host side only), the FP64
# oracle, the mock VapourSynth host and the throughput driver, all with the
# ROCm clang so that one sanitizer runtime serves the whole process. The build
# happens in a scratch copy of the tree, so the in-tree libraries are left as
# they are. CPU only: nothing here runs on a GPU.
set -euo pipefail
SRC=$(cd "$(dirname "$0")/../.." && pwd)
WORK=${VSHIP_SANITIZE_DIR:-/tmp/vship_amd_sanitize}
LLVM=/opt/rocm/llvm/bin
RT=$("$LLVM/clang" -print-file-name=libclang_rt.asan-x86_64.so)
SAN="-fsanitize=address,undefined -fno-omit-frame-pointer -fno-sanitize-recover=undefined"
rm -rf "$WORK"
mkdir -p "$WORK"
# the committed tree (its scripts' outputs and run results stay behind)
git -C "$SRC" ls-files -z | tar -C "$SRC" --null -T - -cf - | tar -C "$WORK" -xf -
JOBS=$(nproc)
make -s -j "$JOBS" -C "$WORK/vship_amd/csrc" \
    EXTRA_HIPFLAGS="-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-omit-frame-pointer" \
    DRIVER_CXX="$LLVM/clang++" EXTRA_CXXFLAGS="$SAN -shared-libsan"
make -s -C "$WORK/oracle" CC="$LLVM/clang" EXTRA_CFLAGS="$SAN"
make -s -C "$WORK/tests/vs_host" CXX="$LLVM/clang++" EXTRA_CXXFLAGS="$SAN"
cd "$WORK"
# Python itself is not instrumented: its allocations are not leak-checked
export ASAN_OPTIONS=detect_leaks=0:verify_asan_link_order=0:abort_on_error=0:halt_on_error=1
export UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1
export LD_LIBRARY_PATH="$(dirname "$RT")${LD_LIBRARY_PATH:+:$LD_LIBRARY_PATH}"
LD_PRELOAD="$RT" python3 -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@"
