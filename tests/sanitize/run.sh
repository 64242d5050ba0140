#!/bin/bash
# CPU sanitizer runs of the host code (no GPU): builds tests/sanitize/_build
# (Makefile beside this script; the product library must be built first, its
# class-kernel objects are linked in) and runs
#   1. the whole CPU suite (pytest -m "not gpu") on the AddressSanitizer +
#      UndefinedBehaviorSanitizer builds of the engine library and of the
#      kernel-body emulation;
#   2. the emulation tests (which run the quartet planner on 1-16 threads) on
#      the ThreadSanitizer build.
# Any sanitizer report ends the run with a non-zero status.
set -euo pipefail
ROOT="$(cd "$(dirname "$0")/../.." && pwd)"
cd "$ROOT"
make -j8 -C tests/sanitize
RT="$(/opt/rocm/bin/amdclang++ -print-resource-dir)/lib/linux"
B="$ROOT/tests/sanitize/_build"
echo "== AddressSanitizer + UndefinedBehaviorSanitizer: CPU suite"
HFX_LIBRARY="$B/asan/libhfx.so" HFX_EMU_LIBRARY="$B/asan/libhfx_emu.so" \
ASAN_OPTIONS=detect_leaks=0:halt_on_error=1 UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1 \
LD_PRELOAD="$RT/libclang_rt.asan-x86_64.so" \
    python -m pytest tests -m "not gpu" -x -q -p no:cacheprovider
echo "== ThreadSanitizer: emulation and planner threads"
HFX_EMU_LIBRARY="$B/tsan/libhfx_emu.so" TSAN_OPTIONS=halt_on_error=1 \
LD_PRELOAD="$RT/libclang_rt.tsan-x86_64.so" \
    python -m pytest tests/test_emu.py -x -q -p no:cacheprovider
