#!/bin/bash
# Runs the tests against the host code built with AddressSanitizer and UBSan.
#   check_host.sh [pytest args]        CPU tests; host sources built with g++
#                                      (make asan)
#   check_host.sh --gpu [pytest args]  GPU tests; every source's host half
#                                      built with the sanitizers, device
#                                      code unchanged (make asan-hip)
# python is not linked with a sanitizer runtime, so it is preloaded, with
# libstdc++ for its exception interceptor; any report fails the run.
set -e
here="$(cd "$(dirname "$0")" && pwd)"
repo="$(cd "$here/../.." && pwd)"
if [ "$1" = "--gpu" ]; then
    shift
    target=asan-hip
    lib="$here/_build/asan_hip/libwavey_fdtd.so"
    san_rt="$(g++ -print-file-name=libasan.so)"
    marker=gpu
else
    target=asan
    lib="$here/_build/asan/libwavey_fdtd.so"
    san_rt="$(g++ -print-file-name=libasan.so)"
    marker="not gpu"
fi
make -C "$here" -j8 "$target"
logs="$(dirname "$lib")/logs"
rm -rf "$logs" && mkdir -p "$logs"
cxx_rt="$(g++ -print-file-name=libstdc++.so.6)"
set +e
WAVEY_FDTD_LIB="$lib" \
ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0:log_path=$logs/asan" \
UBSAN_OPTIONS="print_stacktrace=1:log_path=$logs/ubsan" \
LD_PRELOAD="$san_rt:$cxx_rt${LD_PRELOAD:+:$LD_PRELOAD}" \
    python3 -m pytest "$repo/tests" -q -m "$marker" -p no:cacheprovider "$@"
rc=$?
set -e
if ls "$logs"/* > /dev/null 2>&1; then
    echo "sanitizer reports:"; head -40 "$logs"/*
    exit 1
fi
exit $rc
