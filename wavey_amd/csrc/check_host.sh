#!/bin/bash
# Runs the CPU tests against the host code built with AddressSanitizer and
# UBSan (make asan). python is not linked with the sanitizer runtime, so it
# is preloaded, with libstdc++ for its exception interceptor; any report
# fails the run. Usage: wavey_amd/csrc/check_host.sh [pytest args]
set -e
here="$(cd "$(dirname "$0")" && pwd)"
repo="$(cd "$here/../.." && pwd)"
make -C "$here" -j8 asan
logs="$here/_build/asan/logs"
rm -rf "$logs" && mkdir -p "$logs"
asan_rt="$(g++ -print-file-name=libasan.so)"
cxx_rt="$(g++ -print-file-name=libstdc++.so.6)"
set +e
WAVEY_FDTD_LIB="$here/_build/asan/libwavey_fdtd.so" \
ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0:log_path=$logs/asan" \
UBSAN_OPTIONS="print_stacktrace=1:log_path=$logs/ubsan" \
LD_PRELOAD="$asan_rt:$cxx_rt${LD_PRELOAD:+:$LD_PRELOAD}" \
    python3 -m pytest "$repo/tests" -q -m "not gpu" -p no:cacheprovider "$@"
rc=$?
set -e
if ls "$logs"/* > /dev/null 2>&1; then
    echo "sanitizer reports:"; head -40 "$logs"/*
    exit 1
fi
exit $rc
