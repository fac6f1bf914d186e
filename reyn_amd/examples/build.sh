#!/bin/bash
# Builds the C++ host-API example and test program against libreyn_amd.so
# into reyn_amd/bin/ (plain g++, C++20; the HIP code lives in the library).
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
root="$(cd "$here/../.." && pwd)"
mkdir -p "$root/reyn_amd/bin"
flags=(-std=c++20 -O2 -Wall -Wextra -I"$root/include" -L"$root/reyn_amd" -lreyn_amd "-Wl,-rpath,\$ORIGIN/..")
g++ "$here/dambreak.cpp" -o "$root/reyn_amd/bin/dambreak" "${flags[@]}"
g++ "$here/scene_main.cpp" -o "$root/reyn_amd/bin/scene_main" "${flags[@]}"
g++ "$root/tests/cpp/test_host_api.cpp" -o "$root/reyn_amd/bin/test_host_api" "${flags[@]}"
