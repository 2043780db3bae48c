#!/bin/sh
# Builds tests/cpp/libfacade_tests.so against the C++ facade and libfpxg.so.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
g++ -std=c++20 -O2 -fPIC -shared -I "$ROOT/include" -o "$HERE/libfacade_tests.so" "$HERE/facade_tests.cpp" \
    -L "$ROOT/fastpoly_amd" -lfpxg -Wl,-rpath,'$ORIGIN/../../fastpoly_amd'
