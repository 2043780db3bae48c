#!/bin/sh
# ASan + UBSan build of the library's host code (the gfx950 device code is
# built as usual: every -fsanitize= goes to the host pass only) and of
# tests/cpp/san_driver.cpp, into tests/cpp/_san/.
#   sh tests/cpp/build_san.sh && tests/cpp/_san/san_driver
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
OUT="$HERE/_san"
mkdir -p "$OUT"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
# Every translation unit of the library (fpxg.hip and the launch_*.hip kernel
# tables), compiled in parallel, then linked.
OBJS=""
PIDS=""
for SRC in "$ROOT"/fastpoly_amd/csrc/*.hip; do
    OBJ="$OUT/$(basename "$SRC" .hip).o"
    "$HIPCC" --offload-arch=gfx950 -O1 -g -std=c++17 -fPIC -c \
        -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-omit-frame-pointer \
        -Xclang -target-feature -Xclang -packed-fp32-ops \
        -I "$ROOT/include" -o "$OBJ" "$SRC" &
    PIDS="$PIDS $!"
    OBJS="$OBJS $OBJ"
done
for P in $PIDS; do wait "$P"; done
"$HIPCC" --offload-arch=gfx950 -shared -fPIC -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined \
    -o "$OUT/libfpxg_san.so" $OBJS
# The driver is plain C++ with no device code: clang++ (the compiler whose
# sanitizer runtime the library's host code uses), not hipcc.
CLANGXX=${CLANGXX:-/opt/rocm/llvm/bin/clang++}
"$CLANGXX" -O1 -g -std=c++17 -fsanitize=address,undefined -fno-omit-frame-pointer -I "$ROOT/include" \
    -o "$OUT/san_driver" "$HERE/san_driver.cpp" -L "$OUT" -lfpxg_san -Wl,-rpath,'$ORIGIN'
