"""Build libfpxg.so (hipcc, gfx950) in-tree.

    python -m fastpoly_amd.build        # or __graft_entry__.build()
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libfpxg.so")
SOURCES = [os.path.join(HERE, "csrc", "fpxg.hip")]
DEPS = SOURCES + [os.path.join(HERE, "csrc", f) for f in ("device.hpp", "crt.hpp", "ntt2.hpp", "runtime.hpp", "series.hpp", "records.hpp", "ew.hpp", "batch.hpp", "serve.hpp", "graphs.hpp", "ops.hpp")] \
    + [os.path.join(ROOT, "include", "fpxg.h")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 forms are kept out of every kernel (none are needed: integer code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(force: bool = True, verbose: bool = False) -> str:
    if not force and os.path.exists(LIB):
        lib_m = os.path.getmtime(LIB)
        if all(os.path.getmtime(d) <= lib_m for d in DEPS):
            return LIB
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + SOURCES
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libfpxg.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
