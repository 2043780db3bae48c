"""Child process of tests/test_gpu_arena.py: a library whose device arena is
capped at 8 MiB (fpxg_init), so allocation failure can be shown without
touching the arena of the test process.  Prints one JSON line."""
import ctypes
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import fastpoly_amd as fa  # noqa: E402

P = 998244353
CAP = 8 << 20


def main():
    fa.init(0, CAP)
    L = fa.lib()
    out = {"capacity": fa.scratch_memory_stats()["capacity_bytes"]}
    rng = np.random.default_rng(3)
    a = rng.integers(0, P, 1 << 21, dtype=np.uint64).astype(np.uint32)
    sentinel = np.full(1 << 22, 0xDEADBEEF, np.uint32)
    got = ctypes.c_size_t(0)
    rc = L.fpxg_conv_u32(P, fa._ptr(a), a.size, fa._ptr(a), a.size, fa.SIZE_MAX, fa._ptr(sentinel),
                         sentinel.size, ctypes.byref(got))
    out["conv_rc"] = rc
    out["conv_untouched"] = bool((sentinel == 0xDEADBEEF).all())
    a[0] = 1
    rc = L.fpxg_series_u32(fa.OPS["inv"], P, fa._ptr(a), a.size, a.size, fa._ptr(sentinel))
    out["inv_rc"] = rc
    out["inv_untouched"] = bool((sentinel == 0xDEADBEEF).all())
    x = rng.integers(0, P, 3000, dtype=np.uint64).astype(np.uint32)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    out["small_conv_ok"] = bool(np.array_equal(fa.conv(P, x, x[:1000]), po.conv(P, x, x[:1000])))
    lib = ctypes.CDLL(os.path.join(ROOT, "tests", "cpp", "libfacade_tests.so"))
    lib.fpx_facade_nomem_tests.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    buf = ctypes.create_string_buffer(1 << 14)
    out["facade_fails"] = lib.fpx_facade_nomem_tests(buf, len(buf))
    out["facade_report"] = buf.value.decode()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
