"""GPU: the device arena that replaces the reference's scratch pool
(memory.hpp:121-258).  Warm calls make no system allocations, cold series
calls make a constant number (tests/test_memory.cpp:654-708 of the
reference), and an allocation failure is FPXG_E_NOMEM / std::bad_alloc with
the caller's buffers untouched (the strong guarantee of
tests/test_memory.cpp:384-400)."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
P = 998244353
Q = 1000000007


@pytest.fixture(scope="module")
def fa():
    import fastpoly_amd
    fastpoly_amd.init(0, 0)
    return fastpoly_amd


def test_warm_calls_make_no_system_allocations(fa, oracle):
    rng = np.random.default_rng(8)
    D = fa.DeviceArray
    n = 1 << 16
    a = D.from_host(rng.integers(0, P, n, dtype=np.uint64).astype(np.uint32))
    s_in = rng.integers(0, P, n, dtype=np.uint64).astype(np.uint32)
    s_in[0] = fa.from_int(P, 1)
    inv_in, log_in = D.from_host(s_in), D.from_host(s_in)
    s_in[0] = 0
    exp_in = D.from_host(s_in)
    sq_in = D(n)
    fa.conv_dev(P, inv_in, n, inv_in, n, sq_in, n)
    out, prod = D(n), D(2 * n - 1)
    host_q = rng.integers(0, Q, 5000, dtype=np.uint64).astype(np.uint32)

    def calls():
        fa.conv_dev(P, a, n, a, n, prod)
        for op, x in (("inv", inv_in), ("log", log_in), ("exp", exp_in), ("sqrt", sq_in)):
            fa.series_dev(op, P, x, n, n, out)
        fa.conv(Q, host_q, host_q[:3000])
        fa.inv(P, fa.from_int(P, np.arange(1, 4098)), 4097)
        fa.synchronize()

    for _ in range(3):  # eager, graph capture, graph replay
        calls()
    before = fa.scratch_memory_stats()
    for _ in range(3):
        calls()
    after = fa.scratch_memory_stats()
    assert after["system_allocations"] == before["system_allocations"], (before, after)
    assert after["cache_hits"] > before["cache_hits"]
    assert after["in_use_bytes"] == before["in_use_bytes"]


def _geometric_cases(fa, n, ratio=7):
    """Closed forms independent of the library (tests/test_memory.cpp:656-672
    of the reference): g = sum r^i x^i; inv(g) = 1 - r x; exp(sum r^i x^i / i)
    = g; g^2 has coefficients (i + 1) r^i."""
    pw = [pow(ratio, i, P) for i in range(n)]
    geo = fa.from_int(P, np.array(pw, dtype=np.uint64))
    sq = fa.from_int(P, np.array([(i + 1) * pw[i] % P for i in range(n)], dtype=np.uint64))
    ex = fa.from_int(P, np.array([0] + [pw[i] * pow(i, P - 2, P) % P for i in range(1, n)], dtype=np.uint64))
    inv_want = np.zeros(n, np.uint32)
    inv_want[0], inv_want[1] = fa.from_int(P, 1), fa.from_int(P, P - ratio)
    return geo, sq, ex, inv_want


@pytest.mark.parametrize("n", [4097, 4105])
def test_cold_series_allocations_are_constant(fa, n):
    geo, sq, ex, inv_want = _geometric_cases(fa, n)
    fa.set_graphs(False)  # captured graphs hold arena blocks; drop them
    try:
        for op, x, want, most in (("inv", geo, inv_want, 2), ("exp", ex, geo, 3), ("sqrt", sq, None, 2)):
            fa.release_scratch_memory()
            before = fa.scratch_memory_stats()
            got = fa.series(op, P, x, n)
            after = fa.scratch_memory_stats()
            assert after["system_allocations"] - before["system_allocations"] <= most, (op, before, after)
            if want is not None:
                assert np.array_equal(got, want), op
            else:  # the root is g or -g (modint.hpp:139-166 picks the sign)
                sign = fa.to_int(P, int(got[0]))
                assert sign in (1, P - 1)
                want = geo if sign == 1 else fa.from_int(P, (P - fa.to_int(P, geo).astype(np.int64)) % P)
                assert np.array_equal(got, want)
    finally:
        fa.set_graphs(True)


def test_tiny_arena_fails_cleanly():
    """Separate process: its library is initialised with an 8 MiB arena."""
    env = dict(os.environ)
    r = subprocess.run([sys.executable, os.path.join(HERE, "arena_child.py")], capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, r.stderr[-3000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["capacity"] == 8 << 20
    assert out["conv_rc"] == 3 and out["conv_untouched"], out  # FPXG_E_NOMEM
    assert out["inv_rc"] == 3 and out["inv_untouched"], out
    assert out["small_conv_ok"], out
    assert out["facade_fails"] == 0, out["facade_report"]
