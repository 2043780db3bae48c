"""GPU: series calls replayed from the graph cache give, on every replay, the
same limbs as the eager path and the oracle, with other work between replays,
new data under the same key, more keys than the cache holds, and after the
cache is dropped."""
import numpy as np
import pytest

import golden_inputs as gi

pytestmark = pytest.mark.gpu

P = 998244353
Q = 1000000007


@pytest.fixture(scope="module")
def fa():
    import fastpoly_amd
    fastpoly_amd.init(0, 0)
    return fastpoly_amd


def _input(oracle, op, p, n, seed):
    a = oracle.MT(seed).rnd(p, n)
    if op == "exp":
        a[0] = 0
    elif op in ("log", "sqrt"):
        a[0] = oracle.from_int(p, 1)
    elif a[0] == 0:
        a[0] = 1
    return a


@pytest.mark.parametrize("p", [P, Q])
@pytest.mark.parametrize("op", ["inv", "log", "exp", "sqrt"])
def test_replays_match_oracle(fa, oracle, p, op):
    n = 3000
    D = fa.DeviceArray
    a = _input(oracle, op, p, n, 11)
    want = oracle.series(op, p, a, n)
    da, out = D.from_host(a), D(n)
    x = D.from_host(_input(oracle, "inv", p, 500, 12))
    y = D(1000)
    for rep in range(4):
        fa.series_dev(op, p, da, n, n, out)
        assert np.array_equal(out.to_host(), want), (op, p, rep)
        # eager work between replays, touching other buffers
        fa.conv_dev(p, x, 500, x, 500, y)
        fa.synchronize()


@pytest.mark.parametrize("op", ["exp", "sqrt"])
def test_large_step_replays(fa, oracle, op):
    """Newton steps of 2^14 points and more (split products): eager first
    call, captured second call, replays with new data under the same key,
    then the same without graphs."""
    n = 70000
    D = fa.DeviceArray
    da, out = D(n), D(n)
    try:
        for graphs in (True, False):
            fa.set_graphs(graphs)
            for seed in (31, 32, 33, 34):
                a = _input(oracle, op, P, n, seed)
                assert fa.lib().fpxg_memcpy_h2d(da.ptr, a.ctypes.data, 4 * n) == 0
                fa.series_dev(op, P, da, n, n, out)
                assert np.array_equal(out.to_host(), oracle.series(op, P, a, n)), (op, graphs, seed)
    finally:
        fa.set_graphs(True)


def test_same_key_new_data(fa, oracle):
    """New input data under the same pointers and scan signature replays the
    cached graph, which must read the new data."""
    D = fa.DeviceArray
    for op, n in (("inv", 5000), ("exp", 4000), ("log", 3000)):
        da, out = D(n), D(n)
        for seed in (21, 22, 23):
            a = _input(oracle, op, P, n, seed)
            if op == "inv":
                a[0] = 987654
            assert fa.lib().fpxg_memcpy_h2d(da.ptr, a.ctypes.data, 4 * n) == 0
            fa.series_dev(op, P, da, n, n, out)
            assert np.array_equal(out.to_host(), oracle.series(op, P, a, n)), (op, seed)


def test_same_key_new_constant_term(fa, oracle):
    """The replay key knows only whether a[0] is 0, 1 or other; a replayed
    inverse with another a[0] must still use that a[0] (regression: a schedule
    that baked 1/a[0] in on the host)."""
    D = fa.DeviceArray
    for n in (40, 3000, 70000):
        da, out = D(n), D(n)
        for seed, a0 in ((41, 12345), (42, 987654), (43, 5), (44, 998244352)):
            a = _input(oracle, "inv", P, n, seed)
            a[0] = oracle.from_int(P, a0)
            assert fa.lib().fpxg_memcpy_h2d(da.ptr, a.ctypes.data, 4 * n) == 0
            fa.series_dev("inv", P, da, n, n, out)
            assert np.array_equal(out.to_host(), oracle.series("inv", P, a, n)), (n, a0)


def test_many_keys_and_release(fa, oracle):
    D = fa.DeviceArray
    outs = []
    for i in range(80):  # more keys than the cache keeps
        n = 40 + 7 * i
        a = _input(oracle, "inv", P, n, 100 + i)
        da, out = D.from_host(a), D(n)
        fa.series_dev("inv", P, da, n, n, out)
        outs.append((a, n, da, out))
    for a, n, da, out in outs:
        assert np.array_equal(out.to_host(), oracle.series("inv", P, a, n)), n
    fa.release_scratch_memory()
    for a, n, da, out in outs[:5]:
        fa.series_dev("inv", P, da, n, n, out)
        assert np.array_equal(out.to_host(), oracle.series("inv", P, a, n)), n


def test_autotune_keeps_results_exact(fa, oracle):
    """After the autotuner has picked pass splits, transforms and products
    still give the oracle's limbs."""
    plans = fa.autotune(10, 18, 2)
    assert set(plans) == set(range(10, 19))
    for lg, (cols, us) in plans.items():
        assert 0 <= cols <= 12 and us > 0
    for lg in (12, 13, 15, 17):
        n = 1 << lg
        x = oracle.MT(1 + n).rnd(167772161, n)
        assert np.array_equal(fa.ntt_forward(167772161, x, root=3), oracle.ntt(167772161, x, root=3)), lg
    # the fused products' own splits (tuned from 2^13 up)
    for lg in range(13, 19):
        pp = fa.product_plan_info(1 << lg)
        assert pp["row_log"] + pp["col_levels"] == lg and pp["row_log"] <= 14, (lg, pp)
        assert pp["col_levels"] == 0 or pp["col_levels"] >= 4, (lg, pp)
        # one workgroup's product (no column levels) only up to 2^13 points
        assert lg <= 13 or pp["col_levels"] >= 4, (lg, pp)
    a = oracle.MT(5).rnd(P, 70000)
    b = oracle.MT(6).rnd(P, 50001)
    assert np.array_equal(fa.conv(P, a, b), oracle.conv(P, a, b))
    for la, lb in ((3000, 2000), (5000, 4000), (9000, 9000), (40000, 30000)):
        assert np.array_equal(fa.conv(P, a[:la], b[:lb]), oracle.conv(P, a[:la], b[:lb])), (la, lb)
    for op in ("exp", "inv", "sqrt", "log"):
        s = _input(oracle, op, P, 40000, 9)
        assert np.array_equal(fa.series(op, P, s, 40000), oracle.series(op, P, s, 40000)), op


@pytest.mark.parametrize("op", ["inv", "log", "exp", "sqrt"])
def test_replays_with_changing_scan(fa, oracle, op):
    """One call site (same buffers) whose input alternates between scan
    signatures: the graph cache must pick the schedule of the current scan
    (fewer nonzero coefficients, a leading zero for sqrt, another leading
    coefficient for sqrt) and raise the domain error with the output
    untouched for inputs that violate the domain, on every call."""
    n = 20000
    D = fa.DeviceArray
    da, out = D(n), D(n)
    base = _input(oracle, op, P, n, 61)

    def variant(kind):
        a = base.copy()
        if kind == "short":  # a shorter polynomial: other last nonzero index
            a[n // 3:] = 0
        elif kind == "bad":  # domain error for every op
            if op == "inv":
                a[0] = 0
            elif op == "exp":
                a[0] = oracle.from_int(P, 1)
            elif op == "log":
                a[0] = oracle.from_int(P, 2)
            else:
                a[0] = 0
                a[1] = oracle.from_int(P, 1)  # odd valuation
        elif kind == "shift" and op == "sqrt":  # valuation 2: another schedule
            a[2:] = base[:-2]
            a[:2] = 0
        elif kind == "lead" and op == "sqrt":  # another square root taken on the host
            a[0] = oracle.from_int(P, 4)
        return a

    kinds = ["base", "base", "base", "short", "base", "bad", "base", "short", "short", "short", "bad", "base",
             "base", "shift", "lead", "base", "lead", "lead", "base"]
    sentinel = np.full(n, 7, dtype=np.uint32)
    for i, kind in enumerate(kinds):
        a = variant(kind)
        assert fa.lib().fpxg_memcpy_h2d(da.ptr, a.ctypes.data, 4 * n) == 0
        if kind == "bad":
            assert fa.lib().fpxg_memcpy_h2d(out.ptr, sentinel.ctypes.data, 4 * n) == 0
            with pytest.raises(fa.DomainError):
                fa.series_dev(op, P, da, n, n, out)
            assert np.array_equal(out.to_host(), sentinel), (op, i, kind)
            continue
        fa.series_dev(op, P, da, n, n, out)
        assert np.array_equal(out.to_host(), oracle.series(op, P, a, n)), (op, i, kind)
