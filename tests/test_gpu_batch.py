"""GPU: merged batches give, request by request, exactly what the single calls
give (and what the oracle gives), including error statuses, mixed moduli,
uneven sizes and the device-pointer submit/wait form."""
import ctypes

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


def _single(fa, r):
    try:
        if r.op == "conv":
            return fa.conv(r.p, r.a, r.b, None if r.n == fa.SIZE_MAX else r.n)
        return fa.series(r.op, r.p, r.a, r.n)
    except Exception as e:  # the error the single call raises
        return e


def _same(x, y):
    if isinstance(x, Exception) or isinstance(y, Exception):
        return type(x) is type(y)
    return np.array_equal(x, y)


def _mixed_requests(fa, po, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for p in (P, Q, 167772161, 7340033, 65537, 2, 3):
        for la, lb in ((1, 1), (5, 300), (64, 64), (1000, 3000), (4096, 4097), (20000, 9000), (70000, 1)):
            a = rng.integers(0, p, la, dtype=np.uint32)
            b = rng.integers(0, p, lb, dtype=np.uint32)
            reqs.append(fa.Request("conv", p, a, b))
            reqs.append(fa.Request("conv", p, a, b, n=max(1, (la + lb) // 3)))
    for p in (P, Q, 167772161, 65537):
        for n in (1, 2, 9, 33, 100, 1000, 5000, 40000):
            one = po.from_int(p, 1)
            a = rng.integers(0, p, n, dtype=np.uint32)
            ai = a.copy(); ai[0] = int(rng.integers(1, p))
            al = a.copy(); al[0] = one
            ae = a.copy(); ae[0] = 0
            reqs += [fa.Request("inv", p, ai, n=n), fa.Request("log", p, al, n=n),
                     fa.Request("exp", p, ae, n=n), fa.Request("sqrt", p, al, n=n)]
    # Error cases: each must fail alone, exactly as the single call does.
    z = np.zeros(10, np.uint32)
    reqs += [fa.Request("inv", P, z, n=5), fa.Request("log", P, z, n=5), fa.Request("exp", P,
             np.full(4, 7, np.uint32), n=4), fa.Request("conv", 1000000000, z, z),
             fa.Request("sqrt", P, np.array([0, 5, 1], np.uint32), n=4),
             fa.Request("sqrt", 2, np.array([1, 1], np.uint32), n=2),
             fa.Request("sqrt", 2, np.array([1, 0, 1], np.uint32), n=3),
             fa.Request("log", 65537, gi.series_input(po, 65537, 70000), n=70000),
             fa.Request("conv", P, np.zeros(0, np.uint32), z)]
    return reqs


def test_batch_matches_single_calls(fa, oracle):
    reqs = _mixed_requests(fa, oracle, 1)
    got = fa.batch(reqs)
    for r, g in zip(reqs, got):
        want = _single(fa, r)
        assert _same(g, want), (r.op, r.p, r.a.size, None if r.b is None else r.b.size, r.n, g, want)


def test_batch_matches_oracle(fa, oracle):
    rng = np.random.default_rng(2)
    reqs, wants = [], []
    for i in range(40):
        p = (P, Q)[i % 2]
        la, lb = int(rng.integers(1, 3000)), int(rng.integers(1, 3000))
        a = rng.integers(0, p, la, dtype=np.uint32)
        b = rng.integers(0, p, lb, dtype=np.uint32)
        reqs.append(fa.Request("conv", p, a, b))
        wants.append(oracle.conv(p, a, b))
    for i, op in enumerate(("inv", "log", "exp", "sqrt") * 4):
        p = (P, Q)[i % 2]
        n = 100 + 97 * i
        a = gi.series_input(oracle, p, n, random_a0=(op == "inv"))
        if op == "exp":
            a[0] = 0
        reqs.append(fa.Request(op, p, a, n=n))
        wants.append(oracle.series(op, p, a, n))
    for r, g, w in zip(reqs, fa.batch(reqs), wants):
        assert np.array_equal(g, w), (r.op, r.p, r.a.size)


def test_batch_split_products(fa, oracle):
    """Products past the prime's largest transform (one spectrum per piece,
    spectral block sums, accumulated inverses) recorded into the merged executor."""
    rng = np.random.default_rng(5)
    reqs, wants = [], []
    for p, la, lb in ((65537, 70000, 50000), (65537, 3 * 32768 + 7, 5), (786433, 300000, 200001),
                      (P, 3000, 4000), (65537, 100, 200), (786433, 5, 262145)):
        a = rng.integers(0, p, la, dtype=np.uint32)
        b = rng.integers(0, p, lb, dtype=np.uint32)
        reqs.append(fa.Request("conv", p, a, b))
        wants.append(oracle.conv(p, a, b))
    for r, g, w in zip(reqs, fa.batch(reqs), wants):
        assert np.array_equal(g, w), (r.p, r.a.size, r.b.size)


def test_batch_uneven_sample(fa):
    """A scaled-down draw of the uneven workload (log-uniform lengths, two moduli)."""
    from fastpoly_amd.workloads import uneven_batch
    tuples = uneven_batch(n_conv=192, n_series=48, seed=7, conv_log=(6, 16), series_log=(8, 14))
    reqs = [fa.Request(t[0], t[1], t[2], b=t[3]) if t[0] == "conv" else fa.Request(t[0], t[1], t[2], n=t[3])
            for t in tuples]
    got = fa.batch(reqs)
    for r, g in zip(reqs, got):
        assert _same(g, _single(fa, r)), (r.op, r.p, r.a.size, r.n)


def test_batch_device_submit_wait(fa, oracle):
    rng = np.random.default_rng(3)
    items = []
    for i in range(24):
        p = (P, Q, 167772161)[i % 3]
        la, lb = int(rng.integers(1, 20000)), int(rng.integers(1, 20000))
        items.append(("conv", p, rng.integers(0, p, la, dtype=np.uint32), rng.integers(0, p, lb, dtype=np.uint32)))
    for i, op in enumerate(("inv", "exp") * 6):
        p = (P, Q)[i % 2]
        n = int(rng.integers(200, 30000))
        a = rng.integers(0, p, n, dtype=np.uint32)
        a[0] = 0 if op == "exp" else int(rng.integers(1, p))
        items.append((op, p, a, n))
    D = fa.DeviceArray
    arr = (fa.Req * len(items))()
    keep = []
    for i, it in enumerate(items):
        q = arr[i]
        q.op, q.p = fa.OPS[it[0]], it[1]
        da = D.from_host(it[2])
        q.a, q.la = da.ptr, it[2].size
        if it[0] == "conv":
            db = D.from_host(it[3])
            q.b, q.lb, q.n = db.ptr, it[3].size, fa.SIZE_MAX
            cap = it[2].size + it[3].size - 1
            keep.append(db)
        else:
            q.n = cap = it[3]
        out = D(cap)
        q.out, q.out_cap = out.ptr, cap
        keep += [da, out]
        items[i] = it + (out,)
    t = fa.batch_submit(arr)
    fa.wait(t)
    fa.wait(t)  # waiting twice is allowed
    for i, it in enumerate(items):
        assert arr[i].status == 0
        got = it[-1].to_host(arr[i].out_len)
        want = oracle.conv(it[1], it[2], it[3]) if it[0] == "conv" else oracle.series(it[0], it[1], it[2], it[3])
        assert np.array_equal(got, want), (it[0], it[1], it[2].size)


def test_batch_empty(fa):
    assert fa.batch([]) == []


def test_batch_short_side_and_tail_products(fa, oracle):
    """Products with a short side (direct, k_conv_direct_batch) and products a
    few coefficients past a power of two (cyclic product plus tail repair,
    k_tail_fix_batch) in one merged batch and through the server: every
    request against the oracle, across the boundaries of both routes (short
    side 128 / 129 over NTT primes, 256 / 257 over 1e9+7; tails of 1 .. 41)."""
    rng = np.random.default_rng(77)
    reqs, wants = [], []
    for p in (P, Q, 3):
        for ns in (1, 2, 40, 127, 128, 129, 255, 256, 257):
            for nl in (ns, 1000, 5000):
                a = rng.integers(0, p, nl, dtype=np.uint32)
                b = rng.integers(0, p, ns, dtype=np.uint32)
                reqs.append(fa.Request("conv", p, a, b))
                wants.append(oracle.conv(p, a, b))
                reqs.append(fa.Request("conv", p, b, a, n=max(1, nl // 2)))
                wants.append(oracle.conv(p, b, a)[:max(1, nl // 2)])
    for p in (P, Q):
        for N in (1024, 4096):
            for t in (1, 2, 17, 40, 41):
                la = N // 2 + t // 2 + 1
                lb = N + t + 1 - la
                a = rng.integers(0, p, la, dtype=np.uint32)
                b = rng.integers(0, p, lb, dtype=np.uint32)
                reqs.append(fa.Request("conv", p, a, b))
                wants.append(oracle.conv(p, a, b))
                reqs.append(fa.Request("conv", p, a, b, n=N + t // 2))
                wants.append(oracle.conv(p, a, b)[:N + t // 2])
    got = fa.batch(reqs)
    for i, (g, w) in enumerate(zip(got, wants)):
        assert np.array_equal(g, w), (i, reqs[i].p, len(reqs[i].a), len(reqs[i].b))
    with fa.Server() as srv:
        ids = [srv.submit(r) for r in reqs]
        served = [srv.wait(i)[0] for i in ids]
    for i, (g, w) in enumerate(zip(served, wants)):
        assert np.array_equal(g, w), (i, reqs[i].p, len(reqs[i].a), len(reqs[i].b))
