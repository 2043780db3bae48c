"""GPU: the continuous-batching server returns, for every request, exactly the
single call's result, whether the request ran alone or joined a running mix;
errors stay with their request; a tight HBM budget only delays admission."""
import threading
import time

import numpy as np
import pytest

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
    except Exception as e:
        return e


def _same(x, y):
    if isinstance(x, Exception) or isinstance(y, Exception):
        return type(x) is type(y)
    return np.array_equal(x, y)


def _stream(fa, count, seed, lo=4, hi=15):
    from fastpoly_amd.workloads import serving_stream
    reqs = []
    for _, t in serving_stream(count, 1000.0, seed=seed, length_log=(lo, hi)):
        if t[0] == "conv":
            reqs.append(fa.Request("conv", t[1], t[2], b=t[3]))
        else:
            reqs.append(fa.Request(t[0], t[1], t[2], n=t[3]))
    return reqs


def test_serve_identical_alone_and_in_mix(fa):
    reqs = _stream(fa, 160, seed=11)
    alone = [_single(fa, r) for r in reqs]
    with fa.Server() as srv:
        # Alone: one request at a time.
        solo = []
        for r in reqs[:20]:
            solo.append(srv.wait(srv.submit(r)))
        # In the mix: submitted from another thread with small gaps, so
        # requests arrive while others are running.
        ids = [None] * len(reqs)

        def feeder():
            for i, r in enumerate(reqs):
                ids[i] = srv.submit(r)
                if i % 7 == 0:
                    time.sleep(0.0005)

        th = threading.Thread(target=feeder)
        th.start()
        th.join()
        mixed = [srv.wait(i) for i in ids]
        st = srv.stats()
    for r, (g, _), w in zip(reqs[:20], solo, alone[:20]):
        assert _same(g, w), (r.op, r.p, r.a.size)
    for r, (g, _), w in zip(reqs, mixed, alone):
        assert _same(g, w), (r.op, r.p, r.a.size, r.n)
    assert st["completed"] == st["submitted"] == 20 + len(reqs)
    # Continuous batching happened: some request was admitted while another was running.
    spans = sorted((t["admit_s"], t["done_s"]) for _, t in mixed)
    overlap = any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    assert overlap and st["max_active"] > 1


def test_serve_errors_stay_with_their_request(fa):
    z = np.zeros(8, np.uint32)
    good = fa.Request("conv", P, np.arange(1, 100, dtype=np.uint32), b=np.arange(5, 50, dtype=np.uint32))
    bad = [fa.Request("inv", P, z, n=4), fa.Request("log", Q, z, n=4), fa.Request("conv", 12, z, b=z),
           fa.Request("sqrt", P, np.array([0, 3, 1], np.uint32), n=5)]
    with fa.Server() as srv:
        ids = [srv.submit(r) for r in [good] + bad + [good]]
        res = [srv.wait(i) for i in ids]
    want = _single(fa, good)
    assert np.array_equal(res[0][0], want) and np.array_equal(res[-1][0], want)
    for r, (g, _) in zip(bad, res[1:-1]):
        assert _same(g, _single(fa, r)) and isinstance(g, Exception)


def test_serve_tight_budget_and_poll(fa):
    reqs = _stream(fa, 40, seed=12, lo=10, hi=17)
    with fa.Server(hbm_budget_bytes=48 << 20, levels_per_iteration=1) as srv:
        ids = [srv.submit(r) for r in reqs]
        got = {}
        deadline = time.time() + 60
        while len(got) < len(ids) and time.time() < deadline:
            for i in ids:
                if i not in got:
                    x = srv.poll(i)
                    if x is not None:
                        got[i] = x
            time.sleep(0.001)
        assert len(got) == len(ids)
    for r, i in zip(reqs, ids):
        assert _same(got[i][0], _single(fa, r)), (r.op, r.p, r.a.size)


def test_serve_rejected_submit_returns_a_code(fa):
    """A request the server cannot stage comes back as an error code from
    submit (never an abort), and the server keeps serving afterwards."""
    import ctypes
    z = np.zeros(8, np.uint32)
    good = fa.Request("conv", P, np.arange(1, 100, dtype=np.uint32), b=np.arange(5, 50, dtype=np.uint32))
    with fa.Server() as srv:
        q = fa.Req()
        q.op, q.p = fa.OPS["conv"], P
        # 2^33 + 2^33 input words: refused by size before any copy is made
        q.a, q.la, q.b, q.lb = z.ctypes.data, 1 << 33, z.ctypes.data, 1 << 33
        q.n, q.out, q.out_cap = 1 << 34, z.ctypes.data, 8
        rid = ctypes.c_uint64(0)
        rc = fa.lib().fpxg_server_submit(srv._h, ctypes.byref(q), ctypes.byref(rid))
        assert rc == fa.E_SIZE
        q2 = fa.Req()
        q2.op, q2.p, q2.a, q2.la, q2.n = fa.OPS["inv"], P, None, 5, 4  # null input with la > 0
        assert fa.lib().fpxg_server_submit(srv._h, ctypes.byref(q2), ctypes.byref(rid)) == fa.E_ARG
        got = srv.wait(srv.submit(good))[0]
    assert np.array_equal(got, _single(fa, good))
