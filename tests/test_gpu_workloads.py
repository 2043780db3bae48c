"""GPU: the full batched and serving workloads against per-request checksums
written by the reference library (tests/golden/workload_checksums.json,
make_workload_checksums.py).

  * the uneven batch (8192 convs, half mod 998244353 and half mod 1e9+7,
    operand lengths log-uniform in [2^6, 2^18], plus 1024 inv/exp at n
    log-uniform in [2^8, 2^16]) in one fpxg_batch_submit over device inputs;
  * the serving stream (1000 Poisson-timed conv / inv / log / exp / sqrt
    requests, lengths log-uniform in [2^10, 2^20], 70% mod 998244353)
    through the continuous-batching server at 1000 requests/s.
"""
import json
import os
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def fa():
    import fastpoly_amd
    fastpoly_amd.init(0, 0)
    return fastpoly_amd


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(HERE, "golden", "workload_checksums.json")) as fh:
        return json.load(fh)


def _hex(oracle, a):
    return "%016x" % oracle.checksum(a)


def test_uneven_batch_full_size(fa, oracle, golden):
    import ctypes  # noqa: F401
    from fastpoly_amd.workloads import uneven_batch
    g = golden["uneven_batch"]
    reqs = uneven_batch(g["n_conv"], g["n_series"], seed=g["seed"])
    assert len(reqs) == len(g["checksums"]) == 9216
    sizes_in = [(r[2].size, r[3].size if r[0] == "conv" else 0) for r in reqs]
    sizes_out = [r[2].size + r[3].size - 1 if r[0] == "conv" else r[3] for r in reqs]
    offs_in = np.concatenate([[0], np.cumsum([x + y for x, y in sizes_in])])
    offs_out = np.concatenate([[0], np.cumsum(sizes_out)])
    host_in = np.empty(int(offs_in[-1]), np.uint32)
    for i, r in enumerate(reqs):
        o = int(offs_in[i])
        host_in[o:o + sizes_in[i][0]] = r[2]
        if r[0] == "conv":
            host_in[o + sizes_in[i][0]:o + sizes_in[i][0] + sizes_in[i][1]] = r[3]
    din = fa.DeviceArray.from_host(host_in)
    dout = fa.DeviceArray(int(offs_out[-1]))
    arr = (fa.Req * len(reqs))()
    for i, r in enumerate(reqs):
        q = arr[i]
        q.op, q.p = fa.OPS[r[0]], r[1]
        q.a, q.la = din.ptr + 4 * int(offs_in[i]), sizes_in[i][0]
        if r[0] == "conv":
            q.b, q.lb, q.n = din.ptr + 4 * (int(offs_in[i]) + sizes_in[i][0]), sizes_in[i][1], fa.SIZE_MAX
        else:
            q.n = r[3]
        q.out, q.out_cap = dout.ptr + 4 * int(offs_out[i]), sizes_out[i]
    fa.wait(fa.batch_submit(arr))
    out = dout.to_host()
    bad = []
    for i in range(len(reqs)):
        assert arr[i].status == 0 and arr[i].out_len == sizes_out[i], i
        got = _hex(oracle, out[int(offs_out[i]):int(offs_out[i + 1])])
        if got != g["checksums"][i]:
            bad.append(i)
    assert not bad, f"{len(bad)} requests differ from the reference, first {bad[:10]}"


def test_serving_stream_full_size(fa, oracle, golden):
    from fastpoly_amd.workloads import serving_stream
    g = golden["serving_stream"]
    stream = serving_stream(g["count"], 1000.0, seed=g["seed"], length_log=tuple(g["length_log"]))
    reqs = []
    for t, r in stream:
        if r[0] == "conv":
            reqs.append((t, fa.Request("conv", r[1], r[2], b=r[3])))
        else:
            reqs.append((t, fa.Request(r[0], r[1], r[2], n=r[3])))
    ids = [None] * len(reqs)
    with fa.Server() as srv:
        t0 = time.perf_counter()

        def feed():  # Poisson arrivals at 1000 requests/s
            for i, (t, r) in enumerate(reqs):
                delay = t - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                ids[i] = srv.submit(r)

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        results = [srv.wait(i)[0] for i in ids]
        st = srv.stats()
    bad = [i for i, res in enumerate(results)
           if isinstance(res, Exception) or _hex(oracle, res) != g["checksums"][i]]
    assert not bad, f"{len(bad)} requests differ from the reference, first {bad[:10]}"
    assert st["max_active"] > 1  # requests overlapped on the device
