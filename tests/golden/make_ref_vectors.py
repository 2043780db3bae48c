"""Write reference-derived vectors for cases the published checksums leave open.

Uses the reference library itself, compiled in place from the reference
checkout by `make -C oracle ref` (oracle/_ref/libfpref_*.so; never committed,
never shipped).  Output: ref_vectors.json with
  * checksums of the workload sizes (conv 2^23 mod 998244353 = 2^24-point aux
    NTTs, conv 2^22 mod 1e9+7, inv at 2^20 with a random constant term, ...);
  * Mont::sqrt root choices (modint.hpp:139-166) for random residues;
  * small exact input/output vectors for series with random constant terms.
Inputs follow bench/bench.cpp:216-373; the random constant term is
a0 = from_int(1 + mt19937_64(5 + n)() % (p - 1)).

    make -C oracle ref && python tests/golden/make_ref_vectors.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import pyoracle as po  # noqa: E402

P, Q = 998244353, 1000000007


def conv_inputs(p, n):
    g = po.MT(2 + n)
    return g.rnd(p, n), g.rnd(p, n)


def series_input(p, n, random_a0=False):
    g = po.MT(3 + n)
    a = g.rnd(p, n)
    if random_a0:
        a[0] = po.from_int(p, 1 + po.MT(5 + n).next() % (p - 1))
    else:
        a[0] = po.from_int(p, 1)
    return a


def main():
    assert po.ref() is not None, "build the reference shim first: make -C oracle ref"
    out = {"backend": po.ref().fpref_backend().decode(), "checksums": {}, "msqrt": {}, "series": {}}
    ck = out["checksums"]

    def record(key, arr):
        ck[key] = "%016x" % po.checksum(arr)
        print(key, ck[key], flush=True)

    t = time.time()
    a, b = conv_inputs(P, 1 << 23)
    record("998244353/conv/8388608", po.ref_conv(P, a, b))
    a, b = conv_inputs(Q, 1 << 22)
    record("1000000007/conv/4194304", po.ref_conv(Q, a, b))
    for p in (P, Q):
        n = 1 << 20
        record(f"{p}/inv-random-a0/{n}", po.ref_series("inv", p, series_input(p, n, True), n))
    for p in (P, Q):
        for n in (1024, 65536):
            record(f"{p}/inv-random-a0/{n}", po.ref_series("inv", p, series_input(p, n, True), n))
    n = 1 << 20
    a = series_input(Q, n)
    record(f"{Q}/inv/{n}", po.ref_series("inv", Q, a, n))
    record(f"{Q}/log/{n}", po.ref_series("log", Q, a, n))
    e = a.copy(); e[0] = 0
    record(f"{Q}/exp/{n}", po.ref_series("exp", Q, e, n))
    record(f"{Q}/sqrt/{n}", po.ref_series("sqrt", Q, po.ref_conv(Q, a, a, n), n))
    print("large vectors: %.1f s" % (time.time() - t))

    # Mont::sqrt root choice for random residues (squares and non-squares).
    for p in (P, Q, 167772161, 469762049, 754974721, 1224736769, 100003, 3):
        g = po.MT(11 + p)
        rows = []
        for _ in range(64):
            x = g.next() % p
            limb = po.from_int(p, x)
            sq = po.mont_mul(p, limb, limb)
            rows.append([sq, po.ref_msqrt(p, sq)])
            rows.append([limb, po.ref_msqrt(p, limb)])
        out["msqrt"][str(p)] = rows

    # Small exact vectors with random constant terms and valuations.
    for p in (P, Q, 1224736769, 100003):
        for n in (7, 40, 100, 257):
            g = po.MT(17 + n + p)
            a = g.rnd(p, n)
            a0 = po.from_int(p, 1 + g.next() % (p - 1))
            cases = {}
            ai = a.copy(); ai[0] = a0
            cases["inv"] = (ai, po.ref_series("inv", p, ai, n))
            al = a.copy(); al[0] = po.from_int(p, 1)
            cases["log"] = (al, po.ref_series("log", p, al, n))
            ae = a.copy(); ae[0] = 0
            cases["exp"] = (ae, po.ref_series("exp", p, ae, n))
            d = a.copy(); d[0] = a0
            sq = po.ref_conv(p, d, d, n)
            sq_shift = [0, 0] + list(sq[: n - 2])
            cases["sqrt"] = (sq, po.ref_series("sqrt", p, sq, n))
            cases["sqrt-valued"] = (sq_shift, po.ref_series("sqrt", p, sq_shift, n))
            for k, (inp, res) in cases.items():
                out["series"][f"{p}/{k}/{n}"] = {"a": [int(x) for x in inp], "out": [int(x) for x in res]}
    with open(os.path.join(HERE, "ref_vectors.json"), "w") as fh:
        json.dump(out, fh, indent=0, sort_keys=True)
    print("wrote ref_vectors.json")


if __name__ == "__main__":
    main()
