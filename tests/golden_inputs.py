"""Deterministic inputs of the reference benchmark (bench/bench.cpp:216-373)
and of the workload configurations, generated with std::mt19937_64.

Used by the tests and bench.py; the generator itself is the oracle's
restatement (oracle/fpx_oracle.c), pinned by the published checksums.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def reference_checksums():
    with open(os.path.join(GOLDEN, "reference_checksums.json")) as fh:
        return json.load(fh)["checksums"]


def ref_vectors():
    with open(os.path.join(GOLDEN, "ref_vectors.json")) as fh:
        return json.load(fh)


def bench_case(po, p, op, n):
    """(callable inputs, expected-output-producing description) for one bench row.

    Returns (kind, args): kind in {"ntt-forward", "ntt-inverse", "conv", "series"}.
    """
    if op in ("ntt-forward", "ntt-inverse"):
        g = po.MT(1 + n)
        return g.rnd(p, n)
    if op in ("conv", "conv-trunc"):
        g = po.MT(2 + n)
        a = g.rnd(p, n)
        return a, g.rnd(p, n)
    if op == "conv-tail":
        g = po.MT(2 + n)
        a = g.rnd(p, n + 1)
        return a, g.rnd(p, n + 1)
    if op == "square":
        return po.MT(2 + n).rnd(p, n)
    if op == "skinny":
        g = po.MT(4 + n)
        a = g.rnd(p, 40)
        return a, g.rnd(p, n)
    if op in ("inv", "log", "exp", "sqrt", "pow"):
        a = po.MT(3 + n).rnd(p, n)
        a[0] = po.from_int(p, 1)
        return a
    raise KeyError(op)


def rotl64(x, r):
    return ((x << r) | (x >> (64 - r))) & ((1 << 64) - 1)


def division_case(po, p, n):
    """div / mod / divmod bench inputs (bench/bench.cpp:255-263): a of n, b of max(1, n/2)."""
    g = po.MT(4 + n)
    a = g.rnd(p, n)
    return a, g.rnd(p, max(1, n // 2))


def division_checksum(po, op, q, r):
    """checksum(q) ^ rotl(checksum(r), 17) with the operand div or mod does not return empty."""
    empty = np.zeros(0, np.uint32)
    if op == "div":
        r = empty
    if op == "mod":
        q = empty
    return format(po.checksum(q) ^ rotl64(po.checksum(r), 17), "016x")


POW_K = 1000003  # bench/bench.cpp:372-373


def series_input(po, p, n, random_a0=False):
    a = po.MT(3 + n).rnd(p, n)
    if random_a0:
        a[0] = po.from_int(p, 1 + po.MT(5 + n).next() % (p - 1))
    else:
        a[0] = po.from_int(p, 1)
    return a


def hexsum(po, arr):
    return "%016x" % po.checksum(np.asarray(arr, dtype=np.uint32))
