"""Synthetic workloads of the batched and serving paths (numpy only).

Uneven batch: independent convolutions, half mod 998244353 and half mod
1000000007 (auxiliary-prime path), each operand length round(2^U) with
U ~ Uniform[6, 18] drawn independently per operand, coefficients uniform in
[0, p); plus Newton series instances (inv with a0 uniform in [1, p), exp with
b0 = 0) at n = round(2^U), U ~ Uniform[8, 16].

Serving stream: Poisson arrivals of conv 50% / inv 20% / log 10% / exp 10% /
sqrt 10% requests, lengths round(2^U) with U ~ Uniform[10, 20], moduli
998244353 (70%) or 1000000007 (30%).

Every instance draws its coefficients from its own numpy PCG64 generator
seeded (master_seed, index), so any instance can be regenerated alone.
"""
from __future__ import annotations

import numpy as np

P_NTT = 998244353
P_CRT = 1000000007
MONT_SHIFT = 32


def mont_one(p: int) -> int:
    return (1 << MONT_SHIFT) % p


def _logu(rng, lo: int, hi: int) -> int:
    return int(round(2.0 ** rng.uniform(lo, hi)))


def _coeffs(p: int, n: int, rng) -> np.ndarray:
    return rng.integers(0, p, n, dtype=np.uint32)


def series_input(op: str, p: int, n: int, rng) -> np.ndarray:
    """A valid input for op at length n, other coefficients uniform: inv has
    a0 uniform in [1, p), log a0 = 1, exp a0 = 0, sqrt a0 = 1 (a residue, so
    the power-series square root exists)."""
    a = _coeffs(p, n, rng)
    if op == "inv":
        a[0] = int(rng.integers(1, p))
    elif op == "log":
        a[0] = mont_one(p)
    elif op == "exp":
        a[0] = 0
    elif op == "sqrt":
        a[0] = mont_one(p)
    return a


def uneven_batch(n_conv: int = 8192, n_series: int = 1024, seed: int = 20261203,
                 conv_log=(6, 18), series_log=(8, 16)):
    """List of request tuples: ("conv", p, a, b) and (op, p, a, n)."""
    reqs = []
    for i in range(n_conv):
        rng = np.random.default_rng((seed, 0, i))
        p = P_NTT if i % 2 == 0 else P_CRT
        la, lb = _logu(rng, *conv_log), _logu(rng, *conv_log)
        reqs.append(("conv", p, _coeffs(p, la, rng), _coeffs(p, lb, rng)))
    for i in range(n_series):
        rng = np.random.default_rng((seed, 1, i))
        op = "inv" if i % 2 == 0 else "exp"
        p = P_NTT if (i // 2) % 2 == 0 else P_CRT
        n = _logu(rng, *series_log)
        reqs.append((op, p, series_input(op, p, n, rng), n))
    return reqs


SERVE_MIX = (("conv", 0.5), ("inv", 0.2), ("log", 0.1), ("exp", 0.1), ("sqrt", 0.1))


def serving_stream(count: int, rate_per_s: float, seed: int = 20261204, length_log=(10, 20)):
    """(arrival_s, request) pairs, arrivals Poisson at rate_per_s."""
    rng = np.random.default_rng((seed, 2))
    t = 0.0
    out = []
    ops = [o for o, _ in SERVE_MIX]
    probs = [w for _, w in SERVE_MIX]
    for i in range(count):
        t += rng.exponential(1.0 / rate_per_s)
        op = ops[int(rng.choice(len(ops), p=probs))]
        p = P_NTT if rng.random() < 0.7 else P_CRT
        r = np.random.default_rng((seed, 3, i))
        if op == "conv":
            la, lb = _logu(r, *length_log), _logu(r, *length_log)
            out.append((t, ("conv", p, _coeffs(p, la, r), _coeffs(p, lb, r))))
        else:
            n = _logu(r, *length_log)
            out.append((t, (op, p, series_input(op, p, n, r), n)))
    return out
