"""GPU parity: every result of the HIP path (through the C ABI) is bit-identical
to the oracle / the reference's published checksums on the same seeded inputs.
"""
import random

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


def _rnd(po, p, n, seed):
    return po.MT(seed).rnd(p, n)


# ------------------------------------------------------------------ NTT order
@pytest.mark.parametrize("p", [P, 167772161, 469762049, 754974721, 1224736769])
def test_ntt_forward_inverse_exact(fa, oracle, p):
    root = oracle.lib().fpo_primitive_root(p)
    for lg in list(range(1, 15)) + [15, 16, 17, 18, 20]:
        n = 1 << lg
        x = _rnd(oracle, p, n, 1 + n)
        want = oracle.ntt(p, x, root=root)
        got = fa.ntt_forward(p, x, root=root)
        assert np.array_equal(got, want), (p, n)
        back = fa.ntt_inverse(p, want, root=root)
        assert np.array_equal(back, x), (p, n)


@pytest.mark.parametrize("p,lg", [(P, 23), (167772161, 24)])
def test_ntt_largest_sizes(fa, oracle, p, lg):
    """2^23 and the 2^24 transform the roofline times (column pass, then row
    pass, in place through the pass entry point) against the oracle."""
    n = 1 << lg
    root = oracle.lib().fpo_primitive_root(p)
    x = _rnd(oracle, p, n, 99 + lg)
    want = oracle.ntt(p, x, root=root)
    assert np.array_equal(fa.ntt_forward(p, x, root=root), want)
    d = fa.DeviceArray.from_host(x)
    fa.ntt_pass_dev(p, d, 0, root=root)
    fa.ntt_pass_dev(p, d, 1, root=root)
    assert np.array_equal(d.to_host(), want)
    d.free()
    assert np.array_equal(fa.ntt_inverse(p, want, root=root), x)


@pytest.mark.parametrize("p,lg", [(P, 23), (167772161, 24)])
def test_ntt_row_kernel_option(fa, oracle, p, lg):
    """The per-row forward kernel (persistent row pass switched off) gives
    the same limbs as the oracle at the largest sizes."""
    n = 1 << lg
    root = oracle.lib().fpo_primitive_root(p)
    x = _rnd(oracle, p, n, 7 + lg)
    want = oracle.ntt(p, x, root=root)
    fa.set_row_dma(False)
    try:
        assert np.array_equal(fa.ntt_forward(p, x, root=root), want)
    finally:
        fa.set_row_dma(True)
    assert np.array_equal(fa.ntt_forward(p, x, root=root), want)


@pytest.mark.parametrize("p,lg", [(167772161, 16), (167772161, 20), (167772161, 24), (469762049, 20)])
def test_ntt_extreme_inputs(fa, oracle, p, lg):
    """Inputs at the top of the range (p - 1 everywhere, p - 1 / 0
    alternating, p - 1 on a random half) give the forward passes'
    intermediate sums and differences their largest magnitudes."""
    n = 1 << lg
    root = oracle.lib().fpo_primitive_root(p)
    rng = np.random.default_rng(lg)
    pats = [np.full(n, p - 1, dtype=np.uint32),
            np.tile(np.array([p - 1, 0], dtype=np.uint32), n // 2),
            np.where(rng.random(n) < 0.5, p - 1, 0).astype(np.uint32)]
    for i, x in enumerate(pats):
        want = oracle.ntt(p, x, root=root)
        assert np.array_equal(fa.ntt_forward(p, x, root=root), want), (p, lg, i)


def test_ntt_windows(fa, oracle):
    po = oracle
    for lg in (4, 10, 14, 15, 17):
        n = 1 << lg
        x = _rnd(po, P, n, 7 + n)
        plan = fa.NttPlan(fa.Mont(P), n)
        for count in (1, n // 4, n // 2, n // 2 + 3, n):
            pre = x.copy()
            pre[count:] = 0
            assert np.array_equal(plan.forward_prefix(x, count), po.ntt(P, pre)), (n, count)
        spec = po.ntt(P, x)
        full = po.ntt(P, spec, inverse=True)
        for count in (1, n // 4, n // 2 + 1, n):
            got = plan.inverse_prefix(spec, count)
            assert np.array_equal(got[:count], full[:count])
        hh = plan.inverse_highhalf(spec)
        assert np.array_equal(hh[n // 2:], full[n // 2:])


def test_ntt_size_errors(fa):
    with pytest.raises(fa.NttSizeError):
        fa.ntt_forward(P, np.zeros(3, np.uint32))
    with pytest.raises(fa.NttSizeError):
        fa.ntt_forward(Q, np.zeros(4, np.uint32))
    with pytest.raises(fa.UnsupportedModulus):
        fa.conv(1000000000, [1], [1])


@pytest.mark.parametrize("p,op,n,want", [
    (int(m), op, int(s), ck)
    for m, t in gi.reference_checksums().items() for op, sizes in t.items()
    if op == "ntt-forward" for s, ck in sizes.items() if int(s) <= (1 << 20)])
def test_golden_ntt_forward(fa, oracle, p, op, n, want):
    root = oracle.lib().fpo_primitive_root(p)
    x = gi.bench_case(oracle, p, op, n)
    assert gi.hexsum(oracle, fa.ntt_forward(p, x, root=root)) == want


# ------------------------------------------------------------------ conv
def test_config1_conv_65536(fa, oracle):
    a, b = gi.bench_case(oracle, P, "conv", 65536)
    assert gi.hexsum(oracle, fa.conv(P, a, b)) == "841755dfc72581be"


@pytest.mark.parametrize("p", [P, Q, 1224736769, 167772161, 100003, 3, 2])
def test_conv_random_shapes(fa, oracle, p):
    rng = random.Random(p)
    shapes = [(1, 1), (1, 7), (2, 2), (5, 40), (40, 41), (64, 64), (100, 3000), (4097, 4096),
              (8191, 8193), (16384, 16385), (70000, 3), (33000, 40000)]
    for la, lb in shapes:
        a = _rnd(oracle, p, la, rng.randrange(1 << 30))
        b = _rnd(oracle, p, lb, rng.randrange(1 << 30))
        want = oracle.conv(p, a, b)
        assert np.array_equal(fa.conv(p, a, b), want), (p, la, lb)
        n = (la + lb) // 3 + 1
        assert np.array_equal(fa.conv(p, a, b, n), want[:n]), (p, la, lb, n)
    x = _rnd(oracle, p, 5000, 11)
    assert np.array_equal(fa.conv(p, x, x), oracle.conv(p, x, x))
    assert fa.conv(p, [], x).size == 0 and fa.conv(p, x, x, 0).size == 0


@pytest.mark.parametrize("fused_log", [1, 31])
def test_crt_fused_column_pass(fa, oracle, fused_log):
    """Three-prime (CRT) products with Garner in the inverse column pass at
    every size with a column pass (1) and never (31, residues then a Garner
    launch): products over 1e9+7 (three primes), 100003 (two), 3 and 2 (one),
    truncated and squared; Newton series over 1e9+7 (window products,
    negated); a merged batch.  Each against the oracle."""
    fa.set_crt_fused_log(fused_log)
    try:
        for p in (Q, 100003, 3, 2):
            rng = random.Random(p * 7 + fused_log)
            for la, lb in ((4097, 4096), (16384, 16385), (33000, 40000)):
                a = _rnd(oracle, p, la, rng.randrange(1 << 30))
                b = _rnd(oracle, p, lb, rng.randrange(1 << 30))
                want = oracle.conv(p, a, b)
                assert np.array_equal(fa.conv(p, a, b), want), (p, la, lb)
                n = (la + lb) // 3 + 1
                assert np.array_equal(fa.conv(p, a, b, n), want[:n]), (p, la, lb, n)
            x = _rnd(oracle, p, 20000, rng.randrange(1 << 30))
            assert np.array_equal(fa.conv(p, x, x), oracle.conv(p, x, x)), p
        rng = random.Random(fused_log)
        n = 20000
        a = _rnd(oracle, Q, n, rng.randrange(1 << 30))
        a[0] = oracle.from_int(Q, 12345)
        a1 = a.copy(); a1[0] = oracle.from_int(Q, 1)
        a0 = a.copy(); a0[0] = 0
        sq = oracle.conv(Q, a1, a1, n)
        want = {"inv": oracle.series("inv", Q, a, n), "log": oracle.series("log", Q, a1, n),
                "exp": oracle.series("exp", Q, a0, n), "sqrt": oracle.series("sqrt", Q, sq, n)}
        ins = {"inv": a, "log": a1, "exp": a0, "sqrt": sq}
        for op in ("inv", "log", "exp", "sqrt"):
            for _ in range(3):  # eager, then captured, then replayed
                assert np.array_equal(fa.series(op, Q, ins[op], n), want[op]), (op, fused_log)
        b = _rnd(oracle, Q, 30000, 5)
        reqs = [fa.Request("conv", Q, a, b=b), fa.Request("conv", 100003, b, b=b[:9000]),
                fa.Request("exp", Q, a0, n=n), fa.Request("sqrt", Q, sq, n=n)]
        got = fa.batch(reqs)
        assert np.array_equal(got[0], oracle.conv(Q, a, b))
        assert np.array_equal(got[1], oracle.conv(100003, b, b[:9000]))
        assert np.array_equal(got[2], want["exp"]) and np.array_equal(got[3], want["sqrt"])
    finally:
        fa.set_crt_fused_log(0)
    with pytest.raises(fa.ArgumentError):
        fa.set_crt_fused_log(32)


@pytest.mark.parametrize("fused_log", [0, 31])
def test_crt_block_pairs_accumulate(fa, oracle, fused_log):
    """A 1e9+7 product with an operand past 2^23 coefficients runs as block
    pairs accumulated into the output (crt_convolution.hpp:187-205): here a
    2^24-point three-prime block (Garner accumulating in the inverse column
    pass by default, after it with 31) and a one-row block.  Checked at the
    block boundary, the ends and random positions against exact integer sums
    (limb = sum of a_i b_j times 2^-32 mod p)."""
    fa.set_crt_fused_log(fused_log)
    try:
        la, lb = (1 << 23) + 100, 300
        a = _rnd(oracle, Q, la, 77)
        b = _rnd(oracle, Q, lb, 78)
        got = fa.conv(Q, a, b)
        assert got.size == la + lb - 1
        rinv = pow(1 << 32, Q - 2, Q)
        bl = [int(v) for v in b]
        rng = random.Random(fused_log + 3)
        pos = list(range(0, 64)) + list(range((1 << 23) - 40, (1 << 23) + 340)) + \
            list(range(la + lb - 65, la + lb - 1)) + [rng.randrange(la + lb - 1) for _ in range(200)]
        for i in pos:
            lo, hi = max(0, i - la + 1), min(lb - 1, i)
            acc = sum(int(a[i - j]) * bl[j] for j in range(lo, hi + 1))
            assert int(got[i]) == acc % Q * rinv % Q, i
    finally:
        fa.set_crt_fused_log(0)


@pytest.mark.parametrize("p", [P, Q, 1224736769, 3])
def test_conv_short_side_direct(fa, oracle, p):
    """Products with a short side (at most 128 coefficients over an NTT prime,
    256 otherwise) are computed directly (no transform): both orders, both
    boundaries, tile edges, a long side of 2^20, truncation inside and past
    the long side."""
    rng = random.Random(p + 9)
    for ls, ll in ((1, 1), (1, 1025), (2, 3), (40, 1023), (63, 1024), (64, 1025), (65, 5000), (64, 64),
                   (128, 3000), (129, 3000), (256, 5000), (257, 5000), (200, 200), (17, 1 << 20),
                   (64, 300001)):
        a = _rnd(oracle, p, ls, rng.randrange(1 << 30))
        b = _rnd(oracle, p, ll, rng.randrange(1 << 30))
        want = oracle.conv(p, a, b)
        assert np.array_equal(fa.conv(p, a, b), want), (p, ls, ll)
        assert np.array_equal(fa.conv(p, b, a), want), (p, ll, ls)
        for n in (1, ll // 2 + 1, ll + 3):
            assert np.array_equal(fa.conv(p, a, b, n), want[:n]), (p, ls, ll, n)


@pytest.mark.parametrize("p", [P, 167772161, 65537])
def test_conv_tail_repair(fa, oracle, p):
    """Products a few coefficients past a power of two N: the N-point cyclic
    product plus the directly computed high tail (tails 1, 2, 39, 40, 41 at
    several N, operands up to N each, truncation below N, inside the tail and
    past it, squaring)."""
    rng = random.Random(p + 13)
    for N in (256, 4096, 1 << 15):
        for tail in (1, 2, 39, 40, 41):
            full = N + tail
            for la in (N, full // 2 + 1, full - 129):
                lb = full + 1 - la
                if la > N or lb > N or min(la, lb) <= 128:
                    continue
                a = _rnd(oracle, p, la, rng.randrange(1 << 30))
                b = _rnd(oracle, p, lb, rng.randrange(1 << 30))
                want = oracle.conv(p, a, b)
                assert np.array_equal(fa.conv(p, a, b), want), (p, N, tail, la)
                for n in (N - 5, N + tail // 2 + 1, full + 7):
                    assert np.array_equal(fa.conv(p, a, b, n), want[:n]), (p, N, tail, la, n)
        x = _rnd(oracle, p, N // 2 + 10, 3)
        assert np.array_equal(fa.conv(p, x, x), oracle.conv(p, x, x)), (p, N)


@pytest.mark.parametrize("p,shapes", [
    # 65537 = 2^16 + 1: largest transform 2^16, pieces of 2^15
    (65537, [(70000, 50000), (3 * 32768 + 7, 5), (65537, 65537), (40000, 130000), (1 << 16, 1 << 16),
             (1 << 19, 1 << 19)]),  # the last: 16 x 16 pieces, past the split limit (three-prime path)
    # 786433 = 3 * 2^18 + 1: largest transform 2^18, pieces of 2^17; operands
    # of at most 2^18 coefficients take the coset product (cyclic and
    # g-twisted spectra, 786432 > 2^18)
    (786433, [(1 << 18, 1 << 18), (300000, 200001), (5, 262145), (131073, 393215), (200000, 150000),
              (1 << 18, 1000), (140000, 1 << 18)]),
])
def test_conv_split_beyond_max_transform(fa, oracle, p, shapes):
    """Products longer than the prime's largest transform N: operands of at
    most N coefficients as the coset product when N < p - 1 (65537 has no
    coset: p - 1 = N), else one spectrum per half-length piece, block sums
    inverted and accumulated at their offsets; uneven piece counts, truncated
    outputs (inside the low half) and squaring included."""
    rng = random.Random(p + 1)
    for la, lb in shapes:
        a = _rnd(oracle, p, la, rng.randrange(1 << 30))
        b = _rnd(oracle, p, lb, rng.randrange(1 << 30))
        want = oracle.conv(p, a, b)
        assert np.array_equal(fa.conv(p, a, b), want), (p, la, lb)
        big = (p - 1) & -(p - 1)  # the largest transform
        for n in ((la + lb) // 3 + 1, min(la + lb - 1, big + 5)):
            assert np.array_equal(fa.conv(p, a, b, n), want[:n]), (p, la, lb, n)
    for lx in (100000, 200000):
        x = _rnd(oracle, p, lx, 7)
        assert np.array_equal(fa.conv(p, x, x), oracle.conv(p, x, x)), (p, lx)


def _eval_limbs(p, limbs, x0):
    """Sum of v_i x0^i mod p over the values v_i of Montgomery limbs (numpy,
    exact: every product below 2^60, every partial sum below 2^64)."""
    n = len(limbs)
    rinv = pow(1 << 32, p - 2, p)
    v = limbs.astype(np.uint64) * np.uint64(rinv) % np.uint64(p)
    blk = 4096
    lo = np.array([pow(x0, i, p) for i in range(blk)], dtype=np.uint64)
    hi = np.array([pow(x0, blk * j, p) for j in range((n + blk - 1) // blk)], dtype=np.uint64)
    pw = (hi[:, None] * lo[None, :] % np.uint64(p)).reshape(-1)[:n]
    return int((v * pw % np.uint64(p)).sum() % np.uint64(p))


@pytest.mark.gpu
def test_coset_product_full_size(fa):
    """The coset product at the headline prime's largest transform, operands
    of 2^23 and 2^23 - 12345 coefficients: the full product agrees with
    a(x0) b(x0) at two points (a wrong low or high half fails both), and the
    truncations inside and just past the low half are its prefixes.  (The
    symmetric 2^23 x 2^23 product is pinned by the reference's checksum.)"""
    rng = np.random.default_rng(23)
    la, lb = 1 << 23, (1 << 23) - 12345
    a = rng.integers(0, P, la, dtype=np.uint32)
    b = rng.integers(0, P, lb, dtype=np.uint32)
    full = fa.conv(P, a, b)
    assert full.size == la + lb - 1
    for x0 in (3, 123456789):
        assert _eval_limbs(P, full, x0) == _eval_limbs(P, a, x0) * _eval_limbs(P, b, x0) % P, x0
    for n in ((1 << 23) - 1000, (1 << 23) + 777):
        assert np.array_equal(fa.conv(P, a, b, n), full[:n]), n


@pytest.mark.parametrize("p,op,n,want", [
    (int(m), op, int(s), ck)
    for m, t in gi.reference_checksums().items() for op, sizes in t.items()
    if op in ("conv", "conv-trunc", "conv-tail", "square", "skinny") for s, ck in sizes.items()
    if int(s) <= (1 << 20)])
def test_golden_conv(fa, oracle, p, op, n, want):
    from test_oracle import run_conv_op
    assert gi.hexsum(oracle, run_conv_op(oracle, p, op, n, conv=fa.conv)) == want


# ------------------------------------------------------------------ series
SERIES_OPS = ("inv", "log", "exp", "sqrt", "inv-short", "inv-linear", "exp-linear")


@pytest.mark.parametrize("p,op,n,want", [
    (int(m), op, int(s), ck)
    for m, t in gi.reference_checksums().items() for op, sizes in t.items()
    if op in SERIES_OPS for s, ck in sizes.items() if int(s) <= (1 << 20)])
def test_golden_series(fa, oracle, p, op, n, want):
    from test_oracle import run_series_op
    got = run_series_op(oracle, p, op, n, series=fa.series, conv=fa.conv)
    assert gi.hexsum(oracle, got) == want


@pytest.mark.parametrize("key", sorted(gi.ref_vectors()["series"].keys()))
def test_ref_series_vectors(fa, key):
    p, op, n = key.split("/")
    p, n = int(p), int(n)
    case = gi.ref_vectors()["series"][key]
    base = "sqrt" if op.startswith("sqrt") else op
    assert [int(v) for v in fa.series(base, p, case["a"], n)] == case["out"]


@pytest.mark.parametrize("p", [P, Q, 1224736769, 100003])
def test_series_random_vs_oracle(fa, oracle, p):
    rng = random.Random(p + 5)
    for n in (1, 2, 3, 9, 31, 32, 33, 40, 41, 64, 65, 100, 1000, 4097, 20000):
        a = _rnd(oracle, p, n, rng.randrange(1 << 30))
        a[0] = oracle.from_int(p, 1 + rng.randrange(p - 1))
        assert np.array_equal(fa.inv(p, a, n), oracle.series("inv", p, a, n)), ("inv", p, n)
        a1 = a.copy(); a1[0] = oracle.from_int(p, 1)
        if n <= p:
            assert np.array_equal(fa.log(p, a1, n), oracle.series("log", p, a1, n)), ("log", p, n)
            a0 = a.copy(); a0[0] = 0
            assert np.array_equal(fa.exp(p, a0, n), oracle.series("exp", p, a0, n)), ("exp", p, n)
        sq = oracle.conv(p, a, a, n)
        assert np.array_equal(fa.sqrt(p, sq, n), oracle.series("sqrt", p, sq, n)), ("sqrt", p, n)


@pytest.mark.parametrize("p", [3, 5, 7, 13, 97, 65537])
def test_series_at_domain_limit(fa, oracle, p):
    """log and exp up to n = p (the largest the domain allows) over small
    primes: the reciprocal tables 1/(c + i) and 1/i run to the end of the
    field, four elements sharing each Fermat inverse."""
    rng = random.Random(p)
    one = oracle.from_int(p, 1)
    for n in sorted({1, 2, min(p, 4096) - 1, min(p, 4096)} - {0}):
        a = np.array([oracle.from_int(p, rng.randrange(p)) for _ in range(n)], dtype=np.uint32)
        la = a.copy()
        la[0] = one
        assert np.array_equal(fa.series("log", p, la, n), oracle.series("log", p, la, n)), (p, n)
        ea = a.copy()
        ea[0] = 0
        assert np.array_equal(fa.series("exp", p, ea, n), oracle.series("exp", p, ea, n)), (p, n)


def test_series_edge_cases(fa, oracle):
    po = oracle
    one = po.from_int(P, 1)
    for op, a in [("inv", [0, one]), ("inv", []), ("log", [po.from_int(P, 2)]), ("exp", [one]),
                  ("sqrt", [0, one]), ("sqrt", [po.from_int(P, 3)])]:
        with pytest.raises(fa.DomainError):
            fa.series(op, P, a, 10)
    # short inputs, zero tails, valuations (inverse.hpp:169-177, square_root.hpp:26-44)
    cases = [("inv", [one] + [0] * 5), ("inv", [one, po.from_int(P, 5)]), ("inv", [one, 0, 0, one]),
             ("exp", [0]), ("exp", [0, one]), ("exp", []), ("sqrt", [0, 0, 0, 0, po.from_int(P, 4)]),
             ("sqrt", [0] * 6), ("sqrt", [0, 0, one, one]), ("log", [one])]
    for op, a in cases:
        for n in (1, 5, 70, 3000):
            assert np.array_equal(fa.series(op, P, a, n), po.series(op, P, a, n)), (op, a, n)
    assert fa.inv(P, [one], 0).size == 0


# ------------------------------------------------------------------ workload sizes
@pytest.mark.parametrize("key", sorted(gi.ref_vectors()["checksums"].keys()))
def test_workload_checksums(fa, oracle, key):
    po = oracle
    p, op, n = key.split("/")
    p, n = int(p), int(n)
    want = gi.ref_vectors()["checksums"][key]
    if op == "conv":
        a, b = gi.bench_case(po, p, "conv", n)
        got = fa.conv(p, a, b)
    elif op == "inv-random-a0":
        got = fa.inv(p, gi.series_input(po, p, n, True), n)
    else:
        from test_oracle import run_series_op
        got = run_series_op(po, p, op, n, series=fa.series, conv=fa.conv)
    assert gi.hexsum(po, got) == want


# ------------------------------------------------------------------ sqrt past n
def _sqrt_long_cases(po):
    """sqrt inputs longer than n: the valuation v is found over the whole input
    and a[v + i] is read for i < n - v/2 (square_root.hpp:26-44), so
    coefficients at and past x^n matter."""
    f = lambda x: po.from_int(P, x)  # noqa: E731
    cases = [([0, 0, 0, 0, f(4)], 3), ([0, 0, f(1), f(1), f(1), f(1)], 4), ([0, 0, 0, 0, 0, f(4)], 3),
             ([0] * 8 + [f(9), f(5)], 3), ([0, 0, 0, 0, f(3)], 2), ([0] * 12 + [f(16)] + [f(7)] * 9, 5)]
    rng = random.Random(77)
    for n in (5, 70, 600, 3000):
        for half in (1, 3, 40):
            v = 2 * half
            sq = po.MT(1000 * n + v).rnd(P, n)
            sq[0] = f(rng.randrange(1, 1000))
            body = po.conv(P, sq, sq, n)
            a = np.concatenate([np.zeros(v, np.uint32), body, po.MT(n + v).rnd(P, rng.randrange(0, 2 * v + 3))])
            cases.append((a, n))
        cases.append(([0] * (2 * n + 1) + [f(4)], n))  # odd valuation past n
        cases.append(([0] * (2 * n + 2) + [f(4)], n))  # even, half >= n: zeros
    return cases


def _want_sqrt(po, a, n):
    try:
        return po.series("sqrt", P, a, n)
    except po.OracleError:
        return None


def test_sqrt_input_longer_than_n(fa, oracle):
    for a, n in _sqrt_long_cases(oracle):
        want = _want_sqrt(oracle, a, n)
        if want is None:
            with pytest.raises(fa.DomainError):
                fa.sqrt(P, a, n)
        else:
            assert np.array_equal(fa.sqrt(P, a, n), want), (len(a), n)


def test_sqrt_input_longer_than_n_reference(fa, oracle):
    if oracle.ref() is None:
        pytest.skip("oracle/_ref (the reference library built from its headers) is not built")
    f = lambda x: oracle.from_int(P, x)  # noqa: E731
    for a, n in [([0, 0, 0, 0, f(4)], 3), ([0, 0, f(1), f(1), f(1), f(1)], 4)]:
        assert np.array_equal(fa.sqrt(P, a, n), oracle.ref_series("sqrt", P, np.asarray(a, np.uint32), n))
    got = fa.sqrt(P, [0, 0, 0, 0, f(4)], 3)  # x^2 * (+-2): the root's sign is Mont::sqrt's choice
    assert got[0] == 0 and got[1] == 0 and got[2] in (f(2), f(P - 2))


def test_sqrt_input_longer_than_n_dev_batch_server(fa, oracle):
    cases = _sqrt_long_cases(oracle)
    wants = [_want_sqrt(oracle, a, n) for a, n in cases]
    # device pointers
    for (a, n), want in zip(cases, wants):
        d = fa.DeviceArray.from_host(np.asarray(a, np.uint32))
        o = fa.DeviceArray(n)
        if want is None:
            with pytest.raises(fa.DomainError):
                fa.series_dev("sqrt", P, d, len(a), n, o)
        else:
            fa.series_dev("sqrt", P, d, len(a), n, o)
            assert np.array_equal(o.to_host(), want), (len(a), n)
    # merged batch (host arrays) and the server
    reqs = [fa.Request("sqrt", P, a, n=n) for a, n in cases]
    got = fa.batch(reqs)
    with fa.Server() as srv:
        ids = [srv.submit(fa.Request("sqrt", P, a, n=n)) for a, n in cases]
        served = [srv.wait(i)[0] for i in ids]
    for g, s, want, (a, n) in zip(got, served, wants, cases):
        for res in (g, s):
            if want is None:
                assert isinstance(res, fa.DomainError), (len(a), n)
            else:
                assert np.array_equal(res, want), (len(a), n)


def test_series_domain_error_writes_nothing(fa, oracle):
    """A series call that raises the reference's domain error leaves the
    caller's output buffer as it was (the reference returns nothing), on a
    first (eager) call at a fresh site, for every op and for p = 2."""
    f = lambda p, x: oracle.from_int(p, x)  # noqa: E731
    cases = [
        ("inv", P, [0, f(P, 1), f(P, 2)], 5),
        ("log", P, [f(P, 2), f(P, 1)], 5),
        ("exp", P, [f(P, 1), f(P, 1)], 5),
        ("sqrt", P, [0, f(P, 1), f(P, 1)], 5),   # odd valuation
        ("sqrt", P, [f(P, 5), f(P, 1)], 5),      # 5 is a non-residue mod 998244353
        ("sqrt", 2, [1, 1, 0, 1], 3),            # characteristic 2: nonzero odd coefficient
    ]
    for op, p, a, n in cases:
        d = fa.DeviceArray.from_host(np.asarray(a, np.uint32))
        sentinel = np.arange(n, dtype=np.uint32) + 11
        o = fa.DeviceArray.from_host(sentinel)
        with pytest.raises(fa.DomainError):
            fa.series_dev(op, p, d, len(a), n, o)
        assert np.array_equal(o.to_host(), sentinel), (op, p, a)


@pytest.mark.parametrize("step_max", [9, 13])
def test_newton_step_placement(fa, oracle, step_max):
    """Newton steps of 2^10 .. 2^13 points run as one workgroup's launch up to
    2^step_max points and as four-step products beyond (default 2^12): with
    every step split (9) and with one-launch steps up to 2^13 (13), inv / log /
    exp / sqrt over 998244353 give the oracle's limbs, eagerly and replayed."""
    fa.set_step_log_max(step_max)
    try:
        rng = random.Random(step_max)
        for n in (3000, 9000, 20000):
            a = _rnd(oracle, P, n, rng.randrange(1 << 30))
            a[0] = oracle.from_int(P, 777)
            a1 = a.copy(); a1[0] = oracle.from_int(P, 1)
            a0 = a.copy(); a0[0] = 0
            sq = oracle.conv(P, a1, a1, n)
            ins = {"inv": a, "log": a1, "exp": a0, "sqrt": sq}
            for op in ("inv", "log", "exp", "sqrt"):
                want = oracle.series(op, P, ins[op], n)
                d, o = fa.DeviceArray.from_host(ins[op]), fa.DeviceArray(n)
                for rep in range(3):  # eager, captured, replayed
                    fa.series_dev(op, P, d, n, n, o)
                    assert np.array_equal(o.to_host(), want), (op, n, step_max, rep)
    finally:
        fa.set_step_log_max(0)
    with pytest.raises(fa.ArgumentError):
        fa.set_step_log_max(14)
