"""Pin the CPU oracle before trusting it as the checker of the GPU path.

Every oracle function is checked against something independent of its own
formulas:
  * the reference's published output checksums (bench CSVs, FNV-1a of the full
    output for the deterministic bench inputs) -> tests/golden/reference_checksums.json;
  * vectors produced by the reference library compiled in this container
    (tests/golden/ref_vectors.json, written by tests/golden/make_ref_vectors.py);
  * mathematics: brute-force DFT in bit-reversed order, O(n^2) integer
    products, closed forms in Python integers, series identities and
    coefficient recurrences.
"""
import random

import numpy as np
import pytest

import golden_inputs as gi

PRIMES = [998244353, 1000000007, 167772161, 469762049, 754974721, 1224736769, 100003, 3, 2]
SERIES_MAX = 1 << 17
CONV_MAX = 1 << 20


# ---------------------------------------------------------------- field ops
@pytest.mark.parametrize("p", PRIMES)
def test_field_closed_forms(oracle, p):
    po = oracle
    f = po.field(p)
    R = 1 << 32
    assert f.one == (R % p if p != 2 else 1)
    if p != 2:  # -p^-1 mod 2^32 (p = 2 keeps plain one-bit residues)
        assert (f.ninv * p) % R == R - 1
    rng = random.Random(p)
    for _ in range(200):
        x, y = rng.randrange(p), rng.randrange(p)
        lx, ly = po.from_int(p, x), po.from_int(p, y)
        if p != 2:
            assert lx == x * R % p
        assert po.to_int(p, lx) == x
        assert po.to_int(p, po.mont_mul(p, lx, ly)) == x * y % p
        e = rng.randrange(1 << 40)
        assert po.to_int(p, po.mont_pow(p, lx, e)) == pow(x, e, p)


def test_mt19937_64_standard_value(oracle):
    g = oracle.MT(5489)
    for _ in range(9999):
        g.next()
    assert g.next() == 9981545732273789042  # C++ [rand.predef] requirement


@pytest.mark.parametrize("p", [998244353, 1000000007, 167772161, 469762049, 754974721, 1224736769, 100003, 3])
def test_mont_sqrt_root_choice(oracle, p):
    rows = gi.ref_vectors()["msqrt"][str(p)]
    for limb, want in rows:
        got = oracle.mont_sqrt(p, limb)
        assert got == want
        if want is not None:
            assert oracle.mont_mul(p, want, want) == limb


# ---------------------------------------------------------------- NTT
def _brute_dft_bitrev(p, root, x):
    n = len(x)
    lg = n.bit_length() - 1
    w = pow(root, (p - 1) // n, p)
    out = [0] * n
    for f in range(n):
        s = sum(x[j] * pow(w, j * f, p) for j in range(n)) % p
        r = int(format(f, f"0{lg}b")[::-1], 2) if lg else 0
        out[r] = s
    return out


@pytest.mark.parametrize("p", [998244353, 167772161, 469762049, 754974721, 1224736769])
@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64, 128])
def test_ntt_forward_is_bitreversed_dft(oracle, p, n):
    po = oracle
    rng = random.Random(n * 7 + p)
    x = [rng.randrange(p) for _ in range(n)]
    limbs = [po.from_int(p, v) for v in x]
    got = [po.to_int(p, v) for v in po.ntt(p, limbs)]
    root = po.field(p).root
    assert got == _brute_dft_bitrev(p, root, x)
    back = po.ntt(p, po.ntt(p, limbs), inverse=True)
    assert list(back) == limbs


def test_ntt_rejects_bad_sizes(oracle):
    with pytest.raises(oracle.OracleError):
        oracle.ntt(998244353, np.zeros(3, np.uint32))  # not a power of two
    with pytest.raises(oracle.OracleError):
        oracle.ntt(1000000007, np.zeros(4, np.uint32))  # v2(p-1) = 1


def _golden_rows(ops, limit):
    ck = gi.reference_checksums()
    for mod, table in sorted(ck.items()):
        for op, sizes in sorted(table.items()):
            if op not in ops:
                continue
            for size, want in sorted(sizes.items(), key=lambda kv: int(kv[0])):
                if int(size) <= limit:
                    yield int(mod), op, int(size), want


@pytest.mark.parametrize("p,op,n,want", list(_golden_rows({"ntt-forward", "ntt-inverse"}, CONV_MAX)))
def test_golden_ntt(oracle, p, op, n, want):
    po = oracle
    x = gi.bench_case(po, p, op, n)
    root = po.lib().fpo_primitive_root(p)
    y = po.ntt(p, x, root=root)
    if op == "ntt-inverse":
        y = po.ntt(p, y, inverse=True, root=root)
    assert gi.hexsum(po, y) == want


CONV_OPS = {"conv", "conv-trunc", "conv-tail", "square", "skinny"}


def run_conv_op(po, p, op, n, conv=None):
    conv = conv or po.conv
    x = gi.bench_case(po, p, op, n)
    if op == "conv":
        return conv(p, x[0], x[1])
    if op == "conv-trunc":
        return conv(p, x[0], x[1], n)
    if op == "conv-tail":
        return conv(p, x[0], x[1])
    if op == "square":
        return conv(p, x, x)
    if op == "skinny":
        return conv(p, x[0], x[1])
    raise KeyError(op)


@pytest.mark.parametrize("p,op,n,want", list(_golden_rows(CONV_OPS, CONV_MAX)))
def test_golden_conv(oracle, p, op, n, want):
    assert gi.hexsum(oracle, run_conv_op(oracle, p, op, n)) == want


SERIES_OPS = {"inv", "log", "exp", "sqrt", "inv-short", "inv-linear", "exp-linear"}


def run_series_op(po, p, op, n, series=None, conv=None):
    series = series or po.series
    conv = conv or po.conv
    if op == "inv-short":
        return series("inv", p, [po.from_int(p, 1), po.from_int(p, p - 1)], n)
    if op == "inv-linear":
        return series("inv", p, [po.from_int(p, 1), po.from_int(p, p - 7 % p)], n)
    if op == "exp-linear":
        return series("exp", p, [0, po.from_int(p, 1)], n)
    a = gi.bench_case(po, p, op, n)
    if op in ("inv", "log"):
        return series(op, p, a, n)
    if op == "exp":
        b = a.copy()
        b[0] = 0
        return series("exp", p, b, n)
    if op == "sqrt":
        return series("sqrt", p, conv(p, a, a, n), n)
    raise KeyError(op)


@pytest.mark.parametrize("p,op,n,want", list(_golden_rows(SERIES_OPS, SERIES_MAX)))
def test_golden_series(oracle, p, op, n, want):
    assert gi.hexsum(oracle, run_series_op(oracle, p, op, n)) == want


@pytest.mark.slow
@pytest.mark.parametrize("p,op,n,want",
                         [r for r in _golden_rows(SERIES_OPS, 1 << 20) if r[2] > SERIES_MAX])
def test_golden_series_large(oracle, p, op, n, want):
    assert gi.hexsum(oracle, run_series_op(oracle, p, op, n)) == want


# ---------------------------------------------------------------- conv math
@pytest.mark.parametrize("p", [998244353, 1000000007, 1224736769, 100003, 3, 2])
def test_conv_matches_integer_product(oracle, p):
    po = oracle
    rng = random.Random(p)
    for la, lb in [(1, 1), (1, 50), (40, 41), (41, 41), (63, 200), (300, 257), (1000, 999)]:
        a = [rng.randrange(p) for _ in range(la)]
        b = [rng.randrange(p) for _ in range(lb)]
        want = [0] * (la + lb - 1)
        for i, x in enumerate(a):
            for j, y in enumerate(b):
                want[i + j] += x * y
        want = [v % p for v in want]
        got = po.conv(p, [po.from_int(p, v) for v in a], [po.from_int(p, v) for v in b])
        assert [po.to_int(p, v) for v in got] == want


@pytest.mark.parametrize("p", [998244353, 1000000007, 100003, 2])
def test_crt_small_blocks_and_truncation(oracle, p):
    """crt_convolution with a forced small block (test_prime_poly.cpp:110, :117)."""
    po = oracle
    g = po.MT(99 + p)
    a, b = g.rnd(p, 700), g.rnd(p, 333)
    want = po.conv_naive(p, a, b)
    for block in (64, 100, 1 << 23):
        assert np.array_equal(po.conv_crt(p, a, b, want.size, block), want)
    assert np.array_equal(po.conv(p, a, b, 500), want[:500])
    assert po.conv(p, a, b, 0).size == 0
    assert po.conv(p, [], b).size == 0


def test_crt_prime_count_kats(oracle):
    # crt_convolution.hpp:78-83 thresholds: (p-1)^2 * terms vs p1-1, p1 p2 - 1.
    p1, p2 = 167772161, 469762049
    for p in (2, 3, 100003, 998244353, 1000000007):
        sq = (p - 1) ** 2
        for terms in (1, 2, 40, 1000, 1 << 16, 1 << 23):
            want = 1 if sq <= (p1 - 1) // terms else 2 if sq <= (p1 * p2 - 1) // terms else 3
            assert oracle.crt_prime_count(p, terms) == want
    assert oracle.crt_prime_count(998244353, 1 << 23) == 3
    assert oracle.crt_prime_count(1000000007, 1) == 3


# ---------------------------------------------------------------- series math
def _ints(po, p, limbs):
    return [po.to_int(p, int(v)) for v in limbs]


@pytest.mark.parametrize("p", [998244353, 1000000007, 1224736769, 100003])
@pytest.mark.parametrize("n", [1, 2, 9, 33, 64, 100, 513])
def test_series_identities(oracle, p, n):
    po = oracle
    g = po.MT(n * 31 + p)
    a = g.rnd(p, n)
    a[0] = po.from_int(p, 1 + g.next() % (p - 1))
    b = po.series("inv", p, a, n)
    prod = po.conv(p, a, b, n)
    assert _ints(po, p, prod) == [1] + [0] * (n - 1)
    al = a.copy()
    al[0] = po.from_int(p, 1)
    if n <= p:
        lg = po.series("log", p, al, n)
        assert np.array_equal(po.series("exp", p, lg, n), al)
    sq = po.conv(p, a, a, n)
    r = po.series("sqrt", p, sq, n)
    assert np.array_equal(po.conv(p, r, r, n), sq)


@pytest.mark.parametrize("p", [998244353, 1000000007])
def test_series_recurrences(oracle, p):
    """reference_inverse / reference_exp of test_poly.cpp:430-454."""
    po = oracle
    for n in (15, 16, 17, 31, 32, 33, 63, 65, 129, 257, 513):
        g = po.MT(n + 5)
        a = g.rnd(p, n)
        a[0] = po.from_int(p, 3)
        x = _ints(po, p, a)
        inv = [pow(x[0], p - 2, p)]
        for i in range(1, n):
            inv.append(-sum(x[j] * inv[i - j] for j in range(1, i + 1)) * inv[0] % p)
        assert _ints(po, p, po.series("inv", p, a, n)) == inv
        a[0] = 0
        x[0] = 0
        e = [1]
        for i in range(1, n):
            e.append(sum(j * x[j] * e[i - j] for j in range(1, i + 1)) * pow(i, p - 2, p) % p)
        assert _ints(po, p, po.series("exp", p, a, n)) == e


def test_series_kats(oracle):
    po = oracle
    p = 998244353
    one = po.from_int(p, 1)
    # 1/(1-x) = sum x^i (test_poly.cpp:171-173)
    assert _ints(po, p, po.series("inv", p, [one, po.from_int(p, p - 1)], 50)) == [1] * 50
    # log(1+x) = sum (-1)^(i+1) x^i / i (test_poly.cpp:199-203)
    lg = _ints(po, p, po.series("log", p, [one, one], 20))
    assert lg == [0] + [((-1) ** (i + 1)) * pow(i, p - 2, p) % p for i in range(1, 20)]
    # exp(x) = sum x^i / i! (test_poly.cpp:205-211)
    ex = _ints(po, p, po.series("exp", p, [0, one], 20))
    fact = 1
    want = []
    for i in range(20):
        want.append(pow(fact, p - 2, p))
        fact = fact * (i + 1) % p
    assert ex == want


def test_series_domain_errors(oracle):
    po = oracle
    p = 998244353
    one = po.from_int(p, 1)
    for op, a in [("inv", [0, one]), ("inv", []), ("log", [po.from_int(p, 2)]), ("log", []),
                  ("exp", [one]), ("sqrt", [0, one]), ("sqrt", [po.from_int(p, 3)])]:
        with pytest.raises(po.OracleError) as e:
            po.series(op, p, a, 10)
        assert e.value.code == po.E_DOMAIN
    with pytest.raises(po.OracleError):
        po.series("log", 3, [po.from_int(3, 1)], 4)  # n > p
    with pytest.raises(po.OracleError):
        po.series("exp", 3, [0], 4)  # n - 1 >= p
    assert po.series("inv", p, [one], 0).size == 0


@pytest.mark.parametrize("key", sorted(gi.ref_vectors()["series"].keys()))
def test_ref_series_vectors(oracle, key):
    p, op, n = key.split("/")
    p, n = int(p), int(n)
    case = gi.ref_vectors()["series"][key]
    base = "sqrt" if op.startswith("sqrt") else op
    got = oracle.series(base, p, case["a"], n)
    assert [int(v) for v in got] == case["out"]


@pytest.mark.slow
@pytest.mark.parametrize("key", sorted(gi.ref_vectors()["checksums"].keys()))
def test_ref_workload_checksums(oracle, key):
    po = oracle
    p, op, n = key.split("/")
    p, n = int(p), int(n)
    want = gi.ref_vectors()["checksums"][key]
    if op == "conv":
        g = po.MT(2 + n)
        a = g.rnd(p, n)
        got = po.conv(p, a, g.rnd(p, n))
    elif op == "inv-random-a0":
        got = po.series("inv", p, gi.series_input(po, p, n, True), n)
    else:
        got = run_series_op(po, p, op, n)
    assert gi.hexsum(po, got) == want


# ---------------------------------------------------------------- pow, divmod
POW_MAX = 1 << 16


@pytest.mark.parametrize("p,op,n,want", list(_golden_rows({"pow"}, POW_MAX)))
def test_golden_pow(oracle, p, op, n, want):
    a = gi.bench_case(oracle, p, "pow", n)
    assert gi.hexsum(oracle, oracle.poly_pow(p, a, gi.POW_K, n)) == want


@pytest.mark.parametrize("p,op,n,want", list(_golden_rows({"div", "mod", "divmod"}, POW_MAX)))
def test_golden_division(oracle, p, op, n, want):
    a, b = gi.division_case(oracle, p, n)
    q, r = oracle.divmod(p, a, b)
    assert gi.division_checksum(oracle, op, q, r) == want


@pytest.mark.parametrize("p", [998244353, 1000000007, 100003, 3, 2])
def test_pow_matches_repeated_products(oracle, p):
    """a^k by k-1 naive products (every branch of power.hpp: shifts, a
    binomial, binary powers, exp(k log))."""
    rng = random.Random(p)
    for trial in range(30):
        n = rng.randint(1, 60)
        la = rng.randint(0, 12)
        a = np.array([rng.randrange(p) for _ in range(la)], dtype=np.uint32)
        if la and trial % 3 == 0:
            a[: rng.randint(0, la - 1)] = 0          # valuation
        if la > 2 and trial % 5 == 0:
            a[2:] = 0                                 # two-term polynomial
        k = rng.choice([0, 1, 2, 3, 7, 64, 65, 200])
        want = np.zeros(n, np.uint32)
        want[0] = oracle.from_int(p, 1)
        for _ in range(k):
            prod = oracle.conv_naive(p, want, a, n) if la else np.zeros(n, np.uint32)
            want = np.zeros(n, np.uint32)
            want[: len(prod)] = prod[:n]
        assert np.array_equal(oracle.poly_pow(p, a, k, n), want), (p, n, list(a), k)


@pytest.mark.parametrize("p", [998244353, 1000000007, 100003, 2])
def test_divmod_identity(oracle, p):
    """b q + r = a with deg r < deg b, q and r trimmed, on every branch
    (short divisor, long division, reversed-inverse quotient)."""
    rng = random.Random(p + 1)
    for la, lb in [(0, 1), (5, 9), (40, 1), (40, 3), (100, 17), (300, 100), (300, 250), (2000, 60), (5000, 2400)]:
        a = np.array([rng.randrange(p) for _ in range(la)], dtype=np.uint32)
        b = np.array([rng.randrange(p) for _ in range(lb)], dtype=np.uint32)
        if lb and b[-1] == 0:
            b[-1] = 1
        q, r = oracle.divmod(p, a, b)
        assert len(r) < lb and (len(r) == 0 or r[-1] != 0) and (len(q) == 0 or q[-1] != 0)
        bq = oracle.conv_naive(p, b, q, la) if len(q) else np.zeros(0, np.uint32)
        tot = np.zeros(max(la, 1), np.uint32)
        tot[: len(bq)] = bq[:la] if la else bq[:0]
        for i in range(len(r)):
            tot[i] = (int(tot[i]) + int(r[i])) % p if p != 2 else int(tot[i]) ^ int(r[i])
        at = a.copy()
        assert np.array_equal(tot[:la], at), (p, la, lb)
    with pytest.raises(oracle.OracleError):
        oracle.divmod(p, np.ones(3, np.uint32), np.zeros(2, np.uint32))


@pytest.mark.parametrize("p", [998244353, 1000000007, 100003, 2])
def test_pow_divmod_match_reference_library(oracle, p):
    """The restatement against the reference library itself (oracle/_ref)."""
    if oracle.ref() is None or not hasattr(oracle.ref(), "fpref_pow"):
        pytest.skip("reference library not built here")
    rng = random.Random(p + 11)
    for t in range(20):
        n = rng.choice([1, 7, 100, 3000])
        la = rng.choice([0, 1, 2, 5, 100, 4000])
        a = np.array([rng.randrange(p) for _ in range(la)], np.uint32)
        if la and t % 3 == 0:
            a[: rng.randint(0, la - 1)] = 0
        k = rng.choice([0, 1, 2, 5, 64, 65, 1000003, (1 << 64) - 1])
        assert np.array_equal(oracle.poly_pow(p, a, k, n), oracle.ref_pow(p, a, k, n)), (p, n, la, k)
    for la, lb in [(5, 9), (40, 3), (300, 100), (5000, 2400), (20000, 19990)]:
        a = np.array([rng.randrange(p) for _ in range(la)], np.uint32)
        b = np.array([rng.randrange(p) for _ in range(lb)], np.uint32)
        b[-1] = 1
        q1, r1 = oracle.divmod(p, a, b)
        q2, r2 = oracle.ref_divmod(p, a, b)
        assert np.array_equal(q1, q2) and np.array_equal(r1, r2), (p, la, lb)


def test_oracle_matches_workload_checksums_sample(oracle):
    """The per-request checksums of the full workloads (written by the
    reference library) agree with the oracle on the requests small enough for
    the scalar oracle."""
    import json
    import os
    from fastpoly_amd.workloads import serving_stream, uneven_batch
    with open(os.path.join(os.path.dirname(__file__), "golden", "workload_checksums.json")) as fh:
        g = json.load(fh)

    def run(r):
        if r[0] == "conv":
            return oracle.conv(r[1], r[2], r[3])
        return oracle.series(r[0], r[1], r[2], r[3])

    def size(r):
        return r[2].size + (r[3].size if r[0] == "conv" else r[3])

    batch = uneven_batch(g["uneven_batch"]["n_conv"], g["uneven_batch"]["n_series"], seed=g["uneven_batch"]["seed"])
    picked = [i for i, r in enumerate(batch) if size(r) <= 1 << 13][:30]
    picked += [i for i, r in enumerate(batch) if r[0] != "conv" and size(r) <= 1 << 12][:10]
    assert len(picked) >= 30
    for i in picked:
        assert "%016x" % oracle.checksum(run(batch[i])) == g["uneven_batch"]["checksums"][i], i
    s = g["serving_stream"]
    stream = [r for _, r in serving_stream(s["count"], 1000.0, seed=s["seed"], length_log=tuple(s["length_log"]))]
    picked = [i for i, r in enumerate(stream) if size(r) <= 1 << 12][:12]
    assert len(picked) >= 5
    for i in picked:
        assert "%016x" % oracle.checksum(run(stream[i])) == s["checksums"][i], i
