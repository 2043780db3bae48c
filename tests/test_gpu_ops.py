"""GPU: poly::pow and poly::divmod / div / mod through the C ABI, bit-exact
against the reference's published checksums and the oracle."""
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


def _rows(ops, limit):
    for m, t in sorted(gi.reference_checksums().items()):
        for op, sizes in t.items():
            if op in ops:
                for n, want in sizes.items():
                    if int(n) <= limit:
                        yield int(m), op, int(n), want


@pytest.mark.parametrize("p,op,n,want", list(_rows({"pow"}, 1 << 20)))
def test_golden_pow(fa, oracle, p, op, n, want):
    a = gi.bench_case(oracle, p, "pow", n)
    assert gi.hexsum(oracle, fa.pow(p, a, gi.POW_K, n)) == want


@pytest.mark.parametrize("p,op,n,want", list(_rows({"div", "mod", "divmod"}, 1 << 20)))
def test_golden_division(fa, oracle, p, op, n, want):
    a, b = gi.division_case(oracle, p, n)
    if op == "div":
        q, r = fa.div(p, a, b), np.zeros(0, np.uint32)
    elif op == "mod":
        q, r = np.zeros(0, np.uint32), fa.mod(p, a, b)
    else:
        q, r = fa.divmod(p, a, b)
    assert gi.division_checksum(oracle, op, q, r) == want


@pytest.mark.parametrize("p", [P, Q, 100003, 3, 2])
def test_pow_random_vs_oracle(fa, oracle, p):
    rng = random.Random(p)
    for trial in range(40):
        n = rng.choice([1, 2, 5, 33, 100, 1000, 5000])
        la = rng.choice([0, 1, 2, 3, 10, 200, 6000])
        a = np.array([rng.randrange(p) for _ in range(la)], dtype=np.uint32)
        if la and trial % 3 == 0:
            a[: rng.randint(0, la - 1)] = 0
        if la > 2 and trial % 4 == 0:
            a[2:] = 0
        k = rng.choice([0, 1, 2, 3, 5, 64, 65, 1000003, (1 << 64) - 1])
        assert np.array_equal(fa.pow(p, a, k, n), oracle.poly_pow(p, a, k, n)), (p, n, la, k)


@pytest.mark.parametrize("p", [P, Q, 100003, 2])
def test_divmod_random_vs_oracle(fa, oracle, p):
    rng = random.Random(p + 7)
    for la, lb in [(0, 1), (5, 9), (40, 1), (40, 3), (100, 17), (300, 100), (300, 250), (2000, 60),
                   (5000, 2400), (70000, 30000), (70000, 69990)]:
        a = np.array([rng.randrange(p) for _ in range(la)], dtype=np.uint32)
        b = np.array([rng.randrange(p) for _ in range(lb)], dtype=np.uint32)
        if lb > 3:
            b[-3:] = 0  # trailing zeros are trimmed
            b[-4] = 1
        if not b.any():
            b[0] = 1
        q, r = fa.divmod(p, a, b)
        wq, wr = oracle.divmod(p, a, b)
        assert np.array_equal(q, wq) and np.array_equal(r, wr), (p, la, lb)
        assert np.array_equal(fa.div(p, a, b), wq) and np.array_equal(fa.mod(p, a, b), wr)
    with pytest.raises(fa.DomainError):
        fa.divmod(p, np.ones(5, np.uint32), np.zeros(3, np.uint32))
