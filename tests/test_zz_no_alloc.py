"""GPU: fpxg_conv_dev_noalloc (the conv entry point that allocates nothing,
its scratch a static device array) gives fpxg_conv_dev's results on two calls
in a row, the second smaller, without a device allocation; sizes past its
array and unprepared moduli are refused with a clear error."""
import ctypes

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


def _dev(fa, rng, p, n):
    return fa.DeviceArray.from_host(rng.integers(0, p, n, dtype=np.uint64).astype(np.uint32))


def _both(fa, p, a, la, b, lb, n=None):
    cap = la + lb - 1 if n is None else min(la + lb - 1, n)
    o1, o2 = fa.DeviceArray(cap), fa.DeviceArray(cap)
    g1 = fa.conv_dev(p, a, la, b, lb, o1, n)
    before = fa.scratch_memory_stats()
    g2 = fa.conv_dev_noalloc(p, a, la, b, lb, o2, n)
    fa.synchronize()
    after = fa.scratch_memory_stats()
    assert after["system_allocations"] == before["system_allocations"]
    assert after["in_use_bytes"] == before["in_use_bytes"]
    assert g1 == g2 == cap
    return o1.to_host(), o2.to_host()


# (la, lb, n): four-step native products, the largest that fits, a one-launch
# product, a short side (direct), a tail past a power of two, truncation.
SHAPES_P = [(1 << 20, 1 << 20, None), (1 << 21, 1 << 21, None), (3000, 2000, None), (5000, 60, None),
            ((1 << 16) + 3, 1 << 16, None), (1 << 19, 1 << 19, 700001)]
SHAPES_Q = [(1 << 18, 1 << 18, None), (300000, 200000, None), (3000, 2000, None), (5000, 200, None),
            (1 << 17, 1 << 17, 100000)]


@pytest.mark.parametrize("p,shapes", [(P, SHAPES_P), (Q, SHAPES_Q)])
def test_noalloc_matches_conv_dev(fa, p, shapes):
    rng = np.random.default_rng(p % 1000)
    for la, lb, n in shapes:
        a, b = _dev(fa, rng, p, la), _dev(fa, rng, p, lb)
        # twice in a row: first this size, then a smaller one over the array's old contents
        x1, y1 = _both(fa, p, a, la, b, lb, n)
        assert np.array_equal(x1, y1), (p, la, lb, n)
        x2, y2 = _both(fa, p, a, la // 3 + 1, b, lb // 5 + 1)
        assert np.array_equal(x2, y2), (p, la, lb, "smaller")


def test_noalloc_matches_oracle_small(fa, oracle):
    rng = np.random.default_rng(4)
    for p in (P, Q):
        for la, lb in ((5000, 3000), (700, 900), (20000, 150)):
            ha = rng.integers(0, p, la, dtype=np.uint64).astype(np.uint32)
            hb = rng.integers(0, p, lb, dtype=np.uint64).astype(np.uint32)
            a, b, o = fa.DeviceArray.from_host(ha), fa.DeviceArray.from_host(hb), fa.DeviceArray(la + lb - 1)
            fa.conv_dev_noalloc(p, a, la, b, lb, o)
            assert np.array_equal(o.to_host(), oracle.conv(p, ha, hb)), (p, la, lb)


def test_noalloc_refuses_what_it_cannot_serve(fa):
    rng = np.random.default_rng(5)
    big = _dev(fa, rng, P, 1 << 22)
    out = fa.DeviceArray(1 << 23)
    with pytest.raises(fa.NttSizeError):  # 2^23-point spectra: beyond the 32 MiB array
        fa.conv_dev_noalloc(P, big, 1 << 22, big, 1 << 22, out)
    small = _dev(fa, rng, 97, 100)
    o = fa.DeviceArray(199)
    got = ctypes.c_size_t(0)
    rc = fa.lib().fpxg_conv_dev_noalloc(999999937, small.ptr, 100, small.ptr, 100, fa.SIZE_MAX, o.ptr, 199,
                                        ctypes.byref(got))
    assert rc == fa.E_ARG  # no plan was built for this modulus yet
    fa.conv(999999937, [1, 2, 3], [4, 5])  # an ordinary call builds it
    assert fa.lib().fpxg_conv_dev_noalloc(999999937, small.ptr, 100, small.ptr, 100, fa.SIZE_MAX, o.ptr, 199,
                                          ctypes.byref(got)) == fa.OK
